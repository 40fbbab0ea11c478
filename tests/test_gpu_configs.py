"""The benchmark configurations at full size on the GPU.

C2: depth 6 (147,456 unknowns), kappa 80, 5 seeded bumps: the device matvec
against the oracle's (restatement of DFMM::MatVecProduct, GMRES/GMRES.cpp:201-254)
at 1e-10, and GMRES at C2's own tolerance 1e-10 verified through the oracle's
operator.

C3: depth 7 (589,824 unknowns), kappa 160, 8 seeded bumps: the FP32 matvec
within its stated bound of FP64, FP32-inner iterative refinement to 1e-8, and
the 64 plane waves at 2 pi j / 64 through FP32-inner batched GMRES with every
system FP64-verified; the solutions are checked through the oracle's matvec
on the same inputs.
"""
import os
import time

import numpy as np
import pytest

import lippmann_schwinger_solver_amd as L
import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def seeded_angle(seed):
    """First draw of the SplitMix64 stream seeded with 1000 + seed, in [0, 2 pi)."""
    z = (1000 + seed + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    z ^= z >> 31
    return 2 * np.pi * ((z >> 11) * 2.0 ** -53)


def config_inputs(depth, kappa, nbumps, seed=0):
    inp = L.InputsToDFMM(kappa=kappa, nLevelsUniform=depth, L=0.5, nChebNodes=6, degreeOfBases=6, TOL_POW=10,
                         nCones_LFR=16)
    inp.contrast = L.bumps_from_seed(seed, nbumps)
    inp.inc_angle = seeded_angle(seed)
    return inp


def log(msg):
    print(msg, flush=True)


# ------------------------------------------------------------------ C2 ----
@pytest.fixture(scope="module")
def c2():
    inp = config_inputs(6, 80.0, 5)
    t0 = time.time()
    S = L.DFMM(inp)
    S.FMMTreeInitialisation()
    t1 = time.time()
    o = O.Oracle(inp.to_c())
    log(f"C2: device context {t1 - t0:.1f} s, oracle {time.time() - t1:.1f} s, N = {S.N}")
    yield S, o
    S.close()
    o.close()


def test_c2_matvec_matches_oracle(c2):
    S, o = c2
    for seed in (0, 1):
        x = H.rand_vec(S.N, 100 + seed)
        e = H.relerr(S.MatVecProduct(x), o.matvec(x))
        log(f"C2 matvec vs oracle (seed {seed}): {e:.2e}")
        assert e <= 1e-10
    b = S.rhs()
    assert H.relerr(b, o.rhs()) <= 1e-15
    e = H.relerr(S.MatVecProduct(b), o.matvec(b))
    assert e <= 1e-10


def test_c2_gmres_at_its_tolerance(c2):
    """GMRES(50) to C2's tol 1e-10: the device's FP64 residual and the one the
    oracle's operator gives for the same iterate both meet it."""
    S, o = c2
    b = S.rhs()
    x, st = S.gmres_ex(b, mode="fp64", restart_len=50, max_restarts=40, tol=1e-10)
    r_or = H.relerr(o.matvec(x), b)
    log(f"C2 GMRES(50) tol 1e-10: {st}; residual through the oracle {r_or:.3e}")
    assert st["status"] == 0 and st["converged"] == 1 and st["rel_residual"] <= 1e-10
    assert r_or <= 1.05e-10


# ------------------------------------------------------------------ C3 ----
@pytest.fixture(scope="module")
def c3():
    inp = config_inputs(7, 160.0, 8)
    t0 = time.time()
    S = L.DFMM(inp)
    S.FMMTreeInitialisation()
    S.autotune(reps=2)
    log(f"C3: device context {time.time() - t0:.1f} s, N = {S.N}")
    yield inp, S
    S.close()


@pytest.fixture(scope="module")
def c3_oracle(c3):
    inp, _ = c3
    t0 = time.time()
    o = O.Oracle(inp.to_c())
    log(f"C3 oracle built in {time.time() - t0:.1f} s")
    yield o
    o.close()


@pytest.fixture(scope="module")
def c3_solutions(c3):
    """FP32-inner refinement of the seeded-angle system, plain FP64 GMRES(50) of
    the same, and the 64-angle FP32-inner batch (FP32 GMRES(24) inside FP64
    flexible GMRES(30))."""
    _, S = c3
    b = S.rhs()
    x64, s64 = S.gmres_ex(b, mode="fp64", restart_len=50, max_restarts=40, tol=1e-8)
    x32, s32 = S.gmres_ex(b, mode="fp32ir", restart_len=50, max_restarts=1, tol=1e-8, inner_tol=1e-5, max_outer=20)
    angles = 2 * np.pi * np.arange(64) / 64
    B = S.rhs_angles(angles)
    # the default bench line's solver: FP32 GMRES(24) inside FP64 flexible GMRES(30)
    X, sb, rel = S.gmres_batch(B, restart_len=24, max_restarts=1, tol=1e-8, mode="fp32fgmres", inner_tol=1e-5,
                               max_outer=30)
    log(f"C3 plain FP64 {s64}\nC3 FP32-IR {s32}\nC3 batched FP32-inner flexible {sb}, max residual {rel.max():.3e}")
    return dict(b=b, x64=x64, s64=s64, x32=x32, s32=s32, B=B, X=X, sb=sb, rel=rel)


def test_c3_fp32_matvec_bound(c3):
    """FP32 operators, FP32 accumulation inside each pass, against the FP64
    matvec: the fused upward operators carry the inverse of skeleton blocks, so
    FP32 loses about two digits against its epsilon. Stated bounds: 3e-5 for
    random vectors over every node (measured 0.95e-5 and 1.13e-5) and 1e-5 for
    right-hand sides, supported where the contrast is (measured 7.9e-6)."""
    _, S = c3
    for seed in (3, 4):
        x = H.rand_vec(S.N, seed)
        e = H.relerr(S.MatVecProduct_fp32(x), S.MatVecProduct(x))
        log(f"C3 FP32 vs FP64 matvec, random vector: {e:.2e}")
        assert e <= 3e-5
    b = S.rhs()
    e = H.relerr(S.MatVecProduct_fp32(b), S.MatVecProduct(b))
    log(f"C3 FP32 vs FP64 matvec, right-hand side: {e:.2e}")
    assert e <= 1e-5


def test_c3_fp32_refinement(c3, c3_solutions):
    _, S = c3
    d = c3_solutions
    b, x32, s32 = d["b"], d["x32"], d["s32"]
    assert d["s64"]["status"] == 0 and d["s64"]["rel_residual"] <= 1e-8
    assert s32["status"] == 0 and s32["converged"] == 1 and s32["rel_residual"] <= 1e-8 and s32["refinements"] >= 1
    assert H.relerr(S.MatVecProduct(x32), b) <= 1.05e-8
    # both solutions solve the same system to 1e-8: they agree to the
    # conditioning times that residual
    assert H.relerr(x32, d["x64"]) <= 1e-6


def test_c3_batched_64_angles(c3, c3_solutions):
    _, S = c3
    d = c3_solutions
    B, X, sb, rel = d["B"], d["X"], d["sb"], d["rel"]
    assert sb["status"] == 0 and sb["converged"] == 64 and rel.max() <= 1e-8
    worst = max(H.relerr(S.MatVecProduct(X[:, q]), B[:, q]) for q in range(64))
    log(f"C3 batched: recomputed FP64 residual max {worst:.3e}")
    assert worst <= 1.05e-8
    # system 0 of the batch (angle 0) against a single plain FP64 solve of it
    x0, s0 = S.gmres_ex(B[:, 0], mode="fp64", restart_len=50, max_restarts=40, tol=1e-8)
    assert H.relerr(X[:, 0], x0) <= 1e-6


def test_c3_solutions_through_the_oracle(c3, c3_solutions, c3_oracle):
    """The FP32-IR solution and every fourth system of the 64-angle batch,
    checked with the oracle's FP64 matvec of the same inputs (every system is
    checked through the device's FP64 matvec above); and the device matvec
    itself against the oracle's at C3."""
    _, S = c3
    d = c3_solutions
    o = c3_oracle
    x = H.rand_vec(S.N, 9)
    e = H.relerr(S.MatVecProduct(x), o.matvec(x))
    log(f"C3 device matvec vs oracle: {e:.2e}")
    assert e <= 1e-10
    r32 = H.relerr(o.matvec(d["x32"]), d["b"])
    worst = max(H.relerr(o.matvec(d["X"][:, q]), d["B"][:, q]) for q in range(0, 64, 4))
    log(f"C3 residuals through the oracle: FP32-IR {r32:.3e}, batched max {worst:.3e}")
    assert r32 <= 1.05e-8 and worst <= 1.05e-8
