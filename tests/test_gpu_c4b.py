"""The ragged batch at its configured size: 48 independent scatterers, tree
depth uniform in {4, 5, 6}, kappa uniform in [20, 80], 3-8 seeded bumps and a
seeded plane-wave angle each, solved together to tol 1e-8. Every system is
FP64-verified through its own context's FP64 matvec; the depth-4 systems also
through the oracle's operator (the reference solves one system per
RestartedGMRES call, GMRES/GMRES.cpp:307-319)."""
import time

import numpy as np
import pytest

import lippmann_schwinger_solver_amd as L
import oracle as O
from tests import helpers as H
from tests.test_gpu_configs import seeded_angle

pytestmark = pytest.mark.gpu


def ragged_cases(n=48, seed=2024):
    """(depth, kappa, bumps) of scatterer i, from one numpy stream."""
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(4, 7)), float(rng.uniform(20, 80)), int(rng.integers(3, 9))) for _ in range(n)]


def scatterer_inputs(i, depth, kappa, nb):
    inp = L.InputsToDFMM(kappa=kappa, nLevelsUniform=depth, L=0.5, nChebNodes=6, degreeOfBases=6, TOL_POW=10,
                         nCones_LFR=16)
    inp.contrast = L.bumps_from_seed(100 + i, nb)
    inp.inc_angle = seeded_angle(100 + i)
    return inp


@pytest.fixture(scope="module")
def ragged48():
    cases = ragged_cases()
    t0 = time.time()
    solvers, rhs, inps = [], [], []
    for i, (depth, kappa, nb) in enumerate(cases):
        inp = scatterer_inputs(i, depth, kappa, nb)
        S = L.DFMM(inp)
        S.FMMTreeInitialisation()
        solvers.append(S)
        rhs.append(S.rhs())
        inps.append(inp)
    print(f"48 scatterers built in {time.time() - t0:.1f} s: depths "
          f"{[c[0] for c in cases].count(4)}/{[c[0] for c in cases].count(5)}/{[c[0] for c in cases].count(6)} "
          f"(4/5/6), {sum(S.N for S in solvers)} unknowns", flush=True)
    yield cases, solvers, rhs, inps
    for S in solvers:
        S.close()


def test_ragged_batch_48(ragged48):
    cases, solvers, rhs, inps = ragged48
    assert sorted(set(c[0] for c in cases)) == [4, 5, 6]
    t0 = time.perf_counter()
    xs, stats, rc = L.gmres_ragged(solvers, rhs, restart_len=50, max_restarts=20, tol=1e-8)
    t = time.perf_counter() - t0
    print(f"ragged batch: {48 / t:.1f} solves/s, steps {[s['iterations'] for s in stats]}", flush=True)
    assert rc == 0
    worst = 0.0
    for S, b, x, st in zip(solvers, rhs, xs, stats):
        assert st["converged"] == 1 and st["rel_residual"] <= 1e-8
        worst = max(worst, H.relerr(S.MatVecProduct(x), b))
    print(f"recomputed FP64 residuals: max {worst:.3e}", flush=True)
    assert worst <= 1.05e-8
    # the depth-4 scatterers through the oracle's operator of the same inputs
    for (depth, _, _), inp, b, x in zip(cases, inps, rhs, xs):
        if depth == 4:
            o = O.Oracle(inp.to_c())
            assert H.relerr(o.matvec(x), b) <= 1.05e-8
            o.close()
