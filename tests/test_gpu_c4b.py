"""The ragged batch at its configured size: 48 independent scatterers, tree
depth uniform in {4, 5, 6}, kappa uniform in [20, 80], 3-8 seeded bumps and a
seeded plane-wave angle each, solved together to tol 1e-8.

The fused batch (lsg_ragged_*) applies all 48 operators in one launch
sequence and solves all systems at once with per-system Krylov bases; it is
checked against each scatterer's own context (matvec, iterations, solution),
every system is FP64-verified through its context's FP64 matvec, and the
depth-4 systems through the oracle's operator as well (the reference solves
one system per RestartedGMRES call, GMRES/GMRES.cpp:307-319). The concurrent
contexts (lsg_gmres_ragged) solve the same batch."""
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
    hosts, solvers, rhs, inps = [], [], [], []
    for i, (depth, kappa, nb) in enumerate(cases):
        inp = scatterer_inputs(i, depth, kappa, nb)
        h = L.HostBuild(inp)
        S = L.DFMM(inp)
        S.FMMTreeInitialisation(host=h)        # the context uploads the same build
        hosts.append(h)
        solvers.append(S)
        rhs.append(S.rhs())
        inps.append(inp)
    t1 = time.time()
    R = L.RaggedBatch(hosts)
    print(f"48 scatterers built in {t1 - t0:.1f} s, fused in {time.time() - t1:.1f} s: depths "
          f"{[c[0] for c in cases].count(4)}/{[c[0] for c in cases].count(5)}/{[c[0] for c in cases].count(6)} "
          f"(4/5/6), {sum(R.sizes)} unknowns", flush=True)
    yield cases, hosts, solvers, rhs, inps, R
    R.close()
    for S in solvers:
        S.close()


def test_fused_matvec_matches_each_context(ragged48):
    cases, hosts, solvers, rhs, inps, R = ragged48
    assert R.sizes == [S.N for S in solvers] and sorted(set(c[0] for c in cases)) == [4, 5, 6]
    xs = [H.rand_vec(S.N, 300 + i) for i, S in enumerate(solvers)]
    ys = R.matvec(xs)
    worst = max(H.relerr(y, S.MatVecProduct(x)) for S, x, y in zip(solvers, xs, ys))
    R.autotune(reps=2)
    R.use_graph(True)
    ys2 = R.matvec(xs)
    worst2 = max(H.relerr(y, S.MatVecProduct(x)) for S, x, y in zip(solvers, xs, ys2))
    print(f"fused matvec vs contexts: {worst:.2e}, after autotune + graph {worst2:.2e}; "
          f"fused matvec {R.time_matvec(reps=5):.3f} ms", flush=True)
    # the same per-destination arithmetic; the kernel variant of a pass (one
    # workgroup, one wavefront or one lane per destination row) may differ
    # from the context's, which changes the summation order (one running sum
    # per row against 2-256 interleaved partial sums). The fused upward
    # operators carry inverses of skeleton blocks and cancel downstream, so
    # that rounding shows at up to 1.5e-12 (measured after autotuning)
    assert worst <= 1e-11 and worst2 <= 1e-11


def test_fused_ragged_gmres_48(ragged48):
    cases, hosts, solvers, rhs, inps, R = ragged48
    R.gmres(rhs, restart_len=50, max_restarts=20, tol=1e-8)           # warm-up (workspace)
    t0 = time.perf_counter()
    xs, stats, rc = R.gmres(rhs, restart_len=50, max_restarts=20, tol=1e-8)
    t_fused = time.perf_counter() - t0
    assert rc == 0
    t0 = time.perf_counter()
    seq = [S.gmres(b, restart_len=50, max_restarts=20, tol=1e-8) for S, b in zip(solvers, rhs)]
    t_seq = time.perf_counter() - t0
    print(f"fused ragged GMRES: {48 / t_fused:.1f} solves/s; one after the other {48 / t_seq:.1f} solves/s "
          f"({t_seq / t_fused:.2f}x); steps {[s['iterations'] for s in stats]}", flush=True)
    worst = 0.0
    for S, b, x, st, (xq, sq) in zip(solvers, rhs, xs, stats, seq):
        assert st["converged"] == 1 and st["rel_residual"] <= 1e-8
        worst = max(worst, H.relerr(S.MatVecProduct(x), b))
        # the same per-system algorithm: the same steps up to a rounding tie at
        # the stopping test, and the same solution to the conditioning of tol
        assert abs(st["iterations"] - sq["iterations"]) <= 2
        assert H.relerr(x, xq) <= 1e-6
    print(f"recomputed FP64 residuals: max {worst:.3e}", flush=True)
    assert worst <= 1.05e-8
    for (depth, _, _), inp, b, x in zip(cases, inps, rhs, xs):
        if depth == 4:
            o = O.Oracle(inp.to_c())
            assert H.relerr(o.matvec(x), b) <= 1.05e-8
            o.close()


def test_concurrent_contexts_48(ragged48):
    """lsg_gmres_ragged: the contexts solved concurrently, one host thread each."""
    cases, hosts, solvers, rhs, inps, R = ragged48
    t0 = time.perf_counter()
    xs, stats, rc = L.gmres_ragged(solvers, rhs, restart_len=50, max_restarts=20, tol=1e-8)
    t = time.perf_counter() - t0
    print(f"concurrent contexts: {48 / t:.1f} solves/s", flush=True)
    assert rc == 0
    worst = max(H.relerr(S.MatVecProduct(x), b) for S, b, x in zip(solvers, rhs, xs))
    assert all(st["converged"] == 1 for st in stats) and worst <= 1.05e-8
