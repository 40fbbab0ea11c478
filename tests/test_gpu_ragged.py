"""hipGraph-captured matvec, the kernel-variant autotuner, and the ragged batch
(independent scatterers of different depth, wavenumber and contrast)."""
import numpy as np
import pytest

import lippmann_schwinger_solver_amd as L
import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def make(inp):
    S = L.DFMM(inp)
    S.FMMTreeInitialisation()
    return S


def test_graph_replays_match_eager():
    """Several replays of the captured matvec, with eager solver work in between,
    reproduce the eager matvec bit for bit (same kernels, same order)."""
    S = make(H.inputs(depth=4, kappa=30.0, q=6, seed=2))
    xs = [H.rand_vec(S.N, 20 + k) for k in range(3)]
    ref = [S.MatVecProduct(x) for x in xs]
    S.use_graph(True)
    for rep in range(3):
        for x, r in zip(xs, ref):
            assert np.array_equal(S.MatVecProduct(x), r)
        S.gmres(S.rhs(), restart_len=10, max_restarts=1, tol=1e-3)      # eager-free solver work in between
    S.use_graph(False)
    assert np.array_equal(S.MatVecProduct(xs[0]), ref[0])


def test_gmres_with_graph_matches_without():
    S = make(H.inputs(depth=4, kappa=20.0, q=6, seed=3))
    b = S.rhs()
    x0, s0 = S.gmres(b, restart_len=40, max_restarts=10, tol=1e-9)
    S.use_graph(True)
    x1, s1 = S.gmres(b, restart_len=40, max_restarts=10, tol=1e-9)
    assert s0["iterations"] == s1["iterations"] and np.array_equal(x0, x1)


def test_set_contrast_after_graph_capture():
    """A contrast set or removed after use_graph(True) reaches the replayed
    matvec: the graph is captured again around the new contrast buffer."""
    S = make(H.inputs(depth=4, kappa=20.0, q=6, seed=5, find_phi=0))
    x = H.rand_vec(S.N, 6)
    v = S.MatVecProduct(x)                           # V x (no contrast yet)
    qk2 = np.abs(H.rand_vec(S.N, 7).real) * 100.0
    S.use_graph(True)
    S.set_contrast(qk2)
    assert H.relerr(S.MatVecProduct(x), x + qk2 * v) < 1e-13
    S.set_contrast(2.0 * qk2)
    assert H.relerr(S.MatVecProduct(x), x + 2.0 * qk2 * v) < 1e-13
    S.set_contrast(None)
    assert np.array_equal(S.MatVecProduct(x), v)
    S.use_graph(False)
    assert np.array_equal(S.MatVecProduct(x), v)


def test_autotune_keeps_the_operator():
    inp = H.inputs(depth=4, kappa=30.0, q=6, seed=4)
    S = make(inp)
    o = O.Oracle(inp.to_c())
    x = H.rand_vec(S.N, 5)
    y0 = S.MatVecProduct(x)
    n, report = S.autotune(reps=3)
    print(report, flush=True)
    assert report.count("\n") == len(report.strip().splitlines()) and "near_field" in report
    y1 = S.MatVecProduct(x)
    assert H.relerr(y1, y0) < 1e-13 and H.relerr(y1, o.matvec(x)) < 1e-10


def test_ragged_batch():
    """Six scatterers: depth 3-5, kappa 20-60, own bump contrasts and angles;
    solved concurrently (one thread and stream pair per context, graphs on)."""
    cases = [(3, 20.0, 1), (4, 35.0, 2), (5, 50.0, 3), (4, 60.0, 4), (3, 45.0, 5), (5, 25.0, 6)]
    solvers, rhs = [], []
    for depth, kappa, seed in cases:
        S = make(H.inputs(depth=depth, kappa=kappa, q=6, seed=seed, nbumps=3 + seed % 6, angle=0.3 * seed))
        S.autotune(reps=2)
        S.use_graph(True)
        solvers.append(S)
        rhs.append(S.rhs())
    xs, stats, rc = L.gmres_ragged(solvers, rhs, restart_len=50, max_restarts=20, tol=1e-8)
    print([s["iterations"] for s in stats], flush=True)
    assert rc == 0
    for S, b, x, st in zip(solvers, rhs, xs, stats):
        assert st["converged"] == 1 and st["rel_residual"] <= 1e-8
        assert H.relerr(S.MatVecProduct(x), b) <= 1.5e-8
        xs1, s1 = S.gmres(b, restart_len=50, max_restarts=20, tol=1e-8)      # sequential solve, same answer
        assert s1["iterations"] == st["iterations"] and H.relerr(x, xs1) < 1e-12
