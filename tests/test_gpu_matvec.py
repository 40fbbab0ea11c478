"""GPU matvec (HIP kernels through the C ABI) against the oracle."""
import numpy as np
import pytest

import lippmann_schwinger_solver_amd as L
import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

# GPU vs oracle: same skeletons, fused operators and FMA on the device versus
# per-matvec QR solves on the CPU; the gate is 1e-10 relative (FP64).
TOL = 1e-10


def make(inp):
    S = L.DFMM(inp)
    S.FMMTreeInitialisation()
    return S


@pytest.mark.parametrize("depth,kappa,q,find_phi", [
    (3, 10.0, 0, 1), (4, 20.0, 0, 1), (4, 20.0, 0, 0), (4, 30.0, 6, 1), (5, 40.0, 6, 1)])
def test_matvec_matches_oracle(depth, kappa, q, find_phi):
    inp = H.inputs(depth=depth, kappa=kappa, q=q, find_phi=find_phi)
    print(f"case depth={depth} kappa={kappa} q={q} find_phi={find_phi}", flush=True)
    S = make(inp)
    o = O.Oracle(inp.to_c())
    assert S.N == o.N
    for seed in (0, 1):
        x = H.rand_vec(S.N, seed)
        e = H.relerr(S.MatVecProduct(x), o.matvec(x))
        print(f"  seed {seed}: rel err vs oracle {e:.3e}", flush=True)
        assert e < TOL, e
    # the reference's own right-hand side
    b = S.rhs()
    assert np.array_equal(b, o.rhs())
    assert H.relerr(S.MatVecProduct(b), o.matvec(b)) < TOL
    S.close()


# Rank 16: the fused upward operators P = Atilde^-1 Ar reach 5.9e7 at depth 4,
# so their cancellation amplifies rounding. The CPU numpy application of the
# same host plan is 2.05e-10 from the oracle, and rounding the operators to
# FP32 alone leaves 2e-2. Ranks 25 and 36 stay below 1.2e4 and at 1e-14.
@pytest.mark.parametrize("n,depth,kappa,tol64,tol32", [
    (4, 4, 20.0, 5e-10, None), (5, 4, 20.0, TOL, 1e-4), (5, 3, 10.0, TOL, 1e-4)])
def test_matvec_other_node_counts(n, depth, kappa, tol64, tol32):
    """nChebNodes 4 and 5 (ranks 16 and 25, degreeOfBases = nChebNodes)
    against the oracle. This covers the FP64 near field on MFMA at 2 and 4 real
    row tiles, and the passes at other row counts. The check is repeated after
    autotuning, which may switch the near-field kernel; the FP32 operator (FMA
    near field) is checked within FP32 accuracy where the operators allow it."""
    print(f"case nChebNodes={n} depth={depth} kappa={kappa}", flush=True)
    inp = L.InputsToDFMM(kappa=kappa, nLevelsUniform=depth, nChebNodes=n, degreeOfBases=n, TOL_POW=10)
    S = make(inp)
    o = O.Oracle(inp.to_c())
    assert S.N == o.N == 4 ** depth * n * n
    x = H.rand_vec(S.N, 7)
    ref = o.matvec(x)
    e = H.relerr(S.MatVecProduct(x), ref)
    S.autotune(reps=2)
    e_tuned = H.relerr(S.MatVecProduct(x), ref)
    e32 = H.relerr(S.MatVecProduct_fp32(x), ref)
    print(f"  rel err vs oracle {e:.3e}, after autotuning {e_tuned:.3e}, FP32 {e32:.3e}", flush=True)
    assert e < tol64 and e_tuned < tol64
    if tol32 is not None:
        assert e32 < tol32
    S.close()


def test_context_from_host_build():
    """One host build uploaded to two contexts (lsg_create_from_host): both
    apply the operator of a context that builds its own, bit for bit, and
    solve with it; the host build stays usable."""
    inp = H.inputs(depth=4, kappa=20.0, q=6, seed=1)
    hb = L.HostBuild(inp)
    S0 = make(inp)
    x = H.rand_vec(S0.N, 8)
    y0 = S0.MatVecProduct(x)
    ctxs = []
    for _ in range(2):
        S = L.DFMM(inp)
        S.FMMTreeInitialisation(host=hb)
        assert S.N == S0.N and np.array_equal(S.gridPoints, S0.gridPoints)
        assert np.array_equal(S.MatVecProduct(x), y0)
        ctxs.append(S)
    b = ctxs[1].rhs()
    xs, st = ctxs[1].gmres(b, restart_len=50, max_restarts=4, tol=1e-10)
    assert st["converged"] == 1 and H.relerr(S0.MatVecProduct(xs), b) < 2e-10
    assert hb.info()["N"] == S0.N and hb.plan()["ops"].size > 0
    for S in ctxs + [S0]:
        S.close()


def test_matvec_against_dense_small():
    """GPU matvec vs the dense getMatrixEntry2 operator (depth 3, N = 2304).
    With the reference's ACA (absolute exit 1e-8 on the cross vectors) the
    compression leaves 2.3e-8 (Gaussian contrast); with that exit at 1e-16
    and TOL_POW 14 (lsg_inputs.aca_exit_pow) the device matvec matches the
    dense operator to 1e-10, and the oracle built the same way to 1e-12."""
    inp = H.inputs(depth=3, kappa=10.0)
    S = make(inp)
    o = O.Oracle(inp.to_c())
    x = H.rand_vec(S.N, 3)
    assert H.relerr(S.MatVecProduct(x), o.dense_matvec(x)) < 5e-7
    for q in (0, 6):
        inp = H.inputs(depth=3, kappa=10.0, q=q, seed=0, tol_pow=14)
        inp.aca_exit_pow = 16
        S = make(inp)
        o = O.Oracle(inp.to_c())
        y = S.MatVecProduct(x)
        e_dense, e_or = H.relerr(y, o.dense_matvec(x)), H.relerr(y, o.matvec(x))
        print(f"q {q}: TOL_POW 14, ACA exit 1e-16: vs dense {e_dense:.2e}, vs oracle {e_or:.2e}", flush=True)
        assert e_dense < 1e-10 and e_or < 1e-12


def test_block_matvec_equals_single():
    """nrhs >= 2 runs the FP64 MFMA block operator: same operators, different
    summation order, so columns agree with the single-vector path to rounding."""
    S = make(H.inputs(depth=4, kappa=20.0))
    X = np.stack([H.rand_vec(S.N, s) for s in range(3)], axis=1)
    Y = S.matvec_block(X)
    for j in range(3):
        e = H.relerr(Y[:, j], S.MatVecProduct(X[:, j]))
        assert e < 1e-13, e


def test_zero_contrast_is_identity():
    S = make(H.inputs(depth=4, kappa=20.0, q=6, amp_scale=0.0))
    x = H.rand_vec(S.N, 4)
    assert np.array_equal(S.MatVecProduct(x), x)


def test_edge_inputs_rejected():
    S = make(H.inputs(depth=3, kappa=10.0))
    x = H.rand_vec(S.N, 0)
    x[5] = np.nan
    with pytest.raises(L.LsgError) as e:
        S.MatVecProduct(x)
    assert e.value.code == L.LSG_ERR_NUMERIC
    with pytest.raises(ValueError):
        S.MatVecProduct(np.zeros(S.N - 1, complex))
    # every solve wrapper checks lengths before the C side reads N entries
    b = S.rhs()
    short = np.ones(S.N - 1, complex)
    for call in (lambda: S.gmres(short), lambda: S.gmres(b, x0=short), lambda: S.gmres_ex(short),
                 lambda: S.gmres_ex(b, x0=short, mode="fp32ir"), lambda: S.gmres_batch(np.ones((S.N + 1, 2))),
                 lambda: S.gmres_batch(np.stack([b, b], 1), X0=np.zeros((S.N, 3))),
                 lambda: S.MatVecProduct_fp32(short), lambda: L.gmres_ragged([S], [short]),
                 lambda: L.gmres_ragged([S], [b, b])):
        with pytest.raises(ValueError):
            call()
    # a non-finite tolerance is an argument error, not the fixed-cycle mode
    with pytest.raises(L.LsgError) as e:
        S.gmres_ex(b, tol=float("nan"))
    assert e.value.code == L.LSG_ERR_ARG


def test_full_size_c2_linearity_and_determinism():
    """Config C2 (depth 6, 147,456 unknowns, kappa 80, 5 seeded bumps):
    linearity and run-to-run bit determinism of the device matvec."""
    S = make(H.inputs(depth=6, kappa=80.0, q=6, seed=0))
    x, y = H.rand_vec(S.N, 10), H.rand_vec(S.N, 11)
    ax, ay = S.MatVecProduct(x), S.MatVecProduct(y)
    a, b = 0.7 - 0.2j, -1.3 + 0.5j
    assert H.relerr(S.MatVecProduct(a * x + b * y), a * ax + b * ay) < 1e-13
    assert np.array_equal(S.MatVecProduct(x), ax)


@pytest.mark.parametrize("depth,kappa,q,tree", [(4, 30.0, 6, 0), (5, 40.0, 6, 0), (5, 40.0, 0, 1)])
def test_every_pass_kernel_matches_oracle(depth, kappa, q, tree):
    """Each of the three pass kernels (one workgroup per destination, one
    wavefront per destination, one lane per output row) forced on every pass,
    with and without the first three upward passes fused per subtree: the FP64
    matvec against the oracle (uniform trees with bumps, and an adaptive tree
    whose near field is a pass too), the FP32 matvec against the FP64 one, and
    the captured graph against the eager matvec."""
    print(f"case depth={depth} kappa={kappa} q={q} tree_mode={tree}", flush=True)
    inp = H.inputs(depth=depth, kappa=kappa, q=q)
    if tree:
        inp = L.InputsToDFMM(kappa=kappa, nLevelsUniform=depth, Qchoice=q, tree_mode=1, treeAdaptivity=5,
                             TOL_POW=8)
    S = make(inp)
    o = O.Oracle(inp.to_c())
    x = H.rand_vec(S.N, 11)
    ref = o.matvec(x)
    ys = {}
    for k in ("block", "wave", "lane", "block+fused", "wave+fused", "lane+fused"):
        S.set_pass_kernel(k.split("+")[0], fused=k.endswith("fused"))
        ys[k] = S.MatVecProduct(x)
        e = H.relerr(ys[k], ref)
        e32 = H.relerr(S.MatVecProduct_fp32(x), ys[k])
        S.use_graph(True)
        g = S.MatVecProduct(x)
        S.use_graph(False)
        print(f"  {k}: vs oracle {e:.3e}, FP32 vs FP64 {e32:.3e}", flush=True)
        assert e < TOL and e32 < 1e-3
        assert np.array_equal(g, ys[k])
    for k in ys:
        assert H.relerr(ys[k], ys["block"]) < 1e-13
    S.close()
