"""Block (multi right-hand-side) matvec on FP64 MFMA and batched GMRES."""
import numpy as np
import pytest

import lippmann_schwinger_solver_amd as L
import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL = 1e-10     # block matvec vs oracle, as for the single-vector matvec


def make(inp):
    S = L.DFMM(inp)
    S.FMMTreeInitialisation()
    return S


@pytest.mark.parametrize("find_phi,nrhs", [(1, 5), (0, 3), (1, 70)])
def test_block_matvec_matches_oracle(find_phi, nrhs):
    inp = H.inputs(depth=4, kappa=30.0, q=6, find_phi=find_phi)
    print(f"case find_phi={find_phi} nrhs={nrhs}", flush=True)
    S = make(inp)
    o = O.Oracle(inp.to_c())
    X = np.stack([H.rand_vec(S.N, 100 + s) for s in range(nrhs)], axis=1)
    Y = S.matvec_block(X)
    worst_single = worst_oracle = 0.0
    for j in sorted({0, 1, nrhs // 2, nrhs - 1, min(nrhs - 1, 64)}):
        worst_single = max(worst_single, H.relerr(Y[:, j], S.MatVecProduct(X[:, j])))
        worst_oracle = max(worst_oracle, H.relerr(Y[:, j], o.matvec(X[:, j])))
    print(f"  block vs single {worst_single:.2e}, vs oracle {worst_oracle:.2e}", flush=True)
    assert worst_single < 1e-13 and worst_oracle < TOL


def test_rhs_angles_matches_rhs():
    inp = H.inputs(depth=3, kappa=10.0, q=6)
    inp.inc_angle = 0.7
    S = make(inp)
    b = S.rhs_angles([0.7])[:, 0]
    assert H.relerr(b, S.rhs()) < 1e-14


def test_batched_gmres_small():
    inp = H.inputs(depth=4, kappa=20.0, q=6)
    S = make(inp)
    angles = 2 * np.pi * np.arange(8) / 8
    B = S.rhs_angles(angles)
    X, st, rel = S.gmres_batch(B, restart_len=30, max_restarts=10, tol=1e-8)
    print(f"batched: {st}, rel {rel}", flush=True)
    assert st["status"] == 0 and st["converged"] == 8
    for j in range(8):
        r = H.relerr(S.MatVecProduct(X[:, j]), B[:, j])       # independent FP64 residual
        assert rel[j] <= 1e-8 and r <= 1.5e-8, (j, rel[j], r)
        xs, _ = S.gmres(B[:, j], restart_len=30, max_restarts=10, tol=1e-8)
        assert H.relerr(X[:, j], xs) < 1e-6


def test_batched_gmres_c2_64_angles():
    """64 plane waves at 2 pi j / 64 on the C2 geometry (depth 6, kappa 80, 5 bumps)."""
    S = make(H.inputs(depth=6, kappa=80.0, q=6, seed=0))
    angles = 2 * np.pi * np.arange(64) / 64
    B = S.rhs_angles(angles)
    X, st, rel = S.gmres_batch(B, restart_len=50, max_restarts=20, tol=1e-8)
    print(f"C2 x 64 angles: {st['iterations']} steps, {st['restarts']} cycles, {st['seconds']:.3f} s, "
          f"max rel residual {rel.max():.2e}", flush=True)
    assert st["status"] == 0 and st["converged"] == 64 and rel.max() <= 1e-8
    for j in (0, 17, 63):
        assert H.relerr(S.MatVecProduct(X[:, j]), B[:, j]) <= 1.5e-8


def test_batched_gmres_argument_errors():
    S = make(H.inputs(depth=3, kappa=10.0))
    B = np.stack([S.rhs()] * 2, axis=1)
    with pytest.raises(L.LsgError) as e:
        S.gmres_batch(np.zeros((S.N, 65), complex), tol=1e-8)
    assert e.value.code == L.LSG_ERR_ARG
    with pytest.raises(L.LsgError) as e:
        S.gmres_batch(B, tol=0.0)
    assert e.value.code == L.LSG_ERR_ARG
    B[:, 1] = 0
    with pytest.raises(L.LsgError) as e:
        S.gmres_batch(B, tol=1e-8)
    assert e.value.code == L.LSG_ERR_NUMERIC


def test_fp32_block_matvec_and_batched_refinement():
    """FP32 block operator (FP32 MFMA) inside FP64 refinement of the batch: every
    system reaches 1e-8 in FP64, checked again through the single-vector matvec."""
    S = make(H.inputs(depth=5, kappa=40.0, q=6, seed=1))
    angles = 2 * np.pi * np.arange(16) / 16
    B = S.rhs_angles(angles)
    X, st, rel = S.gmres_batch(B, restart_len=50, max_restarts=1, tol=1e-8, mode="fp32ir", inner_tol=1e-5)
    X64, s64, _ = S.gmres_batch(B, restart_len=50, max_restarts=20, tol=1e-8)
    print(f"batched FP32-IR: {st}; FP64: {s64['iterations']} steps", flush=True)
    assert st["status"] == 0 and st["converged"] == 16 and rel.max() <= 1e-8
    assert st["refinements"] >= 1
    for j in (0, 7, 15):
        assert H.relerr(S.MatVecProduct(X[:, j]), B[:, j]) <= 1.5e-8
        assert H.relerr(X[:, j], X64[:, j]) < 1e-6


@pytest.mark.parametrize("m,K,arith", [(24, 30, "split_bf16"), (8, 40, "split_bf16"), (24, 30, "fp32")])
def test_batched_flexible_gmres_fp32_inner(m, K, arith):
    """FP64 flexible GMRES(K) of the batch, each outer step preconditioned by
    one FP32-inner batched GMRES(m) solve (split-BF16 or FP32 MFMA operators,
    FP32 pass state, FP16 basis): every system reaches 1e-8 in FP64
    (recomputed through the single-vector matvec and checked through the
    oracle's operator for two of them), with the same solutions as the FP64
    batch to the conditioning times the tolerance; refinement with the same
    inner solves is run beside it for the step counts."""
    inp = H.inputs(depth=5, kappa=40.0, q=6, seed=1)
    print(f"case m={m} K={K} arith={arith}", flush=True)
    S = make(inp)
    S.set_inner_arith(arith)
    o = O.Oracle(inp.to_c())
    angles = 2 * np.pi * np.arange(16) / 16
    B = S.rhs_angles(angles)
    X, st, rel = S.gmres_batch(B, restart_len=m, max_restarts=1, tol=1e-8, mode="fp32fgmres", inner_tol=1e-5,
                               max_outer=K)
    Xi, si, _ = S.gmres_batch(B, restart_len=m, max_restarts=1, tol=1e-8, mode="fp32ir", inner_tol=1e-5,
                              max_outer=80)
    X64, s64, _ = S.gmres_batch(B, restart_len=50, max_restarts=20, tol=1e-8)
    print(f"m={m} K={K}: flexible {st}; refinement {si['iterations']} inner steps, {si['refinements']} refinements; "
          f"FP64 {s64['iterations']} steps", flush=True)
    assert st["status"] == 0 and st["converged"] == 16 and rel.max() <= 1e-8 and st["refinements"] >= 1
    for j in range(16):
        assert H.relerr(S.MatVecProduct(X[:, j]), B[:, j]) <= 1.05e-8
        assert H.relerr(X[:, j], X64[:, j]) < 1e-6
    for j in (0, 9):
        assert H.relerr(o.matvec(X[:, j]), B[:, j]) <= 1.05e-8
    S.close()
    o.close()


@pytest.mark.parametrize("depth,kappa", [(4, 30.0), (5, 40.0)])
def test_fp32_block_matvec_tiles_per_wave(depth, kappa):
    """FP32 block passes with one and with two right-hand-side tiles per
    wavefront: the same products to FP32 rounding, both within FP32 accuracy of
    the FP64 block matvec, on all 64 columns."""
    S = make(H.inputs(depth=depth, kappa=kappa, q=6, seed=2))
    S.set_inner_arith("fp32")
    X = np.stack([H.rand_vec(S.N, 40 + j) for j in range(64)], axis=1)
    Y64 = S.matvec_block(X)
    ys = {}
    for nt in (1, 2):
        S.set_block_tiles(nt)
        ys[nt] = S.matvec_block_fp32(X)
        e = max(H.relerr(ys[nt][:, j], Y64[:, j]) for j in range(64))
        print(f"tiles {nt}: FP32 vs FP64 block {e:.2e}", flush=True)
        assert e < 1e-3
    assert max(H.relerr(ys[2][:, j], ys[1][:, j]) for j in range(64)) < 1e-5
    S.close()


@pytest.mark.parametrize("depth,kappa,find_phi", [(4, 30.0, 1), (5, 40.0, 1), (4, 30.0, 0)])
def test_split_bf16_block_matvec(depth, kappa, find_phi):
    """Split-BF16 block operator (each FP32 value as a bf16 pair, the pairs'
    products as four BF16 MFMA terms, FP32 sums) against the FP64 block matvec
    on all 64 columns,
    with one and with two right-hand-side tiles per wavefront (the same sums in
    the same order: equal results), and with the near field beside the passes
    (one more FP32 rounding: within 1e-5). The operator amplifies arithmetic error
    about 80-fold (the FP32 block operator, unit roundoff 6e-8, lands near 5e-6
    here); the split values carry about 2^-17 each, which puts it near 1e-4:
    the bound is 1e-3, and it must be less accurate than FP32 (the
    two arithmetics really differ)."""
    S = make(H.inputs(depth=depth, kappa=kappa, q=6, seed=2, find_phi=find_phi))
    print(f"case depth={depth} kappa={kappa} find_phi={find_phi}", flush=True)
    X = np.stack([H.rand_vec(S.N, 40 + j) for j in range(64)], axis=1)
    Y64 = S.matvec_block(X)
    S.set_inner_arith("fp32")
    Y32 = S.matvec_block_fp32(X)
    S.set_inner_arith("split_bf16")
    ys = {}
    for nt in (1, 2, 3, 4):
        S.set_block_tiles(nt)
        ys[nt] = S.matvec_block_fp32(X)
        e = max(H.relerr(ys[nt][:, j], Y64[:, j]) for j in range(64))
        e32 = max(H.relerr(Y32[:, j], Y64[:, j]) for j in range(64))
        print(f"tiles {nt}: split-BF16 vs FP64 block {e:.2e} (FP32 {e32:.2e})", flush=True)
        assert e < 1e-3
        assert e > e32
    assert max(H.relerr(ys[2][:, j], ys[1][:, j]) for j in range(64)) < 1e-6
    # near field beside the passes: the leaf passes are added to the stored
    # FP32 near-field rows (one more FP32 rounding)
    for nt in (3, 4):
        assert max(H.relerr(ys[nt][:, j], ys[1][:, j]) for j in range(64)) < 1e-5
    S.close()


def test_low_precision_batch_with_contrast_per_column():
    """A volume-potential context with a different contrast per right-hand side
    (A_q = I + diag(qk2[:, q]) V): the low-precision block operator applies the
    contrasts after its product (FP64 product rows and a diagonal pass) in every
    layout (FP32 MFMA; split BF16 with one or two tiles per wavefront, the near
    field after or beside the passes), checked column by column against the
    FP64 block matvec; the batch then needs no more inner steps than with FP32
    MFMA plus a margin, and every system reaches 1e-8 in FP64 through its own
    operator. (Flexible GMRES converges even around a wrong inner operator, only
    slowly: the step count and the matvec check are what catch one.)"""
    S = make(H.inputs(depth=4, kappa=20.0, q=6, seed=0, find_phi=0))
    x, y = S.gridPoints[:, 0], S.gridPoints[:, 1]
    qk2 = np.stack([(20.0 ** 2) * (0.5 + 0.1 * j) * np.exp(-((x - 0.05 * j) ** 2 + y ** 2) / 0.02)
                    for j in range(6)], axis=1)
    S.set_contrast(qk2)
    Xr = np.stack([H.rand_vec(S.N, 50 + j) for j in range(6)], axis=1)
    Y64 = S.matvec_block(Xr)
    S.set_inner_arith("fp32")
    e32 = max(H.relerr(a, b) for a, b in zip(S.matvec_block_fp32(Xr).T, Y64.T))
    B = np.stack([-qk2[:, j] * np.exp(1j * 20.0 * (x * np.cos(j) + y * np.sin(j))) for j in range(6)], axis=1)
    _, s32, _ = S.gmres_batch(B, restart_len=20, max_restarts=1, tol=1e-8, mode="fp32fgmres", inner_tol=1e-5,
                              max_outer=30)
    S.set_inner_arith("split_bf16")
    esb = {}
    for t in (1, 2, 3, 4):
        S.set_block_tiles(t)
        esb[t] = max(H.relerr(a, b) for a, b in zip(S.matvec_block_fp32(Xr).T, Y64.T))
    X, st, rel = S.gmres_batch(B, restart_len=20, max_restarts=1, tol=1e-8, mode="fp32fgmres", inner_tol=1e-5,
                               max_outer=30)
    print(f"per-column contrast: block FP32 {e32:.2e}, split BF16 {esb}; batch {st}; FP32 MFMA batch "
          f"{s32['iterations']} inner steps", flush=True)
    assert e32 < 1e-5
    assert all(e < 1e-3 for e in esb.values())
    assert st["status"] == 0 and st["converged"] == 6 and rel.max() <= 1e-8
    assert st["iterations"] <= s32["iterations"] + 40
    for j in (0, 5):
        S.set_contrast(qk2[:, j])
        assert H.relerr(S.MatVecProduct(X[:, j]), B[:, j]) <= 1.05e-8
    S.close()
