"""The product's host build (tree, tables, NCA) against the oracle and the
reference's structural pins. No GPU needed."""
import ctypes
import os
import re

import numpy as np
import pytest

import lippmann_schwinger_solver_amd as L
import oracle as O
from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in ("lsg.h", "lsg_host.h", "lsg_learn.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(lsgn?_[a-z_0-9]+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    decl = declared_symbols()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(n for n, _, _ in L.SYMBOLS) == decl


def test_create_without_gpu_fails_loudly():
    """No CPU fallback: lsg_create refuses to run without a HIP device."""
    import torch  # noqa: F401  (only to learn whether a GPU exists)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    S = L.DFMM(H.inputs(depth=3, kappa=10.0))
    with pytest.raises(L.LsgError):
        S.FMMTreeInitialisation()


def test_context_from_host_build_checks():
    """lsg_create_from_host rejects a tree-only build before it looks for a
    device, and without a GPU it fails loudly like lsg_create."""
    inp = H.inputs(depth=3, kappa=10.0)
    tree_only = L.HostBuild(inp, tree_only=True)
    with pytest.raises(L.LsgError) as e:
        L.DFMM(inp).FMMTreeInitialisation(host=tree_only)
    assert e.value.code == L.LSG_ERR_ARG and "tree" in str(e.value)
    import torch  # noqa: F401  (only to learn whether a GPU exists)
    if not torch.cuda.is_available():
        with pytest.raises(L.LsgError) as e:
            L.DFMM(inp).FMMTreeInitialisation(host=L.HostBuild(inp))
        assert e.value.code == L.LSG_ERR_DEVICE


@pytest.mark.parametrize("depth,kappa,expect", [
    (4, 20.0, (3, [0, 0, 116, 1756, 5628])),
    (6, 80.0, (5, [0, 0, 0, 2600, 20532, 45628, 105276])),
    (7, 160.0, (6, [0, 0, 0, 732, 46424, 103924, 195516, 431676])),
])
def test_tree_structure_pins(depth, kappa, expect):
    """level_LFR and interaction-list pair counts per level (LFR IL and HFR
    per-cone IL, G_FMM2DTree.hpp:1356-1461) of the three uniform configs."""
    h = L.HostBuild(H.inputs(depth=depth, kappa=kappa), tree_only=True)
    inf = h.info()
    assert inf["level_LFR"] == expect[0]
    assert inf["N"] == 36 * 4 ** depth
    assert h.il_pairs().tolist() == expect[1]


@pytest.mark.parametrize("depth,kappa,q", [(3, 10.0, 0), (4, 20.0, 0), (4, 30.0, 6), (5, 40.0, 6)])
def test_tables_and_skeletons_match_oracle(depth, kappa, q):
    """Leaf tables after the 6-digit round trip and every NCA skeleton (pivot
    lists of getNodes_LFR/HFR and the user check points) are identical."""
    inp = H.inputs(depth=depth, kappa=kappa, q=q)
    h = L.HostBuild(inp)
    o = O.Oracle(inp.to_c())
    assert np.array_equal(h.far_table(), o.far_table())
    assert np.array_equal(h.near_tables(), o.near_tables())
    assert np.array_equal(h.qpinv(), o.qpinv())
    assert np.array_equal(h.skeleton(), o.skeleton())


def test_skeletons_match_oracle_with_a_smaller_aca_exit():
    """lsg_inputs.aca_exit_pow (the ACA's absolute exit 10^-p instead of the
    reference's 1e-8) changes the skeletons the same way in the product's
    build and in the oracle."""
    inp = H.inputs(depth=4, kappa=30.0, q=6, seed=2, tol_pow=14)
    ref = L.HostBuild(inp).skeleton()
    inp.aca_exit_pow = 16
    h = L.HostBuild(inp)
    o = O.Oracle(inp.to_c())
    assert np.array_equal(h.skeleton(), o.skeleton())
    assert h.skeleton().size > ref.size          # the ACAs run longer: larger skeletons


def test_bumps_are_seeded_and_in_range():
    a = L.bumps_from_seed(7, 8)
    b = L.bumps_from_seed(7, 8)
    assert bytes(a) == bytes(b)
    assert a.n_bumps == 8 and a.kind == 6
    for i in range(8):
        assert -0.35 <= a.cx[i] <= 0.35 and -0.35 <= a.cy[i] <= 0.35
        assert 0.03 <= a.width[i] <= 0.10 and 0.5 <= a.amp[i] <= 1.5


def test_invalid_inputs_are_rejected():
    for bad in (dict(nChebNodes=1), dict(kappa=-1.0), dict(TOL_POW=0), dict(nLevelsUniform=1), dict(tree_mode=2)):
        inp = H.inputs(depth=3, kappa=10.0)
        for k, v in bad.items():
            setattr(inp, k, v)
        with pytest.raises(L.LsgError) as e:
            L.HostBuild(inp)
        assert e.value.code == L.LSG_ERR_ARG


@pytest.mark.parametrize("depth,kappa,q,find_phi", [(3, 10.0, 0, 1), (4, 20.0, 0, 1), (4, 20.0, 0, 0), (4, 30.0, 6, 1),
                                                    (5, 40.0, 6, 1)])
def test_fused_plan_matches_oracle(depth, kappa, q, find_phi):
    """The fused device operators (Atilde^-1 Ar upward, K Atilde_in^-1 downward,
    pass-throughs folded into their children) applied by a numpy mirror of the
    HIP kernels reproduce the oracle's matvec."""
    from tests import plan_numpy as P
    inp = H.inputs(depth=depth, kappa=kappa, q=q, find_phi=find_phi)
    h = L.HostBuild(inp)
    o = O.Oracle(inp.to_c())
    x = H.rand_vec(o.N, 0)
    y = P.apply_plan(h.plan(), x, h.info()["n_state"], find_phi=bool(find_phi))
    e = H.relerr(y, o.matvec(x))
    assert e < 1e-10, e


@pytest.mark.parametrize("depth,kappa,q", [(3, 10.0, 0), (4, 30.0, 6)])
def test_entry_hook_matches_oracle_far_entry(depth, kappa, q):
    """lsg_host_entries (host kernel, the one the GPU entry kernel is checked
    against) equals the oracle's getMatrixEntry on well-separated blocks."""
    inp = H.inputs(depth=depth, kappa=kappa, q=q)
    h = L.HostBuild(inp)
    o = O.Oracle(inp.to_c())
    N = h.info()["N"]
    rows = np.array([0, 5, 17, 35, 36, 71])
    cols = np.arange(N - 72, N, 7)
    E = h.entries(rows, cols)
    ref = np.array([[o.far_entry(int(i), int(j)) for j in cols] for i in rows])
    assert np.max(np.abs(E - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_aca_hook_host_rank_and_pivots_are_valid():
    """lsg_host_aca (host path): pivots are distinct positions in range and the
    pivot block K(row pivots, column pivots) is nonsingular."""
    inp = H.inputs(depth=4, kappa=30.0, q=6)
    h = L.HostBuild(inp)
    N = h.info()["N"]
    # check points on the leaf of largest contrast (the kernel rows scale with k^2 q)
    lead = int(np.argmax(h.plan()["qk2"])) // 36
    rows = np.arange(36 * lead, 36 * lead + 36)
    far = 0 if lead >= N // 72 else N // 36 - 1
    cols = np.arange(36 * far, 36 * far + 36)
    k, rp, cp = h.aca(rows, cols)
    assert 0 < k <= 36
    assert len(set(rp.tolist())) == k and len(set(cp.tolist())) == k
    assert rp.max() < rows.size and cp.max() < cols.size
    S = h.entries(rows[rp], cols[cp])
    assert np.linalg.matrix_rank(S) == k
