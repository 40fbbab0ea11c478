"""MI355X-native DAFMM+NCA solver for the 2-D Lippmann-Schwinger equation.

Python mirror of the reference's matvec/GMRES interface (GMRES/GMRES.cpp):

    inputs = InputsToDFMM(nCones_LFR=16, nChebNodes=6, L=0.5, kappa=20.0, ...)
    S = DFMM(inputs)                 # GMRES/GMRES.cpp:152-157
    S.FMMTreeInitialisation()        # GMRES/GMRES.cpp:161-199
    out = S.MatVecProduct(b)         # GMRES/GMRES.cpp:201-254
    phi = RestartedGMRES(restart, S, m, inputs, b)   # GMRES/GMRES.cpp:307-319

Everything runs through the C ABI of liblsg.so (include/lsg.h); there is no
CPU execution path - DFMM.FMMTreeInitialisation raises when no GPU is usable.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblsg.so")
MAX_BUMPS = 16

LSG_OK = 0
LSG_ERR_ARG = -1
LSG_ERR_DEVICE = -2
LSG_ERR_SETUP = -3
LSG_ERR_NUMERIC = -4
LSG_ERR_NOCONV = -5


class LsgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lsg error {code}: {msg}")
        self.code = code


class lsg_contrast(C.Structure):
    _fields_ = [("kind", C.c_int), ("n_bumps", C.c_int), ("cx", C.c_double * MAX_BUMPS),
                ("cy", C.c_double * MAX_BUMPS), ("width", C.c_double * MAX_BUMPS),
                ("amp", C.c_double * MAX_BUMPS)]


class lsg_inputs(C.Structure):
    _fields_ = [("nCones_LFR", C.c_int), ("nChebNodes", C.c_int), ("treeAdaptivity", C.c_int),
                ("L", C.c_double), ("kappa", C.c_double), ("yes2DFMM", C.c_int),
                ("degreeOfBases", C.c_int), ("TOL_POW", C.c_int), ("m", C.c_int), ("restart", C.c_int),
                ("nLevelsUniform", C.c_int), ("Qchoice", C.c_int), ("tree_mode", C.c_int),
                ("find_phi", C.c_int), ("inc_angle", C.c_double), ("contrast", lsg_contrast),
                ("setup_device", C.c_int), ("precomputations_dir", C.c_char_p), ("aca_exit_pow", C.c_int)]


class lsg_info(C.Structure):
    _fields_ = [("N", C.c_int64), ("nLevels", C.c_int), ("level_LFR", C.c_int), ("n_leaves", C.c_int),
                ("E_far", C.c_int64), ("E_near", C.c_int64), ("bytes_matvec", C.c_int64),
                ("flops_matvec", C.c_int64), ("n_state", C.c_int64), ("max_rank", C.c_int),
                ("setup_seconds", C.c_double), ("setup_phase_seconds", C.c_double * 4)]


def _info_dict(inf):
    d = {k: getattr(inf, k) for k, _ in lsg_info._fields_}
    d["setup_phase_seconds"] = list(d["setup_phase_seconds"])
    return d


class lsg_stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("restarts", C.c_int), ("converged", C.c_int),
                ("rel_residual", C.c_double), ("est_residual", C.c_double), ("seconds", C.c_double),
                ("refinements", C.c_int)]


LSG_SOLVE_FP64 = 0
LSG_SOLVE_FP32_IR = 1
LSG_SOLVE_FP32_FGMRES = 2
LSG_ARITH_FP32 = 0
LSG_ARITH_SPLIT_BF16 = 1


class lsg_solve_opts(C.Structure):
    _fields_ = [("mode", C.c_int), ("restart_len", C.c_int), ("max_restarts", C.c_int), ("tol", C.c_double),
                ("inner_tol", C.c_double), ("max_outer", C.c_int), ("device_ptrs", C.c_int)]


# (name, restype, argtypes) of every symbol in include/lsg.h and include/lsg_host.h
_P = C.c_void_p
_D = C.POINTER(C.c_double)
SYMBOLS = [
    ("lsg_bumps_from_seed", C.c_int, [C.c_uint64, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.POINTER(lsg_contrast)]),
    ("lsg_create", C.c_int, [C.POINTER(lsg_inputs), C.POINTER(_P)]),
    ("lsg_destroy", None, [_P]),
    ("lsg_last_error", C.c_char_p, [_P]),
    ("lsg_create_error", C.c_char_p, []),
    ("lsg_get_info", C.c_int, [_P, C.POINTER(lsg_info)]),
    ("lsg_grid_points", C.c_int, [_P, _D, _D]),
    ("lsg_leaf_boxes", C.c_int, [_P, _D, _D, _D]),
    ("lsg_contrast_at_nodes", C.c_int, [_P, _D]),
    ("lsg_incident", C.c_int, [_P, _D]),
    ("lsg_rhs", C.c_int, [_P, _D]),
    ("lsg_interpolate", C.c_int, [_P, _D, _D, _D, C.c_int64, _D]),
    ("lsg_write_precomputations", C.c_int, [_P, C.c_char_p]),
    ("lsg_matvec", C.c_int, [_P, _D, _D, C.c_int, C.c_int64]),
    ("lsg_matvec_device", C.c_int, [_P, _P, _P]),
    ("lsg_sync", C.c_int, [_P]),
    ("lsg_profile_passes", C.c_int, [_P, C.c_int, _D, C.POINTER(C.c_int64), C.c_int]),
    ("lsg_time_matvec", C.c_int, [_P, C.c_int, _D]),
    ("lsg_gmres", C.c_int, [_P, _D, _D, C.c_int, C.c_int, C.c_double, C.POINTER(lsg_stats)]),
    ("lsg_matvec_fp32", C.c_int, [_P, _D, _D, C.c_int, C.c_int64]),
    ("lsg_time_matvec_fp32", C.c_int, [_P, C.c_int, _D]),
    ("lsg_time_matvec_block", C.c_int, [_P, C.c_int, _D]),
    ("lsg_time_matvec_block_fp32", C.c_int, [_P, C.c_int, _D]),
    ("lsg_set_contrast", C.c_int, [_P, _D, C.c_int]),
    ("lsg_autotune", C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    ("lsg_tuning", C.c_char_p, [_P]),
    ("lsg_set_pass_kernel", C.c_int, [_P, C.c_int]),
    ("lsg_set_block_tiles", C.c_int, [_P, C.c_int]),
    ("lsg_set_inner_arith", C.c_int, [_P, C.c_int]),
    ("lsg_use_graph", C.c_int, [_P, C.c_int]),
    ("lsg_gmres_ragged", C.c_int, [C.POINTER(_P), C.c_int, C.POINTER(_P), C.POINTER(_P), C.POINTER(lsg_solve_opts),
                                   C.POINTER(lsg_stats)]),
    # include/lsg_learn.h
    ("lsgn_create", C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_uint64, C.POINTER(_P)]),
    ("lsgn_destroy", None, [_P]),
    ("lsgn_last_error", C.c_char_p, [_P]),
    ("lsgn_num_params", C.c_int64, [_P]),
    ("lsgn_get_params", C.c_int, [_P, C.POINTER(C.c_float)]),
    ("lsgn_set_params", C.c_int, [_P, C.POINTER(C.c_float)]),
    ("lsgn_forward", C.c_int, [_P, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)]),
    ("lsgn_loss_grad", C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float),
                                 C.POINTER(C.c_float)]),
    ("lsgn_set_training_data", C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int]),
    ("lsgn_train_step", C.c_int, [_P, C.POINTER(C.c_int), C.c_int, C.c_float, C.POINTER(C.c_float)]),
    ("lsg_attach_preconditioner", C.c_int, [_P, _P]),
    ("lsg_apply_preconditioner", C.c_int, [_P, _D, _D]),
    ("lsg_grid_resample", C.c_int, [_P, _D, _D, C.c_int]),
    ("lsg_gmres_learned", C.c_int, [_P, _P, _P, C.POINTER(lsg_solve_opts), C.c_int, C.POINTER(lsg_stats)]),
    ("lsg_gmres_ex", C.c_int, [_P, _P, _P, C.POINTER(lsg_solve_opts), C.POINTER(lsg_stats)]),
    ("lsg_gmres_noalloc", C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_double, C.POINTER(lsg_stats)]),
    ("lsg_gmres_batch", C.c_int, [_P, _P, _P, C.c_int, C.c_int64, C.POINTER(lsg_solve_opts), _D,
                                  C.POINTER(lsg_stats)]),
    ("lsg_host_build", C.c_int, [C.POINTER(lsg_inputs), C.POINTER(_P)]),
    ("lsg_create_from_host", C.c_int, [_P, C.POINTER(_P)]),
    ("lsg_host_build_tree", C.c_int, [C.POINTER(lsg_inputs), C.POINTER(_P)]),
    ("lsg_host_destroy", None, [_P]),
    ("lsg_host_error", C.c_char_p, []),
    ("lsg_host_info", C.c_int, [_P, C.POINTER(lsg_info)]),
    ("lsg_host_far_table", C.c_int64, [_P, _D]),
    ("lsg_host_near_tables", C.c_int64, [_P, _D]),
    ("lsg_host_level_boxes", C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]),
    ("lsg_host_leaf_boxes", C.c_int64, [_P, C.POINTER(C.c_int)]),
    ("lsg_host_aca_shapes", C.c_int64, [_P, C.POINTER(C.c_int)]),
    ("lsg_host_skeleton", C.c_int64, [_P, C.POINTER(C.c_int)]),
    ("lsg_host_level_ranks", C.c_int, [_P, C.POINTER(C.c_int), C.c_int]),
    ("lsg_host_entries", C.c_int, [_P, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int, _P]),
    ("lsg_host_aca", C.c_int, [_P, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int,
                               C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("lsg_host_il_pairs", C.c_int, [_P, C.POINTER(C.c_int64), C.c_int]),
    ("lsg_host_qpinv", C.c_int64, [_P, _D]),
    ("lsg_host_plan_ops", C.c_int64, [_P, _D]),
    ("lsg_host_plan_src", C.c_int64, [_P, C.POINTER(C.c_int)]),
    ("lsg_host_plan_dests", C.c_int64, [_P, C.POINTER(C.c_int64)]),
    ("lsg_host_plan_nbr", C.c_int64, [_P, C.POINTER(C.c_int)]),
    ("lsg_host_plan_qk2", C.c_int64, [_P, _D]),
    ("lsg_ragged_create", C.c_int, [C.POINTER(_P), C.c_int, C.POINTER(_P)]),
    ("lsg_ragged_destroy", None, [_P]),
    ("lsg_ragged_error", C.c_char_p, [_P]),
    ("lsg_ragged_size", C.c_int64, [_P, C.c_int]),
    ("lsg_ragged_matvec", C.c_int, [_P, C.POINTER(_P), C.POINTER(_P)]),
    ("lsg_ragged_gmres", C.c_int, [_P, C.POINTER(_P), C.POINTER(_P), C.c_int, C.c_int, C.c_double,
                                   C.POINTER(lsg_stats)]),
    ("lsg_ragged_autotune", C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    ("lsg_ragged_use_graph", C.c_int, [_P, C.c_int]),
    ("lsg_ragged_time_matvec", C.c_int, [_P, C.c_int, _D]),
]

_lib = None


def load_library(path=LIB_PATH):
    """Load liblsg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing; run __graft_entry__.build() or "
                      "python lippmann_schwinger_solver_amd/build.py")
    lib = C.CDLL(path)
    for name, res, args in SYMBOLS:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(_D)


def bumps_from_seed(seed, n, cmin=-0.35, cmax=0.35, wmin=0.03, wmax=0.10, amin=0.5, amax=1.5):
    """Seeded Gaussian-bump contrast (SplitMix64 stream)."""
    lib = load_library()
    c = lsg_contrast()
    rc = lib.lsg_bumps_from_seed(C.c_uint64(seed), n, cmin, cmax, wmin, wmax, amin, amax, C.byref(c))
    if rc != 0:
        raise LsgError(rc, "bad bump arguments")
    return c


class InputsToDFMM:
    """The reference's run-time inputs (GMRES/GMRES.cpp:129-138, 322-346)."""

    def __init__(self, nCones_LFR=16, nChebNodes=6, treeAdaptivity=6, L=0.5, kappa=20.0, yes2DFMM=1,
                 degreeOfBases=6, TOL_POW=10, m=50, restart=1, nLevelsUniform=4, Qchoice=0,
                 tree_mode=0, find_phi=1, inc_angle=0.0, contrast=None, setup_device=0,
                 precomputations_dir=None, aca_exit_pow=0):
        self.nCones_LFR = nCones_LFR
        self.nChebNodes = nChebNodes
        self.treeAdaptivity = treeAdaptivity
        self.L = L
        self.kappa = kappa
        self.yes2DFMM = yes2DFMM
        self.degreeOfBases = degreeOfBases
        self.TOL_POW = TOL_POW
        self.m = m
        self.restart = restart
        self.nLevelsUniform = nLevelsUniform
        self.Qchoice = Qchoice
        self.tree_mode = tree_mode
        self.find_phi = find_phi
        self.inc_angle = inc_angle
        self.contrast = contrast
        self.setup_device = setup_device   # 1: NCA skeletons and operator entries built on the GPU
        # read the tables from <dir>/precomputations (the reference's cache) instead of computing them
        self.precomputations_dir = precomputations_dir
        # 0: the reference's absolute ACA exit 1e-8; p: 10^-p (see lsg_inputs.aca_exit_pow)
        self.aca_exit_pow = aca_exit_pow

    def to_c(self):
        s = lsg_inputs()
        for k in ("nCones_LFR", "nChebNodes", "treeAdaptivity", "L", "kappa", "yes2DFMM", "degreeOfBases",
                  "TOL_POW", "m", "restart", "nLevelsUniform", "Qchoice", "tree_mode", "find_phi", "inc_angle",
                  "setup_device", "aca_exit_pow"):
            setattr(s, k, getattr(self, k))
        if self.precomputations_dir:
            s.precomputations_dir = str(self.precomputations_dir).encode()
        if self.contrast is not None:
            s.contrast = self.contrast
            s.Qchoice = self.contrast.kind
        else:
            s.contrast.kind = self.Qchoice
        return s


def write_precomputations(inputs, directory):
    """Write the table cache of `inputs` under <directory>/precomputations in the
    reference's layout (M tables of levels 0..nLevels of the tree the inputs
    build, neighbour tables of levels 2..10)."""
    lib = load_library()
    c = inputs.to_c()
    rc = lib.lsg_write_precomputations(C.byref(c), str(directory).encode())
    if rc != LSG_OK:
        raise LsgError(rc, lib.lsg_create_error().decode())


class HostBuild:
    """Host-only operator build (tree, tables, NCA) for inspection without a GPU."""

    def __init__(self, inputs, tree_only=False):
        self._lib = load_library()
        self._c = inputs.to_c()
        self._h = C.c_void_p()
        fn = self._lib.lsg_host_build_tree if tree_only else self._lib.lsg_host_build
        rc = fn(C.byref(self._c), C.byref(self._h))
        if rc != 0:
            raise LsgError(rc, self._lib.lsg_host_error().decode())

    def info(self):
        inf = lsg_info()
        self._lib.lsg_host_info(self._h, C.byref(inf))
        return _info_dict(inf)

    def far_table(self):
        n = self._lib.lsg_host_far_table(self._h, None)
        out = np.zeros(2 * n)
        self._lib.lsg_host_far_table(self._h, _dptr(out))
        return out.view(np.complex128)

    def near_tables(self):
        n = self._lib.lsg_host_near_tables(self._h, None)
        out = np.zeros(2 * n)
        self._lib.lsg_host_near_tables(self._h, _dptr(out))
        return out.view(np.complex128)

    def skeleton(self):
        n = self._lib.lsg_host_skeleton(self._h, None)
        out = np.zeros(n, dtype=np.int32)
        self._lib.lsg_host_skeleton(self._h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def entries(self, rows, cols, device=False):
        """K(rows, cols) from the host entry kernel or the GPU one (complex nr x nc)."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        out = np.zeros(2 * r.size * c.size)
        rc = self._lib.lsg_host_entries(self._h, r.ctypes.data_as(C.POINTER(C.c_int)), r.size,
                                        c.ctypes.data_as(C.POINTER(C.c_int)), c.size, 1 if device else 0, _dptr(out))
        if rc != 0:
            raise LsgError(rc, "lsg_host_entries failed")
        return out.view(np.complex128).reshape(c.size, r.size).T

    def aca(self, rows, cols, device=False):
        """One ACA of K(rows, cols): (rank, row pivots, column pivots)."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        k = C.c_int()
        rp = np.zeros(min(r.size, c.size), dtype=np.int32)
        cp = np.zeros(min(r.size, c.size), dtype=np.int32)
        ip = C.POINTER(C.c_int)
        rc = self._lib.lsg_host_aca(self._h, r.ctypes.data_as(ip), r.size, c.ctypes.data_as(ip), c.size,
                                    1 if device else 0, C.byref(k), rp.ctypes.data_as(ip), cp.ctypes.data_as(ip))
        if rc != 0:
            raise LsgError(rc, "lsg_host_aca failed")
        return k.value, rp[:k.value], cp[:k.value]

    def level_ranks(self):
        out = np.zeros(16, dtype=np.int32)
        n = self._lib.lsg_host_level_ranks(self._h, out.ctypes.data_as(C.POINTER(C.c_int)), 16)
        return out[:n]

    def level_boxes(self):
        """(boxes per level, leaf boxes per level), levels 0..nLevels."""
        b = np.zeros(16, dtype=np.int32)
        lf = np.zeros(16, dtype=np.int32)
        ip = C.POINTER(C.c_int)
        n = self._lib.lsg_host_level_boxes(self._h, b.ctypes.data_as(ip), lf.ctypes.data_as(ip), 16)
        return b[:n], lf[:n]

    def leaf_boxes(self):
        """Leaf order: (n_leaves, 2) array of (level, box index in the level)."""
        n = self._lib.lsg_host_leaf_boxes(self._h, None)
        out = np.zeros(2 * n, dtype=np.int32)
        self._lib.lsg_host_leaf_boxes(self._h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out.reshape(-1, 2)

    def aca_shapes(self):
        """Per skeleton owner: (level, box, cone, out rows, out cols, out rank, in rows, in cols, in rank)."""
        n = self._lib.lsg_host_aca_shapes(self._h, None)
        out = np.zeros(n, dtype=np.int32)
        self._lib.lsg_host_aca_shapes(self._h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out.reshape(-1, 9)

    def il_pairs(self):
        out = np.zeros(16, dtype=np.int64)
        n = self._lib.lsg_host_il_pairs(self._h, out.ctypes.data_as(C.POINTER(C.c_int64)), 16)
        return out[:n]

    def plan(self):
        """Flattened device plan: dict(ops, src, dests (n x 7), nbr, qk2, T)."""
        lib, h = self._lib, self._h
        n = lib.lsg_host_plan_ops(h, None)
        ops = np.zeros(2 * n)
        lib.lsg_host_plan_ops(h, _dptr(ops))
        n = lib.lsg_host_plan_src(h, None)
        src = np.zeros(n, dtype=np.int32)
        lib.lsg_host_plan_src(h, src.ctypes.data_as(C.POINTER(C.c_int)))
        n = lib.lsg_host_plan_dests(h, None)
        dests = np.zeros((n, 7), dtype=np.int64)
        lib.lsg_host_plan_dests(h, dests.ctypes.data_as(C.POINTER(C.c_int64)))
        n = lib.lsg_host_plan_nbr(h, None)
        nbr = np.zeros(n, dtype=np.int32)
        lib.lsg_host_plan_nbr(h, nbr.ctypes.data_as(C.POINTER(C.c_int)))
        n = lib.lsg_host_plan_qk2(h, None)
        qk2 = np.zeros(n)
        lib.lsg_host_plan_qk2(h, _dptr(qk2))
        return dict(ops=ops.view(np.complex128), src=src, dests=dests, nbr=nbr.reshape(-1, 9), qk2=qk2,
                    T=self.near_tables())

    def qpinv(self):
        n = self._lib.lsg_host_qpinv(self._h, None)
        out = np.zeros(n)
        self._lib.lsg_host_qpinv(self._h, _dptr(out))
        return out

    def close(self):
        if self._h:
            self._lib.lsg_host_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DFMM:
    """Drop-in for the reference's DFMM facade (GMRES/GMRES.cpp:140-255)."""

    def __init__(self, inputs):
        self.inputs = inputs
        self._lib = load_library()
        self._ctx = None
        self.N = 0
        self.gridPoints = None

    def FMMTreeInitialisation(self, host=None):
        """Build the operator and upload it; with `host` (a HostBuild of the
        same inputs), upload that build instead of building again."""
        ctx = C.c_void_p()
        if host is None:
            c = self.inputs.to_c()
            rc = self._lib.lsg_create(C.byref(c), C.byref(ctx))
        else:
            rc = self._lib.lsg_create_from_host(host._h, C.byref(ctx))
        if rc != 0:
            raise LsgError(rc, self._lib.lsg_create_error().decode())
        self._ctx = ctx
        inf = self.info()
        self.N = int(inf["N"])
        gx = np.zeros(self.N)
        gy = np.zeros(self.N)
        self._lib.lsg_grid_points(self._ctx, _dptr(gx), _dptr(gy))
        self.gridPoints = np.stack([gx, gy], axis=1)

    def _check(self, rc):
        if rc != 0:
            raise LsgError(rc, self._lib.lsg_last_error(self._ctx).decode())

    def _vec(self, v, name):
        """v as a contiguous complex128 copy-or-view of shape (N,); anything else
        raises ValueError (the C side reads and writes exactly N entries)."""
        a = np.ascontiguousarray(v, dtype=np.complex128)
        if a.shape != (self.N,):
            raise ValueError(f"{name} must have shape ({self.N},), got {a.shape}")
        return a

    def _x0(self, x0):
        return np.zeros(self.N, dtype=np.complex128) if x0 is None else np.array(self._vec(x0, "x0"))

    def info(self):
        inf = lsg_info()
        self._check(self._lib.lsg_get_info(self._ctx, C.byref(inf)))
        return _info_dict(inf)

    def MatVecProduct(self, b, output=None):
        """output = A b (complex128, length N)."""
        b = np.ascontiguousarray(b, dtype=np.complex128)
        if b.shape != (self.N,):
            raise ValueError("b must have shape (N,)")
        out = np.empty(self.N, dtype=np.complex128)
        bd = b.view(np.float64)
        od = out.view(np.float64)
        self._check(self._lib.lsg_matvec(self._ctx, _dptr(bd), _dptr(od), 1, self.N))
        if output is not None:
            output[...] = out
        return out

    def matvec_block(self, X):
        """A X for X of shape (N, nrhs)."""
        Xt = np.ascontiguousarray(np.asarray(X, dtype=np.complex128).T)   # nrhs x N, column j at j*N
        Yt = np.empty_like(Xt)
        self._check(self._lib.lsg_matvec(self._ctx, _dptr(Xt.view(np.float64)), _dptr(Yt.view(np.float64)),
                                         Xt.shape[0], self.N))
        return Yt.T

    def matvec_block_fp32(self, X):
        """A32 X for X of shape (N, nrhs), nrhs >= 2: the FP32 block operator."""
        Xt = np.ascontiguousarray(np.asarray(X, dtype=np.complex128).T)
        if Xt.shape[0] < 2:
            raise ValueError("the FP32 block operator needs at least two right-hand sides")
        Yt = np.empty_like(Xt)
        self._check(self._lib.lsg_matvec_fp32(self._ctx, _dptr(Xt.view(np.float64)), _dptr(Yt.view(np.float64)),
                                              Xt.shape[0], self.N))
        return Yt.T

    def rhs(self):
        out = np.empty(self.N, dtype=np.complex128)
        self._check(self._lib.lsg_rhs(self._ctx, _dptr(out.view(np.float64))))
        return out

    def incident(self):
        out = np.empty(self.N, dtype=np.complex128)
        self._check(self._lib.lsg_incident(self._ctx, _dptr(out.view(np.float64))))
        return out

    def contrast_k2(self):
        out = np.empty(self.N)
        self._check(self._lib.lsg_contrast_at_nodes(self._ctx, _dptr(out)))
        return out

    def gmres(self, b, x0=None, restart_len=50, max_restarts=20, tol=1e-10):
        """Device GMRES; returns (x, stats dict). tol <= 0 runs the reference's fixed cycles."""
        b = self._vec(b, "b")
        x = self._x0(x0)
        st = lsg_stats()
        rc = self._lib.lsg_gmres(self._ctx, _dptr(b.view(np.float64)), _dptr(x.view(np.float64)), restart_len,
                                 max_restarts, tol, C.byref(st))
        stats = {k: getattr(st, k) for k, _ in lsg_stats._fields_}
        if rc != 0 and rc != LSG_ERR_NOCONV:
            self._check(rc)
        stats["status"] = rc
        return x, stats

    def gmres_ex(self, b, x0=None, mode="fp64", restart_len=50, max_restarts=20, tol=1e-8, inner_tol=1e-6,
                 max_outer=10, device_ptrs=None):
        """GMRES with solver options. mode "fp64" (= gmres) or "fp32ir" (FP64
        iterative refinement around FP32-operator GMRES). With device_ptrs =
        (b_dev, x_dev) (integer device addresses) nothing is copied and b, x0
        are ignored; returns (x or None, stats)."""
        o = lsg_solve_opts(mode={"fp64": LSG_SOLVE_FP64, "fp32ir": LSG_SOLVE_FP32_IR}[mode], restart_len=restart_len,
                           max_restarts=max_restarts, tol=tol, inner_tol=inner_tol, max_outer=max_outer,
                           device_ptrs=1 if device_ptrs else 0)
        st = lsg_stats()
        if device_ptrs:
            x = None
            rc = self._lib.lsg_gmres_ex(self._ctx, C.c_void_p(device_ptrs[0]), C.c_void_p(device_ptrs[1]), C.byref(o),
                                        C.byref(st))
        else:
            b = self._vec(b, "b")
            x = self._x0(x0)
            rc = self._lib.lsg_gmres_ex(self._ctx, b.ctypes.data_as(_P), x.ctypes.data_as(_P), C.byref(o),
                                        C.byref(st))
        stats = {k: getattr(st, k) for k, _ in lsg_stats._fields_}
        if rc != 0 and rc != LSG_ERR_NOCONV:
            self._check(rc)
        stats["status"] = rc
        return x, stats

    def gmres_noalloc(self, b_dev, x_dev, restart_len=50, max_restarts=20, tol=1e-8):
        """lsg_gmres_noalloc: FP64 GMRES on device vectors (integer addresses of
        N complex128; x holds the initial guess) with the library's scratch array
        as workspace, allocating nothing; returns the stats dict."""
        st = lsg_stats()
        rc = self._lib.lsg_gmres_noalloc(self._ctx, C.c_void_p(b_dev), C.c_void_p(x_dev), restart_len, max_restarts,
                                         tol, C.byref(st))
        stats = {k: getattr(st, k) for k, _ in lsg_stats._fields_}
        if rc != 0 and rc != LSG_ERR_NOCONV:
            self._check(rc)
        stats["status"] = rc
        return stats

    def gmres_batch(self, B, X0=None, restart_len=50, max_restarts=20, tol=1e-8, device_ptrs=None, ld=None,
                    nrhs=None, mode="fp64", inner_tol=1e-5, max_outer=20):
        """Batched GMRES on the columns of B (N x nrhs, nrhs <= 64); returns
        (X, stats, rel_residuals). With device_ptrs = (b_dev, x_dev) the
        column-major blocks (leading dimension ld) stay on the device and X is None.
        mode "fp64", "fp32ir" (FP64 refinement around FP32 GMRES) or
        "fp32fgmres" (FP64 flexible GMRES(max_outer), each outer step
        preconditioned by one FP32 GMRES(restart_len) solve to inner_tol)."""
        o = lsg_solve_opts(mode={"fp64": LSG_SOLVE_FP64, "fp32ir": LSG_SOLVE_FP32_IR,
                                 "fp32fgmres": LSG_SOLVE_FP32_FGMRES}[mode], restart_len=restart_len,
                           max_restarts=max_restarts, tol=tol, inner_tol=inner_tol, max_outer=max_outer,
                           device_ptrs=1 if device_ptrs else 0)
        st = lsg_stats()
        if device_ptrs:
            rel = np.zeros(nrhs)
            rc = self._lib.lsg_gmres_batch(self._ctx, C.c_void_p(device_ptrs[0]), C.c_void_p(device_ptrs[1]), nrhs,
                                           ld or self.N, C.byref(o), _dptr(rel), C.byref(st))
            X = None
        else:
            Ba = np.asarray(B, dtype=np.complex128)
            if Ba.ndim == 1:
                Ba = Ba[:, None]
            if Ba.ndim != 2 or Ba.shape[0] != self.N:
                raise ValueError(f"B must have shape ({self.N}, nrhs), got {Ba.shape}")
            Bt = np.ascontiguousarray(Ba.T)     # nrhs x N
            nrhs = Bt.shape[0]
            if X0 is not None and np.shape(X0) != Ba.shape:
                raise ValueError(f"X0 must have the shape of B {Ba.shape}, got {np.shape(X0)}")
            Xt = np.zeros_like(Bt) if X0 is None else np.ascontiguousarray(np.asarray(X0, dtype=np.complex128).reshape(Ba.shape).T)
            rel = np.zeros(nrhs)
            rc = self._lib.lsg_gmres_batch(self._ctx, Bt.ctypes.data_as(_P), Xt.ctypes.data_as(_P), nrhs, self.N,
                                           C.byref(o), _dptr(rel), C.byref(st))
            X = Xt.T
        stats = {k: getattr(st, k) for k, _ in lsg_stats._fields_}
        if rc != 0 and rc != LSG_ERR_NOCONV:
            self._check(rc)
        stats["status"] = rc
        return X, stats, rel

    def interpolate(self, f, x, y):
        """Field f at the nodes (length N) re-interpolated at the points (x[i], y[i])
        through each leaf's total-degree Chebyshev expansion (hybridTrueValidation_gmres.cpp:186-209)."""
        f = np.ascontiguousarray(f, dtype=np.complex128)
        if f.shape != (self.N,):
            raise ValueError("f must have shape (N,)")
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        if x.shape != y.shape or x.ndim != 1:
            raise ValueError("x and y must be 1-D of equal length")
        out = np.empty(x.size, dtype=np.complex128)
        self._check(self._lib.lsg_interpolate(self._ctx, _dptr(f.view(np.float64)), _dptr(x), _dptr(y), x.size,
                                              _dptr(out.view(np.float64))))
        return out

    def rhs_angles(self, angles):
        """Right-hand sides -k^2 q u_inc for plane waves at the given angles (N x len(angles))."""
        x, y = self.gridPoints[:, 0], self.gridPoints[:, 1]
        qk2 = self.contrast_k2()
        k = self.inputs.kappa
        a = np.asarray(angles, dtype=float)
        return -qk2[:, None] * np.exp(1j * k * (np.outer(x, np.cos(a)) + np.outer(y, np.sin(a))))

    def MatVecProduct_fp32(self, x):
        """y = A32 x: FP32 operators and FP32 accumulation inside each pass, FP64 x and y."""
        x = self._vec(x, "x")
        out = np.empty_like(x)
        self._check(self._lib.lsg_matvec_fp32(self._ctx, _dptr(x.view(np.float64)), _dptr(out.view(np.float64)),
                                              1, self.N))
        return out

    def set_contrast(self, qk2):
        """Volume-potential context only: use A = I + diag(qk2[:, q]) V for right-hand
        side q (qk2 of shape (N,) or (N, ncol), ncol <= 64); None removes it."""
        if qk2 is None:
            self._check(self._lib.lsg_set_contrast(self._ctx, None, 0))
            return
        q = np.asarray(qk2, dtype=np.float64)
        q = q.reshape(self.N, -1)
        qc = np.ascontiguousarray(q.T)        # ncol x N: column j at j*N
        self._check(self._lib.lsg_set_contrast(self._ctx, _dptr(qc), qc.shape[0]))

    def attach_preconditioner(self, net):
        """Attach a learned.LearnedNet (or None) as the learned preconditioner."""
        self._check(self._lib.lsg_attach_preconditioner(self._ctx, net._h if net is not None else None))
        self._net = net

    def apply_preconditioner(self, v):
        v = self._vec(v, "v")
        z = np.empty_like(v)
        self._check(self._lib.lsg_apply_preconditioner(self._ctx, _dptr(v.view(np.float64)), _dptr(z.view(np.float64))))
        return z

    def to_grid(self, v):
        """Leaf-node vector (N complex) -> G x G grid (complex), Chebyshev interpolation per leaf."""
        v = self._vec(v, "v")
        G = self._net.grid
        out = np.empty(G * G, dtype=np.complex128)
        self._check(self._lib.lsg_grid_resample(self._ctx, _dptr(v.view(np.float64)), _dptr(out.view(np.float64)), 1))
        return out.reshape(G, G)

    def to_nodes(self, g):
        """G x G grid (complex) -> leaf nodes, bilinear interpolation."""
        g = np.ascontiguousarray(np.asarray(g, dtype=np.complex128).reshape(-1))
        out = np.empty(self.N, dtype=np.complex128)
        self._check(self._lib.lsg_grid_resample(self._ctx, _dptr(g.view(np.float64)), _dptr(out.view(np.float64)), 0))
        return out

    def gmres_learned(self, b, x0=None, guess=True, right=True, restart_len=50, max_restarts=20, tol=1e-8):
        """Flexible GMRES with the attached network as initial guess and/or right preconditioner."""
        o = lsg_solve_opts(mode=LSG_SOLVE_FP64, restart_len=restart_len, max_restarts=max_restarts, tol=tol)
        b = self._vec(b, "b")
        x = self._x0(x0)
        st = lsg_stats()
        flags = (1 if guess else 0) | (2 if right else 0)
        rc = self._lib.lsg_gmres_learned(self._ctx, b.ctypes.data_as(_P), x.ctypes.data_as(_P), C.byref(o), flags,
                                         C.byref(st))
        stats = {k: getattr(st, k) for k, _ in lsg_stats._fields_}
        if rc != 0 and rc != LSG_ERR_NOCONV:
            self._check(rc)
        stats["status"] = rc
        return x, stats

    def autotune(self, reps=3):
        """Pick the faster kernel variant per pass and for the near field; returns (n_changed, report)."""
        n = C.c_int()
        self._check(self._lib.lsg_autotune(self._ctx, reps, C.byref(n)))
        return n.value, self._lib.lsg_tuning(self._ctx).decode()

    def tuning(self):
        """The autotuner's last report and the block operator's timed choices."""
        return self._lib.lsg_tuning(self._ctx).decode()

    def set_pass_kernel(self, kernel, fused=False):
        """Run every pass with one kernel: "block" (one workgroup per
        destination), "wave" (one wavefront per destination) or "lane" (one
        lane per output row); with fused=True the first three upward passes,
        and the last two small downward passes, each run as one launch of one
        workgroup per group of subtrees. autotune() replaces the choice."""
        k = {"block": 0, "wave": 1, "lane": 2}[kernel] + (3 if fused else 0)
        self._check(self._lib.lsg_set_pass_kernel(self._ctx, k))

    def set_block_tiles(self, tiles):
        """Right-hand-side tiles of 16 per wavefront in the block passes of the
        inner solves' arithmetic (1 or 2; in split BF16 also 3 or 4: 1 or 2
        tiles with the near field beside the upward and downward passes),
        replacing the choice made by timing."""
        self._check(self._lib.lsg_set_block_tiles(self._ctx, tiles))

    def set_inner_arith(self, arith):
        """Arithmetic of the FP32-inner batched solves and of matvec_block_fp32:
        "split_bf16" (default; operators and sources split into bf16 pairs, 16
        significant bits, their products as four BF16 MFMA terms with FP32
        sums) or "fp32" (FP32 MFMA)."""
        modes = {"fp32": LSG_ARITH_FP32, "split_bf16": LSG_ARITH_SPLIT_BF16}
        if arith not in modes:
            raise ValueError(f"inner arithmetic must be one of {sorted(modes)}")
        self._check(self._lib.lsg_set_inner_arith(self._ctx, modes[arith]))

    def use_graph(self, on=True):
        """Replay the FP64 matvec from a captured hipGraph (on) or launch kernel by kernel (off)."""
        self._check(self._lib.lsg_use_graph(self._ctx, 1 if on else 0))

    def time_matvec_block(self, reps=10, fp32=False):
        """Average device time (ms) of one 64-right-hand-side block matvec."""
        ms = C.c_double()
        f = self._lib.lsg_time_matvec_block_fp32 if fp32 else self._lib.lsg_time_matvec_block
        self._check(f(self._ctx, reps, C.byref(ms)))
        return ms.value

    def time_matvec_fp32(self, reps=20):
        ms = C.c_double()
        self._check(self._lib.lsg_time_matvec_fp32(self._ctx, reps, C.byref(ms)))
        return ms.value

    def profile_passes(self, reps=5):
        ms = np.zeros(64)
        by = np.zeros(64, dtype=np.int64)
        n = self._lib.lsg_profile_passes(self._ctx, reps, _dptr(ms), by.ctypes.data_as(C.POINTER(C.c_int64)), 64)
        if n < 0:
            self._check(n)
        names = self._lib.lsg_last_error(self._ctx).decode().strip(",").split(",")
        return [(names[i], float(ms[i]), int(by[i])) for i in range(n)]

    def time_matvec(self, reps=20):
        """Average device time (ms) of one full matvec."""
        ms = C.c_double()
        self._check(self._lib.lsg_time_matvec(self._ctx, reps, C.byref(ms)))
        return ms.value

    def close(self):
        if self._ctx:
            self._lib.lsg_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def GMRES(x0, S, m, inputs, b):
    """One GMRES(m) cycle from x0 (GMRES/GMRES.cpp:257-305); returns the solution."""
    x, _ = S.gmres(b, x0=x0, restart_len=m, max_restarts=1, tol=0.0)
    return x


def RestartedGMRES(restart, S, m, inputs, b):
    """`restart` cycles of GMRES(m) from x0 = 0 (GMRES/GMRES.cpp:307-319)."""
    x, _ = S.gmres(b, x0=None, restart_len=m, max_restarts=restart, tol=0.0)
    return x


class RaggedBatch:
    """Independent scatterers fused into one operator on the device
    (lsg_ragged_*): `hosts` are HostBuild objects of uniform trees (any depth,
    wavenumber and contrast). matvec and gmres take and return one vector per
    system; gmres solves all systems at once, each with its own Krylov basis
    and stopping point."""

    def __init__(self, hosts):
        self._lib = load_library()
        self._hosts = list(hosts)
        arr = (_P * len(self._hosts))(*[h._h for h in self._hosts])
        self._r = C.c_void_p()
        rc = self._lib.lsg_ragged_create(arr, len(self._hosts), C.byref(self._r))
        if rc != 0:
            raise LsgError(rc, self._lib.lsg_create_error().decode())
        self.n = len(self._hosts)
        self.sizes = [int(self._lib.lsg_ragged_size(self._r, i)) for i in range(self.n)]

    def _check(self, rc):
        if rc != 0:
            raise LsgError(rc, self._lib.lsg_ragged_error(self._r).decode())

    def _vecs(self, vs, name):
        if len(vs) != self.n:
            raise ValueError(f"{name}: one vector per system ({self.n}), got {len(vs)}")
        out = []
        for i, (v, N) in enumerate(zip(vs, self.sizes)):
            a = np.ascontiguousarray(v, dtype=np.complex128)
            if a.shape != (N,):
                raise ValueError(f"{name}[{i}] must have shape ({N},), got {a.shape}")
            out.append(a)
        return out

    def matvec(self, xs):
        xs = self._vecs(xs, "x")
        ys = [np.empty(N, dtype=np.complex128) for N in self.sizes]
        self._check(self._lib.lsg_ragged_matvec(self._r, (_P * self.n)(*[x.ctypes.data for x in xs]),
                                                (_P * self.n)(*[y.ctypes.data for y in ys])))
        return ys

    def gmres(self, rhs, x0=None, restart_len=50, max_restarts=20, tol=1e-8):
        """Returns (xs, per-system stats dicts, rc); rc LSG_ERR_NOCONV if a system missed tol."""
        bs = self._vecs(rhs, "rhs")
        xs = [np.zeros(N, dtype=np.complex128) for N in self.sizes] if x0 is None else \
            [np.array(a) for a in self._vecs(x0, "x0")]
        st = (lsg_stats * self.n)()
        rc = self._lib.lsg_ragged_gmres(self._r, (_P * self.n)(*[b.ctypes.data for b in bs]),
                                        (_P * self.n)(*[x.ctypes.data for x in xs]), restart_len, max_restarts, tol, st)
        if rc != 0 and rc != LSG_ERR_NOCONV:
            self._check(rc)
        stats = [{k: getattr(st[i], k) for k, _ in lsg_stats._fields_} for i in range(self.n)]
        return xs, stats, rc

    def autotune(self, reps=3):
        n = C.c_int()
        self._check(self._lib.lsg_ragged_autotune(self._r, reps, C.byref(n)))
        return n.value

    def use_graph(self, on=True):
        self._check(self._lib.lsg_ragged_use_graph(self._r, 1 if on else 0))

    def time_matvec(self, reps=10):
        ms = C.c_double()
        self._check(self._lib.lsg_ragged_time_matvec(self._r, reps, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "_r", None):
            self._lib.lsg_ragged_destroy(self._r)
            self._r = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gmres_ragged(solvers, rhs, restart_len=50, max_restarts=20, tol=1e-8, mode="fp64", inner_tol=1e-5, max_outer=20):
    """Solve one system per DFMM context concurrently (lsg_gmres_ragged): the
    contexts may differ in depth, wavenumber and contrast. Returns (xs, stats list)."""
    n = len(solvers)
    if n < 1 or len(rhs) != n:
        raise ValueError(f"one right-hand side per solver: {len(rhs)} for {n} solvers")
    lib = solvers[0]._lib
    bs = [S._vec(b, f"rhs[{i}]") for i, (S, b) in enumerate(zip(solvers, rhs))]
    xs = [np.zeros(S.N, dtype=np.complex128) for S in solvers]
    ctxs = (C.c_void_p * n)(*[S._ctx for S in solvers])
    bp = (C.c_void_p * n)(*[b.ctypes.data for b in bs])
    xp = (C.c_void_p * n)(*[x.ctypes.data for x in xs])
    o = lsg_solve_opts(mode={"fp64": LSG_SOLVE_FP64, "fp32ir": LSG_SOLVE_FP32_IR}[mode], restart_len=restart_len,
                       max_restarts=max_restarts, tol=tol, inner_tol=inner_tol, max_outer=max_outer, device_ptrs=0)
    st = (lsg_stats * n)()
    rc = lib.lsg_gmres_ragged(ctxs, n, bp, xp, C.byref(o), st)
    stats = [{k: getattr(st[i], k) for k, _ in lsg_stats._fields_} for i in range(n)]
    if rc != 0 and rc != LSG_ERR_NOCONV:
        raise LsgError(rc, "ragged batch failed: " + "; ".join(S._lib.lsg_last_error(S._ctx).decode() for S in solvers))
    return xs, stats, rc
