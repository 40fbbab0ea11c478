"""ORACLE - test infrastructure only.

ctypes wrapper of oracle/_build/liblsg_oracle.so, the CPU restatement of the
reference's DAFMM matvec and GMRES (see lsg_oracle.cpp for the file:line map).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liblsg_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise OSError(f"{LIB} missing: run make -C oracle")
        lib = C.CDLL(LIB)
        P, D, I64 = C.c_void_p, C.POINTER(C.c_double), C.c_int64
        sig = {
            "or_create": (C.c_int, [P, C.POINTER(P)]),
            "or_error": (C.c_char_p, []),
            "or_destroy": (None, [P]),
            "or_n": (I64, [P]),
            "or_matvec": (C.c_int, [P, D, D]),
            "or_dense_matvec": (C.c_int, [P, D, D]),
            "or_dense_entry": (C.c_int, [P, C.c_int, C.c_int, D]),
            "or_far_entry": (C.c_int, [P, C.c_int, C.c_int, D]),
            "or_rhs": (C.c_int, [P, D]),
            "or_grid": (C.c_int, [P, D, D]),
            "or_gmres": (C.c_int, [P, D, D, C.c_int, C.c_int, C.POINTER(C.c_int)]),
            "or_gmres_cycle": (C.c_int, [P, D, D, C.c_int, C.POINTER(C.c_int)]),
            "or_far_table": (I64, [P, D]),
            "or_near_tables": (I64, [P, D]),
            "or_skeleton": (I64, [P, C.POINTER(C.c_int)]),
            "or_neighbor_table": (I64, [P, C.c_int, C.c_int, C.c_int, D]),
            "or_bessel": (C.c_int, [C.c_int, C.c_double, D, D]),
            "or_qpinv": (I64, [P, D]),
            "or_box_center": (C.c_int, [P, C.c_int, C.c_int, D]),
            "or_levels": (C.c_int, [P, C.POINTER(C.c_int)]),
            "or_contrast_k2": (C.c_int, [P, D]),
            "or_interaction_list": (I64, [P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
            "or_gauss_legendre": (C.c_int, [C.c_int, D, D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def bessel(n, x):
    j, y = C.c_double(), C.c_double()
    _load().or_bessel(n, x, C.byref(j), C.byref(y))
    return j.value, y.value


def gauss_legendre(n):
    x, w = np.zeros(n), np.zeros(n)
    _load().or_gauss_legendre(n, _d(x), _d(w))
    return x, w


class Oracle:
    """Reference-structured DAFMM operator built on the CPU from lsg_inputs."""

    def __init__(self, c_inputs):
        lib = _load()
        self._lib = lib
        self._h = C.c_void_p()
        rc = lib.or_create(C.byref(c_inputs), C.byref(self._h))
        if rc != 0:
            raise RuntimeError("oracle setup failed: " + lib.or_error().decode())
        self.N = int(lib.or_n(self._h))

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty(self.N, dtype=np.complex128)
        self._lib.or_matvec(self._h, _d(x.view(np.float64)), _d(y.view(np.float64)))
        return y

    def dense_matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty(self.N, dtype=np.complex128)
        self._lib.or_dense_matvec(self._h, _d(x.view(np.float64)), _d(y.view(np.float64)))
        return y

    def dense_entry(self, i, j):
        v = np.zeros(2)
        self._lib.or_dense_entry(self._h, i, j, _d(v))
        return complex(v[0], v[1])

    def far_entry(self, i, j):
        v = np.zeros(2)
        self._lib.or_far_entry(self._h, i, j, _d(v))
        return complex(v[0], v[1])

    def rhs(self):
        b = np.empty(self.N, dtype=np.complex128)
        self._lib.or_rhs(self._h, _d(b.view(np.float64)))
        return b

    def grid(self):
        x, y = np.zeros(self.N), np.zeros(self.N)
        self._lib.or_grid(self._h, _d(x), _d(y))
        return x, y

    def gmres(self, b, restart, m):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.empty(self.N, dtype=np.complex128)
        it = C.c_int()
        self._lib.or_gmres(self._h, _d(b.view(np.float64)), _d(x.view(np.float64)), restart, m, C.byref(it))
        return x, it.value

    def gmres_cycle(self, b, x0, m):
        """One GMRES(m) cycle from x0 (GMRES/GMRES.cpp:257-305); returns (x, steps)."""
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.array(x0, dtype=np.complex128)
        it = C.c_int()
        self._lib.or_gmres_cycle(self._h, _d(b.view(np.float64)), _d(x.view(np.float64)), m, C.byref(it))
        return x, it.value

    def far_table(self):
        n = self._lib.or_far_table(self._h, None)
        out = np.zeros(2 * n)
        self._lib.or_far_table(self._h, _d(out))
        return out.view(np.complex128)

    def near_tables(self):
        n = self._lib.or_near_tables(self._h, None)
        out = np.zeros(2 * n)
        self._lib.or_near_tables(self._h, _d(out))
        return out.view(np.complex128)

    def neighbor_table(self, level, kind, slot):
        """Near-field table in use at a leaf level (kind 0 colleague, 1 coarse,
        2 fine, 3 separated fine), or None when no leaf uses the slot."""
        n = self._lib.or_neighbor_table(self._h, level, kind, slot, None)
        if n < 0:
            return None
        out = np.zeros(2 * n)
        self._lib.or_neighbor_table(self._h, level, kind, slot, _d(out))
        r = int(round(np.sqrt(n)))
        return out.view(np.complex128).reshape(r, r)

    def skeleton(self):
        n = self._lib.or_skeleton(self._h, None)
        out = np.zeros(n, dtype=np.int32)
        self._lib.or_skeleton(self._h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def qpinv(self):
        n = self._lib.or_qpinv(self._h, None)
        out = np.zeros(n)
        self._lib.or_qpinv(self._h, _d(out))
        return out

    def box_center(self, level, box):
        xy = np.zeros(2)
        leaf = self._lib.or_box_center(self._h, level, box, _d(xy))
        return xy, bool(leaf)

    def contrast_k2(self):
        out = np.zeros(self.N)
        self._lib.or_contrast_k2(self._h, _d(out))
        return out

    def levels(self):
        out = np.zeros(64, dtype=np.int32)
        n = self._lib.or_levels(self._h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return int(out[0]), int(out[1]), out[2:n].tolist()

    def interaction_list(self, level, box, cone=-1):
        n = self._lib.or_interaction_list(self._h, level, box, cone, None)
        out = np.zeros(2 * n, dtype=np.int32)
        self._lib.or_interaction_list(self._h, level, box, cone, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out.reshape(-1, 2)

    def close(self):
        if self._h:
            self._lib.or_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
