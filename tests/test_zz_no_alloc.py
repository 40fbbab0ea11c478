"""lsg_gmres_noalloc (the allocation-free GMRES entry point) against
lsg_gmres_ex with device pointers, both called through the C ABI: two calls in
a row on the library's scratch array, the second on a smaller system, each
giving the same iterates as the allocating entry point."""
import ctypes as C

import numpy as np
import pytest

import lippmann_schwinger_solver_amd as L
from tests import helpers as H

pytestmark = pytest.mark.gpu


class DevVec:
    """N complex128 in device memory (plain hipMalloc, outside the library)."""

    def __init__(self, host):
        self.hip = C.CDLL("libamdhip64.so")
        self.host = np.ascontiguousarray(host, dtype=np.complex128)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.host.nbytes)) == 0
        self.put(self.host)

    def put(self, v):
        v = np.ascontiguousarray(v, dtype=np.complex128)
        assert self.hip.hipMemcpy(self.ptr, v.ctypes.data_as(C.c_void_p), C.c_size_t(v.nbytes), 1) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def solve_both(S, b, m, tol):
    lib = S._lib
    bd, xa, xb = DevVec(b), DevVec(np.zeros_like(b)), DevVec(np.zeros_like(b))
    o = L.lsg_solve_opts(mode=L.LSG_SOLVE_FP64, restart_len=m, max_restarts=20, tol=tol, device_ptrs=1)
    sa, sb = L.lsg_stats(), L.lsg_stats()
    ra = lib.lsg_gmres_ex(S._ctx, bd.ptr, xa.ptr, C.byref(o), C.byref(sa))
    rb = lib.lsg_gmres_noalloc(S._ctx, bd.ptr, xb.ptr, m, 20, tol, C.byref(sb))
    xa_h, xb_h = xa.get(), xb.get()
    for v in (bd, xa, xb):
        v.free()
    return ra, sa, xa_h, rb, sb, xb_h


@pytest.mark.parametrize("depth,kappa,m", [(5, 40.0, 50), (4, 20.0, 30)])
def test_noalloc_matches_allocating_entry_point(depth, kappa, m):
    # first the larger system, then the smaller one: both calls share the scratch array
    S = L.DFMM(H.inputs(depth=depth, kappa=kappa, q=6, seed=depth))
    S.FMMTreeInitialisation()
    b = S.rhs()
    ra, sa, xa, rb, sb, xb = solve_both(S, b, m, 1e-9)
    print(f"depth {depth}: allocating {sa.iterations} steps {sa.rel_residual:.3e}; "
          f"no-alloc {sb.iterations} steps {sb.rel_residual:.3e}", flush=True)
    assert ra == 0 and rb == 0
    assert sb.converged == 1 and sb.rel_residual <= 1e-9
    assert sb.iterations == sa.iterations and sb.restarts == sa.restarts
    # the same kernels in the same order on other addresses: the same iterate
    assert H.relerr(xb, xa) <= 1e-14
    assert H.relerr(S.MatVecProduct(xb), b) <= 1.05e-9
    S.close()


def test_noalloc_rejects_oversized_workspace():
    S = L.DFMM(H.inputs(depth=6, kappa=40.0, q=6, seed=1))
    S.FMMTreeInitialisation()
    b = DevVec(S.rhs())
    x = DevVec(np.zeros(S.N, dtype=complex))
    st = L.lsg_stats()
    rc = S._lib.lsg_gmres_noalloc(S._ctx, b.ptr, x.ptr, 200, 1, 1e-8, C.byref(st))   # 201 x 147456 x 16 B > 128 MiB
    assert rc == L.LSG_ERR_ARG and b"maximum" in S._lib.lsg_last_error(S._ctx)
    rc = S._lib.lsg_gmres_noalloc(S._ctx, b.ptr, x.ptr, 50, 20, 1e-8, C.byref(st))    # fits
    assert rc == 0 and st.rel_residual <= 1e-8
    b.free()
    x.free()
    S.close()
