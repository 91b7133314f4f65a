"""potg_phi_at_noalloc, the point-potential query that allocates nothing,
against potg_phi_at on the same context: two calls in a row, the second
with fewer points, both bit-identical to the allocating entry point (the
same walk kernels run).  The points and results live in device memory
obtained from the HIP runtime libpotg.so itself links."""
import ctypes as C

import numpy as np
import pytest

from gpu_common import big_case, gpu_ctx, oracle_sim

from pot_amd import potg as G

pytestmark = pytest.mark.gpu


def _hip():
    H = C.CDLL("libamdhip64.so.7")  # the runtime libpotg.so is linked against
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, H, nbytes):
        self.H = H
        self.p = C.c_void_p()
        assert H.hipMalloc(C.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.H.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape, dtype):
        out = np.zeros(shape, dtype)
        assert out.nbytes <= self.n
        assert self.H.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.H.hipFree(self.p)
            self.p = C.c_void_p()


@pytest.mark.parametrize("geometry,precision", [(0, G.POTG_FP64), (1, G.POTG_MIXED)],
                         ids=["sphere_fp64", "prolate_mixed"])
def test_phi_noalloc_matches_phi_at(geometry, precision):
    case = big_case(6001, geometry=geometry, A=2.0 if geometry else 1.0, seed=33)
    parts = oracle_sim(case).particles()
    ctx = gpu_ctx(case, parts, precision=precision, q=-40)
    ctx.build_tree()
    R = float(case["args"]["-r"]) * float(case["args"]["-m"])
    rng = np.random.default_rng(5)
    H = _hip()
    n0 = 5000
    dx = DevBuf(H, 3 * 8 * n0)
    dphi = DevBuf(H, 8 * n0)
    try:
        for n in (n0, 777):  # the second call smaller: nothing of the first may leak in
            u = rng.normal(size=(n, 3))
            u /= np.linalg.norm(u, axis=1)[:, None]
            r = R * rng.uniform(0.05, 0.999, size=(n, 1)) ** (1.0 / 3.0)
            xyz = np.ascontiguousarray(R + u * r)
            want = ctx.phi_at(0, xyz)
            dx.put(xyz)
            dphi.put(np.full(n0, np.nan))
            rc = ctx.L.potg_phi_at_noalloc(ctx.h, 0, dx.p, n, dphi.p, None)
            assert rc == 0, ctx.L.potg_last_error()
            ctx.sync()
            got = dphi.get(n, np.float64)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), n
            # beyond n the caller's array is untouched
            rest = dphi.get(n0, np.float64)[n:]
            assert np.all(np.isnan(rest))
    finally:
        dx.free()
        dphi.free()


def test_phi_noalloc_rejects_too_many_points():
    case = big_case(2001, seed=34)
    ctx = gpu_ctx(case, oracle_sim(case).particles())
    ctx.build_tree()
    H = _hip()
    n = G.POTG_NOALLOC_MAX_POINTS + 1
    dx, dphi = DevBuf(H, 3 * 8 * n), DevBuf(H, 8 * n)  # real buffers of that size
    try:
        rc = ctx.L.potg_phi_at_noalloc(ctx.h, 0, dx.p, n, dphi.p, None)
        assert rc == G.POTG_ERR_ARG
        assert b"more than" in ctx.L.potg_last_error()
    finally:
        dx.free()
        dphi.free()
