"""femgpu_apply_noalloc (no counterpart in the reference: the product for callers that cannot
allocate) through the C ABI against femgpu_apply, the entry point it mirrors, on device buffers:
bitwise the same values (the same kernels and sums, every temporary in the library's static
scratch array), on two calls in a row whose second problem is smaller than the first (the array
keeps the first call's contents), in both precisions, for the matrix-free product (box-only and
general Hexa, Tetra_10) and the stored SELL product, with the caller's and an internal node order."""
import ctypes as C
import sys

import numpy as np
import pytest

import lmtpp_amd as L
from lmtpp_amd import _capi as A
from problems import build, jitter_interior, lognormal_E, tetra10_jittered

pytestmark = pytest.mark.gpu

MF = (A.SPMV_MATRIX_FREE, 2, A.LOADS_NT)
FULL = (A.SPMV_FULL, 2, A.LOADS_NT)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


class DeviceBuffer:
    """A device array of n doubles through the HIP runtime libfemgpu.so is bound to."""

    def __init__(self, n):
        A.load()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.n = n
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(8 * n)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n), 1) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _problems():
    c1, e1 = L.make_rect(L.HEXA, (25, 25, 25))               # box-only matrix-free product
    c2, e2 = L.make_rect(L.HEXA, (7, 6, 5))
    c2 = jitter_interior(c2, 1.0 / 6, 0.15, 3)               # general Hexa elements
    c3, e3 = tetra10_jittered(3, seed=2)
    return [("hexa lattice 24^3", L.HEXA, c1, e1, MF), ("jittered hexa 6x5x4", L.HEXA, c2, e2, MF),
            ("tetra10 3^3", L.TETRA10, c3, e3, MF), ("jittered hexa, stored SELL", L.HEXA, c2, e2, FULL)]


def _noalloc(f, precision, x):
    lib = A.load()
    dx, dy = DeviceBuffer(x.size), DeviceBuffer(x.size)
    dx.put(x)
    dy.put(np.full(x.size, np.nan))
    rc = lib.femgpu_apply_noalloc(f.handle, precision, dx.ptr, dy.ptr)
    y = dy.get() if rc == A.OK else None
    dx.free()
    dy.free()
    return rc, y


@pytest.mark.parametrize("node_order", [A.NODE_ORDER_NATURAL, A.NODE_ORDER_RCM])
def test_noalloc_product_equals_apply(node_order):
    lib = A.load()
    lib.femgpu_set_node_order(node_order)
    try:
        for name, kind, c, e, cfg in _problems():           # sizes decrease after the first
            say("apply_noalloc", name, "node order", node_order)
            f, *_ = build(kind, c, e, lognormal_E(len(e), 0.4, 5))
            f.assemble()
            f.set_spmv_config(fp64=cfg, fp32=cfg)
            x = np.sin(np.arange(3 * len(c)) * 0.37)
            for precision in (L.FP64, L.MIXED):
                rc, y0 = _noalloc(f, precision, x)
                if precision == L.FP64:                      # first call after the assembly
                    # the stored FP64 matrix needs nothing built; the matrix-free product's records
                    # are built by its first use, which allocates: refused
                    assert rc == (A.OK if cfg == FULL else A.NOT_READY), rc
                f.apply(x, precision)                        # the entry point it mirrors builds the product
                if rc == A.OK:
                    assert np.array_equal(y0, f.apply(x, precision))
                for xk in (x, 0.5 * x + 1.0):                # twice: the scratch keeps the last call's values
                    rc, y = _noalloc(f, precision, xk)
                    assert rc == A.OK, (rc, lib.femgpu_last_error())
                    assert np.array_equal(y, f.apply(xk, precision)), (name, precision)
            f.close()
    finally:
        lib.femgpu_set_node_order(A.NODE_ORDER_NATURAL)


def test_noalloc_rejects_bad_input():
    lib = A.load()
    c, e = L.make_rect(L.HEXA, (4, 4, 4))
    f, *_ = build(L.HEXA, c, e, 210000.0)
    d = DeviceBuffer(3 * len(c))
    assert lib.femgpu_apply_noalloc(f.handle, 0, d.ptr, d.ptr) == A.NOT_READY   # not assembled
    f.assemble()
    assert lib.femgpu_apply_noalloc(f.handle, 2, d.ptr, d.ptr) == A.BAD_INPUT
    assert lib.femgpu_apply_noalloc(f.handle, 0, None, d.ptr) == A.BAD_INPUT
    d.free()
    f.close()
