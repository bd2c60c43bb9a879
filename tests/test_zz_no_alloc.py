"""hpl_gesl_noalloc: the substitution that allocates nothing (run with -m gpu).

It runs hpl_gesl's kernels on the caller's device memory with the inverses in
the library's fixed device array, so its solution must equal hpl_gesl's bit
for bit on the same factor and right-hand side.  Two calls in a row, the
second on a smaller system (the array keeps the first call's contents), a
leading dimension above n, and the size limit, which is refused on the host
before anything is enqueued.  Device memory comes from the HIP runtime the
library itself is linked against (hipMalloc through ctypes).
"""
import ctypes as C

import numpy as np
import pytest

from hpcc_fpga_amd.linpack import load_library
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libhpl_hip.so uses."""

    def __init__(self):
        load_library()
        self.h = C.CDLL("libamdhip64.so.7")  # already loaded by the library: same runtime
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipSetDevice.argtypes = [C.c_int]
        assert self.h.hipSetDevice(0) == 0

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, a):
        assert self.h.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.h.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0  # synchronous

    def free(self, p):
        assert self.h.hipFree(p) == 0


def _factor(ctx, n, lda, seed):
    """A GEFA factor of the sized system (host, leading dimension lda) and b."""
    A, b, _, _ = O.generate_sized(n, seed, 0)
    F = np.zeros((n, lda), np.float32)
    F[:, :n] = A
    B = 256
    while n % B:
        B //= 2
    ctx._check(ctx._L.hpl_gefa(ctx._c, F.ctypes.data, n, lda, B, None))
    return F, b


def _solve_both(ctx, hip, L, n, lda, seed):
    F, b = _factor(ctx, n, lda, seed)
    x_ref = b.copy()
    ctx._check(L.hpl_gesl(ctx._c, F.ctypes.data, x_ref.ctypes.data, n, lda, None))
    dF = hip.alloc(F.nbytes)
    dx = hip.alloc(b.nbytes)
    try:
        hip.h2d(dF, F)
        hip.h2d(dx, b)
        st = L.hpl_gesl_noalloc(dF, n, lda, dx, None)
        assert st == 0, st
        x = np.empty_like(b)
        hip.d2h(x, dx)
    finally:
        hip.free(dF)
        hip.free(dx)
    return x, x_ref, F


def test_noalloc_equals_hpl_gesl_twice_then_smaller(ctx):
    L = load_library()
    hip = _Hip()
    for n, lda, seed in ((2048, 2048, 3), (1024, 1088, 5), (544, 544, 9)):
        _say(f"gesl_noalloc n={n} lda={lda}")
        x, x_ref, _ = _solve_both(ctx, hip, L, n, lda, seed)
        assert np.array_equal(x.view(np.uint32), x_ref.view(np.uint32)), n
        assert np.max(np.abs(x - 1.0)) < 1e-4


def test_noalloc_refuses_sizes_and_arguments():
    L = load_library()
    hip = _Hip()
    d = hip.alloc(4096)
    try:
        # n * (416 + 8 * 256) bytes exceed the 40 MiB array from n = 17152
        # (the first multiple of 256 past 16896): refused on the host
        # (HPL_ERR_BUDGET), nothing enqueued
        assert L.hpl_gesl_noalloc(d, 32768, 32768, d, None) == 5
        assert L.hpl_gesl_noalloc(d, 17152, 17152, d, None) == 5
        assert L.hpl_gesl_noalloc(None, 64, 64, d, None) == 1
        assert L.hpl_gesl_noalloc(d, 64, 63, d, None) == 1   # lda < n
        assert L.hpl_gesl_noalloc(d, 64, 66, d, None) == 1   # lda not a multiple of 4
        assert L.hpl_gesl_noalloc(d, 48, 48, d, None) == 8   # n not a multiple of 32
    finally:
        hip.free(d)
