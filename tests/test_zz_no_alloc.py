"""sdem_gnn_gemm_noalloc against sdem_gnn_gemm: the same products with no
allocation in the call (its staging and split-K partial sums in the
library's fixed scratch array), through the C ABI on device buffers that
the test allocates with the HIP runtime.  Two calls in a row, the second
with smaller inputs, so the second reads nothing the first left behind."""
import ctypes as C

import numpy as np
import pytest

from sdem_amd import gnn, native

pytestmark = pytest.mark.gpu

hipMemcpyHostToDevice, hipMemcpyDeviceToHost = 1, 2
SDEM_E_CAPACITY = -3


@pytest.fixture(scope="module")
def hip():
    rt = C.CDLL("libamdhip64.so")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


class Dev:
    """A device buffer holding a float32 array (allocated by the test)."""

    def __init__(self, rt, arr):
        self.rt, self.arr = rt, np.ascontiguousarray(arr, np.float32)
        self.p = C.c_void_p()
        assert rt.hipMalloc(C.byref(self.p), max(self.arr.nbytes, 4)) == 0
        assert rt.hipMemcpy(self.p, self.arr.ctypes.data, self.arr.nbytes, hipMemcpyHostToDevice) == 0

    def get(self):
        out = np.empty_like(self.arr)
        assert self.rt.hipMemcpy(out.ctypes.data, self.p, out.nbytes, hipMemcpyDeviceToHost) == 0
        return out

    def free(self):
        self.rt.hipFree(self.p)


def noalloc(lib, rt, a, b, ta, tb, bias, flags, splits):
    f = lib.sdem_gnn_gemm_noalloc
    f.argtypes = [C.c_int32] * 5 + [C.c_void_p] * 3 + [C.c_int32] * 2 + [C.c_void_p] * 3
    f.restype = C.c_int
    M = a.shape[1] if ta else a.shape[0]
    K = a.shape[0] if ta else a.shape[1]
    N = b.shape[0] if tb else b.shape[1]
    da, db = Dev(rt, a), Dev(rt, b)
    dbias = Dev(rt, bias if bias is not None else np.zeros(N, np.float32))
    dc, dpre = Dev(rt, np.full((M, N), np.nan, np.float32)), Dev(rt, np.full((M, N), np.nan, np.float32))
    rc = f(int(ta), int(tb), M, N, K, da.p, db.p, dbias.p if bias is not None else None, flags, splits, dc.p,
           dpre.p, None)
    assert rc == 0, rc
    assert rt.hipDeviceSynchronize() == 0
    out = dc.get(), dpre.get()
    for d in (da, db, dbias, dc, dpre):
        d.free()
    return out


@pytest.mark.parametrize("splits", [1, 4])
def test_noalloc_gemm_matches_gemm(hip_lib, hip, splits):
    print(f"no-alloc GEMM, split-K {splits}: two calls, the second smaller", flush=True)
    rng = np.random.default_rng(5 + splits)
    for (M, N, K, ta, tb, flags) in ((300, 130, 260, 0, 1, 3 | 4), (70, 40, 100, 1, 0, 1)):
        a = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
        b = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
        bias = rng.normal(size=N).astype(np.float32)
        want, want_pre = gnn.gemm(a, b, ta=bool(ta), tb=bool(tb), bias=bias, flags=flags, splits=splits,
                                  lib=hip_lib)
        got, got_pre = noalloc(hip_lib, hip, a, b, ta, tb, bias, flags, splits)
        assert np.array_equal(got, want), np.abs(got - want).max()
        if splits == 1 and flags & 4:
            assert np.array_equal(got_pre, want_pre)


def test_noalloc_gemm_refuses_what_does_not_fit(hip_lib, hip):
    """Staging beyond the fixed scratch is refused before any work."""
    f = hip_lib.sdem_gnn_gemm_noalloc
    f.argtypes = [C.c_int32] * 5 + [C.c_void_p] * 3 + [C.c_int32] * 2 + [C.c_void_p] * 3
    f.restype = C.c_int
    dummy = Dev(hip, np.zeros(4, np.float32))
    M = N = K = 4096  # 2 * 4096^2 * 2 bytes of bf16 copies alone: 64 MiB
    rc = f(0, 0, M, N, K, dummy.p, dummy.p, None, 0, 1, dummy.p, None, None)
    dummy.free()
    assert rc == SDEM_E_CAPACITY
