"""laps_mlp_grad_noalloc (the network gradient with no device allocation) against
laps_mlp_grad on the same samples, through the C ABI: two calls in a row, the
second with fewer samples, so a stale scratch array would show; and the size
limit is an error, not a fault.

The test's own device buffers come from the HIP runtime the library itself links
(libamdhip64.so.7), through ctypes, so that one runtime serves the whole process."""
import ctypes

import numpy as np
import pytest

from laps_amd import runtime
from laps_amd.params import LAPS_EINVAL

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 8192
H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    def __init__(self):
        runtime.lib()   # loads the library, and with it its HIP runtime
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipSetDevice.argtypes = [ctypes.c_int]
        assert self.h.hipSetDevice(0) == 0
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def test_noalloc_grad_equals_allocating_grad():
    hip = Hip()
    L = runtime.lib()
    rng = np.random.default_rng(11)
    P = (0.2 * rng.standard_normal(runtime.MLP_NPARAM)).astype(np.float32)
    try:
        for n in (3000, 700):   # the second call is smaller: nothing may depend on the first's contents
            print(f"noalloc gradient: n = {n}", flush=True)
            X = rng.standard_normal((n, 28)).astype(np.float32)
            T = rng.standard_normal((n, runtime.MLP_NOUT)).astype(np.float32)
            g_ref, l_ref = runtime.mlp_grad(X, T, P)
            nan_g = np.full(runtime.MLP_NPARAM, np.nan, np.float32)
            nan_l = np.full(1, np.nan, np.float32)
            dX, dT, dP = hip.upload(X), hip.upload(T), hip.upload(P)
            dG, dL = hip.upload(nan_g), hip.upload(nan_l)
            rc = L.laps_mlp_grad_noalloc(0, n, dX, dT, dP, dG, dL, None)
            assert rc == 0, L.laps_last_error().decode()
            g = hip.download(dG, nan_g)
            l = float(hip.download(dL, nan_l)[0])
            # same kernels on the same samples, fixed-order sums: the results are identical
            assert np.array_equal(g, g_ref), float(np.abs(g - g_ref).max())
            assert l == l_ref
            hip.free()
    finally:
        hip.free()


def test_noalloc_grad_rejects_oversized_batch():
    hip = Hip()
    L = runtime.lib()
    try:
        p = hip.upload(np.zeros(16, np.float32))
        # rejected on the host before any launch
        assert L.laps_mlp_grad_noalloc(0, NOALLOC_MAX + 1, p, p, p, p, p, None) == LAPS_EINVAL
        assert b"outside" in L.laps_last_error()
        assert L.laps_mlp_grad_noalloc(0, 0, p, p, p, p, p, None) == LAPS_EINVAL
    finally:
        hip.free()
