"""kd_eval_batch_device_noalloc (the entry point that allocates nothing) through
the C-ABI against kd_eval_batch_device on the same device buffers: identical
values and domain flags (same kernels, so bit for bit), on two calls in a row,
the second with fewer points, at an order of the per-configuration kernel and
at one of the split kernel; and the size limit is refused with KD_ERR_RANGE."""
import ctypes as C

import numpy as np
import pytest

from keldy_amd import capi
from tests import gpu_util as U

pytestmark = pytest.mark.gpu
KD_NOALLOC_MAX_POINTS = 32768


_hip = None


def hip():
    """The HIP runtime through ctypes: device buffers for the device-pointer entry points."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return _hip


class DevBuf:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(nbytes, 8)) == 0
        self.nbytes = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device

    def download(self, shape, dtype):
        a = np.empty(shape, dtype=dtype)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, a.nbytes, 2) == 0  # device to host
        return a

    def fill(self, byte):
        assert hip().hipMemset(self.p, byte, self.nbytes) == 0

    def __del__(self):
        if self.p:
            hip().hipFree(self.p)


def _dev_eval(fn, ctx, du, n):
    out = DevBuf(16 * n)
    dom = DevBuf(4 * n)
    out.fill(0xFF)  # NaN pattern
    dom.fill(0xFF)
    assert hip().hipDeviceSynchronize() == 0
    rc = fn(ctx.h, du.p, n, 1, out.p, dom.p)
    assert rc == capi.KD_OK, capi.lib().kd_last_error(ctx.h)
    return out.download((n, 2), np.float64), dom.download((n,), np.int32)


@pytest.mark.parametrize("order", [5, 8])
def test_noalloc_matches_eval_batch_device(flat_g0, order):
    t = 10.0
    ctx = U.charge_ctx(flat_g0, order, t)
    rng = np.random.default_rng(300 + order)
    L = capi.lib()
    for n in (KD_NOALLOC_MAX_POINTS, 1000):  # the limit, then a smaller call on the same scratch
        u = rng.uniform(-0.5, t, size=(n, order))  # some points out of the domain
        du = DevBuf(u.nbytes)
        du.upload(u)
        print(f"order {order}: {n} points", flush=True)
        ref, rdom = _dev_eval(L.kd_eval_batch_device, ctx, du, n)
        got, dom = _dev_eval(L.kd_eval_batch_device_noalloc, ctx, du, n)
        assert np.array_equal(dom, rdom)
        assert np.any(dom == 0) and np.any(dom == 1)
        assert np.array_equal(got, ref)


def test_noalloc_refuses_more_than_its_scratch(flat_g0):
    ctx = U.charge_ctx(flat_g0, 4, 10.0)
    n = KD_NOALLOC_MAX_POINTS + 1
    du = DevBuf(8 * 4 * n)
    out = DevBuf(16 * n)
    rc = capi.lib().kd_eval_batch_device_noalloc(ctx.h, du.p, n, 1, out.p, None)
    assert rc == capi.KD_ERR_RANGE
    assert b"KD_NOALLOC_MAX_POINTS" in capi.lib().kd_last_error(ctx.h)
