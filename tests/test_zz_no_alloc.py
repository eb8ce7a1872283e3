"""grus_cc_noalloc, the allocation-free entry point: through the C-ABI it must
give exactly grus_cc's labels (the existing entry point it shares its
kernels with), call after call, with every device buffer beyond the graph
and the caller's output taken from the library's fixed scratch area.  The
output buffers here are allocated by the test with the HIP runtime."""
import ctypes as C

import numpy as np
import pytest

import grus_framework_amd as grus
from grus_framework_amd import capi

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
    return _hip


class DeviceBuf:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(max(nbytes, 4))) == 0
        self.nbytes = nbytes

    def fill(self, byte):
        # hipMemset runs on the null stream, which does not order with the
        # library's non-blocking stream: wait for it before the library runs,
        # or the fill can land in the middle of the call (garbage labels are
        # parent pointers to CC's kernels)
        assert hip().hipMemset(self.p, C.c_int(byte), C.c_size_t(self.nbytes)) == 0
        assert hip().hipDeviceSynchronize() == 0

    def to_host(self, n):
        out = np.empty(n, np.uint32)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(4 * n), 2) == 0
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = None


def cc_noalloc(g, method, buf):
    st = capi.Stats()
    capi.check(capi.lib().grus_cc_noalloc(g.h, method, buf.p, C.byref(st)))
    return buf.to_host(g.V), st


@pytest.fixture(scope="module")
def dctx():
    c = grus.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("method", [0, 1])
def test_noalloc_equals_cc_twice_then_smaller(dctx, method):
    big = grus.Graph.rmat(dctx, 15)
    small = grus.Graph.rmat(dctx, 11, edge_factor=4)  # many components
    buf = DeviceBuf(4 * big.V)
    try:
        for g in (big, big, small):
            want, _ = grus.cc(g, "hook" if method == 0 else "lp")
            buf.fill(0xAB)  # stale output must be overwritten, not read
            got, st = cc_noalloc(g, method, buf)
            assert (got == want).all()
            assert st.vertices_reached == g.V and st.device_ms > 0
        assert len(np.unique(grus.cc(small)[0])) > 1
    finally:
        buf.free()
        big.close()
        small.close()


def test_noalloc_limits_and_errors(dctx):
    g = grus.Graph.rmat(dctx, 21)  # 2 M vertices: above the method-1 limit
    buf = DeviceBuf(4 * g.V)
    try:
        with pytest.raises(grus.GrusError) as e:
            cc_noalloc(g, 1, buf)
        assert e.value.code == capi.E_BUDGET
        # method 0 needs 4 KiB of scratch at any size
        got, _ = cc_noalloc(g, 0, buf)
        assert (got == grus.cc(g)[0]).all()
        host = np.zeros(g.V, np.uint32)
        rc = capi.lib().grus_cc_noalloc(g.h, 0, host.ctypes.data_as(C.c_void_p), None)
        assert rc == capi.E_INVALID  # a host buffer is refused before any launch
        rc = capi.lib().grus_cc_noalloc(g.h, 2, buf.p, None)
        assert rc == capi.E_INVALID
    finally:
        buf.free()
        g.close()
