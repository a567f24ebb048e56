"""The allocation-free surrogate forward (mh_mlp_forward_noalloc) against
mh_mlp_forward, through the C-ABI: bit for bit, on two calls in a row (the
second with a smaller batch, so the shared scratch array holds the first
call's larger activations), on a caller stream and on the network's own,
and a batch over the scratch limit is refused with MH_ERR_RANGE.

Device buffers come from the HIP runtime libmercury_hip itself links (by
soname, so the process holds one runtime), not from torch."""
import ctypes

import numpy as np
import pytest

import mercury_amd as mh
from mercury_amd.surrogate import MLP

pytestmark = pytest.mark.gpu

MH_ERR_RANGE = -3
SCRATCH_FLOATS = 4 << 20  # MH_MLP_SCRATCH_FLOATS
H2D, D2H = 1, 2           # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    mh.load()  # libmercury_hip first: its runtime is the one the soname resolves to
    rt = ctypes.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    rt.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    rt.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return rt


class DevArray:
    def __init__(self, rt, nbytes):
        self.rt, self.p = rt, ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(self.p), nbytes) == 0

    def put(self, a):
        assert self.rt.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a):
        assert self.rt.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.rt.hipFree(self.p)


def test_mlp_forward_noalloc_matches_forward():
    print("mlp forward no-alloc: 64-256-256-1, batches 3000 then 700", flush=True)
    lib = mh.load()
    rt = hip()
    net = MLP([64, 256, 256, 1], seed=3)
    rng = np.random.default_rng(11)
    stream = ctypes.c_void_p()
    assert rt.hipStreamCreate(ctypes.byref(stream)) == 0
    bufs = []
    try:
        for batch in (3000, 700):
            x = rng.normal(size=(batch, 64)).astype(np.float32)
            want = net.forward(x)
            dx, dy = DevArray(rt, x.nbytes), DevArray(rt, 4 * batch)
            bufs += [dx, dy]
            dx.put(x)
            dy.put(np.full(batch, np.nan, np.float32))
            rc = lib.mh_mlp_forward_noalloc(net._h, dx.p, batch, dy.p, stream)
            assert rc == 0, lib.mh_last_error()
            assert rt.hipStreamSynchronize(stream) == 0
            got = dy.get(np.zeros((batch, 1), np.float32))
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        # the network's own stream (NULL), a third call with a new batch
        x = rng.normal(size=(64, 64)).astype(np.float32)
        dx, dy = DevArray(rt, x.nbytes), DevArray(rt, 4 * 64)
        bufs += [dx, dy]
        dx.put(x)
        assert lib.mh_mlp_forward_noalloc(net._h, dx.p, 64, dy.p, None) == 0
        net.forward(x[:1])  # mh_mlp_forward synchronizes the network's stream
        got = dy.get(np.zeros((64, 1), np.float32))
        assert np.array_equal(got.view(np.uint32), net.forward(x).view(np.uint32))
        # over the limit: refused before any launch (the buffers are never
        # touched, so small ones are passed)
        too_many = SCRATCH_FLOATS // 2 // 256 + 1
        assert lib.mh_mlp_forward_noalloc(net._h, dx.p, too_many, dy.p, None) == MH_ERR_RANGE
    finally:
        rt.hipStreamSynchronize(stream)
        rt.hipStreamDestroy(stream)
        for b in bufs:
            b.free()
