"""The allocation-free entry point (gp_parallel_toplesets_propagation_gpu_noalloc)
against the existing sorted[] driver (gp_parallel_toplesets_propagation_gpu),
FP64 on both: two calls in a row on one GPU, the second with a smaller mesh,
must each give the existing driver's field bit for bit (the scratch array
keeps the first call's contents); oversized meshes and meshes without device
tables are rejected.  Device buffers come from the HIP runtime the library
itself links (hipMalloc through ctypes: test plumbing)."""
import ctypes as C

import numpy as np
import pytest

from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G

pytestmark = pytest.mark.gpu


def _hip():
    L.lib()                                   # the library loads its HIP runtime first
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, h, nbytes):
        self.h, self.p = h, C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), max(nbytes, 8)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert self.h.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return self

    def get(self, a):
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.h.hipFree(self.p)


def _noalloc(C_, src, ts):
    h = _hip()
    _, srt, lim = ts
    src, srt, lim = (np.ascontiguousarray(x, np.uint32) for x in (src, srt, lim))
    bufs = [DevBuf(h, x.nbytes).put(x) for x in (src, lim, srt)]
    out = DevBuf(h, 8 * C_.n_vertices()).put(np.full(C_.n_vertices(), -1.0))
    try:
        rc = L.lib().gp_parallel_toplesets_propagation_gpu_noalloc(C_.handle, bufs[0].p, len(src), bufs[1].p, len(lim),
                                                                   bufs[2].p, out.p, None)
        L.check(rc)
        assert h.hipDeviceSynchronize() == 0
        return out.get(np.zeros(C_.n_vertices()))
    finally:
        for b in bufs + [out]:
            b.free()


def test_noalloc_matches_existing_driver():
    # Two calls in a row: a 160^2 height field with 3 sources, then a smaller
    # icosphere with one source.
    cases = [(G.height_field(160, 31), [100, 9000, 20000]), (G.icosphere(4, 0.01, 32), [7])]
    for (xyz, f), src in cases:
        C_ = P.che(xyz, f)
        src = np.array(src, np.uint32)
        P.ptp(C_, src, precision=L.PREC_F64, device_setup=True)    # builds the mesh's FP64 device tables
        ts = C_.compute_toplesets(src)
        ref, _ = P.parallel_toplesets_propagation_gpu(C_, src, ts)
        got = _noalloc(C_, src, ts)
        print(f"\n[noalloc] V={len(xyz)} sources={src.tolist()} identical={np.array_equal(got, ref)}", flush=True)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_noalloc_rejects():
    xyz, f = G.height_field(24, 33)
    C_ = P.che(xyz, f)
    ts = C_.compute_toplesets([5])
    with pytest.raises(P.GPError) as e:                            # no FP64 device tables yet
        _noalloc(C_, np.array([5], np.uint32), ts)
    assert e.value.status == L.ERR_INVALID
    n = (1 << 20) + 1                                              # one vertex past the limit
    big = np.zeros((n, 3))
    big[:, 0] = np.arange(n)
    big[:, 1] = (np.arange(n) % 2) * 1e-3
    Cb = P.che(big, np.array([[0, 1, 2]], np.uint32))
    h = _hip()
    d = DevBuf(h, 64).put(np.zeros(16, np.uint32))
    rc = L.lib().gp_parallel_toplesets_propagation_gpu_noalloc(Cb.handle, d.p, 1, d.p, 2, d.p, d.p, None)
    d.free()
    assert rc == L.ERR_CAPACITY
