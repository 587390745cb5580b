"""csaw_mdrw_noalloc (the allocation-free multi-dimensional walk, for callers
that cannot allocate device memory) against csaw_mdrw on the same inputs:
identical edges and offsets, on two calls in a row (the second smaller, so
stale scratch from the first would show), plus the size limit and the
device-side seed check.  Buffers come from the HIP runtime the library links
(libamdhip64.so.7)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(nbytes, 4)) == 0
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def to_host(self, p, n, dtype):
        out = np.zeros(max(n, 1), dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, n * out.itemsize, 2) == 0  # device to host
        return out[:n]


def run_noalloc(C, hip, g, seeds, S, FS, D, request=0, key=42, sync=1):
    bufs = [hip.upload(seeds.astype(np.int32)), hip.alloc(4 * S * D), hip.alloc(4 * S * D), hip.alloc(8 * (S + 1))]
    try:
        # poison the outputs: every entry the call leaves must be rewritten
        for b, nb in zip(bufs[1:], (4 * S * D, 4 * S * D, 8 * (S + 1))):
            assert hip.L.hipMemset(b, 0x7f, max(nb, 4)) == 0
        assert hip.L.hipDeviceSynchronize() == 0
        st = C.lib().csaw_mdrw_noalloc(g.h, key, request, S, FS, D, bufs[0], bufs[1], bufs[2], bufs[3], None, sync)
        if st != 0:
            return st, None
        if not sync:  # the work runs on the graph's own (non-blocking) stream
            assert hip.L.hipDeviceSynchronize() == 0
        off = hip.to_host(bufs[3], S + 1, np.int64)
        n = int(off[-1])
        return 0, (hip.to_host(bufs[1], n, np.int32), hip.to_host(bufs[2], n, np.int32), off)
    finally:
        for b in bufs:
            hip.L.hipFree(b)


def test_mdrw_noalloc_matches_mdrw(C, O):
    hip = Hip()
    g = C.Graph.rmat(14, 8)
    off, adj = g.csr()
    og = O.OracleGraph(off, adj)
    # two calls in a row, the second with a smaller input; then the first again
    for (S, FS, D, req) in ((3000, 16, 40, 3), (257, 5, 7, 4), (3000, 16, 40, 3), (1, 1, 1, 0)):
        seeds = g.seeds(S * FS, request=req)
        st, got = run_noalloc(C, hip, g, seeds, S, FS, D, request=req)
        assert st == 0, C.last_error()
        src, dst, offs = C.mdrw(g, seeds, S, FS, D, request=req)
        assert np.array_equal(got[2], offs), (S, FS, D)
        assert np.array_equal(got[0], src) and np.array_equal(got[1], dst), (S, FS, D)
        so, do, oo = O.mdrw(og, seeds, S, FS, D, request=req)
        assert np.array_equal(got[2], oo) and np.array_equal(got[0], so) and np.array_equal(got[1], do)


def test_mdrw_noalloc_limits(C):
    hip = Hip()
    g = C.Graph.rmat(10, 4)
    # empty requests
    st, got = run_noalloc(C, hip, g, np.zeros(1, np.int32), 0, 1, 5)
    assert st == 0 and got[2].tolist() == [0]
    seeds = g.seeds(64)
    st, got = run_noalloc(C, hip, g, seeds, 64, 1, 0)
    assert st == 0 and (got[2] == 0).all()
    # beyond the 32 MiB scratch: rejected before any device work
    S, D = 65536, 128
    st = C.lib().csaw_mdrw_noalloc(g.h, 42, 0, S, 1, D, ctypes.c_void_p(1), ctypes.c_void_p(1), ctypes.c_void_p(1),
                                   ctypes.c_void_p(1), None, 1)
    assert st == 6 and "scratch" in C.last_error()  # CSAW_ERR_CAPACITY
    # a seed outside [0, V): its sample emits nothing, the others are walked,
    # and the synchronous call reports it
    bad = g.seeds(8 * 4)
    bad[4 * 3 + 1] = g.V
    st, _ = run_noalloc(C, hip, g, bad, 8, 4, 6)
    assert st == 1 and "seed" in C.last_error()
    st, got = run_noalloc(C, hip, g, bad, 8, 4, 6, sync=0)
    assert st == 0
    offs = got[2]
    assert offs[4] == offs[3]  # sample 3 emitted nothing
    fixed = bad.copy()
    fixed[4 * 3 + 1] = 0
    src, dst, oo = C.mdrw(g, fixed, 8, 4, 6)
    for s in range(8):  # every other sample is walked as usual
        if s != 3:
            a, b = offs[s], offs[s + 1]
            assert np.array_equal(got[0][a:b], src[oo[s]:oo[s + 1]]) and np.array_equal(got[1][a:b], dst[oo[s]:oo[s + 1]])
