"""csaw_plan_copy_device fills caller-owned device buffers with the same edge
lists csaw_plan_copy returns on the host.  The buffers come from the HIP
runtime the library itself links (libamdhip64.so.7): a framework that hands
its own device memory to the library must share that runtime."""
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

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(nbytes, 4)) == 0
        return p

    def to_host(self, p, n, dtype):
        out = np.zeros(max(n, 1), dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, n * out.itemsize, 2) == 0  # device to host
        return out[:n]


def test_plan_copy_device(C):
    hip = Hip()
    g = C.Graph.rmat(14, 8, weights=True)
    specs = []
    for i, (fan, bias) in enumerate([([5, 3], 0), ([4, 2, 2], 1), ([6], 2), ([3, 0], 1)]):
        specs.append({"seeds": g.seeds(300 + 50 * i, request=i), "fanouts": fan, "bias": bias, "dedup": 1,
                      "key": 42, "request": i})
    for mode in (C.MODE_PLAIN, C.MODE_BATCHED):
        p = C.Plan(g, specs, mode=mode)
        p.run()
        for r, spec in enumerate(specs):
            want = p.result(r)
            n = p.n_edges(r)
            S, D = len(spec["seeds"]), len(spec["fanouts"])
            bufs = [hip.alloc(4 * n), hip.alloc(4 * n), hip.alloc(4 * n), hip.alloc(8 * D * (S + 1))]
            try:
                st = C.lib().csaw_plan_copy_device(p.h, r, bufs[0], bufs[1], bufs[2], bufs[3], None, 1)
                assert st == 0, C.last_error()
                got = {"src": hip.to_host(bufs[0], n, np.int32), "dst": hip.to_host(bufs[1], n, np.int32),
                       "smp": hip.to_host(bufs[2], n, np.int32),
                       "hop_off": hip.to_host(bufs[3], D * (S + 1), np.int64).reshape(D, S + 1)}
                for k in ("src", "dst", "smp", "hop_off"):
                    assert np.array_equal(got[k], want[k]), (mode, r, k)
            finally:
                for b in bufs:
                    hip.L.hipFree(b)
        p.close()
