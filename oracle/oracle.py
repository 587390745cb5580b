"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker or as the timed CPU baseline.  The product library
(c-saw_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libcsaw_oracle.so")

BIAS_UNIFORM, BIAS_DEGREE, BIAS_WEIGHT, BIAS_USER = 0, 1, 2, 3
MAX_HOPS = 8


class Graph(ctypes.Structure):
    _fields_ = [("V", ctypes.c_int64), ("E", ctypes.c_int64),
                ("off", ctypes.c_void_p), ("adj", ctypes.c_void_p), ("cdf", ctypes.c_void_p)]


class Request(ctypes.Structure):
    _fields_ = [("samples", ctypes.c_int32), ("depth", ctypes.c_int32),
                ("fanout", ctypes.c_int32 * MAX_HOPS), ("bias", ctypes.c_int32),
                ("dedup", ctypes.c_int32), ("request", ctypes.c_uint32), ("key", ctypes.c_uint64)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.orc_philox.argtypes = [P, P, P]
        L.orc_rmat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               ctypes.c_uint64, P, P, P, P]
        L.orc_rmat.restype = ctypes.c_int
        L.orc_free.argtypes = [P]
        L.orc_seeds.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int64, ctypes.c_int64, P]
        L.orc_weights.argtypes = [ctypes.c_uint64, ctypes.c_int64, P]
        L.orc_fp32_cdf.argtypes = [ctypes.c_int64, P, P, P]
        L.orc_sample_capacity.argtypes = [P]
        L.orc_sample_capacity.restype = ctypes.c_int64
        L.orc_sample.argtypes = [P, P, P, P, P, P, P, ctypes.c_int]
        L.orc_sample.restype = ctypes.c_int64
        L.orc_walk.argtypes = [P, ctypes.c_int32, ctypes.c_uint64, ctypes.c_uint32, P, ctypes.c_int64,
                               ctypes.c_int32, P, ctypes.c_int]
        L.orc_walk.restype = ctypes.c_int64
        L.orc_mdrw.argtypes = [P, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32,
                               ctypes.c_int32, P, P, P, P]
        L.orc_mdrw.restype = ctypes.c_int64
        L.orc_select_source.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32, P]
        L.orc_select_source.restype = ctypes.c_int32
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def philox(counter, key):
    c = np.asarray(counter, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox(_p(c), _p(k), _p(o))
    return o


def rmat(scale, edge_factor, a=0.57, b=0.19, c=0.19, key=42):
    V, E = ctypes.c_int64(), ctypes.c_int64()
    off_p, adj_p = ctypes.c_void_p(), ctypes.c_void_p()
    rc = lib().orc_rmat(scale, edge_factor, a, b, c, key, ctypes.byref(V), ctypes.byref(E),
                        ctypes.byref(off_p), ctypes.byref(adj_p))
    if rc != 0:
        raise RuntimeError(f"orc_rmat failed: {rc}")
    off = np.ctypeslib.as_array(ctypes.cast(off_p, ctypes.POINTER(ctypes.c_int64)), (V.value + 1,)).copy()
    adj = np.ctypeslib.as_array(ctypes.cast(adj_p, ctypes.POINTER(ctypes.c_int32)), (max(E.value, 1),))[:E.value].copy()
    lib().orc_free(off_p)
    lib().orc_free(adj_p)
    return off, adj


def seeds(V, n, key=42, request=0):
    out = np.zeros(n, dtype=np.int32)
    lib().orc_seeds(key, request, V, n, _p(out))
    return out


def weights(E, key=42):
    w = np.zeros(E, dtype=np.float32)
    lib().orc_weights(key, E, _p(w))
    return w


def fp32_cdf(off, w):
    off = np.ascontiguousarray(off, dtype=np.int64)
    w = np.ascontiguousarray(w, dtype=np.float32)
    cdf = np.zeros_like(w)
    lib().orc_fp32_cdf(len(off) - 1, _p(off), _p(w), _p(cdf))
    return cdf


class OracleGraph:
    """Keeps the numpy arrays alive for the ctypes struct."""

    def __init__(self, off, adj, weights=None, cdf=None):
        self.off = np.ascontiguousarray(off, dtype=np.int64)
        self.adj = np.ascontiguousarray(adj, dtype=np.int32)
        if cdf is None and weights is not None:
            cdf = fp32_cdf(self.off, weights)
        self.cdf = None if cdf is None else np.ascontiguousarray(cdf, dtype=np.float32)
        self.s = Graph(len(self.off) - 1, len(self.adj), _p(self.off).value, _p(self.adj).value,
                       _p(self.cdf).value if self.cdf is not None else None)

    @property
    def V(self):
        return len(self.off) - 1


def make_request(samples, fanouts, bias=BIAS_UNIFORM, dedup=1, key=42, request=0):
    r = Request()
    r.samples = samples
    r.depth = len(fanouts)
    for i, f in enumerate(fanouts):
        r.fanout[i] = f
    r.bias = bias
    r.dedup = dedup
    r.request = request
    r.key = key
    return r


def sample(g, seeds_arr, fanouts, bias=BIAS_UNIFORM, dedup=1, key=42, request=0, nthreads=1):
    """Returns dict(src, dst, smp, hop_off[depth, S+1])."""
    seeds_arr = np.ascontiguousarray(seeds_arr, dtype=np.int32)
    r = make_request(len(seeds_arr), fanouts, bias, dedup, key, request)
    cap = max(lib().orc_sample_capacity(ctypes.byref(r)), 1)
    src = np.zeros(cap, np.int32)
    dst = np.zeros(cap, np.int32)
    smp = np.zeros(cap, np.int32)
    ho = np.zeros(max(len(fanouts), 1) * (len(seeds_arr) + 1), np.int64)
    n = lib().orc_sample(ctypes.byref(g.s), ctypes.byref(r), _p(seeds_arr), _p(src), _p(dst), _p(smp), _p(ho),
                         nthreads)
    if n < 0:
        raise ValueError("orc_sample rejected the request")
    return {"src": src[:n], "dst": dst[:n], "smp": smp[:n],
            "hop_off": ho[:len(fanouts) * (len(seeds_arr) + 1)].reshape(len(fanouts), len(seeds_arr) + 1)}


def walk(g, seeds_arr, length, bias=BIAS_WEIGHT, key=42, request=0, nthreads=1):
    seeds_arr = np.ascontiguousarray(seeds_arr, dtype=np.int32)
    W = np.zeros((len(seeds_arr), length + 1), np.int32)
    n = lib().orc_walk(ctypes.byref(g.s), bias, key, request, _p(seeds_arr), len(seeds_arr), length, _p(W),
                       nthreads)
    if n < 0:
        raise ValueError("orc_walk rejected the request")
    return W, n


def mdrw(g, seeds_arr, samples, fs, depth, key=42, request=0):
    seeds_arr = np.ascontiguousarray(seeds_arr, dtype=np.int32)
    cap = max(samples * depth, 1)
    src = np.zeros(cap, np.int32)
    dst = np.zeros(cap, np.int32)
    off = np.zeros(samples + 1, np.int64)
    n = lib().orc_mdrw(ctypes.byref(g.s), key, request, samples, fs, depth, _p(seeds_arr), _p(src), _p(dst),
                       _p(off))
    if n < 0:
        raise ValueError("orc_mdrw rejected the request")
    return src[:n], dst[:n], off


def select_source(g, v, fanout, bias=BIAS_UNIFORM, dedup=1, key=42, request=0, hop=0, sample=0, pos=0):
    d = int(g.off[v + 1] - g.off[v])
    idx = np.zeros(max(fanout, d, 1), np.int32)
    k = lib().orc_select_source(ctypes.byref(g.s), bias, dedup, fanout, v, key, request, hop, sample, pos,
                                _p(idx))
    return idx[:k]
