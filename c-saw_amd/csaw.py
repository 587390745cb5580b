"""Python binding of the MI355X C-SAW engine (ctypes over include/csaw.h).

Every call goes to the HIP library lib/libcsaw.so; there is no CPU
fallback.  If the library is missing, importing this module raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcsaw.so")

MAX_HOPS = 8
MAX_FANOUT = 64
NUM_VARIANTS = 5
BIAS_UNIFORM, BIAS_DEGREE, BIAS_WEIGHT, BIAS_USER = 0, 1, 2, 3
VAR_THREAD, VAR_WAVE_BITMAP, VAR_WAVE_HASH, VAR_BLOCK_BITMAP, VAR_BLOCK_HASH = range(5)
VARIANT_NAMES = ["thread", "wave_bitmap", "wave_hash", "block_bitmap", "block_hash"]
SCHED_HEURISTIC, SCHED_MLP, SCHED_FORCED, SCHED_RANDOM = 0, 1, 2, 3
SLICE_AUTO, SLICE_COPY, SLICE_DIRECT = 0, 1, 2
MODE_PLAIN, MODE_BATCHED = 0, 1
STATUS = {0: "OK", 1: "INVALID_ARG", 2: "INVALID_GRAPH", 3: "OOM", 4: "BUDGET_TOO_SMALL", 5: "HIP",
          6: "CAPACITY", 7: "NO_DEVICE", 8: "UNSUPPORTED", 9: "IO"}

P = ctypes.c_void_p
i32, i64, u32, u64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64


class CsawError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: {STATUS.get(status, status)} ({last_error()})")


class Params(ctypes.Structure):
    _fields_ = [("samples", i32), ("frontier_size", i32), ("neighbor_size", i32), ("depth", i32),
                ("fanouts", P), ("bias", i32), ("dedup", i32), ("seed", u64), ("request_id", u32),
                ("sched", i32), ("variant", i32), ("n_blocks", i32), ("n_threads", i32)]


class Stats(ctypes.Structure):
    _fields_ = [("edges", i64), ("kernel_ms", ctypes.c_double), ("algo_bytes", ctypes.c_double),
                ("expansions", i64), ("variant_chunks", i64 * NUM_VARIANTS)]


class Request(ctypes.Structure):
    _fields_ = [("samples", i32), ("depth", i32), ("fanouts", i32 * MAX_HOPS), ("bias", i32), ("dedup", i32),
                ("seed", u64), ("request_id", u32), ("seeds", P)]


class Tuning(ctypes.Structure):
    _fields_ = [("count_per_cu", i32), ("expand_per_cu", i32), ("thread_per_cu", i32), ("block_per_cu", i32),
                ("grid_override", i32), ("chunk_min", i32), ("group_min", i32)]


class PlanOpts(ctypes.Structure):
    _fields_ = [("mode", i32), ("sched", i32), ("variant", i32), ("use_graph", i32), ("timing", i32),
                ("log_chunks", i32), ("explore_key", u64), ("mlp", P), ("tuning", P), ("cdf_cache", i32),
                ("serial_expand", i32)]


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (no CPU fallback exists)")

_L = ctypes.CDLL(LIB_PATH)


def _sig(name, args, res=ctypes.c_int):
    f = getattr(_L, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("csaw_last_error", [], ctypes.c_char_p)
_sig("csaw_graph_create", [P, ctypes.c_int, P, P, i64, i64, P])
_sig("csaw_graph_create_rmat", [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, u64,
                                ctypes.c_int, P])
_sig("csaw_graph_load_bin", [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, P])
_sig("csaw_csr_write_bin", [ctypes.c_char_p, ctypes.c_char_p, P, P, i64, i64])
_sig("csaw_snap_to_bin", [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, P, P])
_sig("csaw_graph_info", [P, P, P, P])
_sig("csaw_graph_copy_csr", [P, P, P, P])
_sig("csaw_graph_destroy", [P])
_sig("csaw_seeds", [P, u64, u32, i64, P])
_sig("csaw_sample", [P, P, P, P, P])
_sig("csaw_edges_info", [P, P, P, P])
_sig("csaw_edges_copy", [P, P, P, P, P])
_sig("csaw_edges_free", [P], None)
_sig("csaw_random_walk", [P, i32, u64, u32, P, i64, i32, P, P])
_sig("csaw_walk_plan_create", [P, i32, u64, u32, P, i64, i32, P])
_sig("csaw_walk_plan_run", [P, P, ctypes.c_int])
_sig("csaw_walk_plan_result", [P, P, P, P, P])
_sig("csaw_walk_plan_destroy", [P])
_sig("csaw_plan_create", [P, P, i32, P, P])
_sig("csaw_plan_run", [P, P, ctypes.c_int])
_sig("csaw_plan_stats", [P, P])
_sig("csaw_plan_request_info", [P, i32, P])
_sig("csaw_plan_copy", [P, i32, P, P, P, P])
_sig("csaw_plan_kernel_times", [P, ctypes.c_char_p, ctypes.c_size_t, P, P, P, i32, P])
_sig("csaw_plan_copy_device", [P, i32, P, P, P, P, P, ctypes.c_int])
_sig("csaw_plan_digest", [P, i32, P])
_sig("csaw_plan_reset_times", [P])
_sig("csaw_plan_destroy", [P])
_sig("csaw_plan_chunk_log", [P, P, P, P, P, i64, P])
_sig("csaw_autotune", [P, P, i32, P, P, P])
_sig("csaw_request_mix", [u64, i32, i32, i32, P], i64)
_sig("csaw_mlp_num_params", [])
_sig("csaw_mlp_create", [u64, P])
_sig("csaw_mlp_destroy", [P])
_sig("csaw_mlp_get_params", [P, P])
_sig("csaw_mlp_set_params", [P, P])
_sig("csaw_mlp_forward", [P, P, i64, P])
_sig("csaw_mlp_grad", [P, P, P, P, i64, P, P])
_sig("csaw_mlp_train", [P, P, P, P, i64, i32, i32, ctypes.c_float, u64, P])


def lib():
    return _L


def last_error():
    s = _L.csaw_last_error()
    return s.decode() if s else ""


def _check(st, what):
    if st != 0:
        raise CsawError(st, what)


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


class Graph:
    def __init__(self, handle):
        self.h = handle
        V, E, md = i64(), i64(), i64()
        _check(_L.csaw_graph_info(self.h, ctypes.byref(V), ctypes.byref(E), ctypes.byref(md)), "graph_info")
        self.V, self.E, self.max_degree = V.value, E.value, md.value

    @classmethod
    def from_csr(cls, off, adj, weights=None):
        off = np.ascontiguousarray(off, dtype=np.int64)
        adj = np.ascontiguousarray(adj, dtype=np.int32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        h = P()
        _check(_L.csaw_graph_create(_p(off), 1, _p(adj), _p(w), len(off) - 1, len(adj), ctypes.byref(h)),
               "graph_create")
        return cls(h)

    @classmethod
    def rmat(cls, scale, edge_factor, a=0.57, b=0.19, c=0.19, key=42, weights=False):
        h = P()
        _check(_L.csaw_graph_create_rmat(scale, edge_factor, a, b, c, key, 1 if weights else 0, ctypes.byref(h)),
               "graph_create_rmat")
        return cls(h)

    @classmethod
    def load_bin(cls, beg, csr, weight=None):
        h = P()
        _check(_L.csaw_graph_load_bin(beg.encode(), csr.encode(), weight.encode() if weight else None,
                                      ctypes.byref(h)), "graph_load_bin")
        return cls(h)

    def csr(self, weights=False):
        off = np.zeros(self.V + 1, np.int64)
        adj = np.zeros(max(self.E, 1), np.int32)
        w = np.zeros(max(self.E, 1), np.float32) if weights else None
        _check(_L.csaw_graph_copy_csr(self.h, _p(off), _p(adj), _p(w)), "graph_copy_csr")
        return (off, adj[:self.E]) if not weights else (off, adj[:self.E], w[:self.E])

    def seeds(self, n, key=42, request=0):
        out = np.zeros(max(n, 1), np.int32)
        _check(_L.csaw_seeds(self.h, key, request, n, _p(out)), "seeds")
        return out[:n]

    def close(self):
        if self.h:
            _L.csaw_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sample(g, seeds, fanouts, bias=BIAS_UNIFORM, dedup=1, key=42, request=0, sched=SCHED_HEURISTIC, variant=0,
           n_blocks=0):
    """Sampler(): one request; returns dict(src, dst, smp, hop_off[depth, S+1], stats)."""
    seeds = np.ascontiguousarray(seeds, dtype=np.int32)
    fan = np.ascontiguousarray(fanouts, dtype=np.int32)
    prm = Params(samples=len(seeds), frontier_size=1, neighbor_size=int(fan[0]) if len(fan) else 0,
                 depth=len(fan), fanouts=_p(fan), bias=bias, dedup=dedup, seed=key, request_id=request,
                 sched=sched, variant=variant, n_blocks=n_blocks, n_threads=0)
    e = P()
    st = Stats()
    _check(_L.csaw_sample(g.h, ctypes.byref(prm), _p(seeds), ctypes.byref(e), ctypes.byref(st)), "sample")
    try:
        n, d, s = i64(), i32(), i32()
        _check(_L.csaw_edges_info(e, ctypes.byref(n), ctypes.byref(d), ctypes.byref(s)), "edges_info")
        src = np.zeros(max(n.value, 1), np.int32)
        dst = np.zeros(max(n.value, 1), np.int32)
        smp = np.zeros(max(n.value, 1), np.int32)
        ho = np.zeros(max(d.value * (s.value + 1), 1), np.int64)
        _check(_L.csaw_edges_copy(e, _p(src), _p(dst), _p(smp), _p(ho)), "edges_copy")
    finally:
        _L.csaw_edges_free(e)
    k = n.value
    return {"src": src[:k], "dst": dst[:k], "smp": smp[:k],
            "hop_off": ho[:d.value * (s.value + 1)].reshape(d.value, s.value + 1), "stats": st}


def random_walk(g, seeds, length, bias=BIAS_WEIGHT, key=42, request=0):
    seeds = np.ascontiguousarray(seeds, dtype=np.int32)
    W = np.zeros((len(seeds), length + 1), np.int32)
    st = Stats()
    _check(_L.csaw_random_walk(g.h, bias, key, request, _p(seeds), len(seeds), length, _p(W), ctypes.byref(st)),
           "random_walk")
    return W, st.edges


class WalkPlan:
    def __init__(self, g, seeds, length, bias=BIAS_WEIGHT, key=42, request=0):
        self.seeds = np.ascontiguousarray(seeds, dtype=np.int32)
        self.len = length
        self.h = P()
        _check(_L.csaw_walk_plan_create(g.h, bias, key, request, _p(self.seeds), len(self.seeds), length,
                                        ctypes.byref(self.h)), "walk_plan_create")

    def run(self, stream=None, sync=True):
        _check(_L.csaw_walk_plan_run(self.h, stream, 1 if sync else 0), "walk_plan_run")

    def result(self, walks=True):
        W = np.zeros((len(self.seeds), self.len + 1), np.int32) if walks else None
        steps, ms, byt = i64(), ctypes.c_double(), ctypes.c_double()
        _check(_L.csaw_walk_plan_result(self.h, _p(W), ctypes.byref(steps), ctypes.byref(ms), ctypes.byref(byt)),
               "walk_plan_result")
        self.algo_bytes = byt.value
        return W, steps.value, ms.value

    def close(self):
        if self.h:
            _L.csaw_walk_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_requests(specs):
    """specs: list of dict(seeds, fanouts, bias, dedup, key, request)."""
    arr = (Request * len(specs))()
    keep = []
    for i, s in enumerate(specs):
        sd = np.ascontiguousarray(s["seeds"], dtype=np.int32)
        keep.append(sd)
        r = arr[i]
        r.samples = len(sd)
        r.depth = len(s["fanouts"])
        for h, f in enumerate(s["fanouts"]):
            r.fanouts[h] = f
        r.bias = s.get("bias", BIAS_UNIFORM)
        r.dedup = s.get("dedup", 1)
        r.seed = s.get("key", 42)
        r.request_id = s.get("request", i)
        r.seeds = _p(sd).value
    return arr, keep


def request_mix(n=256, lo=64, hi=65536, mix_key=42):
    arr = (Request * n)()
    total = _L.csaw_request_mix(mix_key, n, lo, hi, arr)
    if total < 0:
        raise CsawError(1, "request_mix")
    return [{"samples": arr[i].samples, "fanouts": [arr[i].fanouts[h] for h in range(arr[i].depth)],
             "bias": arr[i].bias, "dedup": arr[i].dedup, "key": arr[i].seed, "request": arr[i].request_id}
            for i in range(n)]


class Plan:
    def __init__(self, g, specs, mode=MODE_BATCHED, sched=SCHED_HEURISTIC, variant=0, use_graph=False, timing=False,
                 log_chunks=False, explore_key=0, mlp=None, tuning=None, cdf_cache=False, serial_expand=False):
        self.g = g
        self.specs = specs
        self.reqs, self._keep = make_requests(specs)
        self.tuning = tuning
        opts = PlanOpts(mode=mode, sched=sched, variant=variant, use_graph=1 if use_graph else 0,
                        timing=1 if timing else 0, log_chunks=1 if log_chunks else 0, explore_key=explore_key,
                        mlp=mlp.h if mlp is not None else None,
                        tuning=ctypes.cast(ctypes.byref(tuning), P) if tuning is not None else None,
                        cdf_cache=1 if cdf_cache else 0, serial_expand=1 if serial_expand else 0)
        self.h = P()
        _check(_L.csaw_plan_create(g.h, self.reqs, len(specs), ctypes.byref(opts), ctypes.byref(self.h)),
               "plan_create")

    def run(self, stream=None, sync=True):
        _check(_L.csaw_plan_run(self.h, stream, 1 if sync else 0), "plan_run")

    def stats(self):
        s = Stats()
        _check(_L.csaw_plan_stats(self.h, ctypes.byref(s)), "plan_stats")
        return s

    def n_edges(self, r):
        n = i64()
        _check(_L.csaw_plan_request_info(self.h, r, ctypes.byref(n)), "plan_request_info")
        return n.value

    def result(self, r):
        n = self.n_edges(r)
        S = len(self._keep[r])
        D = len(self.specs[r]["fanouts"])
        src = np.zeros(max(n, 1), np.int32)
        dst = np.zeros(max(n, 1), np.int32)
        smp = np.zeros(max(n, 1), np.int32)
        ho = np.zeros(max(D * (S + 1), 1), np.int64)
        _check(_L.csaw_plan_copy(self.h, r, _p(src), _p(dst), _p(smp), _p(ho)), "plan_copy")
        return {"src": src[:n], "dst": dst[:n], "smp": smp[:n], "hop_off": ho[:D * (S + 1)].reshape(D, S + 1)}

    def kernel_times(self, with_bytes=False):
        """{kind: (avg ms per launch, launches[, algorithmic bytes per launch])}"""
        names = ctypes.create_string_buffer(1024)
        avg = np.zeros(32, np.float64)
        cnt = np.zeros(32, np.int64)
        byt = np.zeros(32, np.float64)
        nk = i32()
        _check(_L.csaw_plan_kernel_times(self.h, names, 1024, _p(avg), _p(cnt), _p(byt), 32, ctypes.byref(nk)),
               "plan_kernel_times")
        nm = names.raw.split(b"\0")
        if with_bytes:
            return {nm[i].decode(): (float(avg[i]), int(cnt[i]), float(byt[i])) for i in range(nk.value)}
        return {nm[i].decode(): (float(avg[i]), int(cnt[i])) for i in range(nk.value)}

    def digest(self, r):
        d = u64()
        _check(_L.csaw_plan_digest(self.h, r, ctypes.byref(d)), "plan_digest")
        return d.value

    def reset_times(self):
        _check(_L.csaw_plan_reset_times(self.h), "plan_reset_times")

    def chunk_log(self, max_n=1 << 24):
        n = i64()
        _check(_L.csaw_plan_chunk_log(self.h, None, None, None, None, 0, ctypes.byref(n)), "plan_chunk_log")
        k = n.value
        X = np.zeros((max(k, 1), 16), np.float32)
        var = np.zeros(max(k, 1), np.int32)
        lt = np.zeros(max(k, 1), np.float32)
        keys = np.zeros(max(k, 1), np.uint64)
        _check(_L.csaw_plan_chunk_log(self.h, _p(X), _p(var), _p(lt), _p(keys), k, ctypes.byref(n)),
               "plan_chunk_log")
        return X[:k], var[:k], lt[:k], keys[:k]

    def close(self):
        if self.h:
            _L.csaw_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(g, specs, mode=MODE_BATCHED, sched=SCHED_HEURISTIC, use_graph=True, cdf_cache=False, mlp=None):
    """Launch and tile shapes for these requests: (Tuning, mean ms per run)."""
    reqs, keep = make_requests(specs)
    opts = PlanOpts(mode=mode, sched=sched, use_graph=1 if use_graph else 0, cdf_cache=1 if cdf_cache else 0,
                    mlp=mlp.h if mlp is not None else None)
    t = Tuning()
    ms = ctypes.c_double()
    _check(_L.csaw_autotune(g.h, reqs, len(specs), ctypes.byref(opts), ctypes.byref(t), ctypes.byref(ms)),
           "autotune")
    return t, ms.value


class Mlp:
    N_IN, N_OUT = 16, NUM_VARIANTS

    def __init__(self, init_key=1):
        self.h = P()
        _check(_L.csaw_mlp_create(init_key, ctypes.byref(self.h)), "mlp_create")
        self.n_params = _L.csaw_mlp_num_params()

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(_L.csaw_mlp_get_params(self.h, _p(p)), "mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_L.csaw_mlp_set_params(self.h, _p(p)), "mlp_set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((len(X), self.N_OUT), np.float32)
        _check(_L.csaw_mlp_forward(self.h, _p(X), len(X), _p(Y)), "mlp_forward")
        return Y

    def grad(self, X, T, M):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        M = np.ascontiguousarray(M, dtype=np.float32)
        loss = ctypes.c_float()
        g = np.zeros(self.n_params, np.float32)
        _check(_L.csaw_mlp_grad(self.h, _p(X), _p(T), _p(M), len(X), ctypes.byref(loss), _p(g)), "mlp_grad")
        return loss.value, g

    def train(self, X, T, M, epochs=20, batch=256, lr=0.05, shuffle_key=7):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        M = np.ascontiguousarray(M, dtype=np.float32)
        loss = ctypes.c_float()
        _check(_L.csaw_mlp_train(self.h, _p(X), _p(T), _p(M), len(X), epochs, batch, lr, shuffle_key,
                                 ctypes.byref(loss)), "mlp_train")
        return loss.value

    def close(self):
        if self.h:
            _L.csaw_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------- streaming
class StreamStats(ctypes.Structure):
    _fields_ = [("edges", i64), ("device_ms", ctypes.c_double), ("h2d_bytes", ctypes.c_double),
                ("partitions", i32), ("peak_device_bytes", i64), ("slice_bytes", i64),
                ("expand_ms", ctypes.c_double), ("chunks", i64), ("variant_chunks", i64 * NUM_VARIANTS),
                ("direct_partitions", i32), ("reused_slices", i32), ("copy_ms", ctypes.c_double),
                ("h2d_raw_bytes", ctypes.c_double), ("hop_ms", ctypes.c_double * 8)]


class StreamOpts(ctypes.Structure):
    _fields_ = [("mode", i32), ("sched", i32), ("variant", i32), ("explore_key", u64), ("mlp", P), ("tuning", P),
                ("slice_policy", i32), ("raw_slices", i32), ("direct_factor", ctypes.c_float),
                ("bin_min_entries", i32), ("log_chunks", i32), ("log_cap", i64), ("log_features", P),
                ("log_variant", P), ("log_log2_ticks", P), ("log_keys", P), ("log_n", P)]


_sig("csaw_host_graph_create_rmat", [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     u64, P])
_sig("csaw_host_graph_from_csr", [P, P, i64, i64, P])
_sig("csaw_host_graph_info", [P, P, P, P, P, P])
_sig("csaw_host_graph_destroy", [P])
_sig("csaw_host_graph_pack", [P, i32, P])
_sig("csaw_stream_sample", [P, P, i64, i32, P, P])
_sig("csaw_stream_sample_opts", [P, P, i64, i32, P, P, P])
_sig("csaw_edges_from_host", [i32, i32, i64, P, P, P, P, P])
_sig("csaw_mdrw", [P, u64, u32, i32, i32, i32, P, P, P, P, P])
# device pointers (seeds, src, dst, offsets), stream, sync; allocates nothing
_sig("csaw_mdrw_noalloc", [P, u64, u32, i32, i32, i32, P, P, P, P, P, ctypes.c_int])


class HostGraph:
    """CSR in pinned host memory for the streaming sampler."""

    def __init__(self, handle):
        self.h = handle
        V, E, md = i64(), i64(), i64()
        po, pa = P(), P()
        _check(_L.csaw_host_graph_info(self.h, ctypes.byref(V), ctypes.byref(E), ctypes.byref(md), ctypes.byref(po),
                                       ctypes.byref(pa)), "host_graph_info")
        self.V, self.E, self.max_degree = V.value, E.value, md.value
        self._po, self._pa = po.value, pa.value

    @classmethod
    def rmat(cls, scale, edge_factor, a=0.57, b=0.19, c=0.19, key=42):
        h = P()
        _check(_L.csaw_host_graph_create_rmat(scale, edge_factor, a, b, c, key, ctypes.byref(h)), "host_graph_rmat")
        return cls(h)

    @classmethod
    def from_csr(cls, off, adj):
        off = np.ascontiguousarray(off, dtype=np.int64)
        adj = np.ascontiguousarray(adj, dtype=np.int32)
        h = P()
        _check(_L.csaw_host_graph_from_csr(_p(off), _p(adj), len(off) - 1, len(adj), ctypes.byref(h)),
               "host_graph_from_csr")
        return cls(h)

    def pack(self, threads=0):
        """Packed copy of the neighbour ids for the streamed slice copies; returns its bytes."""
        b = i64()
        _check(_L.csaw_host_graph_pack(self.h, threads, ctypes.byref(b)), "host_graph_pack")
        return b.value

    def csr_views(self):
        """Zero-copy numpy views of the pinned CSR (valid while the graph lives)."""
        off = np.ctypeslib.as_array(ctypes.cast(self._po, ctypes.POINTER(ctypes.c_int64)), (self.V + 1,))
        adj = np.ctypeslib.as_array(ctypes.cast(self._pa, ctypes.POINTER(ctypes.c_int32)), (max(self.E, 1),))
        return off, adj[:self.E]

    def close(self):
        if self.h:
            _L.csaw_host_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_sample(hg, seeds, fanouts, bias=BIAS_DEGREE, dedup=1, key=42, request=0, budget_bytes=4 << 30,
                  partitions=16, mode=MODE_BATCHED, sched=SCHED_HEURISTIC, variant=0, mlp=None, explore_key=0,
                  tuning=None, slice_policy=0, log_cap=0, raw_slices=False, direct_factor=0.0, bin_min_entries=0):
    """Streams one request; mode / sched / variant / mlp choose how each
    partition's entries are expanded (the edge list is the same for all).
    log_cap > 0 also returns the per-chunk cost-model rows as "log":
    (features [n,16], variant, log2 ticks, keys)."""
    reqs, keep = make_requests([{"seeds": seeds, "fanouts": fanouts, "bias": bias, "dedup": dedup, "key": key,
                                 "request": request}])
    e = P()
    st = StreamStats()
    logs = None
    log_n = i64()
    if log_cap > 0:
        logs = (np.zeros((log_cap, 16), np.float32), np.zeros(log_cap, np.int32), np.zeros(log_cap, np.float32),
                np.zeros(log_cap, np.uint64))
    o = StreamOpts(mode=mode, sched=sched, variant=variant, explore_key=explore_key, slice_policy=slice_policy,
                   raw_slices=1 if raw_slices else 0, direct_factor=direct_factor, bin_min_entries=bin_min_entries,
                   log_chunks=1 if logs else 0, log_cap=log_cap,
                   log_features=_p(logs[0]) if logs else None, log_variant=_p(logs[1]) if logs else None,
                   log_log2_ticks=_p(logs[2]) if logs else None, log_keys=_p(logs[3]) if logs else None,
                   log_n=ctypes.cast(ctypes.byref(log_n), P) if logs else None,
                   mlp=mlp.h if mlp is not None else None,
                   tuning=ctypes.cast(ctypes.byref(tuning), P) if tuning is not None else None)
    _check(_L.csaw_stream_sample_opts(hg.h, reqs, budget_bytes, partitions, ctypes.byref(o), ctypes.byref(e),
                                      ctypes.byref(st)), "stream_sample")
    try:
        n, d, s = i64(), i32(), i32()
        _check(_L.csaw_edges_info(e, ctypes.byref(n), ctypes.byref(d), ctypes.byref(s)), "edges_info")
        src = np.zeros(max(n.value, 1), np.int32)
        dst = np.zeros(max(n.value, 1), np.int32)
        smp = np.zeros(max(n.value, 1), np.int32)
        ho = np.zeros(max(d.value * (s.value + 1), 1), np.int64)
        _check(_L.csaw_edges_copy(e, _p(src), _p(dst), _p(smp), _p(ho)), "edges_copy")
    finally:
        _L.csaw_edges_free(e)
    k = n.value
    out = {"src": src[:k], "dst": dst[:k], "smp": smp[:k],
           "hop_off": ho[:d.value * (s.value + 1)].reshape(d.value, s.value + 1), "stats": st}
    if logs:
        m = log_n.value
        if m > log_cap:
            raise CsawError(6, f"stream_sample: {m} chunk rows exceed log_cap {log_cap}")
        out["log"] = tuple(a[:m] for a in logs)
    return out


def mdrw(g, seeds, samples, frontier_size, depth, key=42, request=0):
    seeds = np.ascontiguousarray(seeds, dtype=np.int32)
    src = np.zeros(max(samples * depth, 1), np.int32)
    dst = np.zeros(max(samples * depth, 1), np.int32)
    off = np.zeros(samples + 1, np.int64)
    st = Stats()
    _check(_L.csaw_mdrw(g.h, key, request, samples, frontier_size, depth, _p(seeds), _p(src), _p(dst), _p(off),
                        ctypes.byref(st)), "mdrw")
    n = int(off[-1])
    return src[:n], dst[:n], off


_sig("csaw_seeds_v", [i64, u64, u32, i64, P])


def seeds_v(V, n, key=42, request=0):
    """Seeds for a vertex count V, computed on the host (same values as Graph.seeds)."""
    out = np.zeros(max(n, 1), np.int32)
    _check(_L.csaw_seeds_v(V, key, request, n, _p(out)), "seeds_v")
    return out[:n]


def write_bin(beg_file, csr_file, off, adj):
    """Writes a CSR as the reference's beg.bin / csr.bin (int64 offsets and ids)."""
    off = np.ascontiguousarray(off, dtype=np.int64)
    adj = np.ascontiguousarray(adj, dtype=np.int32)
    _check(_L.csaw_csr_write_bin(beg_file.encode(), csr_file.encode(), _p(off), _p(adj), len(off) - 1, len(adj)),
           "csr_write_bin")


def snap_to_bin(snap_file, beg_file, csr_file, symmetrize=True):
    """SNAP text edge list -> beg.bin / csr.bin; returns (V, E)."""
    V, E = i64(), i64()
    _check(_L.csaw_snap_to_bin(snap_file.encode(), 1 if symmetrize else 0, beg_file.encode(), csr_file.encode(),
                               ctypes.byref(V), ctypes.byref(E)), "snap_to_bin")
    return V.value, E.value
