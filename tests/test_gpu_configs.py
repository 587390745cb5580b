"""The benchmark configurations at full size on the GPU.

C3  RMAT-22 ef16 FP32 walks, 1,048,576 walkers x 80 steps: oracle rows for
    the first 16,384 walkers; start, -1 persistence, dead-end and step-count
    invariants on every row; every step an edge of the graph on 1 row in 8.
C4  RMAT-22, 256 mixed requests (Philox mix key 42), batched + learned
    schedule + CDF cache + autotuned shapes + hipGraph replay: every
    request's digest equals the plain per-request run's, and the first 2% of
    every request's samples equal the oracle (main.cu:21-93 per request).
C5  RMAT-26 ef16 (int64 offsets) streamed under a 4 GB budget in 16
    partitions with the learned schedule: the edge list of all 65,536 seeds
    equals the in-memory run, and the first 256 seeds equal the oracle
    (streaming_sampling.cu:427-728).
"""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELS = os.path.join(ROOT, "c-saw_amd", "models")
KEYS = ("src", "dst", "smp", "hop_off")
NT = min(16, os.cpu_count() or 1)


def cost_model():
    sys.path.insert(0, os.path.join(ROOT, "c-saw_amd"))
    import cost_model as CM
    return CM


@pytest.fixture(scope="module")
def rmat22(C, O):
    g = C.Graph.rmat(22, 16, weights=True)
    off, adj, w = g.csr(weights=True)
    print(f"RMAT-22 V={g.V} E={g.E}", flush=True)
    return g, off, adj, w, O.OracleGraph(off, adj, weights=w)


def test_c3_walks_full_size(C, O, rmat22):
    g, off, adj, w, og = rmat22
    n, L = 1 << 20, 80
    seeds = g.seeds(n)
    W, steps = C.random_walk(g, seeds, L, bias=C.BIAS_WEIGHT)
    assert W.shape == (n, L + 1) and np.array_equal(W[:, 0], seeds)
    neg = W < 0
    assert not (neg[:, :-1] & ~neg[:, 1:]).any()  # -1 persists
    assert steps == int((~neg[:, 1:]).sum())
    deg = np.diff(off)
    # a step ends (-1) exactly at a vertex without out-neighbours
    u, v = W[:, :-1], W[:, 1:]
    live = u >= 0
    assert np.array_equal(v[live] < 0, deg[u[live]] == 0)
    # every step is an edge: (u, v) found in u's sorted list, on 1 row in 8
    us, vs = W[::8, :-1].ravel(), W[::8, 1:].ravel()
    m = vs >= 0
    us, vs = us[m].astype(np.int64), vs[m].astype(np.int64)
    keys = np.repeat(np.arange(g.V, dtype=np.int64), deg) * g.V + adj
    q = us * g.V + vs
    pos = np.searchsorted(keys, q)
    assert (pos < len(keys)).all() and np.array_equal(keys[pos], q)
    Wo, so = O.walk(og, seeds[:16384], L, bias=O.BIAS_WEIGHT, nthreads=NT)
    assert np.array_equal(Wo, W[:16384])


def test_c4_batched_equals_plain_and_oracle(C, O, rmat22):
    g, off, adj, w, og = rmat22
    CM = cost_model()
    specs = []
    for i, r in enumerate(C.request_mix(256, 64, 65536, 42)):
        r = dict(r)
        r["seeds"] = g.seeds(r["samples"], key=42, request=i)
        specs.append(r)
    mlp, _ = CM.load(C, os.path.join(MODELS, "cost_model_gfx950.json"))
    t = C.Tuning()
    for k, v in json.load(open(os.path.join(MODELS, "tuning_gfx950_c4.json"))).items():
        setattr(t, k, int(v))
    plain = C.Plan(g, specs, mode=C.MODE_PLAIN)
    plain.run()
    want = [plain.digest(i) for i in range(len(specs))]
    bat = C.Plan(g, specs, mode=C.MODE_BATCHED, sched=C.SCHED_MLP, mlp=mlp, use_graph=True, cdf_cache=True,
                 tuning=t)
    for rep in range(2):  # graph capture, then a replay
        bat.run()
        got = [bat.digest(i) for i in range(len(specs))]
        assert got == want, f"replay {rep}: {sum(a != b for a, b in zip(got, want))} requests differ"
    edges = sum(plain.n_edges(i) for i in range(len(specs)))
    print(f"C4 edges per step {edges}", flush=True)
    assert edges > 3e8
    for i, s in enumerate(specs):
        m = max(1, int(np.ceil(0.02 * len(s["seeds"]))))
        r = O.sample(og, s["seeds"][:m], s["fanouts"], bias=s["bias"], key=s["key"], request=s["request"],
                     nthreads=NT)
        full = bat.result(i)
        for h in range(len(s["fanouts"])):
            a, b = full["hop_off"][h, 0], full["hop_off"][h, m]
            ra, rb = r["hop_off"][h, 0], r["hop_off"][h, m]
            assert np.array_equal(full["src"][a:b], r["src"][ra:rb]), (i, h)
            assert np.array_equal(full["dst"][a:b], r["dst"][ra:rb]), (i, h)
            assert np.array_equal(full["smp"][a:b], r["smp"][ra:rb]), (i, h)
    bat.close()
    plain.close()


def test_c5_streaming_full_size(C, O):
    CM = cost_model()
    hg = C.HostGraph.rmat(26, 16)
    print(f"RMAT-26 V={hg.V} E={hg.E}", flush=True)
    mlp, _ = CM.load(C, CM.STREAM_MODEL)
    seeds = C.seeds_v(hg.V, 65536)
    fan = [15, 10, 5]
    res = C.stream_sample(hg, seeds, fan, bias=C.BIAS_DEGREE, budget_bytes=4 << 30, partitions=16,
                          mode=C.MODE_BATCHED, sched=C.SCHED_MLP, mlp=mlp)
    st = res["stats"]
    assert st.partitions == 16 and st.peak_device_bytes <= 4 << 30
    print(f"C5 streamed {st.edges} edges in {st.device_ms:.1f} ms", flush=True)
    off, adj = hg.csr_views()
    g = C.Graph.from_csr(off, adj)
    want = C.sample(g, seeds, fan, bias=C.BIAS_DEGREE)
    for k in KEYS:
        assert np.array_equal(res[k], want[k]), k
    g.close()
    og = O.OracleGraph(off, adj)
    m = 256
    r = O.sample(og, seeds[:m], fan, bias=O.BIAS_DEGREE, nthreads=NT)
    for h in range(len(fan)):
        a, b = res["hop_off"][h, 0], res["hop_off"][h, m]
        ra, rb = r["hop_off"][h, 0], r["hop_off"][h, m]
        assert np.array_equal(res["src"][a:b], r["src"][ra:rb]) and np.array_equal(res["dst"][a:b], r["dst"][ra:rb])
    hg.close()
