"""Region-remap mode (dedup = 2, heur_normalize's collision policy,
functions.cuh:743-868) on the GPU against the oracle: plain and batched
executors, with and without the CDF cache, every bias kind, a
collision-heavy hub graph, the hand-derived known answer, and streaming."""
import numpy as np
import pytest

from test_gpu_parity import KEYS, same, star_with_hub

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g13(C, O):
    g = C.Graph.rmat(13, 8, weights=True)
    off, adj, w = g.csr(weights=True)
    return g, O.OracleGraph(off, adj, weights=w)


@pytest.mark.parametrize("bias", [0, 1, 2, 3])
def test_remap_matches_oracle(C, O, g13, bias):
    g, og = g13
    seeds = g.seeds(400, request=bias)
    obias = bias if bias != 3 else O.BIAS_USER
    want = O.sample(og, seeds, [12, 6, 3], bias=obias, dedup=2, request=bias)
    same(C.sample(g, seeds, [12, 6, 3], bias=bias, dedup=2, request=bias), want, "plain")
    same(C.sample(g, seeds, [12, 6, 3], bias=bias, dedup=2, request=bias, sched=C.SCHED_RANDOM), want, "random")
    specs = [{"seeds": seeds, "fanouts": [12, 6, 3], "bias": bias, "dedup": 2, "request": bias},
             {"seeds": seeds[:50], "fanouts": [5, 5], "bias": bias, "dedup": 1, "request": 9}]
    want2 = O.sample(og, seeds[:50], [5, 5], bias=obias, dedup=1, request=9)
    for mode in (C.MODE_PLAIN, C.MODE_BATCHED):
        for cache in (False, True):
            p = C.Plan(g, specs, mode=mode, cdf_cache=cache, use_graph=mode == C.MODE_BATCHED)
            for _ in range(2):
                p.run()
                same(p.result(0), want, f"mode {mode} cache {cache}")
                same(p.result(1), want2, f"canonical beside remap, mode {mode} cache {cache}")
            p.close()


def test_remap_collisions_and_hubs(C, O):
    off, adj = star_with_hub(O, 9000, 4)
    w = O.weights(len(adj), key=9)
    g = C.Graph.from_csr(off, adj, w)
    og = O.OracleGraph(off, adj, weights=w)
    rng = np.random.default_rng(6)
    seeds = np.concatenate([np.zeros(30, np.int32), rng.integers(1, len(off) - 1, 60).astype(np.int32)])
    for bias in (0, 1, 2):
        for fan in ([2, 2], [7, 9], [64, 1]):
            want = O.sample(og, seeds, fan, bias=bias, dedup=2)
            same(C.sample(g, seeds, fan, bias=bias, dedup=2), want, f"b{bias} {fan}")
            p = C.Plan(g, [{"seeds": seeds, "fanouts": fan, "bias": bias, "dedup": 2, "request": 0}],
                       mode=C.MODE_BATCHED, cdf_cache=True)
            p.run()
            same(p.result(0), want, f"batched b{bias} {fan}")
            p.close()


def test_remap_known_answer(C):
    # FP32 weights [0,1,0,0] and F = 3: [1, 3, 0] (derived by hand in
    # tests/test_oracle.py); degree bias: the one neighbour with out-edges
    d = 4
    lists = [list(range(1, d + 1))] + [[] for _ in range(d)]
    lists[2] = list(range(d + 1, d + 6))
    lists += [[] for _ in range(5)]
    off = np.concatenate([[0], np.cumsum([len(l) for l in lists])]).astype(np.int64)
    adj = np.array([x for l in lists for x in l], np.int32)
    wt = np.ones(len(adj), np.float32)
    wt[:d] = [0, 1, 0, 0]
    g = C.Graph.from_csr(off, adj, wt)
    seeds = np.zeros(50, np.int32)
    exp = np.tile(np.array([1, 3, 0], np.int32) + 1, 50)
    for bias in (1, 2):
        assert np.array_equal(C.sample(g, seeds, [3], bias=bias, dedup=2)["dst"], exp), bias
        for mode in (C.MODE_PLAIN, C.MODE_BATCHED):
            p = C.Plan(g, [{"seeds": seeds, "fanouts": [3], "bias": bias, "dedup": 2, "request": 0}], mode=mode,
                       cdf_cache=True)
            p.run()
            assert np.array_equal(p.result(0)["dst"], exp), (bias, mode)
            p.close()


def test_remap_streaming_equals_in_memory(C, O):
    hg = C.HostGraph.rmat(16, 8)
    off, adj = hg.csr_views()
    g = C.Graph.from_csr(off, adj)
    seeds = C.seeds_v(hg.V, 2000)
    want = C.sample(g, seeds, [10, 4], bias=C.BIAS_DEGREE, dedup=2)
    og = O.OracleGraph(off, adj)
    same(want, O.sample(og, seeds, [10, 4], bias=O.BIAS_DEGREE, dedup=2), "in memory vs oracle")
    for mode in (C.MODE_PLAIN, C.MODE_BATCHED):
        r = C.stream_sample(hg, seeds, [10, 4], bias=C.BIAS_DEGREE, dedup=2, budget_bytes=64 << 20, partitions=4,
                            mode=mode, bin_min_entries=1)
        for k in KEYS:
            assert np.array_equal(r[k], want[k]), (mode, k)


def test_remap_rejects_bad_dedup(C, g13):
    g, _ = g13
    with pytest.raises(C.CsawError):
        C.sample(g, g.seeds(4), [2], dedup=3)
