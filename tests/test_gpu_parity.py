"""GPU parity: the HIP kernels against the CPU oracle, bit for bit.

Every variant (thread / wave / workgroup granularity, bitmap / list taken
set), the plain and batched executors and the hipGraph replay must emit the
oracle's edge lists exactly (integer and index work: bit-exact)."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
KEYS = ("src", "dst", "smp", "hop_off")


def same(a, b, what=""):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"


@pytest.fixture(scope="module")
def g14(C, O):
    g = C.Graph.rmat(14, 8, weights=True)
    off, adj, w = g.csr(weights=True)
    return g, O.OracleGraph(off, adj, weights=w), off, adj, w


def test_rmat_matches_oracle(C, O):
    for scale, ef in ((10, 4), (13, 8)):
        g = C.Graph.rmat(scale, ef, weights=True)
        off, adj, w = g.csr(weights=True)
        off_o, adj_o = O.rmat(scale, ef)
        assert np.array_equal(off, off_o) and np.array_equal(adj, adj_o)
        assert np.array_equal(w, O.weights(len(adj)))
        assert np.array_equal(g.seeds(500, request=2), O.seeds(g.V, 500, request=2))


def test_rmat16_digest_on_gpu(C):
    g = C.Graph.rmat(16, 8)
    off, adj = g.csr()
    want = json.load(open(os.path.join(HERE, "golden", "rmat16.json")))
    assert hashlib.sha256(off.tobytes() + adj.tobytes()).hexdigest() == want["sha256"]
    assert g.max_degree == want["max_degree"]


def test_config1_plain_uniform(C, O):
    # RMAT-16 ef 8; 1000 samples; FrontierSize 1, NeighborSize 2, Depth 2; uniform; bitmap; seed 42
    g = C.Graph.rmat(16, 8)
    off, adj = g.csr()
    og = O.OracleGraph(off, adj)
    seeds = g.seeds(1000)
    want = O.sample(og, seeds, [2, 2], bias=O.BIAS_UNIFORM)
    same(C.sample(g, seeds, [2, 2], bias=C.BIAS_UNIFORM), want, "plain")
    same(C.sample(g, seeds, [2, 2], bias=C.BIAS_UNIFORM, sched=C.SCHED_FORCED, variant=C.VAR_WAVE_BITMAP), want,
         "wave bitmap")


@pytest.mark.parametrize("bias", [0, 1, 2, 3])
def test_all_variants_match_oracle(C, O, g14, bias):
    g, og, off, adj, w = g14
    seeds = g.seeds(700, request=bias)
    obias = bias if bias != 3 else O.BIAS_USER
    want = O.sample(og, seeds, [12, 6, 3], bias=obias)
    same(C.sample(g, seeds, [12, 6, 3], bias=bias), want, "plain")
    for v in range(C.NUM_VARIANTS):
        same(C.sample(g, seeds, [12, 6, 3], bias=bias, sched=C.SCHED_FORCED, variant=v), want, f"variant {v}")
    same(C.sample(g, seeds, [12, 6, 3], bias=bias, sched=C.SCHED_RANDOM), want, "random")


@pytest.mark.parametrize("bias", [0, 1, 2])
def test_no_dedup_matches_oracle(C, O, g14, bias):
    g, og, *_ = g14
    seeds = g.seeds(500, request=7)
    want = O.sample(og, seeds, [5, 4], bias=bias, dedup=0)
    same(C.sample(g, seeds, [5, 4], bias=bias, dedup=0), want, "plain")
    for v in range(C.NUM_VARIANTS):
        same(C.sample(g, seeds, [5, 4], bias=bias, dedup=0, sched=C.SCHED_FORCED, variant=v), want, f"v{v}")


def test_config2_degree_bias(C, O):
    # RMAT-20 ef 5; 8192 seeds; fanouts {25,10}; degree bias; no repeats per source
    g = C.Graph.rmat(20, 5)
    off, adj = g.csr()
    og = O.OracleGraph(off, adj)
    seeds = g.seeds(8192)
    want = O.sample(og, seeds, [25, 10], bias=O.BIAS_DEGREE, nthreads=8)
    same(C.sample(g, seeds, [25, 10], bias=C.BIAS_DEGREE), want, "plain")
    same(C.sample(g, seeds, [25, 10], bias=C.BIAS_DEGREE, sched=C.SCHED_RANDOM), want, "random variants")


def star_with_hub(O, hub_deg, extra):
    """vertex 0 has hub_deg neighbours with varied degrees (hub-cache path)."""
    rng = np.random.default_rng(hub_deg)
    edges = []
    nxt = hub_deg + 1
    for i in range(1, hub_deg + 1):
        edges.append((0, i))
        for _ in range(int(rng.integers(0, extra))):
            edges.append((i, nxt))
            nxt += 1
    und = sorted(edges + [(b, a) for a, b in edges])
    V = nxt
    src = np.array([a for a, _ in und], np.int64)
    adj = np.array([b for _, b in und], np.int32)
    off = np.zeros(V + 1, np.int64)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off), adj


@pytest.mark.parametrize("hub_deg", [1500, 9000, 20000])
def test_hub_vertices(C, O, hub_deg):
    off, adj = star_with_hub(O, hub_deg, 4)
    w = O.weights(len(adj), key=9)
    g = C.Graph.from_csr(off, adj, w)
    og = O.OracleGraph(off, adj, weights=w)
    seeds = np.zeros(64, np.int32)
    for bias in (0, 1, 2):
        want = O.sample(og, seeds, [64, 2], bias=bias)
        same(C.sample(g, seeds, [64, 2], bias=bias), want, f"plain b{bias}")
        for v in range(C.NUM_VARIANTS):
            same(C.sample(g, seeds, [64, 2], bias=bias, sched=C.SCHED_FORCED, variant=v), want, f"b{bias} v{v}")


def test_edge_cases(C, O):
    # zero-degree seeds, d == F, d == F + 1, isolated vertices, asymmetric
    # graph with zero-degree neighbours (all-zero degree weights -> uniform)
    off = np.array([0, 3, 5, 5, 6, 7, 7, 9], np.int64)
    adj = np.array([1, 2, 6, 0, 3, 4, 5, 2, 5], np.int32)
    w = np.array([0.5, 0.25, 1.0, 1.0, 0.5, 1.0, 1.0, 0.0, 0.0], np.float32)
    g = C.Graph.from_csr(off, adj, w)
    og = O.OracleGraph(off, adj, weights=w)
    seeds = np.array([0, 1, 2, 3, 4, 5, 6, 0, 6, 2], np.int32)
    for bias in (0, 1, 2):
        for fan in ([1], [2], [3], [2, 2, 2], [64, 1]):
            for dedup in (0, 1):
                want = O.sample(og, seeds, fan, bias=bias, dedup=dedup)
                same(C.sample(g, seeds, fan, bias=bias, dedup=dedup), want, f"{bias} {fan} {dedup}")
                same(C.sample(g, seeds, fan, bias=bias, dedup=dedup, sched=C.SCHED_RANDOM), want, "random")


def test_empty_requests(C, O, g14):
    g, og, *_ = g14
    r = C.sample(g, np.zeros(0, np.int32), [3, 2])
    assert len(r["src"]) == 0 and r["hop_off"].shape == (2, 1)
    r = C.sample(g, g.seeds(10), [])
    assert len(r["src"]) == 0
    r = C.sample(g, g.seeds(10), [0, 3])
    assert len(r["src"]) == 0


def test_invalid_arguments(C, g14):
    g, *_ = g14
    with pytest.raises(C.CsawError):
        C.sample(g, np.array([g.V], np.int32), [2])
    with pytest.raises(C.CsawError):
        C.sample(g, np.array([0], np.int32), [65])
    g2 = C.Graph.rmat(8, 4)
    with pytest.raises(C.CsawError):
        C.sample(g2, np.array([0], np.int32), [2], bias=C.BIAS_WEIGHT)
    with pytest.raises(C.CsawError):
        C.Graph.from_csr(np.array([0, 2, 1], np.int64), np.array([1, 0], np.int32))
    with pytest.raises(C.CsawError):
        C.Graph.from_csr(np.array([0, 1, 2], np.int64), np.array([1, 5], np.int32))
    # request ids share Philox counter word 3 with the domain tag (bits 28-31)
    with pytest.raises(C.CsawError):
        C.random_walk(g, g.seeds(4), 3, bias=0, request=1 << 28)
    with pytest.raises(C.CsawError):
        C.mdrw(g, g.seeds(8), 4, 2, 3, request=1 << 28)
    with pytest.raises(C.CsawError):
        C.sample(g, g.seeds(4), [2], request=1 << 20)


def test_batched_equals_plain_and_oracle(C, O, g14):
    g, og, *_ = g14
    specs = []
    rng = np.random.default_rng(3)
    for i in range(40):
        fan = [25, 10] if i % 2 else [15, 10, 5]
        specs.append({"seeds": g.seeds(int(rng.integers(1, 200)), request=i), "fanouts": fan, "bias": i % 3,
                      "request": i})
    plain = C.Plan(g, specs, mode=C.MODE_PLAIN)
    plain.run()
    for sched in (C.SCHED_HEURISTIC, C.SCHED_RANDOM):
        bat = C.Plan(g, specs, mode=C.MODE_BATCHED, sched=sched, explore_key=11)
        bat.run()
        for i, s in enumerate(specs):
            want = O.sample(og, s["seeds"], s["fanouts"], bias=s["bias"], request=i)
            same(plain.result(i), want, f"plain req {i}")
            same(bat.result(i), want, f"batched req {i}")


def test_small_hops_unbinned(C, O, g14):
    # batched plans run hops of at most 65,536 frontier entries without
    # binning (one whole-hop kernel per bias kind, each skipping the other
    # kinds' entries): mixed kinds, the user hooks, region remap, with and
    # without the CDF cache, heuristic and learned schedules, graph replays
    import os, sys
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "c-saw_amd"))
    import cost_model as CM
    mlp, _ = CM.load(C, CM.DEFAULT_MODEL)
    g, og, *_ = g14
    specs = [{"seeds": g.seeds(300, request=i), "fanouts": [8, 4], "bias": i % 4, "dedup": 2 if i == 5 else 1,
              "request": i} for i in range(8)]
    want = [O.sample(og, s["seeds"], s["fanouts"], bias=s["bias"] if s["bias"] != 3 else O.BIAS_USER,
                     dedup=s["dedup"], request=i) for i, s in enumerate(specs)]
    for cache in (False, True):
        for sched in (C.SCHED_HEURISTIC, C.SCHED_MLP):
            p = C.Plan(g, specs, mode=C.MODE_BATCHED, cdf_cache=cache, use_graph=True, sched=sched, mlp=mlp)
            for rep in range(2):
                p.run()
                for i in range(len(specs)):
                    same(p.result(i), want[i], f"cache {cache} sched {sched} replay {rep} req {i}")
            p.close()


@pytest.mark.parametrize("cdf_cache,serial", [(False, False), (True, False), (True, True)])
def test_hipgraph_replays(C, O, g14, cdf_cache, serial):
    # the captured hops fork each hop's variant kernels onto side streams
    # (unless serial) and join them; every replay must equal the oracle
    g, og, *_ = g14
    specs = [{"seeds": g.seeds(300, request=i), "fanouts": [10, 5] if i % 2 else [6, 4, 3], "bias": i % 3,
              "request": i} for i in range(6)]
    want = [O.sample(og, s["seeds"], s["fanouts"], bias=s["bias"], request=i) for i, s in enumerate(specs)]
    p = C.Plan(g, specs, mode=C.MODE_BATCHED, use_graph=True, cdf_cache=cdf_cache, serial_expand=serial,
               sched=C.SCHED_RANDOM, explore_key=3)
    for rep in range(4):
        p.run()
        for i in range(len(specs)):
            same(p.result(i), want[i], f"replay {rep} req {i}")
        # eager work between replays
        C.sample(g, specs[0]["seeds"], [3])


def _cache_specs(seeds_of, fans, biases, dedups):
    specs = []
    for b in biases:
        for f in fans:
            for dd in dedups:
                specs.append({"seeds": seeds_of(len(specs)), "fanouts": f, "bias": b, "dedup": dd,
                              "request": len(specs)})
    return specs


def _check_cache_plans(C, O, g, og, specs, what, cdf_cache=True):
    """Every forced variant, the heuristic and random schedules, and the
    plain (one request at a time) sampler, with the CDF cache (index
    searches, sub-wave groups) against the oracle."""
    want = [O.sample(og, s["seeds"], s["fanouts"], bias=s["bias"] if s["bias"] != 3 else O.BIAS_USER,
                     dedup=s["dedup"], request=s["request"]) for s in specs]
    runs = [(C.MODE_BATCHED, C.SCHED_FORCED, v) for v in range(C.NUM_VARIANTS)]
    runs += [(C.MODE_BATCHED, C.SCHED_HEURISTIC, 0), (C.MODE_BATCHED, C.SCHED_RANDOM, 0),
             (C.MODE_PLAIN, C.SCHED_HEURISTIC, 0)]
    for mode, sched, v in runs:
        p = C.Plan(g, specs, mode=mode, sched=sched, variant=v, cdf_cache=cdf_cache, explore_key=5)
        p.run()
        for i, s in enumerate(specs):
            same(p.result(i), want[i],
                 f"{what} mode {mode} sched {sched} variant {v} req {i} {s['fanouts']} b{s['bias']}")
        p.close()


def test_cdf_cache_variants_match_oracle(C, O, g14):
    # fanouts spanning the group sizes 8 / 16 / 32 and the whole-wave path
    g, og, *_ = g14
    fans = [[25, 10], [15, 10, 5], [3, 2], [40, 2], [64]]
    specs = _cache_specs(lambda i: g.seeds(150, request=i), fans, [0, 1, 2, 3], [1, 0])
    _check_cache_plans(C, O, g, og, specs, "rmat14")


def test_cdf_cache_collisions_and_hubs(C, O):
    # d == F + 1 and small d force repeated draws (ordered slow path, cyclic
    # fallback); the hub has degree 20000 (multi-level index search)
    off, adj = star_with_hub(O, 20000, 4)
    w = O.weights(len(adj), key=9)
    g = C.Graph.from_csr(off, adj, w)
    og = O.OracleGraph(off, adj, weights=w)
    rng = np.random.default_rng(4)
    pool = np.concatenate([np.zeros(40, np.int32), rng.integers(1, len(off) - 1, 200).astype(np.int32)])
    fans = [[2, 2], [5, 3], [7, 9], [31, 4], [64, 1]]
    specs = _cache_specs(lambda i: np.roll(pool, i)[:60], fans, [0, 1, 2, 3], [1])
    _check_cache_plans(C, O, g, og, specs, "hub")


def test_cdf_cache_edge_cases(C, O):
    off = np.array([0, 3, 5, 5, 6, 7, 7, 9], np.int64)
    adj = np.array([1, 2, 6, 0, 3, 4, 5, 2, 5], np.int32)
    w = np.array([0.5, 0.25, 1.0, 1.0, 0.5, 1.0, 1.0, 0.0, 0.0], np.float32)
    g = C.Graph.from_csr(off, adj, w)
    og = O.OracleGraph(off, adj, weights=w)
    seeds = np.array([0, 1, 2, 3, 4, 5, 6, 0, 6, 2], np.int32)
    specs = _cache_specs(lambda i: seeds, [[1], [2], [3], [2, 2, 2], [64, 1]], [0, 1, 2], [0, 1])
    _check_cache_plans(C, O, g, og, specs, "edge")


@pytest.mark.parametrize("w,F,want", [([0, 1, 0, 0], 3, [1, 2, 3]), ([0, 0, 0, 1], 3, [3, 0, 1]),
                                      ([0, 0, 0, 0, 0, 1, 0], 4, [5, 6, 0, 1])])
def test_fallback_known_answers(C, O, w, F, want):
    # one positive weight at index p: slot 0 takes p, every later slot misses
    # its 400 attempts and takes the first untaken index cyclically above p
    d = len(w)
    p = w.index(1)
    lists = [list(range(1, d + 1))] + [[] for _ in range(d)]
    lists[1 + p] = list(range(d + 1, d + 6))  # degree bias: the only neighbour with out-edges
    lists += [[] for _ in range(5)]
    off = np.concatenate([[0], np.cumsum([len(l) for l in lists])]).astype(np.int64)
    adj = np.array([x for l in lists for x in l], np.int32)
    wt = np.zeros(len(adj), np.float32)
    wt[:d] = w
    wt[d:] = 1.0
    g = C.Graph.from_csr(off, adj, wt)
    seeds = np.zeros(100, np.int32)
    exp = np.tile(np.array(want, np.int32) + 1, 100)
    for bias in (1, 2):
        assert np.array_equal(C.sample(g, seeds, [F], bias=bias)["dst"], exp), ("plain", bias)
        for v in range(C.NUM_VARIANTS):
            assert np.array_equal(C.sample(g, seeds, [F], bias=bias, sched=C.SCHED_FORCED, variant=v)["dst"], exp)
        for mode in (C.MODE_PLAIN, C.MODE_BATCHED):
            pl = C.Plan(g, [{"seeds": seeds, "fanouts": [F], "bias": bias, "request": 0}], mode=mode, cdf_cache=True)
            pl.run()
            assert np.array_equal(pl.result(0)["dst"], exp), ("cache", mode, bias)
            pl.close()


@pytest.mark.parametrize("bias", [0, 1, 2, 3])
def test_walk_matches_oracle(C, O, g14, bias):
    g, og, *_ = g14
    seeds = g.seeds(3000, request=5)
    W, steps = C.random_walk(g, seeds, 40, bias=bias)
    Wo, so = O.walk(og, seeds, 40, bias=bias)
    assert np.array_equal(W, Wo) and steps == so


@pytest.mark.parametrize("weights", ["random", "one_big"])
def test_fp32_cdf_of_a_hub(C, O, weights):
    # a star whose centre has 5,000 neighbours (a hub: its FP32 CDF is folded
    # by one wave, the leaves' by one thread each); with one big weight
    # first, 1e8 + 1 rounds back to 1e8 at every step of a left-to-right
    # fold, so any other summation order changes the total and the draws
    d = 5000
    off = np.concatenate([[0], d + np.arange(0, d + 1)]).astype(np.int64)
    adj = np.concatenate([np.arange(1, d + 1), np.zeros(d)]).astype(np.int32)
    rng = np.random.default_rng(3)
    w = rng.uniform(1e-3, 1.0, 2 * d).astype(np.float32)
    if weights == "one_big":
        w[:d] = 1.0
        w[0] = 1e8
    g = C.Graph.from_csr(off, adj, w)
    og = O.OracleGraph(off, adj, weights=w)
    seeds = np.zeros(20000, np.int32)
    W, steps = C.random_walk(g, seeds, 2, bias=2)
    Wo, so = O.walk(og, seeds, 2, bias=2)
    assert np.array_equal(W, Wo) and steps == so
    if weights == "one_big":
        assert (W[:, 1] == 1).all()  # C[0] = 1e8 = the total: every draw picks neighbour 0
    got = C.sample(g, seeds[:300], [20], bias=2)
    want = O.sample(og, seeds[:300], [20], bias=2)
    same(got, want, "hub fp32")


def test_walk_more_walkers_than_lanes(C, O, g14):
    # 600,000 walkers outnumber the lanes of one launch (256 CUs x 8 blocks
    # x 256), so blocks take a second slice of walkers; a length of 37
    # leaves a partial tile of staged steps; lengths 0 and 1 and no walkers
    g, og, *_ = g14
    seeds = g.seeds(600000, request=6)
    W, steps = C.random_walk(g, seeds, 37, bias=2)
    Wo, so = O.walk(og, seeds, 37, bias=2, nthreads=8)
    assert np.array_equal(W, Wo) and steps == so
    assert (W[:, 1:] < 0).any() and (W[:, -1] >= 0).any()  # both ended and full walks
    for length in (0, 1):
        W, steps = C.random_walk(g, seeds[:5000], length, bias=1)
        Wo, so = O.walk(og, seeds[:5000], length, bias=1)
        assert np.array_equal(W, Wo) and steps == so
    W, steps = C.random_walk(g, seeds[:0], 10, bias=0)
    assert W.shape == (0, 11) and steps == 0


def test_mdrw_matches_oracle(C, O, g14):
    g, og, *_ = g14
    S, FS, D = 64, 16, 50
    seeds = g.seeds(S * FS, request=9)
    so, do, oo = O.mdrw(og, seeds, S, FS, D)
    src = np.zeros(S * D, np.int32)
    dst = np.zeros(S * D, np.int32)
    off = np.zeros(S + 1, np.int64)
    st = C.Stats()
    rc = C.lib().csaw_mdrw(g.h, 42, 0, S, FS, D, C._p(seeds), C._p(src), C._p(dst), C._p(off), C.ctypes.byref(st))
    assert rc == 0
    n = off[-1]
    assert np.array_equal(off, oo) and np.array_equal(src[:n], so) and np.array_equal(dst[:n], do)


@pytest.mark.parametrize("cdf_cache", [False, True])
def test_limits_fanout_64_depth_8(C, O, g14, cdf_cache):
    # the largest fanout (64, a whole wave per slot set) and the deepest
    # request (8 hops), with every variant and the heuristic, beside an empty
    # request and a single-seed one
    g, og, *_ = g14
    specs = [
        {"seeds": g.seeds(40, request=0), "fanouts": [64, 3], "bias": 1, "dedup": 1, "request": 0},
        {"seeds": g.seeds(20, request=1), "fanouts": [3, 2, 2, 2, 2, 2, 2, 2], "bias": 2, "dedup": 1, "request": 1},
        {"seeds": g.seeds(0, request=2), "fanouts": [5, 5], "bias": 0, "dedup": 1, "request": 2},
        {"seeds": g.seeds(1, request=3), "fanouts": [64, 64], "bias": 0, "dedup": 0, "request": 3},
        {"seeds": g.seeds(30, request=4), "fanouts": [2, 64, 2], "bias": 3, "dedup": 1, "request": 4},
    ]
    _check_cache_plans(C, O, g, og, specs, "limits", cdf_cache=cdf_cache)
