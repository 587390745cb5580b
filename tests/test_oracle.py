"""CPU tests that pin the oracle (no GPU).

Pins, by part of the oracle:
  Philox4x32-10      published Random123 known answers + ATen vectors (golden fixture)
  seeds / weights    closed forms of the documented mappings, computed with Python big ints
  integer ITS        numpy.searchsorted (library routine) over numpy.cumsum, exhaustive small cases
  FP32 CDF           numpy float32 cumsum (sequential) and an FP64 sum within d*eps*T
  selection law      brute-force exact successive-sampling probabilities on tiny graphs, chi-square
  multi-hop          structural invariants; sequential == OpenMP restatement
  RMAT               symmetric, sorted, unique, loop-free; committed digest of the scale-16 CSR
  walk               continuity invariants and step-frequency chi-square
  MDRW               exact law of 2-step edge sequences by enumeration, chi-square (and the
                     same data rejecting a uniform-slot law); all-zero-bias frontiers
  reselect fallback  closed-form answers with a single positive weight (FP32 and degree bias)
"""
import hashlib
import json
import os

import numpy as np
import pytest
from scipy import stats

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def t64(o):
    return (int(o[1]) << 32) | int(o[0])


def mulhi(a, b):
    return (a * b) >> 64


# ----------------------------------------------------------------- philox
def test_philox_known_answers(O):
    kat = json.load(open(os.path.join(HERE, "golden", "philox_kat.json")))
    assert len(kat["vectors"]) >= 3
    for v in kat["vectors"]:
        got = ["%08x" % x for x in O.philox(v["counter"], v["key"])]
        assert got == v["out"], v


# ---------------------------------------------------------- seeds / weights
def test_seeds_closed_form(O):
    V = 1000003
    s = O.seeds(V, 64, key=42, request=3)
    for i in range(64):
        o = O.philox([i, 0, 3, 1 << 28], [42, 0])
        assert s[i] == mulhi(t64(o), V)
    assert s.min() >= 0 and s.max() < V


def test_weights_closed_form(O):
    w = O.weights(100, key=42)
    for e in range(100):
        o = O.philox([e, 0, 0, 3 << 28], [42, 0])
        assert w[e] == np.float32(((int(o[0]) >> 8) + 1) / 16777216.0)
    assert (w > 0).all() and (w <= 1).all()


# -------------------------------------------------------------- RMAT
def check_csr(off, adj):
    V = len(off) - 1
    assert off[0] == 0 and off[-1] == len(adj) and (np.diff(off) >= 0).all()
    src = np.repeat(np.arange(V, dtype=np.int64), np.diff(off))
    assert (adj >= 0).all() and (adj < V).all()
    assert (src != adj).all(), "self loop"
    key = src * V + adj
    assert (np.diff(key) > 0).all(), "lists not sorted/unique"
    rev = np.sort(adj.astype(np.int64) * V + src)
    assert np.array_equal(rev, key), "not symmetric"


@pytest.mark.parametrize("scale,ef", [(8, 8), (10, 4), (12, 16)])
def test_rmat_invariants(O, scale, ef):
    off, adj = O.rmat(scale, ef)
    assert len(off) == (1 << scale) + 1
    assert len(adj) <= 2 * ef << scale
    check_csr(off, adj)


def test_rmat16_digest(O):
    off, adj = O.rmat(16, 8)
    check_csr(off, adj)
    got = {"V": len(off) - 1, "E": int(len(adj)), "max_degree": int(np.diff(off).max()),
           "sha256": hashlib.sha256(off.tobytes() + adj.tobytes()).hexdigest()}
    want = json.load(open(os.path.join(HERE, "golden", "rmat16.json")))
    assert got == want


def test_rmat_degree_skew(O):
    # a=.57 concentrates edges on low ids: vertex 0 is the largest hub
    off, adj = O.rmat(14, 8)
    d = np.diff(off)
    assert d[0] == d.max()
    assert d.max() > 20 * d.mean()


# --------------------------------------------------------- ITS closed forms
def star_graph(weights):
    """vertex 0 -> neighbours 1..d; neighbour i gets degree weights[i-1] via leaves."""
    d = len(weights)
    nxt = d + 1
    edges = []
    for i, w in enumerate(weights, start=1):
        edges.append((0, i))
        for _ in range(int(w) - 1):
            edges.append((i, nxt))
            nxt += 1
    und = edges + [(b, a) for a, b in edges]
    V = nxt
    und.sort()
    src = np.array([a for a, _ in und], np.int64)
    adj = np.array([b for _, b in und], np.int32)
    off = np.zeros(V + 1, np.int64)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off), adj


def test_integer_its_matches_searchsorted(O):
    rng = np.random.default_rng(5)
    for trial in range(30):
        d = int(rng.integers(2, 9))
        w = rng.integers(1, 6, size=d)
        off, adj = star_graph(w)
        g = O.OracleGraph(off, adj)
        C = np.cumsum(w.astype(np.uint64))
        for s in range(20):
            idx = O.select_source(g, 0, 1, bias=O.BIAS_DEGREE, dedup=0, sample=s, hop=trial)
            o = O.philox([0, 0, s, (4 << 28) | trial], [42, 0])
            x = mulhi(t64(o), int(C[-1]))
            assert idx[0] == np.searchsorted(C, x, side="right")


def test_uniform_its_closed_form(O):
    off, adj = star_graph([1] * 7)
    g = O.OracleGraph(off, adj)
    for s in range(50):
        idx = O.select_source(g, 0, 3, bias=O.BIAS_UNIFORM, dedup=0, sample=s)
        for j in range(3):
            o = O.philox([j << 16, 0, s, 4 << 28], [42, 0])
            assert idx[j] == mulhi(t64(o), 7)


def test_fp32_cdf_sequential(O):
    off, adj = O.rmat(10, 8)
    w = O.weights(len(adj))
    cdf = O.fp32_cdf(off, w)
    for v in range(0, len(off) - 1, 7):
        a, b = off[v], off[v + 1]
        if a == b:
            continue
        seq = np.cumsum(w[a:b], dtype=np.float32)
        assert np.array_equal(cdf[a:b], seq)
        exact = np.cumsum(w[a:b].astype(np.float64))
        assert np.all(np.abs(cdf[a:b] - exact) <= (b - a) * np.finfo(np.float32).eps * exact[-1])


def test_fp32_its_matches_searchsorted(O):
    off, adj = O.rmat(9, 8)
    w = O.weights(len(adj))
    g = O.OracleGraph(off, adj, weights=w)
    d = np.diff(off)
    for v in np.nonzero(d > 3)[0][:40]:
        c = g.cdf[off[v]:off[v + 1]]
        for s in range(5):
            idx = O.select_source(g, int(v), 1, bias=O.BIAS_WEIGHT, dedup=0, sample=s)
            o = O.philox([0, 0, s, 4 << 28], [42, 0])
            x = np.float32(np.float32((int(o[0]) >> 8) / 16777216.0) * c[-1])
            assert idx[0] == np.searchsorted(c, x, side="right")


# ------------------------------------------- selection law (brute force)
def successive_pair_probs(w):
    T = float(sum(w))
    P = {}
    for a in range(len(w)):
        for b in range(len(w)):
            if a != b:
                P[(a, b)] = w[a] / T * w[b] / (T - w[a])
    return P


@pytest.mark.parametrize("weights", [[1, 2, 3, 4], [5, 1, 1, 2, 1], [1, 1, 1, 1, 1, 1]])
def test_without_replacement_law(O, weights):
    off, adj = star_graph(weights)
    g = O.OracleGraph(off, adj)
    N = 60000
    r = O.sample(g, np.zeros(N, np.int32), [2], bias=O.BIAS_DEGREE)
    assert len(r["dst"]) == 2 * N
    a = r["dst"][0::2] - 1
    b = r["dst"][1::2] - 1
    assert (a != b).all()
    P = successive_pair_probs(weights)
    keys = sorted(P)
    obs = np.array([np.sum((a == i) & (b == j)) for i, j in keys], float)
    exp = np.array([P[k] * N for k in keys])
    assert obs.sum() == N
    assert stats.chisquare(obs, exp).pvalue > 1e-4


def test_with_replacement_law(O):
    w = [3, 1, 2]
    off, adj = star_graph(w)
    g = O.OracleGraph(off, adj)
    N = 30000
    r = O.sample(g, np.zeros(N, np.int32), [2], bias=O.BIAS_DEGREE, dedup=0)
    a = r["dst"][0::2] - 1
    b = r["dst"][1::2] - 1
    T = sum(w)
    obs = np.array([np.sum((a == i) & (b == j)) for i in range(3) for j in range(3)], float)
    exp = np.array([w[i] * w[j] / T / T * N for i in range(3) for j in range(3)])
    assert stats.chisquare(obs, exp).pvalue > 1e-4


def directed_star(d, extra=None):
    """vertex 0 -> 1..d; extra: {neighbour index: out-degree} gives neighbour
    1+i that many leaves of its own (the leaves come after 1..d)."""
    lists = [list(range(1, d + 1))] + [[] for _ in range(d)]
    nxt = d + 1
    for i, k in (extra or {}).items():
        lists[1 + i] = list(range(nxt, nxt + k))
        lists.extend([] for _ in range(k))
        nxt += k
    off = np.concatenate([[0], np.cumsum([len(l) for l in lists])]).astype(np.int64)
    adj = np.array([x for l in lists for x in l], np.int32)
    return off, adj


# Closed-form answers of the reselect fallback.  With a single positive
# weight at index p, every attempt of every slot draws p: slot 0 takes p,
# each later slot misses all 400 attempts and takes the first untaken index
# scanning upward cyclically from p.
@pytest.mark.parametrize("w,F,want", [
    ([0, 1, 0, 0], 3, [1, 2, 3]),
    ([0, 0, 1, 0, 0], 3, [2, 3, 4]),
    ([0, 0, 0, 1], 3, [3, 0, 1]),
    ([0, 0, 0, 0, 0, 1, 0], 4, [5, 6, 0, 1]),
])
def test_fallback_known_answers(O, w, F, want):
    d = len(w)
    off, adj = directed_star(d)
    # FP32 weights: only vertex 0's entries carry weights
    g = O.OracleGraph(off, adj, weights=np.array(w, np.float32))
    for s in range(20):
        assert list(O.select_source(g, 0, F, bias=O.BIAS_WEIGHT, sample=s, pos=s)) == want
    r = O.sample(g, np.zeros(30, np.int32), [F], bias=O.BIAS_WEIGHT)
    assert (r["dst"].reshape(30, F) == np.array(want) + 1).all()
    # integer (degree) bias: the neighbour at index p is the only one with
    # out-neighbours, so its degree is the only positive bias
    p = w.index(1)
    off, adj = directed_star(d, {p: 5})
    g = O.OracleGraph(off, adj)
    r = O.sample(g, np.zeros(30, np.int32), [F], bias=O.BIAS_DEGREE)
    assert (r["dst"].reshape(30, F) == np.array(want) + 1).all()


# Region-remap mode (dedup = 2, heur_normalize).  Known answer, derived by
# hand from the definition in IEEE arithmetic: FP32 weights [0,1,0,0] (p =
# [0,1,1,1]), F = 3.  Slot 0 draws index 1.  Slot 1 collides at 1: A = 0,
# B = 1, lambda = 0/0 = NaN, so r = NaN, NaN < A is false, r = NaN + 1 and
# lb = 1; no p_i exceeds NaN, so the pick clamps to 3, which is free.  Slot
# 2 (lb = 0, hb = 1 again) remaps to 3 (taken); from then on lb = hb = 1:
# lambda = -1/-1 = 1, delta = 1/0 = inf, r = inf, pick 3 (taken), for all
# 400 attempts; the cyclic fallback from 3 takes 0.
def test_remap_known_answer(O):
    off, adj = directed_star(4)
    g = O.OracleGraph(off, adj, weights=np.array([0, 1, 0, 0], np.float32))
    for s in range(10):
        assert list(O.select_source(g, 0, 3, bias=O.BIAS_WEIGHT, dedup=2, sample=s, pos=s)) == [1, 3, 0]
    # same law through the integer path (degree bias, one neighbour with edges)
    off, adj = directed_star(4, {1: 5})
    g = O.OracleGraph(off, adj)
    r = O.sample(g, np.zeros(20, np.int32), [3], bias=O.BIAS_DEGREE, dedup=2)
    assert (r["dst"].reshape(20, 3) == np.array([1, 3, 0]) + 1).all()


@pytest.mark.parametrize("bias", [0, 1, 2])
def test_remap_invariants_and_first_pick(O, bias):
    off, adj = O.rmat(11, 8)
    w = O.weights(len(adj), key=3)
    g = O.OracleGraph(off, adj, weights=w)
    seeds = O.seeds(len(off) - 1, 300)
    r = O.sample(g, seeds, [6, 4], bias=bias, dedup=2)
    check_sample_invariants(off, adj, seeds, [6, 4], r)
    # slot 0 never collides: its pick is the first i with p_i > u, p the
    # normalised CDF, computed here with numpy (searchsorted over float32 of
    # float64 quotients)
    deg = np.diff(off)
    for s in range(50):
        v = int(seeds[s])
        d = int(deg[v])
        if d <= 6:
            continue
        nb = adj[off[v]:off[v + 1]]
        if bias == 0:
            p = (np.arange(1, d + 1, dtype=np.float64) / d).astype(np.float32)
        elif bias == 1:
            C = np.cumsum(deg[nb].astype(np.uint64))
            p = (C.astype(np.float64) / float(C[-1])).astype(np.float32) if C[-1] else None
        else:
            c = np.cumsum(w[off[v]:off[v + 1]], dtype=np.float32)
            p = (c.astype(np.float64) / float(c[-1])).astype(np.float32)
        if p is None:
            continue
        o = O.philox([0, 0, s, (4 << 28) | 0], [42, 0])
        u = np.float32(int(o[0]) >> 8) * np.float32(2.0 ** -24)
        i = min(int(np.searchsorted(p, u, side="right")), d - 1)
        a = r["hop_off"][0, s]
        assert r["dst"][a] == nb[i], (s, v)


def test_fallback_path_distinct(O):
    # hub neighbour with 99.9% of the weight: later slots exhaust the 400
    # attempts and take the cyclic fallback; picks stay distinct, k = F
    off, adj = star_graph([2000, 1, 1])
    g = O.OracleGraph(off, adj)
    r = O.sample(g, np.zeros(500, np.int32), [3], bias=O.BIAS_DEGREE)
    d = r["dst"].reshape(500, 3)
    assert all(len(set(row)) == 3 for row in d)


# ---------------------------------------------------- multi-hop invariants
def check_sample_invariants(off, adj, seeds, fanouts, r, dedup=1):
    S = len(seeds)
    ho = r["hop_off"]
    assert ho.shape == (len(fanouts), S + 1)
    assert ho[0, 0] == 0 and ho[-1, -1] == len(r["src"])
    deg = np.diff(off)
    front = [[int(x)] for x in seeds]
    for h, F in enumerate(fanouts):
        assert h == 0 or ho[h, 0] == ho[h - 1, -1]
        nxt = []
        for s in range(S):
            e0, e1 = ho[h, s], ho[h, s + 1]
            src, dst = r["src"][e0:e1], r["dst"][e0:e1]
            assert (r["smp"][e0:e1] == s).all()
            pos = 0
            for v in front[s]:
                k = 0 if deg[v] == 0 else (min(F, deg[v]) if dedup else F)
                assert (src[pos:pos + k] == v).all()
                nb = dst[pos:pos + k]
                lst = adj[off[v]:off[v + 1]]
                assert np.isin(nb, lst).all()
                if dedup:
                    assert len(np.unique(nb)) == k
                pos += k
            assert pos == e1 - e0
            nxt.append([int(x) for x in dst])
        front = nxt


@pytest.mark.parametrize("bias", [0, 1, 2])
def test_multihop_invariants_and_omp(O, bias):
    off, adj = O.rmat(12, 8)
    w = O.weights(len(adj))
    g = O.OracleGraph(off, adj, weights=w)
    seeds = O.seeds(len(off) - 1, 300)
    r = O.sample(g, seeds, [6, 3, 2], bias=bias)
    check_sample_invariants(off, adj, seeds, [6, 3, 2], r)
    r2 = O.sample(g, seeds, [6, 3, 2], bias=bias, nthreads=4)
    for k in ("src", "dst", "smp", "hop_off"):
        assert np.array_equal(r[k], r2[k])


def test_multihop_no_dedup(O):
    off, adj = O.rmat(11, 8)
    g = O.OracleGraph(off, adj)
    seeds = O.seeds(len(off) - 1, 100)
    r = O.sample(g, seeds, [4, 3], bias=O.BIAS_DEGREE, dedup=0)
    check_sample_invariants(off, adj, seeds, [4, 3], r, dedup=0)


def test_sample_rejects_bad_seed(O):
    off, adj = O.rmat(8, 4)
    g = O.OracleGraph(off, adj)
    with pytest.raises(ValueError):
        O.sample(g, np.array([len(off)], np.int32), [2])


# ------------------------------------------------------------ walks, MDRW
def test_walk_invariants(O):
    off, adj = O.rmat(12, 8)
    w = O.weights(len(adj))
    g = O.OracleGraph(off, adj, weights=w)
    seeds = O.seeds(len(off) - 1, 400)
    for bias in (0, 1, 2):
        W, steps = O.walk(g, seeds, 30, bias=bias)
        assert (W[:, 0] == seeds).all()
        assert steps == int((W[:, 1:] >= 0).sum())
        for row in W[:50]:
            for t in range(30):
                a, b = row[t], row[t + 1]
                if a < 0:
                    assert b < 0
                elif b >= 0:
                    assert b in adj[off[a]:off[a + 1]]
                else:
                    assert off[a + 1] == off[a]
        W2, s2 = O.walk(g, seeds, 30, bias=bias, nthreads=4)
        assert np.array_equal(W, W2) and s2 == steps


def test_walk_step_law(O):
    w_nb = np.array([0.1, 0.2, 0.3, 0.4], np.float32)
    off = np.array([0, 4, 5, 6, 7, 8], np.int64)
    adj = np.array([1, 2, 3, 4, 0, 0, 0, 0], np.int32)
    weights = np.concatenate([w_nb, np.ones(4, np.float32)])
    g = O.OracleGraph(off, adj, weights=weights)
    N = 40000
    W, _ = O.walk(g, np.zeros(N, np.int32), 1, bias=O.BIAS_WEIGHT)
    obs = np.bincount(W[:, 1] - 1, minlength=4).astype(float)
    assert stats.chisquare(obs, w_nb / w_nb.sum() * N).pvalue > 1e-4


# ----------------------------------------------------------- MDRW law (exact)
MDRW_LISTS = [[1, 2], [0, 2, 3, 4], [3], [], [0, 1, 2]]  # degrees 2, 4, 1, 0, 3


def mdrw_law(lists, front, depth, slot_law="bias"):
    """Exact probability of every emitted edge sequence of a multi-dimensional
    walk, by enumeration: per step slot q with probability VertexBias(front[q])
    / sum (VertexBias = out-degree; "uniform" for the wrong law 1/FS), then a
    uniform neighbour nb of u = front[q]; emit (u, nb); front[q] = nb."""
    if depth == 0:
        return {(): 1.0}
    degs = [len(lists[v]) for v in front]
    tot = sum(degs)
    out = {}
    if tot == 0:  # uniform slot, degree 0: the step is consumed, nothing emitted
        for seq, p in mdrw_law(lists, front, depth - 1, slot_law).items():
            out[seq] = out.get(seq, 0.0) + p
        return out
    for q, v in enumerate(front):
        pq = degs[q] / tot if slot_law == "bias" else 1.0 / len(front)
        if degs[q] == 0:
            continue
        for nb in lists[v]:
            nf = list(front)
            nf[q] = nb
            for seq, p in mdrw_law(lists, nf, depth - 1, slot_law).items():
                k = ((v, nb),) + seq
                out[k] = out.get(k, 0.0) + pq / degs[q] * p
    return out


def _chi2_cells(law, counts, N):
    keys = sorted(law)
    exp = np.array([law[k] * N for k in keys])
    obs = np.array([counts.get(k, 0) for k in keys], float)
    assert sum(counts.values()) == N and set(counts) <= set(law)
    small = exp < 5  # pool sparse cells
    if small.any():
        exp = np.append(exp[~small], exp[small].sum())
        obs = np.append(obs[~small], obs[small].sum())
    return stats.chisquare(obs, exp * obs.sum() / exp.sum()).pvalue


def test_mdrw_law_exact(O):
    # slot chosen proportionally to VertexBias (ITS_MDRW, functions.cuh:1014-1060),
    # neighbour uniform (rand_integer, functions.cuh:210-216), slot replaced
    lists = MDRW_LISTS
    off = np.concatenate([[0], np.cumsum([len(l) for l in lists])]).astype(np.int64)
    adj = np.array([x for l in lists for x in l], np.int32)
    g = O.OracleGraph(off, adj)
    front, D, N = [0, 1, 4], 2, 60000
    src, dst, offs = O.mdrw(g, np.tile(np.array(front, np.int32), N), N, len(front), D)
    counts = {}
    for s in range(N):
        e = range(offs[s], offs[s + 1])
        k = tuple((int(src[i]), int(dst[i])) for i in e)
        counts[k] = counts.get(k, 0) + 1
    assert _chi2_cells(mdrw_law(lists, front, D), counts, N) > 1e-4
    # the same data rejects a uniform slot law (the test has the power to
    # catch that mistake)
    wrong = mdrw_law(lists, front, D, slot_law="uniform")
    assert _chi2_cells(wrong, {k: c for k, c in counts.items() if k in wrong}, sum(
        c for k, c in counts.items() if k in wrong)) < 1e-12


def test_mdrw_all_zero_bias_consumes_steps(O):
    # a frontier of degree-0 vertices: uniform slot, no neighbour, no edge
    lists = MDRW_LISTS
    off = np.concatenate([[0], np.cumsum([len(l) for l in lists])]).astype(np.int64)
    adj = np.array([x for l in lists for x in l], np.int32)
    g = O.OracleGraph(off, adj)
    src, dst, offs = O.mdrw(g, np.full(3 * 10, 3, np.int32), 10, 3, 6)
    assert len(src) == 0 and (offs == 0).all()
    # one positive-bias slot among zeros: every step takes it
    src, dst, offs = O.mdrw(g, np.tile(np.array([3, 2, 3], np.int32), 10), 10, 3, 1)
    assert list(src) == [2] * 10 and list(dst) == [3] * 10


def test_mdrw_invariants(O):
    off, adj = O.rmat(11, 8)
    g = O.OracleGraph(off, adj)
    S, FS, D = 50, 8, 20
    seeds = O.seeds(len(off) - 1, S * FS)
    src, dst, offs = O.mdrw(g, seeds, S, FS, D)
    assert offs[0] == 0 and offs[-1] == len(src)
    for s in range(S):
        front = list(seeds[s * FS:(s + 1) * FS])
        for e in range(offs[s], offs[s + 1]):
            assert src[e] in front
            assert dst[e] in adj[off[src[e]]:off[src[e] + 1]]
            front[front.index(src[e])] = dst[e]
