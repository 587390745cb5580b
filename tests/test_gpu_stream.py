"""Streaming sampler: partitions streamed through a budgeted device must give
exactly the in-memory (and oracle) edge lists."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "c-saw_amd"))
KEYS = ("src", "dst", "smp", "hop_off")


def same(a, b, what=""):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"


@pytest.fixture(scope="module")
def graphs(C, O):
    g = C.Graph.rmat(16, 8)
    off, adj = g.csr()
    return g, C.HostGraph.from_csr(off, adj), O.OracleGraph(off, adj)


@pytest.mark.parametrize("bias", [0, 1, 3])
@pytest.mark.parametrize("parts", [1, 7, 16])
def test_stream_equals_in_memory(C, graphs, bias, parts):
    g, hg, _ = graphs
    seeds = g.seeds(2000, request=bias)
    want = C.sample(g, seeds, [15, 10, 5], bias=bias)
    got = C.stream_sample(hg, seeds, [15, 10, 5], bias=bias, partitions=parts, budget_bytes=1 << 30)
    same(got, want, f"bias {bias} parts {parts}")
    assert got["stats"].partitions == parts
    assert got["stats"].peak_device_bytes <= 1 << 30


def _schedules(C):
    # bin_min_entries=1: every partition is binned, however small
    yield "plain", dict(mode=C.MODE_PLAIN)
    yield "heuristic", dict(mode=C.MODE_BATCHED, sched=C.SCHED_HEURISTIC, bin_min_entries=1)
    for v in range(C.NUM_VARIANTS):
        yield f"forced{v}", dict(mode=C.MODE_BATCHED, sched=C.SCHED_FORCED, variant=v, bin_min_entries=1)
    yield "random", dict(mode=C.MODE_BATCHED, sched=C.SCHED_RANDOM, explore_key=5, bin_min_entries=1)
    # an untrained model picks arbitrary valid variants
    yield "mlp", dict(mode=C.MODE_BATCHED, sched=C.SCHED_MLP, mlp=C.Mlp(3), bin_min_entries=1)
    # the default threshold: small partitions unbinned
    yield "heuristic-default", dict(mode=C.MODE_BATCHED, sched=C.SCHED_HEURISTIC)


@pytest.mark.parametrize("bias", [0, 1, 3])
def test_stream_schedules_equal_in_memory(C, graphs, bias):
    g, hg, _ = graphs
    seeds = g.seeds(1500, request=10 + bias)
    for dedup in (1, 0):
        want = C.sample(g, seeds, [15, 10, 5], bias=bias, dedup=dedup, request=2)
        for name, kw in _schedules(C):
            got = C.stream_sample(hg, seeds, [15, 10, 5], bias=bias, dedup=dedup, request=2, partitions=6,
                                  budget_bytes=1 << 30, **kw)
            same(got, want, f"bias {bias} dedup {dedup} {name}")
            st = got["stats"]
            vc = list(st.variant_chunks)
            if name == "plain":
                assert st.chunks == 0 and sum(vc) == 0
            elif name != "heuristic-default":
                assert st.chunks > 0 and sum(vc) <= st.chunks
            if name == "forced0" and bias == 0:
                # THREAD is valid for every uniform chunk with fanout <= 32
                assert vc[0] == sum(vc)
            if name in ("forced3", "forced4") and bias == 1:
                assert vc[int(name[-1])] == sum(vc)


def _star(hub_deg, extra, key):
    """vertex 0 has hub_deg neighbours with 0..extra-1 further leaves each."""
    rng = np.random.default_rng(key)
    src, dst = [], []
    nxt = hub_deg + 1
    for i in range(1, hub_deg + 1):
        src += [0, i]
        dst += [i, 0]
        for _ in range(int(rng.integers(0, extra))):
            src += [i, nxt]
            dst += [nxt, i]
            nxt += 1
    src, dst = np.array(src, np.int64), np.array(dst, np.int32)
    o = np.lexsort((dst, src))
    src, dst = src[o], dst[o]
    off = np.zeros(nxt + 1, np.int64)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off), dst


@pytest.mark.parametrize("hub_deg", [1500, 9000, 20000])
def test_stream_hub_schedules(C, hub_deg):
    # integer-bias hubs above the workgroup CDF limits come from the
    # per-partition hub pool; the hub sits in partition 0 of 3
    off, adj = _star(hub_deg, 4, hub_deg)
    g = C.Graph.from_csr(off, adj)
    hg = C.HostGraph.from_csr(off, adj)
    seeds = np.concatenate([np.zeros(48, np.int32), np.arange(1, 17, dtype=np.int32)])
    for bias in (1, 3):
        want = C.sample(g, seeds, [64, 2], bias=bias)
        for name, kw in _schedules(C):
            got = C.stream_sample(hg, seeds, [64, 2], bias=bias, partitions=3, budget_bytes=1 << 28, **kw)
            same(got, want, f"hub {hub_deg} bias {bias} {name}")


@pytest.mark.parametrize("bias", [0, 1, 3])
def test_stream_slice_policies(C, graphs, bias):
    # copied slices, in-place reads of the pinned host lists, and the
    # automatic per-(hop, partition) choice give the same edges
    g, hg, _ = graphs
    seeds = g.seeds(1200, request=20 + bias)
    want = C.sample(g, seeds, [15, 10, 5], bias=bias, request=4)
    parts = 7
    for pol in (C.SLICE_COPY, C.SLICE_DIRECT, C.SLICE_AUTO):
        for name, kw in (("plain", dict(mode=C.MODE_PLAIN)), ("heuristic", dict(bin_min_entries=1)),
                         ("forced0", dict(sched=C.SCHED_FORCED, variant=0, bin_min_entries=1)),
                         ("forced3", dict(sched=C.SCHED_FORCED, variant=3, bin_min_entries=1))):
            got = C.stream_sample(hg, seeds, [15, 10, 5], bias=bias, request=4, partitions=parts,
                                  budget_bytes=1 << 30, slice_policy=pol, **kw)
            same(got, want, f"bias {bias} policy {pol} {name}")
            st = got["stats"]
            if pol == C.SLICE_COPY:
                assert st.direct_partitions == 0
                # hops alternate direction: the next hop starts on the two slices left
                assert st.reused_slices >= 2
            if pol == C.SLICE_DIRECT:
                assert st.h2d_bytes == 0 and st.reused_slices == 0
                assert st.direct_partitions >= parts


def test_stream_auto_reads_sparse_hop_in_place(C):
    # 4 seeds on a 2^15-vertex graph: the first hop touches a few lists of
    # each partition, so no slice is copied for it
    g = C.Graph.rmat(15, 16)
    off, adj = g.csr()
    hg = C.HostGraph.from_csr(off, adj)
    seeds = g.seeds(4, request=1)
    got = C.stream_sample(hg, seeds, [3], bias=C.BIAS_DEGREE, partitions=8, budget_bytes=1 << 28)
    same(got, C.sample(g, seeds, [3], bias=C.BIAS_DEGREE))
    assert got["stats"].h2d_bytes == 0 and got["stats"].direct_partitions >= 1


def test_stream_matches_oracle(C, O, graphs):
    g, hg, og = graphs
    seeds = g.seeds(500, request=3)
    want = O.sample(og, seeds, [8, 4], bias=O.BIAS_DEGREE, request=3)
    got = C.stream_sample(hg, seeds, [8, 4], bias=C.BIAS_DEGREE, request=3, partitions=5, budget_bytes=1 << 30)
    same(got, want)


def test_stream_rmat_host_graph(C, O):
    hg = C.HostGraph.rmat(12, 8)
    off, adj = hg.csr_views()
    off_o, adj_o = O.rmat(12, 8)
    assert np.array_equal(off, off_o) and np.array_equal(adj, adj_o)


def test_stream_budget_too_small(C, graphs):
    g, hg, _ = graphs
    with pytest.raises(C.CsawError) as ei:
        C.stream_sample(hg, g.seeds(100), [5, 5], budget_bytes=1 << 20, partitions=4)
    assert ei.value.status == 4


def test_stream_rejects_bad_opts(C, graphs):
    g, hg, _ = graphs
    for kw in (dict(mode=5), dict(sched=9), dict(sched=C.SCHED_FORCED, variant=C.NUM_VARIANTS),
               dict(sched=C.SCHED_MLP, mlp=None), dict(slice_policy=7)):
        with pytest.raises(C.CsawError) as ei:
            C.stream_sample(hg, g.seeds(10), [5], budget_bytes=1 << 28, **kw)
        assert ei.value.status == 1


def test_stream_rejects_weights(C, graphs):
    g, hg, _ = graphs
    with pytest.raises(C.CsawError) as ei:
        C.stream_sample(hg, g.seeds(10), [5], bias=C.BIAS_WEIGHT)
    assert ei.value.status == 8


def test_stream_chunk_log(C, graphs):
    g, hg, _ = graphs
    seeds = g.seeds(800, request=5)
    for v in (0, 3):
        r = C.stream_sample(hg, seeds, [15, 10], bias=C.BIAS_DEGREE, partitions=4, budget_bytes=1 << 30,
                            sched=C.SCHED_FORCED, variant=v, log_cap=1 << 16, bin_min_entries=1)
        X, var, lt, keys = r["log"]
        st = r["stats"]
        assert 0 < len(keys) <= st.chunks
        assert X.shape == (len(keys), 16) and np.all(np.isfinite(X)) and np.all(lt >= 0)
        assert len(np.unique(keys)) == len(keys)
        assert set((keys >> 40).tolist()) <= {0, 1} and set(((keys >> 32) & 0xff).tolist()) <= {0, 1, 2, 3}
        # degree bias with every drawing vertex's CDF built: both variants valid for every chunk
        assert np.all(var == v)


def test_stream_learned_scheduler_beats_heuristic(C):
    import cost_model as CM
    hg = C.HostGraph.rmat(18, 16)
    fan = [15, 10, 5]
    train = CM.collect_stream(C, hg, C.seeds_v(hg.V, 4096, key=7), fan, C.BIAS_DEGREE, partitions=8,
                              budget_bytes=1 << 30)
    test = CM.collect_stream(C, hg, C.seeds_v(hg.V, 4096, key=11), fan, C.BIAS_DEGREE, partitions=8,
                             budget_bytes=1 << 30)
    mlp, loss = CM.train(C, train)
    r = CM.regret(test, mlp)
    print("streaming held-out regret", r)
    assert r["chunks"] > 100
    assert r["mlp"] < r["heuristic"], r


@pytest.mark.parametrize("bias", [0, 1, 3])
def test_stream_packed_slices(C, graphs, bias):
    # slices copied packed and unpacked on the device give the same edges as
    # unpacked copies and as the in-memory sampler
    g, _, _ = graphs
    off, adj = g.csr()
    hg = C.HostGraph.from_csr(off, adj)
    packed = hg.pack(threads=4)
    assert 0 < packed < 4 * len(adj)
    seeds = g.seeds(1500, request=30 + bias)
    want = C.sample(g, seeds, [15, 10, 5], bias=bias, request=6)
    for raw in (False, True):
        for name, kw in (("plain", dict(mode=C.MODE_PLAIN)), ("heuristic", dict(bin_min_entries=1)),
                         ("forced1", dict(sched=C.SCHED_FORCED, variant=1, bin_min_entries=1))):
            got = C.stream_sample(hg, seeds, [15, 10, 5], bias=bias, request=6, partitions=5, budget_bytes=1 << 30,
                                  slice_policy=C.SLICE_COPY, raw_slices=raw, **kw)
            same(got, want, f"bias {bias} raw {raw} {name}")
            st = got["stats"]
            if raw:
                assert st.h2d_bytes == st.h2d_raw_bytes
            else:
                assert 0 < st.h2d_bytes < st.h2d_raw_bytes


def test_stream_packed_multi_piece(C):
    # partitions of ~16M entries: each slice crosses in two staged pieces
    g = C.Graph.rmat(20, 16)
    off, adj = g.csr()
    hg = C.HostGraph.from_csr(off, adj)
    hg.pack()
    seeds = g.seeds(3000, request=9)
    want = C.sample(g, seeds, [10, 4], bias=C.BIAS_DEGREE, request=9)
    got = C.stream_sample(hg, seeds, [10, 4], bias=C.BIAS_DEGREE, request=9, partitions=2, budget_bytes=2 << 30,
                          slice_policy=C.SLICE_COPY)
    same(got, want, "two pieces per slice")
    assert got["stats"].h2d_bytes < got["stats"].h2d_raw_bytes
