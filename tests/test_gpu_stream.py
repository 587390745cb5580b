"""Streaming sampler: partitions streamed through a budgeted device must give
exactly the in-memory (and oracle) edge lists."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
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


def test_stream_rejects_weights(C, graphs):
    g, hg, _ = graphs
    with pytest.raises(C.CsawError) as ei:
        C.stream_sample(hg, g.seeds(10), [5], bias=C.BIAS_WEIGHT)
    assert ei.value.status == 8
