"""The 64-bit integer CDF cache paths (kept when some per-vertex bias total
reaches 2^32; forced here with CSAW_INT_CDF_KEYS=64 on small graphs, whose
totals fit 32 bits): every variant, the plain sampler and walks over the
uint64 cache and its 16-ary index against the oracle."""
import numpy as np
import pytest

from test_gpu_parity import _cache_specs, _check_cache_plans, star_with_hub

pytestmark = pytest.mark.gpu


@pytest.fixture
def wide(monkeypatch):
    monkeypatch.setenv("CSAW_INT_CDF_KEYS", "64")  # read when a graph is created


def test_wide_cache_variants(C, O, wide):
    g = C.Graph.rmat(13, 8, weights=True)
    off, adj, w = g.csr(weights=True)
    og = O.OracleGraph(off, adj, weights=w)
    fans = [[25, 10], [15, 10, 5], [3, 2], [40, 2]]
    specs = _cache_specs(lambda i: g.seeds(120, request=i), fans, [1, 3], [1, 0])
    _check_cache_plans(C, O, g, og, specs, "wide")
    seeds = g.seeds(2000, request=5)
    for bias in (1, 3):
        W, steps = C.random_walk(g, seeds, 30, bias=bias)
        Wo, so = O.walk(og, seeds, 30, bias=bias)
        assert np.array_equal(W, Wo) and steps == so


def test_wide_cache_hub(C, O, wide):
    off, adj = star_with_hub(O, 9000, 4)
    g = C.Graph.from_csr(off, adj)
    og = O.OracleGraph(off, adj)
    rng = np.random.default_rng(2)
    pool = np.concatenate([np.zeros(30, np.int32), rng.integers(1, len(off) - 1, 100).astype(np.int32)])
    specs = _cache_specs(lambda i: np.roll(pool, i)[:50], [[2, 2], [7, 9], [64, 1]], [1, 3], [1])
    _check_cache_plans(C, O, g, og, specs, "wide hub")
