"""The autotuner (csaw_autotune): the shapes it returns are within the
documented bounds, and a plan built with them emits the same edges as the
default plan and the oracle; out-of-range shapes are rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
KEYS = ("src", "dst", "smp", "hop_off")


def test_autotune_shapes_and_parity(C, O):
    g = C.Graph.rmat(13, 8, weights=True)
    off, adj, w = g.csr(weights=True)
    og = O.OracleGraph(off, adj, weights=w)
    specs = []
    for i, r in enumerate(C.request_mix(12, 16, 512, 5)):
        r = dict(r)
        r["seeds"] = g.seeds(r["samples"], request=i)
        specs.append(r)
    t, ms = C.autotune(g, specs, cdf_cache=True)
    assert ms > 0
    for k in ("count_per_cu", "expand_per_cu", "thread_per_cu", "block_per_cu"):
        assert getattr(t, k) in (1, 2, 4, 8, 16), (k, getattr(t, k))
    assert t.chunk_min in (8, 16, 32, 64, 128, 256) and t.group_min in (4, 8, 16), (t.chunk_min, t.group_min)
    tuned = C.Plan(g, specs, mode=C.MODE_BATCHED, use_graph=True, cdf_cache=True, tuning=t)
    default = C.Plan(g, specs, mode=C.MODE_BATCHED, use_graph=True, cdf_cache=True)
    tuned.run()
    default.run()
    for i, s in enumerate(specs):
        want = O.sample(og, s["seeds"], s["fanouts"], bias=s["bias"], key=s["key"], request=s["request"])
        a, b = tuned.result(i), default.result(i)
        for k in KEYS:
            assert np.array_equal(a[k], want[k]) and np.array_equal(b[k], want[k]), (i, k)


@pytest.mark.parametrize("chunk_min,group_min", [(8, 4), (1024, 32), (64, 8), (16, 16)])
def test_tile_shapes_keep_the_edges(C, O, chunk_min, group_min):
    g = C.Graph.rmat(12, 8, weights=True)
    off, adj, w = g.csr(weights=True)
    og = O.OracleGraph(off, adj, weights=w)
    specs = [{"seeds": g.seeds(200, request=i), "fanouts": f, "bias": b, "request": i}
             for i, (f, b) in enumerate([([5, 3], 0), ([10, 2], 1), ([3, 3, 2], 2), ([25], 3)])]
    t = C.Tuning(chunk_min=chunk_min, group_min=group_min)
    for mode in (C.MODE_PLAIN, C.MODE_BATCHED):
        p = C.Plan(g, specs, mode=mode, cdf_cache=True, tuning=t, sched=C.SCHED_RANDOM, explore_key=2)
        p.run()
        for i, s in enumerate(specs):
            want = O.sample(og, s["seeds"], s["fanouts"], bias=s["bias"] if s["bias"] != 3 else O.BIAS_USER,
                            request=i)
            got = p.result(i)
            for k in KEYS:
                assert np.array_equal(got[k], want[k]), (mode, i, k)
        p.close()


def test_tile_shapes_rejected_out_of_range(C):
    g = C.Graph.rmat(8, 4)
    specs = [{"seeds": g.seeds(4), "fanouts": [2], "bias": 0, "request": 0}]
    for bad in (dict(chunk_min=4), dict(chunk_min=24), dict(chunk_min=2048), dict(group_min=2),
                dict(group_min=33), dict(expand_per_cu=-1)):
        with pytest.raises(C.CsawError):
            C.Plan(g, specs, tuning=C.Tuning(**bad))
