"""The reference's on-disk CSR read back onto the GPU (csaw_graph_load_bin,
the counterpart of non-stream/graph.hpp:18-74): the device CSR equals the
files as numpy reads them, with and without a weight file, and sampling and
walks over the loaded graph equal the oracle run over the numpy arrays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
KEYS = ("src", "dst", "smp", "hop_off")


def write_files(C, tmp_path, off, adj, w=None, extra_ids=0):
    beg, csr = str(tmp_path / "beg.bin"), str(tmp_path / "csr.bin")
    C.write_bin(beg, csr, off, adj)
    if extra_ids:  # the reader takes E = beg[V] ids and ignores what follows (graph.hpp:31, 56)
        with open(csr, "ab") as f:
            np.arange(extra_ids, dtype=np.int64).tofile(f)
    wf = None
    if w is not None:
        wf = str(tmp_path / "weight.bin")
        np.asarray(w, np.float32).tofile(wf)
    return beg, csr, wf


@pytest.mark.parametrize("with_weights,extra", [(True, 0), (False, 0), (True, 17)])
def test_load_bin_equals_files(C, tmp_path, with_weights, extra):
    g0 = C.Graph.rmat(12, 8, weights=True, key=5)
    off, adj, w = g0.csr(weights=True)
    beg, csr, wf = write_files(C, tmp_path, off, adj, w if with_weights else None, extra)
    f_off = np.fromfile(beg, np.int64)
    f_ids = np.fromfile(csr, np.int64)[:f_off[-1]]
    g = C.Graph.load_bin(beg, csr, wf)
    assert (g.V, g.E) == (len(f_off) - 1, int(f_off[-1]))
    if with_weights:
        l_off, l_adj, l_w = g.csr(weights=True)
        assert np.array_equal(l_w, np.fromfile(wf, np.float32))
    else:
        l_off, l_adj = g.csr()
        with pytest.raises(C.CsawError):
            g.csr(weights=True)
    assert np.array_equal(l_off, f_off) and np.array_equal(l_adj, f_ids.astype(np.int32))


def test_sampling_over_loaded_graph_matches_oracle(C, O, tmp_path):
    g0 = C.Graph.rmat(13, 8, weights=True, key=6)
    off, adj, w = g0.csr(weights=True)
    beg, csr, wf = write_files(C, tmp_path, off, adj, w)
    g = C.Graph.load_bin(beg, csr, wf)
    # the oracle reads the files itself
    f_off = np.fromfile(beg, np.int64)
    og = O.OracleGraph(f_off, np.fromfile(csr, np.int64).astype(np.int32), weights=np.fromfile(wf, np.float32))
    seeds = g.seeds(800, request=3)
    for bias in (0, 1, 2, 3):
        want = O.sample(og, seeds, [10, 4], bias=bias if bias != 3 else O.BIAS_USER, request=3)
        got = C.sample(g, seeds, [10, 4], bias=bias, request=3)
        for k in KEYS:
            assert np.array_equal(got[k], want[k]), (bias, k)
    W, steps = C.random_walk(g, seeds, 25, bias=2)
    Wo, so = O.walk(og, seeds, 25, bias=2)
    assert np.array_equal(W, Wo) and steps == so
