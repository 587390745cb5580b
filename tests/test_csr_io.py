"""On-disk CSR (beg.bin int64 offsets, csr.bin int64 ids; non-stream/graph.hpp:18-74)
and the SNAP edge-list converter: host-only C-ABI calls, no GPU needed."""
import os

import numpy as np
import pytest

SNAP = """# Directed graph (each unordered pair of nodes is saved once): toy.txt
# Nodes: 7 Edges: 9
# FromNodeId\tToNodeId
0\t1
0\t2
2\t0
1\t1
3\t5  17.5
3\t5
 % a comment after blanks

5,6
4 0\r
"""


def read_bin(beg, csr):
    return np.fromfile(beg, np.int64), np.fromfile(csr, np.int64)


def test_snap_directed_known_answer(C, tmp_path):
    src = tmp_path / "toy.txt"
    src.write_text(SNAP)
    beg, csr = str(tmp_path / "beg.bin"), str(tmp_path / "csr.bin")
    V, E = C.snap_to_bin(str(src), beg, csr, symmetrize=False)
    off, ids = read_bin(beg, csr)
    # self loop 1-1 and the repeated 3-5 dropped; lists sorted; V = 6 + 1
    assert (V, E) == (7, 6)
    assert off.tolist() == [0, 2, 2, 3, 4, 5, 6, 6]
    assert ids.tolist() == [1, 2, 0, 5, 0, 6]


def test_snap_symmetrized_known_answer(C, tmp_path):
    src = tmp_path / "toy.txt"
    src.write_text(SNAP)
    beg, csr = str(tmp_path / "beg.bin"), str(tmp_path / "csr.bin")
    V, E = C.snap_to_bin(str(src), beg, csr, symmetrize=True)
    off, ids = read_bin(beg, csr)
    assert (V, E) == (7, 10)
    assert off.tolist() == [0, 3, 4, 5, 6, 7, 9, 10]
    assert ids.tolist() == [1, 2, 4, 0, 0, 5, 0, 3, 6, 5]


def test_snap_matches_numpy_on_random_lists(C, tmp_path):
    # a random edge list with duplicates and self loops, long enough to cross
    # the converter's 1 MiB read blocks; expected CSR from numpy's unique
    rng = np.random.default_rng(5)
    u = rng.integers(0, 5000, 200000)
    v = rng.integers(0, 5000, 200000)
    src = tmp_path / "r.txt"
    with open(src, "w") as f:
        f.write("# random\n")
        for a, b in zip(u, v):
            f.write(f"{a}\t{b}\n")
    for sym in (False, True):
        beg, csr = str(tmp_path / "b.bin"), str(tmp_path / "c.bin")
        V, E = C.snap_to_bin(str(src), beg, csr, symmetrize=sym)
        a, b = (np.concatenate([u, v]), np.concatenate([v, u])) if sym else (u, v)
        keep = a != b
        pairs = np.unique(a[keep].astype(np.int64) << 32 | b[keep].astype(np.int64))
        ps, pd = pairs >> 32, pairs & 0xffffffff
        want_off = np.zeros(max(u.max(), v.max()) + 2, np.int64)
        np.add.at(want_off, ps + 1, 1)
        off, ids = read_bin(beg, csr)
        assert V == len(want_off) - 1 and E == len(pairs)
        assert np.array_equal(off, np.cumsum(want_off))
        assert np.array_equal(ids, pd)


def test_snap_rejects_malformed(C, tmp_path):
    src = tmp_path / "bad.txt"
    src.write_text("0 1\n2 x\n")
    with pytest.raises(C.CsawError) as ei:
        C.snap_to_bin(str(src), str(tmp_path / "b"), str(tmp_path / "c"))
    assert ei.value.status == 2 and "line 2" in C.last_error()
    with pytest.raises(C.CsawError) as ei:
        C.snap_to_bin(str(tmp_path / "missing.txt"), str(tmp_path / "b"), str(tmp_path / "c"))
    assert ei.value.status == 9


def test_write_bin_round_trip(C, tmp_path):
    off = np.array([0, 2, 2, 5], np.int64)
    adj = np.array([1, 2, 0, 1, 2], np.int32)
    beg, csr = str(tmp_path / "beg.bin"), str(tmp_path / "csr.bin")
    C.write_bin(beg, csr, off, adj)
    o, i = read_bin(beg, csr)
    assert np.array_equal(o, off) and np.array_equal(i, adj.astype(np.int64))
    assert os.path.getsize(beg) == 8 * 4 and os.path.getsize(csr) == 8 * 5
    with pytest.raises(C.CsawError):
        C.write_bin(beg, csr, off, np.array([1, 2, 0, 1, 3], np.int32))  # id 3 >= V


def test_load_bin_rejects_malformed_files(C, tmp_path):
    # every check here runs on the host before anything is uploaded
    beg, csr, wf = str(tmp_path / "beg.bin"), str(tmp_path / "csr.bin"), str(tmp_path / "w.bin")

    def files(off, ids, nbytes_beg=None, w=None):
        b = np.asarray(off, np.int64).tobytes()
        open(beg, "wb").write(b[:nbytes_beg] if nbytes_beg is not None else b)
        np.asarray(ids, np.int64).tofile(csr)
        if w is not None:
            np.asarray(w, np.float32).tofile(wf)

    files([0, 1, 2], [1, 0], nbytes_beg=20)  # not a multiple of 8 bytes
    with pytest.raises(C.CsawError):
        C.Graph.load_bin(beg, csr)
    files([0, 1, 3], [1, 0])  # beg[V] = 3 ids, csr.bin holds 2
    with pytest.raises(C.CsawError):
        C.Graph.load_bin(beg, csr)
    files([0, 1, 2], [1, 2])  # id 2 outside [0, V = 2)
    with pytest.raises(C.CsawError):
        C.Graph.load_bin(beg, csr)
    files([0, 1, 2], [1, -1])
    with pytest.raises(C.CsawError):
        C.Graph.load_bin(beg, csr)
    files([0, 1, 2], [1, 0], w=[0.5])  # fewer than E weights
    with pytest.raises(C.CsawError):
        C.Graph.load_bin(beg, csr, wf)
    with pytest.raises(C.CsawError):
        C.Graph.load_bin(str(tmp_path / "missing.bin"), csr)
