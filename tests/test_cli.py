"""The command line mirroring the reference's driver (mpi_main.cpp:17-49):
argument checks on the CPU; on the GPU a run over beg.bin / csr.bin files
whose edge count equals the library's for the same Sampler() parameters."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "c-saw_amd", "lib", "csaw_cli")
needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="csaw_cli is not built")


@needs_cli
def test_cli_usage_and_gpu_count():
    r = subprocess.run([CLI, "WG", "beg", "csr"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Input: ./exe <dataset name>" in r.stdout
    r = subprocess.run([CLI, "WG", "beg", "csr", "400", "128", "10", "1", "2", "2", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "#GPUs must be 1" in r.stderr


@needs_cli
@pytest.mark.gpu
def test_cli_samples_bin_files(C, tmp_path):
    g = C.Graph.rmat(12, 8)
    off, adj = g.csr()
    beg, csr = str(tmp_path / "beg.bin"), str(tmp_path / "csr.bin")
    C.write_bin(beg, csr, off, adj)
    r = subprocess.run([CLI, "toy", beg, csr, "400", "128", "300", "1", "5", "2", "1"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    edges, name, tmin, tmax = r.stdout.strip().split(",")
    assert name == "toy" and float(tmin) == float(tmax) > 0
    # Sampler(): seeds csaw_seeds(key 42, request = rank 0), user hooks,
    # neighbour sampling without repeats, fanout NeighborSize at every hop
    want = C.sample(g, g.seeds(300, key=42, request=0), [5, 5], bias=C.BIAS_USER, dedup=1, key=42, request=0)
    assert int(edges) == len(want["src"]) > 0
