"""Host round trip of the packed neighbour-id format streamed slices use
(c-saw_amd/csrc/pack.hpp; the device decoder's output is checked by the
streamed sampling tests, which compare every packed run with the unpacked one
and the in-memory sampler)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "pack_check.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pack_round_trip(tmp_path):
    exe = str(tmp_path / "pack_check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address", SRC, "-o", exe], check=True,
                   capture_output=True, timeout=240)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches=0" in r.stdout and "blocks=0" not in r.stdout
