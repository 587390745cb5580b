"""Host check of the 16-ary CDF search index against std::upper_bound.

The device kernels search cached CDFs through c-saw_amd/csrc/cdf_index.hpp;
its search functions are host-callable, so tests/native/cdf_index_check.cpp
compiles them with g++ (AddressSanitizer on) and compares every search with
std::upper_bound over the vertex's CDF range (non-stream/functions.cuh:11-43
is the binary search these replace).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "cdf_index_check.cpp")
ROCM_INC = "/opt/rocm/include"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(ROCM_INC), reason="needs g++ and HIP headers")
def test_cdf_index_matches_upper_bound(tmp_path):
    exe = str(tmp_path / "cdf_index_check")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM_INC}",
           "-x", "c++", SRC, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, timeout=240)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches=0" in r.stdout
