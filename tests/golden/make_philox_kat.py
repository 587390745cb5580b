"""Writes tests/golden/philox_kat.json: Philox4x32-10 known answers.

Three vectors are the published Random123 known answers (kat_vectors,
philox4x32 with 10 rounds).  The rest are produced here by ATen's
BSD-licensed host implementation (ATen/core/PhiloxRNGEngine.h), an
implementation independent of both the oracle's and the library's, whose
engine maps (seed, subsequence, offset) to key = seed and counter =
(offset lo, offset hi, subsequence lo, subsequence hi).  The header only
exists in this build container, so the vectors are committed as data.
"""
import json
import os
import random
import subprocess
import tempfile

import torch.utils.cpp_extension as ext

HERE = os.path.dirname(os.path.abspath(__file__))

PROG = r"""
#include <ATen/core/PhiloxRNGEngine.h>
#include <cstdio>
#include <cstdlib>
int main(int argc, char** argv) {
  for (int i = 1; i + 2 < argc; i += 3) {
    unsigned long long seed = strtoull(argv[i], 0, 0), sub = strtoull(argv[i+1], 0, 0), off = strtoull(argv[i+2], 0, 0);
    at::Philox4_32 eng(seed, sub, off);
    unsigned a = eng(), b = eng(), c = eng(), d = eng();
    printf("%08x %08x %08x %08x\n", a, b, c, d);
  }
  return 0;
}
"""

PUBLISHED = [
    # counter (4 x u32), key (2 x u32), output -- Random123 kat_vectors, philox4x32 10 rounds
    ([0, 0, 0, 0], [0, 0], ["6627e8d5", "e169c58d", "bc57ac4c", "9b00dbd8"]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, ["408f276d", "41c83b0e", "a20bc7c6", "6d5451fd"]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     ["d16cfe09", "94fdcceb", "5001e420", "24126ea1"]),
]


def main():
    rng = random.Random(20201117)
    cases = [(c, k) for c, k, _ in PUBLISHED]
    for _ in range(61):
        cases.append(([rng.getrandbits(32) for _ in range(4)], [rng.getrandbits(32) for _ in range(2)]))
    inc = " ".join("-I" + p for p in ext.include_paths())
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "kat.cpp")
        exe = os.path.join(td, "kat")
        open(src, "w").write(PROG)
        subprocess.check_call(f"g++ -O1 -std=c++17 {inc} {src} -o {exe}", shell=True)
        args = []
        for c, k in cases:
            seed = k[0] | (k[1] << 32)
            sub = c[2] | (c[3] << 32)
            off = c[0] | (c[1] << 32)
            args += [str(seed), str(sub), str(off)]
        out = subprocess.check_output([exe] + args, text=True).split("\n")
    vectors = []
    for (c, k), line in zip(cases, out):
        vectors.append({"counter": c, "key": k, "out": line.split()})
    for (c, k, want), got in zip(PUBLISHED, vectors):
        assert got["out"] == want, (c, k, got["out"], want)
    json.dump({"source": "Random123 kat_vectors (first 3) + ATen PhiloxRNGEngine.h", "vectors": vectors},
              open(os.path.join(HERE, "philox_kat.json"), "w"), indent=0)
    print("wrote", len(vectors), "vectors")


if __name__ == "__main__":
    main()
