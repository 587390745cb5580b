// Round trip of the packed neighbour-id format (c-saw_amd/csrc/pack.hpp):
// sorted random neighbour lists of mixed degrees (small, large, ids near
// 2^31), concatenated as a CSR index array, packed in 64-entry blocks and
// decoded again; every id must come back, and the packed size must be what
// pack_block reported.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "../../c-saw_amd/csrc/pack.hpp"

using namespace csaw;

int main() {
  std::mt19937_64 rng(7);
  long mismatches = 0, blocks = 0;
  for (int trial = 0; trial < 12; ++trial) {
    const int64_t V = trial < 6 ? 1000 : 2147483647;  // id range
    std::vector<int32_t> adj;
    const int nv = 1 + (int)(rng() % 200);
    for (int v = 0; v < nv; ++v) {
      const int r = (int)(rng() % 10);
      int d = r < 4 ? (int)(rng() % 8) : (r < 8 ? (int)(rng() % 200) : (int)(rng() % 5000));
      if (V == 1000) d = std::min(d, 1000);
      std::vector<int32_t> l;
      while ((int)l.size() < d) {
        l.push_back((int32_t)(rng() % (uint64_t)V));
        if ((int)l.size() == d) {
          std::sort(l.begin(), l.end());
          l.erase(std::unique(l.begin(), l.end()), l.end());
        }
      }
      adj.insert(adj.end(), l.begin(), l.end());
    }
    const int64_t E = (int64_t)adj.size();
    const int64_t nblk = (E + kPackBlock - 1) / kPackBlock;
    std::vector<int64_t> boff(nblk + 1, 0);
    for (int64_t b = 0; b < nblk; ++b)
      boff[b + 1] = boff[b] + pack_block(adj.data() + b * kPackBlock, (int)std::min<int64_t>(kPackBlock, E - b * kPackBlock),
                                         nullptr);
    std::vector<uint8_t> data(boff[nblk] + 1);
    for (int64_t b = 0; b < nblk; ++b) {
      const int n = (int)std::min<int64_t>(kPackBlock, E - b * kPackBlock);
      const int64_t got = pack_block(adj.data() + b * kPackBlock, n, data.data() + boff[b]);
      if (got != boff[b + 1] - boff[b] || got > kPackMaxBlockBytes) ++mismatches;
    }
    std::vector<int32_t> out(E);
    for (int64_t b = 0; b < nblk; ++b) {
      unpack_block(data.data() + boff[b], (int)std::min<int64_t>(kPackBlock, E - b * kPackBlock), out.data() + b * kPackBlock);
      ++blocks;
    }
    for (int64_t i = 0; i < E; ++i) mismatches += out[i] != adj[i];
  }
  printf("blocks=%ld mismatches=%ld\n", blocks, mismatches);
  return mismatches != 0;
}
