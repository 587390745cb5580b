// Host check of the 16-ary CDF search index (c-saw_amd/csrc/cdf_index.hpp):
// for random concatenations of per-vertex inclusive CDFs (uint64 and FP32,
// with zero weights and so repeated values), every search through the index,
// single and in lockstep, must equal std::upper_bound over the vertex's range,
// at every CDF value, just below it, at 0 and at random targets.  Arrays are
// sized exactly as the library allocates them (CDF padded to 16 entries,
// index of cdf_index_size(E) keys), so an out-of-bounds read shows up under
// AddressSanitizer.  Exit status 0 = all equal.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "../../c-saw_amd/csrc/cdf_index.hpp"

using namespace csaw;

// index levels exactly as the library's k_index_level builds them
template <typename T>
static void build_index(const std::vector<T>& C, int64_t E, std::vector<T>& K) {
  K.assign(cdf_index_size(E), T(0));
  int64_t at = 0;
  for (int l = 1; l <= cdf_index_levels(E); ++l) {
    const int64_t n = (E + ((int64_t)1 << (4 * l)) - 1) >> (4 * l);
    for (int64_t b = 0; b < n; ++b) K[at + b] = C[std::min(((b + 1) << (4 * l)) - 1, E - 1)];
    at += cdf_index_level_size(E, l);
  }
}

template <typename T>
static long run(uint64_t seed, int64_t E, long* checks) {
  std::mt19937_64 r(seed);
  std::vector<int64_t> off{0};
  while (off.back() < E) {
    const int64_t d = (r() % 3 == 0) ? (int64_t)(r() % 5000) : (int64_t)(r() % 40);
    off.push_back(std::min(E, off.back() + d));
  }
  std::vector<T> C(pad16(E), T(0));
  for (size_t v = 0; v + 1 < off.size(); ++v) {
    T acc = 0;
    for (int64_t i = off[v]; i < off[v + 1]; ++i) {
      T w = (r() % 4 == 0) ? T(0) : (sizeof(T) == 8 ? (T)(r() % 1000) : (T)((r() % 1000) / 7.0));
      acc += w;
      C[i] = acc;
    }
  }
  std::vector<T> K;
  build_index(C, E, K);
  long bad = 0;
  for (size_t v = 0; v + 1 < off.size(); ++v) {
    const int64_t o = off[v];
    const int32_t d = (int32_t)(off[v + 1] - o);
    if (d == 0 || !(C[o + d - 1] > T(0))) continue;
    const T tot = C[o + d - 1];
    std::vector<T> xs{T(0)};
    for (int i = 0; i < d; ++i) {
      xs.push_back(C[o + i]);
      if (C[o + i] > T(0)) xs.push_back(C[o + i] - (sizeof(T) == 8 ? T(1) : T(0.5)));
    }
    for (int k = 0; k < 20; ++k) xs.push_back((T)(tot * ((r() % 1000000) / 1e6)));
    for (T x : xs) {
      if (!(x < tot) || x < T(0)) continue;
      const int32_t ref = (int32_t)(std::upper_bound(C.begin() + o, C.begin() + o + d, x) - (C.begin() + o));
      const int32_t got = cdf_index_search<T>(C.data(), K.data(), E, o, d, x);
      T xq[4] = {x, T(0), x, x};
      int32_t out[4];
      cdf_index_search_n<T, 4>(C.data(), K.data(), E, o, d, xq, 3, out);
      const int32_t ref0 = (int32_t)(std::upper_bound(C.begin() + o, C.begin() + o + d, T(0)) - (C.begin() + o));
      ++*checks;
      if (got != ref || out[0] != ref || out[1] != ref0 || out[2] != ref) {
        if (bad < 5)
          std::printf("mismatch E=%lld v=%zu o=%lld d=%d x=%g ref=%d got=%d lockstep=%d\n", (long long)E, v,
                      (long long)o, d, (double)x, ref, got, out[0]);
        ++bad;
      }
    }
  }
  return bad;
}

int main() {
  long bad = 0, checks = 0;
  for (int64_t E : {1, 15, 16, 17, 255, 256, 257, 4096, 4097, 70000, 1200000}) {
    bad += run<uint64_t>((uint64_t)E * 7 + 1, E, &checks);
    bad += run<float>((uint64_t)E * 3 + 2, E, &checks);
  }
  std::printf("checks=%ld mismatches=%ld\n", checks, bad);
  return bad != 0;
}
