// Implicit 16-ary search index over a concatenated array of per-vertex
// inclusive CDFs C[0 .. E) (vertex v owns C[off[v] .. off[v] + deg[v]),
// nondecreasing inside each vertex's range).
//
// Level l >= 1 holds one key per aligned block of 16^l entries of C:
//   K_l[b] = C[min(16^l (b + 1) - 1, E - 1)]   (the block's last entry),
// padded with zeros to a multiple of 16 keys; the levels are stored one
// after another (level 1 first) in one array, and C itself is allocated
// padded to a multiple of 16 entries.  Levels 1 .. n exist where n is the
// smallest with 16^(n+1) >= E.
//
// upper_bound(v, x) = first i in [0, d) with C[o + i] > x, given x < C[o+d-1]:
// let L = ceil(bitlength(o xor (o+d-1)) / 4), the lowest level whose one
// block b = o >> 4L holds the whole range.  Descending l = L-1 .. 0, the 16
// keys of the sub-blocks of b (level l, or C itself for l = 0) are read with
// one aligned 16-key load and b becomes 16 b + #{sub-blocks c : end_c < o, or
// end_c <= o+d-1 and key_c <= x}.  Sub-blocks ending before o belong to
// earlier vertices (always passed), those ending past the range are never
// passed, and inside the range the keys are the vertex's own nondecreasing
// CDF values, so the passed set is a prefix and the answer's block is the
// next one.  At l = 0 the count gives the answer's absolute position, so the
// result equals a binary search bit for bit, using L reads instead of
// ceil(log2 d) dependent probes.
#pragma once
#include <string.h>

#include "common.hpp"

namespace csaw {

constexpr int kIdxMaxLevels = 8;  // covers E <= 16^9

__host__ __device__ __forceinline__ int64_t pad16(int64_t n) { return (n + 15) & ~(int64_t)15; }

// keys of level l (padded)
__host__ __device__ __forceinline__ int64_t cdf_index_level_size(int64_t E, int l) {
  return pad16((E + ((int64_t)1 << (4 * l)) - 1) >> (4 * l));
}

__host__ __device__ __forceinline__ int cdf_index_levels(int64_t E) {
  int n = 0;
  while (n < kIdxMaxLevels && ((int64_t)1 << (4 * (n + 1))) < E) ++n;
  return n;
}

// total keys of the index of a CDF array of E entries
__host__ __device__ __forceinline__ int64_t cdf_index_size(int64_t E) {
  int64_t t = 0;
  const int n = cdf_index_levels(E);
  for (int l = 1; l <= n; ++l) t += cdf_index_level_size(E, l);
  return t;
}

__host__ __device__ __forceinline__ int cdf_index_reads(int64_t o, int32_t d) {
  const uint64_t diff = (uint64_t)o ^ (uint64_t)(o + d - 1);
  return diff ? (67 - __builtin_clzll(diff)) >> 2 : 0;
}

template <typename T>
__host__ __device__ __forceinline__ void load16(const T* __restrict__ p, T* v) {
  if constexpr (sizeof(T) == 8) {
    const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const ulonglong2 t = q[i];
      v[2 * i] = (T)t.x;
      v[2 * i + 1] = (T)t.y;
    }
  } else {
    const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint4 t = q[i];
      const uint32_t u[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) memcpy(&v[4 * i + k], &u[k], 4);
    }
  }
}

// #{c : end_c < o or (end_c <= last and v[c] <= x)}, end_c = ((16 b + c + 1) << sh) - 1.
// end_c < o  <=>  c < (o >> sh) - 16 b, and end_c <= last  <=>  c < ((last + 1) >> sh) - 16 b.
template <typename T>
__host__ __device__ __forceinline__ int count16(const T* v, int64_t b, int sh, int64_t o, int64_t last, T x) {
  const int64_t lo64 = (o >> sh) - 16 * b, hi64 = ((last + 1) >> sh) - 16 * b;
  const int lo = (int)(lo64 < 0 ? 0 : (lo64 > 16 ? 16 : lo64));
  const int hi = (int)(hi64 < 0 ? 0 : (hi64 > 16 ? 16 : hi64));
  int c = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) c += (i < lo) | ((i < hi) & (v[i] <= x));
  return c;
}

// upper_bound over C[o .. o+d) through the index K of the array C (E entries).
template <typename T>
__host__ __device__ __forceinline__ int32_t cdf_index_search(const T* __restrict__ C, const T* __restrict__ K, int64_t E,
                                                    int64_t o, int32_t d, T x) {
  const int64_t last = o + d - 1;
  const int L = cdf_index_reads(o, d);
  if (L == 0) return 0;
  int64_t at = 0;  // offset of level L-1 in K
  for (int m = 1; m <= L - 2; ++m) at += cdf_index_level_size(E, m);
  int64_t b = o >> (4 * L);
  T v[16];
  for (int l = L - 1; l >= 1; --l) {
    load16<T>(K + at + 16 * b, v);
    b = 16 * b + count16<T>(v, b, 4 * l, o, last, x);
    if (l > 1) at -= cdf_index_level_size(E, l - 1);
  }
  load16<T>(C + 16 * b, v);
  b = 16 * b + count16<T>(v, b, 0, o, last, x);
  return (int32_t)(b - o);
}

// The 16-key block a search over C[o .. o+d) reads first (the same for every
// target), so a caller can load it ahead of the draws.
template <typename T>
__host__ __device__ __forceinline__ const T* cdf_index_first_block(const T* C, const T* K, int64_t E, int64_t o,
                                                                   int32_t d) {
  const int L = cdf_index_reads(o, d);
  if (L <= 1) return C + 16 * (o >> 4);
  int64_t at = 0;
  for (int m = 1; m <= L - 2; ++m) at += cdf_index_level_size(E, m);
  return K + at + 16 * (o >> (4 * L));
}

// NQ searches over the same range in lockstep (one round of loads per level
// for all of them); queries q >= nq are skipped.
template <typename T, int NQ>
__host__ __device__ __forceinline__ void cdf_index_search_n(const T* __restrict__ C, const T* __restrict__ K, int64_t E,
                                                   int64_t o, int32_t d, const T* x, int nq, int32_t* out) {
  const int64_t last = o + d - 1;
  const int L = cdf_index_reads(o, d);
  if (L == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[q] = 0;
    return;
  }
  int64_t at = 0;
  for (int m = 1; m <= L - 2; ++m) at += cdf_index_level_size(E, m);
  int64_t b[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) b[q] = o >> (4 * L);
  for (int l = L - 1; l >= 0; --l) {
    const T* base = l ? K + at : C;
    T v[NQ][16];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      if (q < nq) load16<T>(base + 16 * b[q], v[q]);
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      if (q < nq) b[q] = 16 * b[q] + count16<T>(v[q], b[q], 4 * l, o, last, x[q]);
    if (l > 1) at -= cdf_index_level_size(E, l - 1);
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) out[q] = (int32_t)(b[q] - o);
}
}  // namespace csaw
