// Device-side inverse-transform selection for one frontier entry, in three
// granularities (one lane, one wave64, one 256-lane workgroup) and two
// taken-set representations (LDS bitmap or list).  Every path implements
// the same canonical per-source selection, so all variants emit identical
// edges:
//   d == 0                      -> nothing            (main.cu:99-104)
//   dedup && d <= F             -> all d neighbours   (functions.cuh:859-867)
//   otherwise slot j = 0..F-1 in order; attempt a draws
//     philox(key; (j<<16)|a, pos-in-sample, sample, DOM_SELECT|req|hop)
//   and is accepted unless taken (bitmap_search, functions.cuh:45-70);
//   after kReselect attempts (functions.cuh:844) the first untaken index
//   scanning upward cyclically from the last draw is taken.
// The ITS target: uniform i = mulhi64(t, d); integer biases an exact
// uint64 inclusive CDF C and i = min{i: C_i > mulhi64(t, T)}; FP32 weights
// the per-vertex sequential CDF c and i = min{i: c_i > fl32(u24 * T)}
// (ITS + binary_search, functions.cuh:160-176, 11-43).
#pragma once
#include "cdf_index.hpp"
#include "common.hpp"

namespace csaw {

struct Entry {
  int32_t v;
  int32_t d;
  int64_t o;
  int32_t F;
  int32_t dedup;
  int32_t bias;   // csaw_bias
  int32_t g;      // global sample id (written to the sample column)
  int64_t pos;    // first output slot
  uint32_t c1, c2, c3, k0, k1;
  uint32_t T;     // CDF total from the vertex record (has_T)
  bool has_T;
};

__device__ __forceinline__ Entry load_entry(const GraphDev& G, const ReqDev* __restrict__ R,
                                            const int32_t* __restrict__ gs2req, const HopDev& H, int64_t e) {
  Entry en;
  en.v = H.fv[e];
  en.g = H.fs[e];
  const ReqDev& r = R[gs2req[en.g]];
  en.F = r.fanout[H.hop];
  en.dedup = r.dedup;
  en.bias = r.bias;
  // offset, degree and CDF total from the bias's vertex record when the view
  // has one, else offset and degree from one pair of offsets (one line)
  const int rk = en.bias == CSAW_BIAS_DEGREE ? 0 : (en.bias == CSAW_BIAS_USER ? 1 : 2);
  const VertexRecord* vr = en.bias == CSAW_BIAS_UNIFORM ? nullptr : G.vrec[rk];
  en.has_T = vr != nullptr;
  en.T = 0;
  if (vr) {
    const VertexRecord rec = vr[en.v];
    en.o = rec.off;
    en.d = rec.deg;
    en.T = rec.total;
  } else {
    en.o = G.off[en.v];
    en.d = (int32_t)(G.off[en.v + 1] - en.o);
  }
  en.pos = (int64_t)H.grp_excl[e >> kGroupShift] + H.blk_excl[e >> kBlockShift] + H.local_excl[e];
  en.c1 = (uint32_t)(e - H.fo[en.g]);
  en.c2 = (uint32_t)(en.g - r.gbase);
  en.c3 = select_c3(r.request_id, H.hop);
  en.k0 = r.key_lo;
  en.k1 = r.key_hi;
  return en;
}

__device__ __forceinline__ U4 draw(const Entry& en, int j, int a) {
  return philox4x32_10(((uint32_t)j << 16) | (uint32_t)a, en.c1, en.c2, en.c3, en.k0, en.k1);
}

__device__ __forceinline__ uint32_t int_bias(const GraphDev& G, const csaw_graph_view& hv, int32_t bias, int32_t u) {
  return bias == CSAW_BIAS_USER ? EdgeBias(u, &hv) : (uint32_t)G.deg[u];
}

// ITS state of one entry: the entry's CDF is C[o .. o+d) (o = 0 for LDS).
enum PickMode {
  kPickUniform = 0,
  kPickLds = 1,       // integer CDF built in LDS, binary search
  kPickInt = 2,       // integer CDF in global memory (hub cache), binary search
  kPickF32 = 3,       // FP32 CDF, binary search
  kPickIntIdx = 4,    // integer CDF cache through its 16-ary index K
  kPickF32Idx = 5,    // FP32 CDF through its 16-ary index K
  kPickInt32Idx = 6,  // uint32 integer CDF cache through its 16-ary index K (T < 2^32)
};

struct Picker {
  int mode;
  int32_t d;
  uint64_t T;
  float Tf;
  const void* C;
  const void* K;
  int64_t o;

  // E: entries of the CDF array (sizes the index levels).  IDX = false
  // compiles the index modes out (callers whose graph view has no index).
  template <bool IDX>
  __device__ __forceinline__ int32_t pick(int64_t E, const U4& r) const {
    const uint64_t t = philox_t64(r);
    if (mode == kPickUniform) return (int32_t)mulhi64_32(t, (uint32_t)d);
    if (mode == kPickF32 || mode == kPickF32Idx) {
      const float x = __fmul_rn(philox_u24(r), Tf);
      const float* Cf = static_cast<const float*>(C);
      if constexpr (IDX)
        if (mode == kPickF32Idx) return cdf_index_search<float>(Cf, static_cast<const float*>(K), E, o, d, x);
      return (int32_t)upper_bound_dev<float>(Cf + o, d, x);
    }
    const uint64_t x = mulhi64(t, T);
    if constexpr (IDX) {
      if (mode == kPickInt32Idx)
        return cdf_index_search<uint32_t>(static_cast<const uint32_t*>(C), static_cast<const uint32_t*>(K), E, o, d,
                                          (uint32_t)x);
      if (mode == kPickIntIdx)
        return cdf_index_search<uint64_t>(static_cast<const uint64_t*>(C), static_cast<const uint64_t*>(K), E, o, d, x);
    } else {
      // without the index, a uint32 cache is still searched with its own key
      // width (x < T < 2^32)
      if (mode == kPickInt32Idx)
        return (int32_t)upper_bound_dev<uint32_t>(static_cast<const uint32_t*>(C) + o, d, (uint32_t)x);
    }
    return (int32_t)upper_bound_dev<uint64_t>(static_cast<const uint64_t*>(C) + o, d, x);
  }
};

// Picker for a bias that needs no per-entry build (uniform, FP32, CDF cache,
// hub cache).  Returns false when an integer CDF must be built by the caller.
__device__ __forceinline__ bool picker_global(const GraphDev& G, const Entry& en, int kind, int cap, Picker& pk) {
  pk.d = en.d;
  pk.mode = kPickUniform;
  pk.o = en.o;
  pk.K = nullptr;
  if (kind == kUniform) return true;
  if (kind == kF32) {
    pk.C = G.cdf;
    pk.K = G.cdf_idx;
    pk.Tf = en.has_T ? __uint_as_float(en.T) : G.cdf[en.o + en.d - 1];
    if (pk.Tf > 0.0f) pk.mode = G.cdf_idx ? kPickF32Idx : kPickF32;
    return true;
  }
  const int ck = en.bias == CSAW_BIAS_USER ? 1 : 0;
  if (G.int_cdf32[ck]) {
    pk.C = G.int_cdf32[ck];
    pk.K = G.int_idx32[ck];
    pk.T = en.has_T ? en.T : G.int_cdf32[ck][en.o + en.d - 1];
    if (pk.T > 0) pk.mode = kPickInt32Idx;
    return true;
  }
  const uint64_t* C;
  if (G.int_cdf[ck]) {
    C = G.int_cdf[ck];
    pk.K = G.int_idx[ck];
  } else {
    if (en.d <= cap) return false;
    const int32_t h = hub_find(G.hub_vid, G.n_hubs, en.v);
    C = G.hub_cdf[ck];
    pk.o = G.hub_pool[h];
  }
  pk.C = C;
  pk.T = C[pk.o + en.d - 1];
  if (pk.T > 0) pk.mode = pk.K ? kPickIntIdx : kPickInt;
  return true;
}

__device__ __forceinline__ void write_edge(const GraphDev& G, const csaw_graph_view& hv, const HopDev& H,
                                           const Entry& en, int64_t p, int32_t idx) {
  H.src[p] = en.v;
  H.dst[p] = Update(&hv, G.adj[en.o + idx], en.v);
  H.smp[p] = en.g;
}

__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// ---------------------------------------------------------------------------
// Wave-level selection: lanes j < F hold slot j.  Slots resolve in order;
// a colliding slot evaluates 64 further attempts at once (lane a -> attempt
// a0+a) and takes the lowest free one, which is exactly the first free
// attempt of the sequential definition.  Bitmap: LDS bitmap of d bits
// (bm != nullptr); list: the already-resolved lanes' registers.
// Returns the lane's resolved index (or -1 for lanes >= F).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool wave_taken_lane(int32_t ci, int32_t res, int jj, const uint32_t* bm) {
  if (bm) return (bm[ci >> 5] >> (ci & 31)) & 1u;
  bool hit = false;
  for (int m = 0; m < jj; ++m) hit |= (ci == readlane(res, m));
  return hit;
}

template <bool IDX>
__device__ int32_t wave_select(int64_t E, const Entry& en, const Picker& pk, uint32_t* bm) {
  const int lane = lane_id();
  const int F = en.F;
  const int32_t d = en.d;
  int32_t res = -1;
  if (!en.dedup) {
    if (lane < F) res = pk.template pick<IDX>(E, draw(en, lane, 0));
    return res;
  }
  int32_t cand = 0;
  if (lane < F) cand = pk.template pick<IDX>(E, draw(en, lane, 0));
  // Fast path: up to the first slot whose attempt-0 draw repeats an earlier
  // slot's draw, the sequential definition accepts every attempt-0 draw, so
  // those slots are final; only the rest go through the ordered loop.
  bool dup = false;
  for (int i = 0; i + 1 < F; ++i) {
    const int32_t ci = readlane(cand, i);
    dup |= (lane > i) & (lane < F) & (cand == ci);
  }
  const uint64_t dm = ballot(dup);
  if (dm == 0ull) return lane < F ? cand : -1;
  const int j0 = __builtin_ctzll(dm);
  res = lane < j0 ? cand : -1;
  if (bm) {
    const int words = (d + 31) >> 5;
    for (int i = lane; i < words; i += kWave) bm[i] = 0u;
    wave_fence();
    if (lane < j0) atomicOr(&bm[cand >> 5], 1u << (cand & 31));
    wave_fence();
  }
  for (int jj = j0; jj < F; ++jj) {
    int32_t c = readlane(cand, jj);
    bool taken = bm ? ((bm[c >> 5] >> (c & 31)) & 1u) : (ballot(res == c) != 0ull);
    if (taken) {
      int32_t chosen = -1, last = c;
      for (int a0 = 1; a0 < kReselect; a0 += kWave) {
        const int a = a0 + lane;
        bool fr = false;
        int32_t ci = 0;
        if (a < kReselect) {
          ci = pk.template pick<IDX>(E, draw(en, jj, a));
          fr = !wave_taken_lane(ci, res, jj, bm);
        }
        const uint64_t m = ballot(fr);
        if (m) {
          chosen = readlane(ci, __builtin_ctzll(m));
          break;
        }
        const int ll = kReselect - 1 - a0;
        if (ll < kWave) last = readlane(ci, ll);
      }
      if (chosen < 0) {
        for (int q = 0; q < d; q += kWave) {
          const int32_t i = (int32_t)(((int64_t)last + 1 + q + lane) % d);
          const bool fr = (q + lane < d) && !wave_taken_lane(i, res, jj, bm);
          const uint64_t m = ballot(fr);
          if (m) {
            chosen = readlane(i, __builtin_ctzll(m));
            break;
          }
        }
      }
      c = chosen;
    }
    if (bm) {
      if (lane == 0) bm[c >> 5] |= 1u << (c & 31);
      wave_fence();
    }
    if (lane == jj) res = c;
  }
  return res;
}

// ---------------------------------------------------------------------------
// Region-remap mode (dedup == 2): heur_normalize's collision policy
// (functions.cuh:743-868).  Draws live in the normalised CDF
// p_i = fl32(fl64(C_i) / fl64(T)) (uniform: fl32((i + 1) / d)); r picks the
// first i with p_i > r, clamped to d - 1.  Per slot, bounds lb = 0, hb = 1;
// each attempt draws r; on a collision with the picked interval [A, B) r is
// remapped with the reference's float sequence (A = p_{i-1}, 0 for i = 0;
// B = p_i) and picks again; after kReselect attempts the cyclic fallback.
// Every lane runs the same sequence (the results agree across lanes); lane
// j keeps slot j.  Taken set: the LDS bitmap, or the resolved lanes.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float remap_p(const Picker& pk, int32_t i) {
  switch (pk.mode) {
    case kPickUniform:
      return (float)((double)(i + 1) / (double)pk.d);
    case kPickF32:
    case kPickF32Idx:
      return (float)((double)static_cast<const float*>(pk.C)[pk.o + i] / (double)pk.Tf);
    case kPickInt32Idx:
      return (float)((double)static_cast<const uint32_t*>(pk.C)[pk.o + i] / (double)pk.T);
    default:  // uint64 CDFs: in LDS, the hub cache or the 64-bit cache
      return (float)((double)static_cast<const uint64_t*>(pk.C)[pk.o + i] / (double)pk.T);
  }
}

__device__ __forceinline__ int32_t remap_pick(const Picker& pk, float r) {
  int32_t lo = 0, hi = pk.d;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (remap_p(pk, mid) > r) hi = mid; else lo = mid + 1;
  }
  return lo < pk.d ? lo : pk.d - 1;
}

__device__ int32_t wave_select_remap(const Entry& en, const Picker& pk, uint32_t* bm) {
  const int lane = lane_id();
  const int F = en.F;
  const int32_t d = en.d;
  int32_t res = -1;
  if (bm) {
    const int words = (d + 31) >> 5;
    for (int i = lane; i < words; i += kWave) bm[i] = 0u;
    wave_fence();
  }
  auto taken = [&](int32_t c, int jj) -> bool {
    return bm ? ((bm[c >> 5] >> (c & 31)) & 1u) != 0u : ballot(res == c) != 0ull;
  };
  for (int jj = 0; jj < F; ++jj) {
    float lb = 0.0f, hb = 1.0f;
    int32_t i = 0, chosen = -1;
    for (int a = 0; a < kReselect; ++a) {
      float r = philox_u24(draw(en, jj, a));
      i = remap_pick(pk, r);
      if (!taken(i, jj)) {
        chosen = i;
        break;
      }
      const float A = i > 0 ? remap_p(pk, i - 1) : 0.0f;
      const float B = remap_p(pk, i);
      const float lambda = __fdiv_rn(__fsub_rn(A, lb), __fsub_rn(__fadd_rn(__fsub_rn(A, lb), hb), B));
      const float delta = __fdiv_rn(__fsub_rn(B, A), __fsub_rn(hb, lb));
      r = __fdiv_rn(r, lambda);
      if (r < A) {
        hb = A;
      } else {
        r = __fadd_rn(r, delta);
        lb = B;
      }
      i = remap_pick(pk, r);
      if (!taken(i, jj)) {
        chosen = i;
        break;
      }
    }
    if (chosen < 0) {
      for (int q = 0; q < d; q += kWave) {
        const int32_t c = (int32_t)(((int64_t)i + 1 + q + lane) % d);
        const bool fr = (q + lane < d) && !wave_taken_lane(c, res, jj, bm);
        const uint64_t m = ballot(fr);
        if (m) {
          chosen = readlane(c, __builtin_ctzll(m));
          break;
        }
      }
    }
    if (bm) {
      if (lane == 0) bm[chosen >> 5] |= 1u << (chosen & 31);
      wave_fence();
    }
    if (lane == jj) res = chosen;
  }
  return res;
}

// Builds the integer CDF of entry en in LDS with one wave (d <= cap).
__device__ __forceinline__ uint64_t wave_build_int_cdf(const GraphDev& G, const csaw_graph_view& hv, const Entry& en,
                                                       uint64_t* C) {
  const int lane = lane_id();
  uint64_t acc = 0;
  for (int i0 = 0; i0 < en.d; i0 += kWave) {
    const int i = i0 + lane;
    uint64_t w = 0;
    if (i < en.d) w = int_bias(G, hv, en.bias, G.adj[en.o + i]);
    const uint64_t s = wave_incl_scan_u64(w);
    if (i < en.d) C[i] = acc + s;
    acc += readlane_u64(s, kWave - 1);
  }
  wave_fence();
  return acc;
}

// Copies all d <= F neighbours (the d <= NeighborSize path) with one wave.
__device__ __forceinline__ void wave_copy_all(const GraphDev& G, const csaw_graph_view& hv, const HopDev& H,
                                              const Entry& en) {
  for (int i = lane_id(); i < en.d; i += kWave) write_edge(G, hv, H, en, en.pos + i, i);
}

// Full wave expansion of one entry whose ITS state was prefetched (build:
// the integer CDF must be built here).  cdf: LDS region of kWaveIntCap u64,
// bm: LDS bitmap of kWaveBitmapCap bits (nullptr: list taken-set).
template <int KIND, bool IDX>
__device__ void expand_entry_wave(const GraphDev& G, const csaw_graph_view& hv, const HopDev& H, const Entry& en,
                                  Picker pk, bool build, uint64_t* cdf, uint32_t* bm) {
  if (en.d == 0 || en.F == 0) return;
  if (en.dedup && en.d <= en.F) {
    wave_copy_all(G, hv, H, en);
    return;
  }
  if (build) {
    pk.C = cdf;
    pk.o = 0;
    pk.T = wave_build_int_cdf(G, hv, en, cdf);
    pk.mode = pk.T > 0 ? kPickLds : kPickUniform;
  }
  uint32_t* tb = (bm && en.d <= kWaveBitmapCap) ? bm : nullptr;
  const int32_t res = en.dedup == 2 ? wave_select_remap(en, pk, tb) : wave_select<IDX>(G.E, en, pk, tb);
  if (lane_id() < en.F) write_edge(G, hv, H, en, en.pos + lane_id(), res);
}

__device__ __forceinline__ int64_t readlane_i64(int64_t v, int l) { return (int64_t)readlane_u64((uint64_t)v, l); }

__device__ __forceinline__ Entry bcast_entry(const Entry& m, int j) {
  Entry e;
  e.v = readlane(m.v, j);
  e.d = readlane(m.d, j);
  e.o = readlane_i64(m.o, j);
  e.F = readlane(m.F, j);
  e.dedup = readlane(m.dedup, j);
  e.bias = readlane(m.bias, j);
  e.g = readlane(m.g, j);
  e.pos = readlane_i64(m.pos, j);
  e.c1 = readlane_u(m.c1, j);
  e.c2 = readlane_u(m.c2, j);
  e.c3 = readlane_u(m.c3, j);
  e.k0 = readlane_u(m.k0, j);
  e.k1 = readlane_u(m.k1, j);
  e.T = readlane_u(m.T, j);
  e.has_T = readlane((int32_t)m.has_T, j) != 0;
  return e;
}

__device__ __forceinline__ Picker bcast_picker(const Picker& m, int j) {
  Picker p;
  p.mode = readlane(m.mode, j);
  p.d = readlane(m.d, j);
  p.T = readlane_u64(m.T, j);
  p.Tf = __int_as_float(readlane(__float_as_int(m.Tf), j));
  p.C = (const void*)readlane_u64((uint64_t)m.C, j);
  p.K = (const void*)readlane_u64((uint64_t)m.K, j);
  p.o = readlane_i64(m.o, j);
  return p;
}

// Expands `count` (<= 64) entries with one wave: lane i loads entry i's
// metadata and ITS state (one round of dependent loads for all of them),
// then the wave expands the entries in order from broadcast values.
template <int KIND, bool IDX>
__device__ void wave_expand_batch(const GraphDev& G, const csaw_graph_view& hv, const ReqDev* __restrict__ R,
                                  const int32_t* __restrict__ gs2req, const HopDev& H,
                                  const int32_t* __restrict__ order, int64_t first, int count, uint64_t* cdf,
                                  uint32_t* bm) {
  const int lane = lane_id();
  Entry mine;
  Picker pm;
  pm.mode = 0;
  pm.d = 0;
  pm.T = 0;
  pm.Tf = 0.f;
  pm.C = nullptr;
  pm.K = nullptr;
  pm.o = 0;
  int build = 0;
  mine.d = 0;
  if (lane < count) {
    mine = load_entry(G, R, gs2req, H, order ? (int64_t)order[first + lane] : first + lane);
    if (H.only_kind && bias_kind(mine.bias) + 1 != H.only_kind) mine.d = 0;  // another kind's launch
    if (mine.d > 0 && mine.F > 0 && !(mine.dedup && mine.d <= mine.F))
      build = picker_global(G, mine, KIND, kWaveIntCap, pm) ? 0 : 1;
  }
  for (int j = 0; j < count; ++j) {
    const Entry en = bcast_entry(mine, j);
    const Picker pk = bcast_picker(pm, j);
    expand_entry_wave<KIND, IDX>(G, hv, H, en, pk, readlane(build, j) != 0, cdf, bm);
  }
}

// ---------------------------------------------------------------------------
// Sub-wave groups, for entries with F <= GS < 64: the wave splits into 64/GS
// groups of GS lanes; each round, group g expands entry e0 + g of the batch
// with lane j of the group holding slot j.  The selection is wave_select's
// list taken-set definition with ballots masked to the group and shuffles
// inside the group in place of readlane, so the edges are the same, with
// 64/GS entries in flight per wave instead of one.  Every shuffle reads lanes
// of the shuffling group only, and all lanes of a group follow the same path.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t shfl_abs(int32_t v, int src) { return __builtin_amdgcn_ds_bpermute(src << 2, v); }
__device__ __forceinline__ uint32_t shfl_abs_u(uint32_t v, int src) { return (uint32_t)shfl_abs((int32_t)v, src); }
__device__ __forceinline__ uint64_t shfl_abs_u64(uint64_t v, int src) {
  return (uint64_t)shfl_abs_u((uint32_t)v, src) | ((uint64_t)shfl_abs_u((uint32_t)(v >> 32), src) << 32);
}

// One frontier entry as the group path keeps it (16 dwords, so a batch of 64
// costs 16 registers per lane, and one shuffle per dword to fetch).
struct GEntry {
  int32_t v, d;
  int64_t o;
  int32_t F, g;
  int64_t pos;
  uint32_t c1, c2, c3, k0, k1;
  uint32_t flags;  // bit 0 dedup, bit 1 user-bias cache, bits 4-7 PickMode
  uint64_t T;      // integer CDF total, or the FP32 total's bits
};

// Entry + its ITS state.  For KIND == kInt the CDF cache of the entry's bias
// must exist (without it the entry is left on the uniform picker).
template <int KIND>
__device__ __forceinline__ GEntry make_gentry(const GraphDev& G, const Entry& en) {
  GEntry ge;
  ge.v = en.v;
  ge.d = en.d;
  ge.o = en.o;
  ge.F = en.F;
  ge.g = en.g;
  ge.pos = en.pos;
  ge.c1 = en.c1;
  ge.c2 = en.c2;
  ge.c3 = en.c3;
  ge.k0 = en.k0;
  ge.k1 = en.k1;
  const uint32_t ck = en.bias == CSAW_BIAS_USER ? 1u : 0u;
  uint32_t mode = kPickUniform;
  ge.T = 0;
  if (en.d > 0 && en.F > 0 && !(en.dedup && en.d <= en.F)) {
    if (KIND == kF32) {
      const float Tf = en.has_T ? __uint_as_float(en.T) : G.cdf[en.o + en.d - 1];
      ge.T = __float_as_uint(Tf);
      if (Tf > 0.0f) mode = G.cdf_idx ? kPickF32Idx : kPickF32;
    } else if (KIND == kInt) {
      if (G.int_cdf32[ck]) {
        ge.T = en.has_T ? en.T : G.int_cdf32[ck][en.o + en.d - 1];
        if (ge.T > 0) mode = kPickInt32Idx;
      } else if (G.int_cdf[ck]) {
        ge.T = G.int_cdf[ck][en.o + en.d - 1];
        if (ge.T > 0) mode = G.int_idx[ck] ? kPickIntIdx : kPickInt;
      }
    }
  }
  ge.flags = (en.dedup ? 1u : 0u) | (ck << 1) | (mode << 4);
  return ge;
}

__device__ __forceinline__ U4 gdraw(const GEntry& en, int j, int a) {
  return philox4x32_10(((uint32_t)j << 16) | (uint32_t)a, en.c1, en.c2, en.c3, en.k0, en.k1);
}

// Key widths a group kernel is compiled for: any, or 32-bit keys only (every
// integer cache of the view is uint32; the uint64 searches are compiled out,
// which frees their registers).
enum GroupKeys { kKeysAny = 0, kKeys32 = 1 };

template <int KIND, int KEYS = kKeysAny>
__device__ __forceinline__ int32_t gpick(const GraphDev& G, const GEntry& en, const U4& r) {
  const uint32_t mode = (en.flags >> 4) & 15u;
  const uint64_t t = philox_t64(r);
  if (KIND == kUniform || mode == kPickUniform) return (int32_t)mulhi64_32(t, (uint32_t)en.d);
  if constexpr (KIND == kF32) {
    const float x = __fmul_rn(philox_u24(r), __uint_as_float((uint32_t)en.T));
    if (mode == kPickF32Idx) return cdf_index_search<float>(G.cdf, G.cdf_idx, G.E, en.o, en.d, x);
    return (int32_t)upper_bound_dev<float>(G.cdf + en.o, en.d, x);
  }
  const bool user = (en.flags >> 1) & 1u;
  if (KEYS == kKeys32 || mode == kPickInt32Idx)  // uint32 cache: T < 2^32
    return cdf_index_search<uint32_t>(user ? G.int_cdf32[1] : G.int_cdf32[0], user ? G.int_idx32[1] : G.int_idx32[0],
                                      G.E, en.o, en.d, (uint32_t)mulhi64_32(t, (uint32_t)en.T));
  const uint64_t x = mulhi64(t, en.T);
  const uint64_t* C = user ? G.int_cdf[1] : G.int_cdf[0];
  if (mode == kPickIntIdx)
    return cdf_index_search<uint64_t>(C, user ? G.int_idx[1] : G.int_idx[0], G.E, en.o, en.d, x);
  return (int32_t)upper_bound_dev<uint64_t>(C + en.o, en.d, x);
}

__device__ __forceinline__ GEntry shfl_gentry(const GEntry& m, int src) {
  GEntry e;
  e.v = shfl_abs(m.v, src);
  e.d = shfl_abs(m.d, src);
  e.o = (int64_t)shfl_abs_u64((uint64_t)m.o, src);
  e.F = shfl_abs(m.F, src);
  e.g = shfl_abs(m.g, src);
  e.pos = (int64_t)shfl_abs_u64((uint64_t)m.pos, src);
  e.c1 = shfl_abs_u(m.c1, src);
  e.c2 = shfl_abs_u(m.c2, src);
  e.c3 = shfl_abs_u(m.c3, src);
  e.k0 = shfl_abs_u(m.k0, src);
  e.k1 = shfl_abs_u(m.k1, src);
  e.flags = shfl_abs_u(m.flags, src);
  e.T = shfl_abs_u64(m.T, src);
  return e;
}

__device__ __forceinline__ void gwrite_edge(const GraphDev& G, const csaw_graph_view& hv, const HopDev& H,
                                            const GEntry& en, int64_t p, int32_t idx) {
  H.src[p] = en.v;
  H.dst[p] = Update(&hv, G.adj[en.o + idx], en.v);
  H.smp[p] = en.g;
}

// ci equal to one of the group's resolved slots 0 .. jj-1 (res of lane gbase+m)
__device__ __forceinline__ bool group_taken(int32_t ci, int32_t res, int jj, int gbase) {
  bool hit = false;
#pragma unroll 8
  for (int m = 0; m < jj; ++m) hit |= (ci == shfl_abs(res, gbase + m));
  return hit;
}

// Small-degree entries with 32-bit CDF keys: the group copies the entry's CDF
// into its LDS slot with coalesced loads and every draw is an upper_bound in
// LDS, with no further global access.  The wave's LDS words are shared out
// among its groups, so the slot (the largest degree taking this path) is
// kGroupLdsWords / (64 / GS) words: 64 for groups of 4, 85 for fanout 5,
// 170 for fanout 10, 256 for 16 and 512 for 32 lanes.
constexpr int kGroupSmallCap = 64;  // slot words at the smallest group size

template <int KIND>
__device__ __forceinline__ int32_t gpick_small(const GEntry& en, const uint32_t* lc, const U4& r) {
  if constexpr (KIND == kF32) {
    const float x = __fmul_rn(philox_u24(r), __uint_as_float((uint32_t)en.T));
    return (int32_t)upper_bound_dev<float>(reinterpret_cast<const float*>(lc), en.d, x);
  } else {
    const uint32_t x = (uint32_t)mulhi64_32(philox_t64(r), (uint32_t)en.T);  // uint32 cache: T < 2^32
    return (int32_t)upper_bound_dev<uint32_t>(lc, en.d, x);
  }
}

// lc: the group's LDS slot of cap words (null: no small path)
// gs: group size (4 .. 32, or 64), a run-time value so one build serves all;
// the group is lanes gbase .. gbase + gs - 1 (gmask)
template <int KIND, int KEYS>
__device__ int32_t group_select(const GraphDev& G, const GEntry& en, int gs, int gl, int gbase, uint64_t gmask,
                                uint32_t* lc, int cap, int32_t* cl) {
  const int F = en.F;
  const int32_t d = en.d;
  const bool act = gl < F;
  const uint32_t mode = (en.flags >> 4) & 15u;
  const bool small = KIND != kUniform && lc && d <= cap &&
                     (mode == kPickInt32Idx || mode == kPickF32 || mode == kPickF32Idx);
  if (small) {
    const uint32_t* C = KIND == kF32 ? reinterpret_cast<const uint32_t*>(G.cdf)
                                     : ((en.flags >> 1) & 1u ? G.int_cdf32[1] : G.int_cdf32[0]);
    for (int i = gl; i < d; i += gs) lc[i] = C[en.o + i];
    wave_fence();
  }
  // one draw (slot j, attempt a); only `valid` lanes' results are used
  auto pick = [&](int j, int a, bool valid) -> int32_t {
    if (!valid) return 0;
    if (small) return gpick_small<KIND>(en, lc, gdraw(en, j, a));
    return gpick<KIND, KEYS>(G, en, gdraw(en, j, a));
  };
  int32_t cand = pick(gl, 0, act);
  if (!act) cand = -1;
  if (!(en.flags & 1u)) return cand;
  // duplicate check: every lane writes its candidate into the group's
  // 16-byte-aligned LDS slot, then reads the group's candidates back four
  // at a time (one broadcast ds_read_b128 per four slots instead of one
  // ds_bpermute per slot) and compares with the slots before its own
  cl[gl] = cand;
  wave_fence();
  bool dup = false;
  for (int i0 = 0; i0 + 1 < F; i0 += 4) {
    const int4 q = *reinterpret_cast<const int4*>(cl + i0);
    dup |= (gl > i0 && cand == q.x) | (gl > i0 + 1 && cand == q.y) | (gl > i0 + 2 && cand == q.z) |
           (gl > i0 + 3 && cand == q.w);
  }
  dup &= act;
  wave_fence();  // every lane has read the slot before the next round writes it
  const uint64_t dm = ballot(dup) & gmask;
  if (dm == 0ull) return cand;
  const int j0 = __builtin_ctzll(dm) - gbase;
  int32_t res = gl < j0 ? cand : -1;
  // Taken set: for d <= 128 a 128-bit mask held identically by the group's
  // lanes (O(1) tests, no shuffles; where collisions are frequent), else the
  // resolved lanes' registers.
  const bool use_mask = d <= 128;
  uint64_t mlo = 0, mhi = 0;
  if (use_mask) {
    for (int m = 0; m < j0; ++m) {  // the resolved slots 0 .. j0-1
      const int32_t cm = shfl_abs(cand, gbase + m);
      if (cm < 64) mlo |= 1ull << cm;
      else mhi |= 1ull << (cm - 64);
    }
  }
  auto in_mask = [&](int32_t ci) -> bool { return ((ci < 64 ? mlo >> ci : mhi >> (ci - 64)) & 1ull) != 0ull; };
  for (int jj = j0; jj < F; ++jj) {
    int32_t c = shfl_abs(cand, gbase + jj);
    const bool c_taken = use_mask ? in_mask(c) : ((ballot(res == c) & gmask) != 0ull);
    if (c_taken) {
      int32_t chosen = -1, last = c;
      for (int a0 = 1; a0 < kReselect; a0 += gs) {
        const int a = a0 + gl;
        int32_t ci = pick(jj, a, a < kReselect);
        if (a >= kReselect) ci = 0;
        const bool tk = use_mask ? in_mask(ci) : group_taken(ci, res, jj, gbase);
        const uint64_t m = ballot((a < kReselect) && !tk) & gmask;
        if (m) {
          chosen = shfl_abs(ci, __builtin_ctzll(m));
          break;
        }
        const int ll = kReselect - 1 - a0;
        if (ll < gs) last = shfl_abs(ci, gbase + ll);
      }
      if (chosen < 0) {
        for (int q = 0; q < d; q += gs) {
          const int32_t i = (int32_t)(((int64_t)last + 1 + q + gl) % d);
          const bool tk = use_mask ? in_mask(i) : group_taken(i, res, jj, gbase);
          const uint64_t m = ballot((q + gl < d) && !tk) & gmask;
          if (m) {
            chosen = shfl_abs(i, __builtin_ctzll(m));
            break;
          }
        }
      }
      c = chosen;
    }
    if (use_mask) {
      if (c < 64) mlo |= 1ull << c;
      else mhi |= 1ull << (c - 64);
    }
    if (gl == jj) res = c;
  }
  return res;
}

// Expands `count` (<= 64) entries, all with F <= gs, in groups of gs lanes
// (4 <= gs <= 32, or 64): 64 / gs groups, the remaining 64 mod gs lanes idle.
// Needs no per-entry CDF build: uniform, FP32, or integer biases with the CDF
// cache.  lds: the wave's kGroupLdsWords words for small-degree CDFs (null:
// searches only).
constexpr int kGroupLdsWords = (kWave / kGroupMin) * kGroupSmallCap;
// per-wave candidate slots: NG groups of gs lanes rounded up to four words,
// NG * round4(gs) <= 64 + 3 * NG <= 112 for gs >= 4 (the idle remainder
// lanes never select, so their slot pointer is never dereferenced)
constexpr int kCandSlotWords = 128;

template <int KIND, int KEYS = kKeysAny>
__device__ void group_expand_batch(const GraphDev& G, const csaw_graph_view& hv, const ReqDev* __restrict__ R,
                                   const int32_t* __restrict__ gs2req, const HopDev& H,
                                   const int32_t* __restrict__ order, int64_t first, int count, int gs,
                                   uint32_t* lds) {
  const int NG = kWave / gs;
  const int lane = lane_id();
  const int gi = lane / gs;  // group; gi == NG: the idle remainder lanes
  const int gl = lane - gi * gs, gbase = gi * gs;
  const int cap = kGroupLdsWords / NG;  // LDS slot words per group
  uint32_t* lc = lds && gi < NG ? lds + gi * cap : nullptr;
  // candidate slots of the duplicate check: per wave, each group's gs words
  // rounded up to four (16-byte aligned); the idle remainder lanes' slot
  // follows the groups'
  __shared__ __align__(16) int32_t cand_slots[4][kCandSlotWords];
  const int gstride = (gs + 3) & ~3;
  int32_t* gcl = &cand_slots[(threadIdx.x / kWave) & 3][gi * gstride];
  const uint64_t gmask = gs >= kWave ? ~0ull : (((1ull << gs) - 1ull) << gbase);
  GEntry mine;
  mine.d = 0;
  mine.F = 0;
  uint32_t warm = 0;
  if (lane < count) {
    const Entry en0 = load_entry(G, R, gs2req, H, order ? (int64_t)order[first + lane] : first + lane);
    mine = make_gentry<KIND>(G, en0);
    if (H.only_kind && bias_kind(en0.bias) + 1 != H.only_kind) {  // another kind's launch
      mine.d = 0;
      mine.flags &= ~0xF0u;  // no search state, no warm-up load
    }
    // Every entry's first search block (or, for small degrees, its CDF) is
    // loaded now, for all 64 entries at once, so the rounds below find it in
    // cache instead of each waiting for memory.
    const uint32_t mode = (mine.flags >> 4) & 15u;
    if (KIND == kF32 && mode == kPickF32Idx)
      warm = *reinterpret_cast<const uint32_t*>(
          mine.d <= cap ? G.cdf + mine.o : cdf_index_first_block<float>(G.cdf, G.cdf_idx, G.E, mine.o, mine.d));
    if (KIND == kInt && mode == kPickInt32Idx) {
      const bool user = (mine.flags >> 1) & 1u;
      const uint32_t* C = user ? G.int_cdf32[1] : G.int_cdf32[0];
      warm = mine.d <= cap
                 ? C[mine.o]
                 : *cdf_index_first_block<uint32_t>(C, user ? G.int_idx32[1] : G.int_idx32[0], G.E, mine.o, mine.d);
    }
  }
  // Entries that copy all their d <= F neighbours take no group round: the
  // wave writes their edges in a flat pass, lane s handling copied edge s of
  // the batch (its entry found by a binary search over the running edge
  // counts), and the selecting entries are compacted so that every group
  // round expands one.
  const bool live = lane < count && mine.d > 0 && mine.F > 0;
  const bool cp = live && (mine.flags & 1u) && mine.d <= mine.F;
  const uint32_t kc = cp ? (uint32_t)mine.d : 0u;
  const uint32_t kci = wave_incl_scan_u32(kc);
  const uint32_t ncopy = readlane_u(kci, kWave - 1);
  for (uint32_t s0 = 0; s0 < ncopy; s0 += kWave) {
    const uint32_t s = s0 + (uint32_t)lane;
    int L = 0;  // first lane whose running count exceeds s
#pragma unroll
    for (int b = 32; b >= 1; b >>= 1)
      if (shfl_abs_u(kci, L + b - 1) <= s) L += b;
    L = L < kWave ? L : kWave - 1;
    const uint32_t before = shfl_abs_u(kci, L) - shfl_abs_u(kc, L);
    const int32_t cv = shfl_abs(mine.v, L), cg = shfl_abs(mine.g, L);
    const int64_t co = (int64_t)shfl_abs_u64((uint64_t)mine.o, L);
    const int64_t cpos = (int64_t)shfl_abs_u64((uint64_t)mine.pos, L);
    if (s < ncopy) {
      const int64_t i = (int64_t)(s - before);
      H.src[cpos + i] = cv;
      H.dst[cpos + i] = Update(&hv, G.adj[co + i], cv);
      H.smp[cpos + i] = cg;
    }
  }
  const bool selects = live && !cp;
  const uint64_t sm = ballot(selects);
  const int nsel = __popcll(sm);
  // lane r receives the lane of the r-th selecting entry (ds_permute pushes
  // every lane to a distinct destination: selecting lanes to their rank,
  // the others after them)
  const int rank = __popcll(sm & ((1ull << lane) - 1ull));
  const int dest = selects ? rank : nsel + (lane - rank);
  const int sel_lane = __builtin_amdgcn_ds_permute(dest << 2, lane);
  // Software pipeline: a round's neighbour load is consumed (written out) only
  // after the next round's selection, so its latency overlaps that round's
  // searches instead of adding to them.
  bool have = false;
  int32_t pv = 0, pg = 0, pnb = 0;
  int64_t pp = 0;
  for (int e0 = 0; e0 < nsel; e0 += NG) {
    const int e = gi < NG ? e0 + gi : nsel;
    const int src = shfl_abs(sel_lane, e < nsel ? e : 0);
    const GEntry en = shfl_gentry(mine, src);
    int32_t idx = -1;
    if (e < nsel) {
      const int32_t res = group_select<KIND, KEYS>(G, en, gs, gl, gbase, gmask, lc, cap, gcl);
      idx = gl < en.F ? res : -1;
    }
    if (have) {
      H.src[pp] = pv;
      H.dst[pp] = pnb;
      H.smp[pp] = pg;
    }
    have = idx >= 0;
    if (have) {
      pv = en.v;
      pg = en.g;
      pp = en.pos + gl;
      pnb = Update(&hv, G.adj[en.o + idx], en.v);
    }
  }
  if (have) {
    H.src[pp] = pv;
    H.dst[pp] = pnb;
    H.smp[pp] = pg;
  }
  // consume the warm-up loads where their wait costs nothing (never true:
  // count >= 0; the compiler cannot drop the loads)
  if (count < 0 && warm == 0x5eedu) atomicOr(H.flags, 0u);
}

// ---------------------------------------------------------------------------
// Lane-level selection (THREAD variant).  Valid for F <= kThreadFanCap and,
// for integer biases, d <= kThreadIntCap (weights cached in LDS, transposed
// per lane).  Taken set: a 64-bit mask when d <= 64, else a per-lane LDS list.
// ---------------------------------------------------------------------------
template <int KIND>
__device__ void expand_entry_thread(const GraphDev& G, const csaw_graph_view& hv, const HopDev& H, const Entry& en,
                                    uint32_t* lds, int stride) {
  if (en.d == 0 || en.F == 0) return;
  const int32_t d = en.d;
  if (en.dedup && d <= en.F) {
    for (int32_t i = 0; i < d; ++i) write_edge(G, hv, H, en, en.pos + i, i);
    return;
  }
  // picker (PickMode; kPickLds here is a linear scan over LDS weights).
  // Cf / Cg point at the entry's CDF, Kf / Kg at the index of the whole array.
  int mode = kPickUniform;
  uint64_t T = 0;
  float Tf = 0.f;
  const float* Cf = nullptr;
  const uint64_t* Cg = nullptr;
  const float* Kf = nullptr;
  const uint64_t* Kg = nullptr;
  const int ck = en.bias == CSAW_BIAS_USER ? 1 : 0;
  // Small degrees with a 32-bit CDF in memory: the lane copies it into its LDS
  // column once and every draw (repeated ones included) searches LDS.
  constexpr int kPickLdsCdf = 7;
  const uint32_t* small = nullptr;
  if (d <= kThreadIntCap) {
    if (KIND == kF32) small = reinterpret_cast<const uint32_t*>(G.cdf);
    else if (KIND == kInt) small = G.int_cdf32[ck];
  }
  if (small) {
    for (int32_t i = 0; i < d; ++i) lds[i * stride] = small[en.o + i];
    const uint32_t last = lds[(d - 1) * stride];
    if (KIND == kF32) {
      Tf = __uint_as_float(last);
      if (Tf > 0.0f) mode = kPickLdsCdf;
    } else {
      T = last;
      if (T > 0) mode = kPickLdsCdf;
    }
  } else if (KIND == kF32) {
    Cf = G.cdf + en.o;
    Kf = G.cdf_idx;
    Tf = Cf[d - 1];
    if (Tf > 0.0f) mode = Kf ? kPickF32Idx : kPickF32;
  } else if (KIND == kInt && G.int_cdf32[ck]) {
    T = G.int_cdf32[ck][en.o + d - 1];
    if (T > 0) mode = kPickInt32Idx;
  } else if (KIND == kInt && G.int_cdf[ck]) {
    Cg = G.int_cdf[ck] + en.o;
    Kg = G.int_idx[ck];
    T = Cg[d - 1];
    if (T > 0) mode = Kg ? kPickIntIdx : kPickInt;
  } else if (KIND == kInt) {
    for (int32_t i = 0; i < d; ++i) {
      const uint32_t w = int_bias(G, hv, en.bias, G.adj[en.o + i]);
      lds[i * stride] = w;
      T += w;
    }
    if (T > 0) mode = kPickLds;
  }
  auto pick = [&](const U4& o) -> int32_t {
    const uint64_t t = philox_t64(o);
    if (mode == kPickUniform) return (int32_t)mulhi64_32(t, (uint32_t)d);
    if (mode == kPickLdsCdf) {
      // upper_bound over the lane's LDS column
      const float xf = KIND == kF32 ? __fmul_rn(philox_u24(o), Tf) : 0.f;
      const uint32_t xu = KIND == kF32 ? 0u : (uint32_t)mulhi64(t, T);
      int32_t lo = 0, hi = d;
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        const uint32_t v = lds[mid * stride];
        const bool gt = KIND == kF32 ? __uint_as_float(v) > xf : v > xu;
        if (gt) hi = mid; else lo = mid + 1;
      }
      return lo;
    }
    // (each kind compiles only its own modes: the code stays small)
    if constexpr (KIND == kF32) {
      if (mode == kPickF32) return (int32_t)upper_bound_dev<float>(Cf, d, __fmul_rn(philox_u24(o), Tf));
      return cdf_index_search<float>(G.cdf, Kf, G.E, en.o, d, __fmul_rn(philox_u24(o), Tf));
    }
    if constexpr (KIND == kInt) {
      const uint64_t x = mulhi64(t, T);
      if (mode == kPickInt) return (int32_t)upper_bound_dev<uint64_t>(Cg, d, x);
      if (mode == kPickIntIdx) return cdf_index_search<uint64_t>(G.int_cdf[ck], Kg, G.E, en.o, d, x);
      if (mode == kPickInt32Idx)
        return cdf_index_search<uint32_t>(G.int_cdf32[ck], G.int_idx32[ck], G.E, en.o, d, (uint32_t)x);
      uint64_t acc = 0;
      int32_t i = 0;
      for (; i < d - 1; ++i) {
        acc += lds[i * stride];
        if (acc > x) break;
      }
      return i;
    }
    return 0;
  };
  // Four slots at a time: their attempt-0 searches advance in lockstep, so
  // each lane keeps up to four independent loads in flight.
  constexpr int G4 = 4;
  auto pick4 = [&](int j0, int nq, int32_t* out) {
    // (the index searches are instantiated only for their own kind)
    if (KIND == kInt && mode == kPickInt32Idx) {
      if constexpr (KIND == kInt) {
        uint32_t xi[G4];
#pragma unroll
        for (int q = 0; q < G4; ++q) xi[q] = q < nq ? (uint32_t)mulhi64(philox_t64(draw(en, j0 + q, 0)), T) : 0u;
        cdf_index_search_n<uint32_t, G4>(G.int_cdf32[ck], G.int_idx32[ck], G.E, en.o, d, xi, nq, out);
      }
    } else if (KIND == kInt && mode == kPickIntIdx) {
      if constexpr (KIND == kInt) {
        // 64-bit keys: two pairs in lockstep (16 keys of a search take 32 registers)
        uint64_t xi[G4];
#pragma unroll
        for (int q = 0; q < G4; ++q) xi[q] = q < nq ? mulhi64(philox_t64(draw(en, j0 + q, 0)), T) : 0;
        cdf_index_search_n<uint64_t, 2>(G.int_cdf[ck], Kg, G.E, en.o, d, xi, min(nq, 2), out);
        if (nq > 2) cdf_index_search_n<uint64_t, 2>(G.int_cdf[ck], Kg, G.E, en.o, d, xi + 2, nq - 2, out + 2);
      }
    } else if (KIND == kF32 && mode == kPickF32Idx) {
      if constexpr (KIND == kF32) {
        float xf[G4];
#pragma unroll
        for (int q = 0; q < G4; ++q) xf[q] = q < nq ? __fmul_rn(philox_u24(draw(en, j0 + q, 0)), Tf) : 0.f;
        cdf_index_search_n<float, G4>(G.cdf, Kf, G.E, en.o, d, xf, nq, out);
      }
    } else if (KIND != kUniform && (mode == kPickInt || mode == kPickF32)) {
      int32_t lo[G4], hi[G4];
      uint64_t xi[G4];
      float xf[G4];
#pragma unroll
      for (int q = 0; q < G4; ++q) {
        lo[q] = 0;
        hi[q] = q < nq ? d : 0;
        xi[q] = 0;
        xf[q] = 0.f;
        if (q < nq) {
          const U4 o = draw(en, j0 + q, 0);
          if (mode == kPickInt) xi[q] = mulhi64(philox_t64(o), T);
          else xf[q] = __fmul_rn(philox_u24(o), Tf);
        }
      }
      for (;;) {
        bool any = false;
        uint64_t vi[G4];
        float vf[G4];
        int32_t mid[G4];
#pragma unroll
        for (int q = 0; q < G4; ++q) {
          mid[q] = (lo[q] + hi[q]) >> 1;
          if (lo[q] < hi[q]) {
            any = true;
            if (mode == kPickInt) vi[q] = Cg[mid[q]];
            else vf[q] = Cf[mid[q]];
          }
        }
        if (!any) break;
#pragma unroll
        for (int q = 0; q < G4; ++q) {
          if (lo[q] < hi[q]) {
            const bool gt = mode == kPickInt ? vi[q] > xi[q] : vf[q] > xf[q];
            if (gt) hi[q] = mid[q]; else lo[q] = mid[q] + 1;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < G4; ++q) out[q] = lo[q];
    } else {
      // one copy of pick's code (an unrolled loop would inline it four times)
#pragma unroll 1
      for (int q = 0; q < G4; ++q) out[q] = q < nq ? pick(draw(en, j0 + q, 0)) : 0;
    }
  };
  auto emit4 = [&](int j0, int nq, const int32_t* idx) {
    int32_t nb[G4];
#pragma unroll
    for (int q = 0; q < G4; ++q) nb[q] = q < nq ? G.adj[en.o + idx[q]] : 0;
#pragma unroll
    for (int q = 0; q < G4; ++q)
      if (q < nq) {
        const int64_t p = en.pos + j0 + q;
        H.src[p] = en.v;
        H.dst[p] = Update(&hv, nb[q], en.v);
        H.smp[p] = en.g;
      }
  };
  const bool use_mask = d <= 64;
  uint64_t mask = 0;
  int n_taken = 0;
  auto taken = [&](int32_t i) -> bool {
    if (use_mask) return (mask >> i) & 1ull;
    for (int m = 0; m < n_taken; ++m)
      if ((int32_t)lds[m * stride] == i) return true;
    return false;
  };
  // one loop for both modes, so pick4 and emit4 are inlined once
  for (int j0 = 0; j0 < en.F; j0 += G4) {
    const int nq = min(G4, en.F - j0);
    int32_t idx[G4];
    pick4(j0, nq, idx);
    // not unrolled: the reselect loop (with pick inlined) stays one copy, so the
    // kernel's hot code fits the instruction cache
#pragma unroll 1
    for (int q = 0; q < G4; ++q) {
      if (q >= nq || !en.dedup) break;
      int32_t i = idx[q];
      if (taken(i)) {
        bool found = false;
        for (int a = 1; a < kReselect; ++a) {
          i = pick(draw(en, j0 + q, a));
          if (!taken(i)) {
            found = true;
            break;
          }
        }
        if (!found) {
          do {
            i = (i + 1 == d) ? 0 : i + 1;
          } while (taken(i));
        }
      }
      if (use_mask) mask |= 1ull << i;
      else lds[(n_taken++) * stride] = (uint32_t)i;
      idx[q] = i;
    }
    emit4(j0, nq, idx);
  }
}

}  // namespace csaw
