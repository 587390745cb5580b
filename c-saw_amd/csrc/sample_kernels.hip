// Per-hop sampling pipeline (all sizes device-resident, fixed persistent grids,
// so a whole multi-hop run can be captured in one hipGraph):
//
//   k_count<BIN>    entry -> k (edges it emits), block-local exclusive offsets,
//                   block sums; with BIN also a per-block counting sort of the
//                   entries by (bias kind, ceil log2 degree) into chunks.
//   k_scan_blocks   exclusive scan of block sums within groups of 1024 blocks
//   k_scan_groups   exclusive scan of group sums; hop total
//   k_offsets       per-sample edge offsets of the hop (= next frontier offsets)
//   k_choose        variant per chunk (heuristic / forced / random / MLP),
//                   appended to per-variant queues with wave-aggregated atomics
//   k_expand<V,K>   persistent: pulls chunks of its queue and expands them
//   k_expand_plain  one wave per entry in frontier order (the plain sampler)
//
// The reference runs all hops in one persistent warp loop whose output slots
// and queue positions come from atomics (main.cu:21-93, functions.cuh:971-1011);
// here every output position comes from the prefix sums, so the edge list is
// identical whatever order chunks run in.
#include "sample_kernels.hpp"
#include "select.hpp"

namespace csaw {

// ------------------------------------------------------------------ count
// LIST: bins only (no counts) the n_list entries list[0 .. n_list) of one
// partition; order then holds entry ids by list position.
template <bool BIN, bool LIST = false>
__global__ __launch_bounds__(256) void k_count(GraphDev G, const ReqDev* __restrict__ R, const int32_t* __restrict__ gs2req,
                                               HopDev H, const int32_t* __restrict__ list = nullptr,
                                               int64_t n_list = 0) {
  __shared__ uint32_t wsum[4][4];  // [item row][wave] totals of the block scan
  __shared__ uint32_t hist[kNumBuckets];
  __shared__ uint32_t bstart[kNumBuckets];
  __shared__ uint32_t fsum[kNumBuckets];
  __shared__ uint32_t fmax[kNumBuckets];
  __shared__ uint32_t dmax[kNumBuckets];
  __shared__ unsigned long long dsum[kNumBuckets];
  __shared__ uint32_t lsum[kNumBuckets];
  __shared__ uint32_t nz_cnt[2];
  __shared__ uint32_t chunk_base;
  const int64_t n = LIST ? n_list : *H.n_ptr;
  const int64_t nblk = (n + kBlockEntries - 1) >> kBlockShift;
  const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
  for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int64_t base = blk << kBlockShift;
    if (BIN) {
      for (int b = tid; b < kNumBuckets; b += 256) {
        hist[b] = 0; fsum[b] = 0; fmax[b] = 0; dmax[b] = 0; dsum[b] = 0; lsum[b] = 0;
      }
      __syncthreads();
    }
    // item i of thread tid is block position i * 256 + tid: a wave's 64 lanes
    // hold 64 consecutive entries, so the bucket slots one wave instruction
    // hands out keep runs of neighbouring entries (siblings, whose edges
    // are neighbours in the output) together in a chunk
    uint32_t k[4];
    int bk[4];
    uint32_t slot[4];
    int32_t eid[4];
    unsigned long long tbytes = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t pos = base + i * 256 + tid;
      k[i] = 0;
      bk[i] = -1;
      eid[i] = 0;
      if (pos < n) {
        const int64_t e = LIST ? (int64_t)list[pos] : pos;
        eid[i] = (int32_t)e;
        const int32_t v = H.fv[e];
        const ReqDev& r = R[gs2req[H.fs[e]]];
        const int32_t F = r.fanout[H.hop];
        const int32_t d = G.deg[v];
        k[i] = (uint32_t)entry_k(d, F, r.dedup);
        if (!BIN && H.bytes) {
          const int kind = bias_kind(r.bias);
          const int ck = r.bias == CSAW_BIAS_USER ? 1 : 0;
          const bool built = kind == kInt && !G.int_cdf[ck] && !G.int_cdf32[ck] && d <= kWaveIntCap;
          tbytes += k[i] == 0 ? 20ull
                              : expansion_bytes(kind, built, 1, d, k[i], ceil_log2_u32((uint32_t)d),
                                                r.dedup && d <= F);
        }
        if (BIN && k[i] > 0) {
          bk[i] = entry_bucket(d, F, r.dedup, r.bias);
          slot[i] = atomicAdd(&hist[bk[i]], 1u);
          atomicAdd(&fsum[bk[i]], (uint32_t)F);
          atomicMax(&fmax[bk[i]], (uint32_t)F);
          atomicMax(&dmax[bk[i]], (uint32_t)d);
          atomicAdd(&dsum[bk[i]], (unsigned long long)d);
          atomicAdd(&lsum[bk[i]], (uint32_t)ceil_log2_u32((uint32_t)d));
        }
      }
    }
    if (!BIN && H.bytes) {
      for (int off = 32; off >= 1; off >>= 1) tbytes += __shfl_xor(tbytes, off, 64);
      if (lane == 0 && tbytes) atomicAdd(&H.bytes[kKernPlain], tbytes);
    }
    if (!LIST) {
      // block exclusive scan of the per-entry counts in position order:
      // row i (positions i * 256 .. i * 256 + 255) after rows 0 .. i-1
      uint32_t incl[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        incl[i] = wave_incl_scan_u32(k[i]);
        if (lane == kWave - 1) wsum[i][wid] = incl[i];
      }
      __syncthreads();
      uint32_t row = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        uint32_t pre = row + incl[i] - k[i];
        for (int w = 0; w < wid; ++w) pre += wsum[i][w];
        const int64_t e = base + i * 256 + tid;
        if (e < n) H.local_excl[e] = pre;
        row += wsum[i][0] + wsum[i][1] + wsum[i][2] + wsum[i][3];
      }
      if (tid == 0) H.blk_excl[blk] = row;  // block total, scanned later
    } else {
      __syncthreads();  // the histogram is complete
    }
    if (BIN) {
      // exclusive scan of the bucket histogram (threads 0..127 own one bucket)
      uint32_t hv = 0, hs = 0;
      if (tid < kNumBuckets) {
        hv = hist[tid];
        hs = wave_incl_scan_u32(hv);
        if (lane == kWave - 1) nz_cnt[wid] = hs;  // wave totals
      }
      __syncthreads();
      if (tid < kNumBuckets) bstart[tid] = hs - hv + (wid == 1 ? nz_cnt[0] : 0u);
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (bk[i] >= 0) H.order[base + bstart[bk[i]] + slot[i]] = eid[i];
      // chunk descriptors: each non-empty bucket in pieces of at most cmax
      // entries (more, smaller work items when the hop is small); a piece's
      // sums are its share of the bucket's, its maxima the bucket's
      const uint32_t cmax = chunk_max_for(n, H.expand_waves, H.chunk_min);
      uint32_t nck = tid < kNumBuckets ? (hv + cmax - 1) / cmax : 0u;
      const uint32_t cincl = wave_incl_scan_u32(nck);
      __syncthreads();
      if (tid < kNumBuckets && lane == kWave - 1) nz_cnt[wid] = cincl;
      __syncthreads();
      if (tid == 0) chunk_base = atomicAdd(H.n_chunks, nz_cnt[0] + nz_cnt[1]);
      __syncthreads();
      if (tid < kNumBuckets && hv > 0) {
        const uint32_t r0 = (wid == 1 ? nz_cnt[0] : 0u) + cincl - nck;
        for (uint32_t pc = 0; pc < nck; ++pc) {
          const uint32_t ci = chunk_base + r0 + pc;
          const uint32_t len = min(cmax, hv - pc * cmax);
          const float share = (float)len / (float)hv;
          if (ci < H.chunk_cap) {
            Chunk c;
            c.start = (uint32_t)(base + bstart[tid] + pc * cmax);
            c.len = (uint16_t)len;
            c.bucket = (uint8_t)tid;
            c.hop = (uint8_t)H.hop;
            c.sum_f = (uint32_t)((float)fsum[tid] * share + 0.5f);
            c.max_d = dmax[tid];
            c.sum_d = (float)dsum[tid] * share;
            c.key = (uint32_t)((blk * kNumBuckets + tid) * (kBlockEntries / kChunkMin) + pc);
            c.max_f = (uint8_t)fmax[tid];
            c.pad = 0;
            c.sum_l = (uint16_t)min((uint32_t)((float)lsum[tid] * share + 0.5f), 65535u);
            H.chunks[ci] = c;
          } else {
            atomicOr(H.flags, 1u);
          }
        }
      }
    }
    __syncthreads();  // every wave is done with wsum and the histograms before the next block
  }
}

// -------------------------------------------------------------- scans
__global__ __launch_bounds__(256) void k_scan_blocks(HopDev H) {
  __shared__ uint32_t wsum[4];
  const int64_t n = *H.n_ptr;
  const int64_t nblk = (n + kBlockEntries - 1) >> kBlockShift;
  const int64_t ngrp = (nblk + 1023) >> 10;
  const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
  for (int64_t grp = blockIdx.x; grp < ngrp; grp += gridDim.x) {
    const int64_t b0 = grp << 10;
    uint32_t x[4], s = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t b = b0 + tid * 4 + i;
      x[i] = b < nblk ? H.blk_excl[b] : 0u;
      s += x[i];
    }
    const uint32_t incl = wave_incl_scan_u32(s);
    if (lane == kWave - 1) wsum[wid] = incl;
    __syncthreads();
    uint32_t pre = incl - s;
    for (int w = 0; w < wid; ++w) pre += wsum[w];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t b = b0 + tid * 4 + i;
      if (b < nblk) H.blk_excl[b] = pre;
      pre += x[i];
    }
    if (tid == 255) H.grp_excl[grp] = pre;  // group total
    __syncthreads();
  }
}

// one workgroup: exclusive scan of group totals; writes the hop total
__global__ __launch_bounds__(256) void k_scan_groups(HopDev H, int64_t* __restrict__ total) {
  __shared__ uint64_t wsum[4];
  __shared__ uint64_t carry;
  const int64_t n = *H.n_ptr;
  const int64_t nblk = (n + kBlockEntries - 1) >> kBlockShift;
  const int64_t ngrp = (nblk + 1023) >> 10;
  const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int64_t g0 = 0; g0 < ngrp; g0 += 256) {
    const int64_t gi = g0 + tid;
    const uint64_t x = gi < ngrp ? H.grp_excl[gi] : 0ull;
    const uint64_t incl = wave_incl_scan_u64(x);
    if (lane == kWave - 1) wsum[wid] = incl;
    __syncthreads();
    uint64_t pre = carry + incl - x;
    for (int w = 0; w < wid; ++w) pre += wsum[w];
    if (gi < ngrp) H.grp_excl[gi] = pre;
    __syncthreads();
    if (tid == 255) carry = pre + x;
    __syncthreads();
  }
  if (tid == 0) *total = (int64_t)carry;
}

// per-sample edge offsets for samples [0, gb] of this hop
__global__ void k_offsets(HopDev H, const int64_t* __restrict__ total) {
  const int64_t n = *H.n_ptr;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g <= H.gb; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = H.fo[g];
    H.eo[g] = x >= n ? *total
                     : (int64_t)H.grp_excl[x >> kGroupShift] + H.blk_excl[x >> kBlockShift] + H.local_excl[x];
  }
}

// ------------------------------------------------------------ digest
// Order-sensitive digest of an edge range: sum over positions of a 64-bit
// mix of (position, src, dst, sample); the sum makes it independent of how
// the reduction is scheduled.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

__global__ void k_digest(const int32_t* __restrict__ src, const int32_t* __restrict__ dst,
                         const int32_t* __restrict__ smp, int64_t lo, int64_t hi, int64_t pos0, int64_t gbase,
                         unsigned long long* __restrict__ out) {
  uint64_t acc = 0;
  for (int64_t i = lo + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hi; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t a = (uint64_t)(uint32_t)src[i] | ((uint64_t)(uint32_t)dst[i] << 32);
    const uint64_t b = (uint64_t)(pos0 + i - lo) * 0x9E3779B97F4A7C15ull ^ (uint64_t)(uint32_t)(smp[i] - gbase);
    acc += mix64(a ^ mix64(b));
  }
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane_id() == 0 && acc) atomicAdd(out, (unsigned long long)acc);
}

int launch_digest(hipStream_t st, int num_cu, const int32_t* src, const int32_t* dst, const int32_t* smp, int64_t lo,
                  int64_t hi, int64_t pos0, int64_t gbase, unsigned long long* out) {
  if (hi <= lo) return CSAW_OK;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((hi - lo + 255) / 256, (int64_t)num_cu * 8));
  k_digest<<<grid, 256, 0, st>>>(src, dst, smp, lo, hi, pos0, gbase, out);
  CSAW_LAUNCH_CHECK();
  return CSAW_OK;
}

// --------------------------------------------------------- fills / resets
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// ----------------------------------------------------------- expand plain
// gs: group size for the hop's fanout (group_size_for; 0 = the whole-wave
// path), used for the kinds that need no per-entry CDF build (uniform, FP32,
// and integer biases when the CDF cache is on).  IDX: the graph view carries
// the CDF cache and its 16-ary index (searched through the index).  KEYS:
// kKeys32 when no integer cache of the view is 64-bit (the 64-bit searches
// compiled out of the group path).
template <int KIND, bool IDX, int KEYS>
__global__ __launch_bounds__(256) void k_expand_plain(GraphDev G, csaw_graph_view hv, const ReqDev* __restrict__ R,
                                                      const int32_t* __restrict__ gs2req, HopDev H, int gs) {
  constexpr bool kGroupOk = KIND != kInt || IDX;            // no per-entry CDF build needed
  constexpr bool kGroupLds = KIND == kF32 || (KIND == kInt && IDX);  // small cached CDFs staged in LDS
  __shared__ uint64_t cdf[KIND == kInt ? 4 * kWaveIntCap : 1];
  __shared__ uint32_t bms[4 * (kWaveBitmapCap / 32)];
  __shared__ uint32_t glds[kGroupLds ? 4 * kGroupLdsWords : 1];
  const int64_t n = *H.n_ptr;
  const int wid = threadIdx.x / kWave;
  const int64_t nw = (int64_t)gridDim.x * 4;
  // batch size adapts to the frontier: one entry (one group round) per wave
  // when it is small (all waves busy), up to 64 entries per metadata round
  // when it is large
  const bool grp = kGroupOk && gs > 0;
  const int64_t per = grp ? kWave / gs : 1;
  const int64_t bs = min((int64_t)kWave, max(per, (n + nw - 1) / nw));
  uint32_t* gl = kGroupLds ? glds + wid * kGroupLdsWords : nullptr;
  for (int64_t e0 = (blockIdx.x * 4 + wid) * bs; e0 < n; e0 += nw * bs) {
    const int cnt = (int)min(bs, n - e0);
    if constexpr (kGroupOk) {
      if (grp) {
        group_expand_batch<KIND, KEYS>(G, hv, R, gs2req, H, nullptr, e0, cnt, gs, gl);
        continue;
      }
    }
    wave_expand_batch<KIND, IDX>(G, hv, R, gs2req, H, nullptr, e0, cnt, KIND == kInt ? cdf + wid * kWaveIntCap : nullptr,
                                 bms + wid * (kWaveBitmapCap / 32));
  }
}

// one wave per 64 listed entries (the streaming path: entries of one partition)
template <int KIND>
__global__ __launch_bounds__(256) void k_expand_list(GraphDev G, csaw_graph_view hv, const ReqDev* __restrict__ R,
                                                     const int32_t* __restrict__ gs2req, HopDev H,
                                                     const int32_t* __restrict__ list, int64_t n) {
  __shared__ uint64_t cdf[KIND == kInt ? 4 * kWaveIntCap : 1];
  __shared__ uint32_t bms[4 * (kWaveBitmapCap / 32)];
  const int wid = threadIdx.x / kWave;
  const int64_t nw = (int64_t)gridDim.x * 4;
  const int64_t bs = min((int64_t)kWave, max((int64_t)1, (n + nw - 1) / nw));
  for (int64_t e0 = (blockIdx.x * 4 + wid) * bs; e0 < n; e0 += nw * bs) {
    const int cnt = (int)min(bs, n - e0);
    wave_expand_batch<KIND, false>(G, hv, R, gs2req, H, list, e0, cnt, KIND == kInt ? cdf + wid * kWaveIntCap : nullptr,
                            bms + wid * (kWaveBitmapCap / 32));
  }
}

int launch_expand_list(const LaunchCfg& L, int kind, const GraphDev& G, const csaw_graph_view& hv, const ReqDev* R,
                       const int32_t* gs2req, const HopDev& H, const int32_t* list, int64_t n) {
  if (n <= 0) return CSAW_OK;
  const int64_t nwg = (n + 4 * kWave - 1) / (4 * kWave);
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(nwg, (int64_t)L.num_cu * L.expand_per_cu));
  if (kind == kUniform) k_expand_list<kUniform><<<grid, 256, 0, L.stream>>>(G, hv, R, gs2req, H, list, n);
  else if (kind == kInt) k_expand_list<kInt><<<grid, 256, 0, L.stream>>>(G, hv, R, gs2req, H, list, n);
  else k_expand_list<kF32><<<grid, 256, 0, L.stream>>>(G, hv, R, gs2req, H, list, n);
  CSAW_LAUNCH_CHECK();
  return CSAW_OK;
}

// ------------------------------------------------------------- choose
// THREAD: fanout <= 32 and, for integer biases without the CDF cache,
// degree <= 32 (per-lane LDS weights).  BLOCK only pays for integer biases
// (256 lanes build the CDF); for uniform and FP32 it would idle 3 waves.
__device__ __forceinline__ bool variant_valid(int var, int kind, const Chunk& c, bool int_cache) {
  if (var == CSAW_VAR_THREAD)
    return c.max_f <= kThreadFanCap && (kind != kInt || int_cache || c.max_d <= (uint32_t)kThreadIntCap);
  if (var == CSAW_VAR_BLOCK_BITMAP || var == CSAW_VAR_BLOCK_HASH) return kind == kInt;
  return true;
}

// The fixed heuristic (C-SAW degree thresholds SML_MID=32, MID_LRG=1024,
// util.h:19-20, with the reference's default bitmap dedup, main.cu:26).
__device__ int heuristic_variant(int kind, const Chunk& c, bool cache) {
  const uint32_t d = c.max_d;
  if (d < 32 && variant_valid(CSAW_VAR_THREAD, kind, c, cache)) return CSAW_VAR_THREAD;
  if (d < 1024 || !variant_valid(CSAW_VAR_BLOCK_BITMAP, kind, c, cache)) return CSAW_VAR_WAVE_BITMAP;
  return CSAW_VAR_BLOCK_BITMAP;
}

// Learned choice: argmin of the predicted log2 time over the valid variants
// (wl: the BF16 weights widened to FP32 in LDS).
__device__ int mlp_choose(const ChooseArgs& A, const float* wl, const Chunk& c, int kind, const HopDev& H) {
  float f[kMlpIn], o[kMlpOut];
  chunk_features(c.len, c.sum_d, c.max_d, c.sum_f, c.max_f, kind, c.hop, c.bucket, *H.n_chunks, f, A.in_place);
  mlp_forward1_f32(wl, f, o);
  int best = -1;
  float bv = 3.0e38f;
  for (int v = 0; v < kMlpOut; ++v)
    if (variant_valid(v, kind, c, A.int_cache) && o[v] < bv) {
      bv = o[v];
      best = v;
    }
  return best >= 0 ? best : heuristic_variant(kind, c, A.int_cache);
}

__global__ __launch_bounds__(256) void k_choose(ChooseArgs A, HopDev H) {
  __shared__ float wl[kMlpParams];
  const bool use_mlp = A.sched == CSAW_SCHED_MLP && A.mlp_w;
  if (use_mlp) {
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) wl[i] = bf16_to_f32(A.mlp.w[i]);
    __syncthreads();
  }
  const uint32_t nc = min(*H.n_chunks, H.chunk_cap);
  for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci - threadIdx.x < nc; ci += gridDim.x * blockDim.x) {
    const bool act = ci < nc;
    int q = -1;
    // algorithmic bytes of the chunk for the kernel kind that will run it
    // (timing runs only), summed per wave before one atomic per kind
    unsigned long long cb = 0;
    int ck = -1;
    if (act) {
      const Chunk c = H.chunks[ci];
      if (c.bucket == kBucketCopy) {
        q = kQueueCopy;
        if (H.bytes) {
          ck = kKernCopy;
          cb = (unsigned long long)expansion_bytes(0, false, c.len, (int64_t)c.sum_d, 0, 0, true);
        }
      } else {
        const int kind = bucket_kind(c.bucket);
        int var = heuristic_variant(kind, c, A.int_cache);
        if (c.bucket >= kBucketRemap0) {
          var = CSAW_VAR_WAVE_BITMAP;  // region-remap entries: one wave each, their own queue
        } else if (A.sched == CSAW_SCHED_FORCED) {
          if (variant_valid(A.variant, kind, c, A.int_cache)) var = A.variant;
        } else if (A.sched == CSAW_SCHED_RANDOM) {
          const U4 o = philox4x32_10(c.key, (uint32_t)c.hop, 0u, kDomExplore << 28, (uint32_t)A.explore_key,
                                     (uint32_t)(A.explore_key >> 32));
          int cand = (int)(o.x % CSAW_NUM_VARIANTS);
          for (int t = 0; t < CSAW_NUM_VARIANTS && !variant_valid(cand, kind, c, A.int_cache); ++t)
            cand = (cand + 1) % CSAW_NUM_VARIANTS;
          var = cand;
        } else if (use_mlp) {
          var = mlp_choose(A, wl, c, kind, H);
        }
        q = c.bucket >= kBucketRemap0 ? kQueueRemap0 + kind : var * 3 + kind;
        if (A.chosen) A.chosen[ci] = (uint8_t)var;
        if (H.bytes) {
          // mean probes per draw: the chunk's mean ceil(log2 d)
          const int l = (c.sum_l + c.len / 2) / max((int)c.len, 1);
          const int kk = var == CSAW_VAR_THREAD ? kKernThread : (var <= CSAW_VAR_WAVE_HASH ? kKernWave : kKernBlock);
          const int cap_l = var == CSAW_VAR_THREAD ? 5 : (var <= CSAW_VAR_WAVE_HASH ? 10 : 13);
          const bool built = kind == kInt && !A.int_cache && class_max_log(bucket_class(c.bucket)) <= cap_l;
          ck = kk;
          cb = (unsigned long long)expansion_bytes(kind, built, c.len, (int64_t)c.sum_d, c.sum_f, l, false);
        }
      }
    }
    if (H.bytes) {
      const int kinds[4] = {kKernCopy, kKernThread, kKernWave, kKernBlock};
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        unsigned long long x = ck == kinds[t] ? cb : 0ull;
        for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane_id() == 0 && x) atomicAdd(&H.bytes[kinds[t]], x);
      }
    }
    // wave-aggregated append to queue q
    uint64_t active = ballot(act && q >= 0);
    while (active) {
      const int leader = __builtin_ctzll(active);
      const int lq = readlane(q, leader);
      const uint64_t m = ballot(act && q == lq);
      uint32_t base = 0;
      if (lane_id() == leader) base = atomicAdd(&A.qlen[lq], (uint32_t)__popcll(m));
      base = readlane_u(base, leader);
      if (act && q == lq) A.queues[(int64_t)lq * A.queue_cap + base + __popcll(m & ((1ull << lane_id()) - 1ull))] = ci;
      active &= ~m;
    }
  }
}

// ------------------------------------------------------------- expand
struct ChunkCursor {
  uint32_t ci;     // chunk index
  bool valid;
};

__device__ __forceinline__ ChunkCursor next_chunk(const ExpandArgs& A, int q, uint32_t* slot) {
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0xFFFFFFFFu;
    const uint32_t len = A.qlen[q];
    // an exhausted queue is seen without an atomic on its head
    if (__hip_atomic_load(&A.qhead[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < len) {
      const uint32_t i = atomicAdd(&A.qhead[q], 1u);
      if (i < len) s = A.queues[(int64_t)q * A.queue_cap + i];
    }
    *slot = s;
  }
  __syncthreads();
  ChunkCursor c;
  c.ci = *slot;
  c.valid = c.ci != 0xFFFFFFFFu;
  return c;
}

// Wave-level work queue.  Every claim is an atomic on the queue's one head
// word, and atomics on one address serialise in its L2 channel, so a wave
// claims a run of chunks at once while the queue is long: half of its share
// of what is left (at most kClaimMax), down to single chunks at the tail
// (C4: a quarter of the share 27.2-27.6 ms per step, half 25.2-25.8, all of
// it 25.9-26.7).  A head already past the end is seen with a plain (L2) load and
// costs no atomic.  Lane 0 claims; every lane gets the same answer (no
// workgroup barrier).  Which wave expands which chunk never changes the
// edges (their positions come from the prefix sums).
constexpr uint32_t kClaimMax = 16;
struct WaveClaim {
  uint32_t next = 0, end = 0;
};

__device__ __forceinline__ ChunkCursor next_chunk_wave(const ExpandArgs& A, int q, WaveClaim& wc) {
  ChunkCursor c;
  c.valid = false;
  c.ci = 0xFFFFFFFFu;
  if (wc.next >= wc.end) {
    uint32_t b = 0, e = 0;
    if (lane_id() == 0) {
      const uint32_t len = A.qlen[q];
      const uint32_t h = __hip_atomic_load(&A.qhead[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (h < len) {
        const uint32_t nw = gridDim.x * (blockDim.x / kWave);
        uint32_t k = (len - h) / (2u * nw);
        k = k < 1u ? 1u : (k > kClaimMax ? kClaimMax : k);
        b = atomicAdd(&A.qhead[q], k);
        e = b < len ? min(b + k, len) : b;
      }
    }
    wc.next = readlane_u(b, 0);
    wc.end = readlane_u(e, 0);
    if (wc.next >= wc.end) return c;
  }
  uint32_t ci = 0;
  if (lane_id() == 0) ci = A.queues[(int64_t)q * A.queue_cap + wc.next];
  c.ci = readlane_u(ci, 0);
  c.valid = true;
  ++wc.next;
  return c;
}

__device__ __forceinline__ void log_chunk_wave(const ExpandArgs& A, const HopDev& H, uint32_t ci, int var,
                                               uint64_t t0) {
  if (A.log && lane_id() == 0) {
    ChunkLog L;
    L.key = H.chunks[ci].key;
    L.hop = (uint8_t)H.hop;
    L.variant = (uint8_t)var;
    L.ticks = (uint32_t)(wall_clock64() - t0);
    L.chunk = ci;
    A.log[ci] = L;
  }
}

__device__ __forceinline__ void log_chunk(const ExpandArgs& A, const HopDev& H, uint32_t ci, int var, uint64_t t0) {
  if (A.log && threadIdx.x == 0) {
    const uint64_t t1 = wall_clock64();
    ChunkLog L;
    L.key = H.chunks[ci].key;
    L.hop = (uint8_t)H.hop;
    L.variant = (uint8_t)var;
    L.ticks = (uint32_t)(t1 - t0);
    L.chunk = ci;
    A.log[ci] = L;
  }
}

template <int KIND>
__global__ __launch_bounds__(256) void k_expand_thread(GraphDev G, csaw_graph_view hv, const ReqDev* __restrict__ R,
                                                       const int32_t* __restrict__ gs2req, HopDev H, ExpandArgs A) {
  __shared__ uint32_t lds[kThreadIntCap * 256];
  const int q = CSAW_VAR_THREAD * 3 + KIND;
  // each wave claims its own chunks; lane k handles entries k, k+64, ...
  WaveClaim wc;
  for (;;) {
    const ChunkCursor cc = next_chunk_wave(A, q, wc);
    if (!cc.valid) break;
    const uint64_t t0 = wall_clock64();
    const Chunk c = H.chunks[cc.ci];
    for (int k = lane_id(); k < c.len; k += kWave) {
      const Entry en = load_entry(G, R, gs2req, H, H.order[c.start + k]);
      expand_entry_thread<KIND>(G, hv, H, en, lds + threadIdx.x, 256);
    }
    log_chunk_wave(A, H, cc.ci, CSAW_VAR_THREAD, t0);
  }
}

// G32: group-only build for views whose CDF keys are all 32-bit (uniform,
// FP32, integer biases with uint32 caches): every chunk runs in sub-wave
// groups (GS = 64 for fanouts above 32) and the whole-wave path, LDS CDF
// builds and uint64 searches are compiled out.
template <int KIND, bool BITMAP, bool G32>
__global__ __launch_bounds__(256) void k_expand_wave(GraphDev G, csaw_graph_view hv, const ReqDev* __restrict__ R,
                                                     const int32_t* __restrict__ gs2req, HopDev H, ExpandArgs A) {
  __shared__ uint64_t cdf[!G32 && KIND == kInt ? 4 * kWaveIntCap : 1];
  __shared__ uint32_t bms[!G32 && BITMAP ? 4 * (kWaveBitmapCap / 32) : 1];
  // small-degree CDF slots (the generic integer kernel reuses its CDF region)
  __shared__ uint32_t glds[KIND == kF32 || (G32 && KIND == kInt) ? 4 * kGroupLdsWords : 1];
  const int var = BITMAP ? CSAW_VAR_WAVE_BITMAP : CSAW_VAR_WAVE_HASH;
  const int q = var * 3 + KIND;
  const int wid = threadIdx.x / kWave;
  // each wave claims whole chunks and expands them 64 entries per batch;
  // chunks whose fanouts fit a sub-wave group (and need no CDF build) run
  // 64 / GS entries at a time
  const bool groups_ok = KIND != kInt || A.int_cache;
  uint32_t* gl = nullptr;
  if (KIND == kF32 || (G32 && KIND == kInt)) gl = glds + wid * kGroupLdsWords;
  else if (KIND == kInt) gl = reinterpret_cast<uint32_t*>(cdf + wid * kWaveIntCap);
  WaveClaim wc;
  for (;;) {
    const ChunkCursor cc = next_chunk_wave(A, q, wc);
    if (!cc.valid) break;
    const uint64_t t0 = wall_clock64();
    const Chunk c = H.chunks[cc.ci];
    const int gs = group_size_for(c.max_f, H.group_min);
    for (int k0 = 0; k0 < c.len; k0 += kWave) {
      const int64_t first = (int64_t)c.start + k0;
      const int cnt = min(kWave, (int)c.len - k0);
      if constexpr (G32) {
        group_expand_batch<KIND, kKeys32>(G, hv, R, gs2req, H, H.order, first, cnt, gs, gl);
      } else {
        if (groups_ok && gs < kWave) group_expand_batch<KIND>(G, hv, R, gs2req, H, H.order, first, cnt, gs, gl);
        else
          wave_expand_batch<KIND, true>(G, hv, R, gs2req, H, H.order, first, cnt,
                                        KIND == kInt ? cdf + wid * kWaveIntCap : nullptr,
                                        BITMAP ? bms + wid * (kWaveBitmapCap / 32) : nullptr);
      }
    }
    log_chunk_wave(A, H, cc.ci, var, t0);
  }
}

// Region-remap entries (dedup == 2, heur_normalize's collision policy): one
// wave per entry, whole-wave selection (wave_select_remap); the integer CDF
// is built in LDS (d <= kWaveIntCap) or read from the hub cache or the CDF
// cache, the taken set is an LDS bitmap.
template <int KIND>
__global__ __launch_bounds__(256) void k_expand_remap(GraphDev G, csaw_graph_view hv, const ReqDev* __restrict__ R,
                                                      const int32_t* __restrict__ gs2req, HopDev H, ExpandArgs A) {
  __shared__ uint64_t cdf[KIND == kInt ? 4 * kWaveIntCap : 1];
  __shared__ uint32_t bms[4 * (kWaveBitmapCap / 32)];
  const int wid = threadIdx.x / kWave;
  WaveClaim wc;
  for (;;) {
    const ChunkCursor cc = next_chunk_wave(A, kQueueRemap0 + KIND, wc);
    if (!cc.valid) break;
    const uint64_t t0 = wall_clock64();
    const Chunk c = H.chunks[cc.ci];
    for (int k0 = 0; k0 < c.len; k0 += kWave)
      wave_expand_batch<KIND, true>(G, hv, R, gs2req, H, H.order, (int64_t)c.start + k0, min(kWave, (int)c.len - k0),
                                    KIND == kInt ? cdf + wid * kWaveIntCap : nullptr, bms + wid * (kWaveBitmapCap / 32));
    log_chunk_wave(A, H, cc.ci, CSAW_VAR_WAVE_BITMAP, t0);
  }
}

// Workgroup granularity: 256 lanes build the integer CDF (up to kBlockIntCap
// entries in LDS), wave 0 resolves the slots.  G32: the group-only build (see
// k_expand_wave), for views whose CDF keys are all 32-bit.
template <int KIND, bool BITMAP, bool G32>
__global__ __launch_bounds__(256) void k_expand_block(GraphDev G, csaw_graph_view hv, const ReqDev* __restrict__ R,
                                                      const int32_t* __restrict__ gs2req, HopDev H, ExpandArgs A) {
  // (G32: only the four waves' small-CDF slots, 4 x kGroupLdsWords words)
  __shared__ uint64_t cdf[KIND == kInt ? (G32 ? 2 * kGroupLdsWords : kBlockIntCap) : 1];
  __shared__ uint32_t bms[BITMAP && !G32 ? kBlockBitmapCap / 32 : 1];
  __shared__ uint64_t wsum[4];
  __shared__ uint64_t carry;
  __shared__ uint32_t slot;
  const int var = BITMAP ? CSAW_VAR_BLOCK_BITMAP : CSAW_VAR_BLOCK_HASH;
  const int q = var * 3 + KIND;
  const int tid = threadIdx.x, lane = lane_id(), wid = tid / kWave;
  // Entries that need no CDF build (the CDF cache is on) have nothing for a
  // workgroup to share: its four waves then split the chunk into 64-entry
  // batches expanded in sub-wave groups.
  const bool groups_ok = KIND != kInt || A.int_cache;
  uint32_t* gl = KIND == kInt ? reinterpret_cast<uint32_t*>(cdf) + wid * kGroupLdsWords : nullptr;
  for (;;) {
    const ChunkCursor cc = next_chunk(A, q, &slot);
    if (!cc.valid) break;
    const uint64_t t0 = wall_clock64();
    const Chunk c = H.chunks[cc.ci];
    if (G32 || groups_ok) {
      const int gs = group_size_for(c.max_f, H.group_min);
      for (int k0 = wid * kWave; k0 < c.len; k0 += 4 * kWave) {
        const int64_t first = (int64_t)c.start + k0;
        const int cnt = min(kWave, (int)c.len - k0);
        if constexpr (G32) group_expand_batch<KIND, kKeys32>(G, hv, R, gs2req, H, H.order, first, cnt, gs, gl);
        else group_expand_batch<KIND>(G, hv, R, gs2req, H, H.order, first, cnt, gs, gl);
      }
      __syncthreads();
      log_chunk(A, H, cc.ci, var, t0);
      continue;
    }
    // (G32 builds never get here: the workgroup-wide build below is dead code for them)
    for (int k = 0; k < c.len; ++k) {
      const Entry en = load_entry(G, R, gs2req, H, H.order[c.start + k]);
      if (en.d == 0 || en.F == 0) continue;
      if (en.dedup && en.d <= en.F) {
        for (int i = tid; i < en.d; i += 256) write_edge(G, hv, H, en, en.pos + i, i);
        continue;
      }
      Picker pk;
      bool build = !picker_global(G, en, KIND, kBlockIntCap, pk);
      const bool use_bm = BITMAP && en.dedup && en.d <= kBlockBitmapCap;
      if (build) {
        // block-wide inclusive scan of the integer weights
        if (tid == 0) carry = 0;
        __syncthreads();
        for (int i0 = 0; i0 < en.d; i0 += 256) {
          const int i = i0 + tid;
          uint64_t w = 0;
          if (i < en.d) w = int_bias(G, hv, en.bias, G.adj[en.o + i]);
          const uint64_t s = wave_incl_scan_u64(w);
          if (lane == kWave - 1) wsum[wid] = s;
          __syncthreads();
          uint64_t pre = carry;
          for (int w2 = 0; w2 < wid; ++w2) pre += wsum[w2];
          if (i < en.d) cdf[i] = pre + s;
          __syncthreads();
          if (tid == 255) carry = pre + s;
          __syncthreads();
        }
        pk.C = cdf;
        pk.o = 0;
        pk.T = carry;
        pk.mode = pk.T > 0 ? kPickLds : kPickUniform;
      }
      __syncthreads();
      if (wid == 0) {
        const int32_t res = wave_select<true>(G.E, en, pk, use_bm ? bms : nullptr);
        if (lane < en.F) write_edge(G, hv, H, en, en.pos + lane, res);
      }
      __syncthreads();
    }
    log_chunk(A, H, cc.ci, var, t0);
  }
}

// copy-all chunks (d <= F without replacement): sub-wave groups, lane i of a
// group copying neighbour i (coalesced reads and writes)
__global__ __launch_bounds__(256) void k_expand_copy(GraphDev G, csaw_graph_view hv, const ReqDev* __restrict__ R,
                                                     const int32_t* __restrict__ gs2req, HopDev H, ExpandArgs A) {
  WaveClaim wc;
  for (;;) {
    const ChunkCursor cc = next_chunk_wave(A, kQueueCopy, wc);
    if (!cc.valid) break;
    const Chunk c = H.chunks[cc.ci];
    const int gs = group_size_for(c.max_f, H.group_min);
    for (int k0 = 0; k0 < c.len; k0 += kWave)
      group_expand_batch<kUniform, kKeys32>(G, hv, R, gs2req, H, H.order, (int64_t)c.start + k0,
                                            min(kWave, (int)c.len - k0), gs, nullptr);
  }
}

// ------------------------------------------------------------ launchers
static int grid_of(const LaunchCfg& L, int per_cu) {
  return L.num_cu * per_cu;
}

int launch_count(const LaunchCfg& L, const GraphDev& G, const ReqDev* R, const int32_t* gs2req, const HopDev& H, bool bin,
                 int64_t max_entries) {
  int64_t nb = (max_entries + kBlockEntries - 1) / kBlockEntries;
  int grid = (int)std::max<int64_t>(1, std::min<int64_t>(nb, grid_of(L, L.count_per_cu)));
  if (bin) k_count<true><<<grid, 256, 0, L.stream>>>(G, R, gs2req, H);
  else k_count<false><<<grid, 256, 0, L.stream>>>(G, R, gs2req, H);
  CSAW_LAUNCH_CHECK();
  return CSAW_OK;
}

int launch_bin_list(const LaunchCfg& L, const GraphDev& G, const ReqDev* R, const int32_t* gs2req, const HopDev& H,
                    const int32_t* list, int64_t n) {
  if (n <= 0) return CSAW_OK;
  int64_t nb = (n + kBlockEntries - 1) / kBlockEntries;
  int grid = (int)std::max<int64_t>(1, std::min<int64_t>(nb, grid_of(L, L.count_per_cu)));
  k_count<true, true><<<grid, 256, 0, L.stream>>>(G, R, gs2req, H, list, n);
  CSAW_LAUNCH_CHECK();
  return CSAW_OK;
}

int launch_scans(const LaunchCfg& L, const HopDev& H, int64_t* total, int64_t max_entries) {
  int64_t nb = (max_entries + kBlockEntries - 1) / kBlockEntries;
  int64_t ng = (nb + 1023) / 1024;
  int grid = (int)std::max<int64_t>(1, std::min<int64_t>(ng, grid_of(L, 4)));
  k_scan_blocks<<<grid, 256, 0, L.stream>>>(H);
  CSAW_LAUNCH_CHECK();
  k_scan_groups<<<1, 256, 0, L.stream>>>(H, total);
  CSAW_LAUNCH_CHECK();
  int64_t gs = H.gb + 1;
  int g2 = (int)std::max<int64_t>(1, std::min<int64_t>((gs + 255) / 256, grid_of(L, 8)));
  k_offsets<<<g2, 256, 0, L.stream>>>(H, total);
  CSAW_LAUNCH_CHECK();
  return CSAW_OK;
}

int launch_fill(const LaunchCfg& L, uint32_t* p, uint32_t v, int64_t n) {
  if (n <= 0) return CSAW_OK;
  int grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, grid_of(L, 4)));
  k_fill_u32<<<grid, 256, 0, L.stream>>>(p, v, n);
  CSAW_LAUNCH_CHECK();
  return CSAW_OK;
}

int launch_expand_plain(const LaunchCfg& L, int kind, const GraphDev& G, const csaw_graph_view& hv, const ReqDev* R,
                        const int32_t* gs2req, const HopDev& H, int64_t max_entries, int gs) {
  int64_t nwg = (max_entries + 3) / 4;
  int grid = (int)std::max<int64_t>(1, std::min<int64_t>(nwg, grid_of(L, L.expand_per_cu)));
  if (L.grid_override > 0) grid = L.grid_override;
  // the CDF cache (and its index) is in the view: plain sampling with cdf_cache
  const bool idx = G.cdf_idx || G.int_idx32[0] || G.int_idx32[1] || G.int_idx[0] || G.int_idx[1];
  const bool wide = G.int_cdf[0] || G.int_cdf[1];  // a 64-bit integer cache in the view
  if (kind == kUniform) k_expand_plain<kUniform, false, kKeys32><<<grid, 256, 0, L.stream>>>(G, hv, R, gs2req, H, gs);
  else if (kind == kInt && idx && !wide)
    k_expand_plain<kInt, true, kKeys32><<<grid, 256, 0, L.stream>>>(G, hv, R, gs2req, H, gs);
  else if (kind == kInt && idx) k_expand_plain<kInt, true, kKeysAny><<<grid, 256, 0, L.stream>>>(G, hv, R, gs2req, H, gs);
  else if (kind == kInt) k_expand_plain<kInt, false, kKeysAny><<<grid, 256, 0, L.stream>>>(G, hv, R, gs2req, H, 0);
  else if (idx) k_expand_plain<kF32, true, kKeys32><<<grid, 256, 0, L.stream>>>(G, hv, R, gs2req, H, gs);
  else k_expand_plain<kF32, false, kKeys32><<<grid, 256, 0, L.stream>>>(G, hv, R, gs2req, H, gs);
  CSAW_LAUNCH_CHECK();
  return CSAW_OK;
}

int launch_choose(const LaunchCfg& L, const ChooseArgs& A, const HopDev& H, int64_t max_chunks) {
  int grid = (int)std::max<int64_t>(1, std::min<int64_t>((max_chunks + 255) / 256, grid_of(L, 8)));
  k_choose<<<grid, 256, 0, L.stream>>>(A, H);
  CSAW_LAUNCH_CHECK();
  return CSAW_OK;
}

int launch_expand_chunks(const LaunchCfg& L, const GraphDev& G, const csaw_graph_view& hv, const ReqDev* R,
                         const int32_t* gs2req, const HopDev& H, const ExpandArgs& A, const bool* kinds_present,
                         TimingHook* th, const hipStream_t* streams, int n_streams) {
  int next = 0;
  auto go = [&](int which, auto kern, int per_cu) -> int {
    const hipStream_t s = n_streams > 0 ? streams[next++ % n_streams] : L.stream;
    int grid = grid_of(L, per_cu);
    if (th) th->begin(which, s);
    kern<<<grid, 256, 0, s>>>(G, hv, R, gs2req, H, A);
    CSAW_LAUNCH_CHECK();
    if (th) th->end(which, s);
    return CSAW_OK;
  };
  // integer caches of the view all uint32: the group-only builds
  const bool g32 = A.int_cache && !G.int_cdf[0] && !G.int_cdf[1];
  // WAVE kernels (most of the work) first, so that with several streams
  // they start side by side.  Within them, the variant that carries most of
  // each kind's work starts first, one per stream: the taken-set bitmap for
  // the biased kinds (their high-degree, collision-heavy entries), the list
  // for uniform entries.  Their second variants queue behind them on the
  // four hardware queues instead of delaying a heavy kernel's start (on 256
  // mixed requests over RMAT-22 the FP32 bitmap kernel otherwise started
  // 1.2-1.8 ms into hops 2 and 3).
  if (kinds_present[kInt]) {
    if (g32) CSAW_TRY(go(kKernWave, k_expand_wave<kInt, true, true>, L.expand_per_cu));
    else CSAW_TRY(go(kKernWave, k_expand_wave<kInt, true, false>, L.expand_per_cu));
  }
  if (kinds_present[kF32]) CSAW_TRY(go(kKernWave, k_expand_wave<kF32, true, true>, L.expand_per_cu));
  if (kinds_present[kUniform]) CSAW_TRY(go(kKernWave, k_expand_wave<kUniform, false, true>, L.expand_per_cu));
  if (kinds_present[kInt]) {
    if (g32) CSAW_TRY(go(kKernWave, k_expand_wave<kInt, false, true>, L.expand_per_cu));
    else CSAW_TRY(go(kKernWave, k_expand_wave<kInt, false, false>, L.expand_per_cu));
  }
  if (kinds_present[kF32]) CSAW_TRY(go(kKernWave, k_expand_wave<kF32, false, true>, L.expand_per_cu));
  if (kinds_present[kUniform]) CSAW_TRY(go(kKernWave, k_expand_wave<kUniform, true, true>, L.expand_per_cu));
  for (int kind = 0; kind < 3; ++kind) {
    if (!kinds_present[kind]) continue;
    if (kind == kUniform) {
      CSAW_TRY(go(kKernThread, k_expand_thread<kUniform>, L.thread_per_cu));
      CSAW_TRY(go(kKernBlock, k_expand_block<kUniform, true, true>, L.block_per_cu));
      CSAW_TRY(go(kKernBlock, k_expand_block<kUniform, false, true>, L.block_per_cu));
    } else if (kind == kInt) {
      CSAW_TRY(go(kKernThread, k_expand_thread<kInt>, L.thread_per_cu));
      if (g32) {
        CSAW_TRY(go(kKernBlock, k_expand_block<kInt, true, true>, L.block_per_cu));
        CSAW_TRY(go(kKernBlock, k_expand_block<kInt, false, true>, L.block_per_cu));
      } else {
        CSAW_TRY(go(kKernBlock, k_expand_block<kInt, true, false>, L.block_per_cu));
        CSAW_TRY(go(kKernBlock, k_expand_block<kInt, false, false>, L.block_per_cu));
      }
    } else {
      CSAW_TRY(go(kKernThread, k_expand_thread<kF32>, L.thread_per_cu));
      CSAW_TRY(go(kKernBlock, k_expand_block<kF32, true, true>, L.block_per_cu));
      CSAW_TRY(go(kKernBlock, k_expand_block<kF32, false, true>, L.block_per_cu));
    }
  }
  CSAW_TRY(go(kKernCopy, k_expand_copy, L.expand_per_cu));
  if (A.remap) {
    if (kinds_present[kUniform]) CSAW_TRY(go(kKernWave, k_expand_remap<kUniform>, L.expand_per_cu));
    if (kinds_present[kInt]) CSAW_TRY(go(kKernWave, k_expand_remap<kInt>, L.expand_per_cu));
    if (kinds_present[kF32]) CSAW_TRY(go(kKernWave, k_expand_remap<kF32>, L.expand_per_cu));
  }
  return CSAW_OK;
}

}  // namespace csaw
