// Multi-dimensional random walk (FrontierSize > 1): the reference's
// check_layer (non-stream/main.cu:95-194) with ITS_MDRW (functions.cuh:1014-1060).
// Per sample and step: pick frontier slot q with probability proportional to
// VertexBias(frontier[q]) (draw philox(key; t, 0, sample, DOM_MDRW|request)),
// then a uniform neighbour nb of u = frontier[q] (draw (t, 1, ...)); a degree-0
// pick consumes the step (main.cu:162-163); emit (u, nb); frontier[q] = nb.
//
// One wave per sample; frontier and its VertexBias values live in LDS, the
// total T is kept incrementally and the slot search streams a wave scan over
// the bias values (no stored CDF).  Unlike the reference (D11: only thread 0
// of the grid commits), every sample is walked.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "graph.hpp"
#include "select.hpp"

namespace csaw {

constexpr int kMdrwMaxFs = 4096;

// bad: null when the host has validated the seeds; otherwise a sample with
// a seed outside [0, V) emits nothing and sets *bad (seeds read on the device).
__global__ __launch_bounds__(64) void k_mdrw(GraphDev G, csaw_graph_view hv, uint32_t k0, uint32_t k1, uint32_t request,
                                            int32_t S, int32_t fs, int32_t depth, const int32_t* __restrict__ seeds,
                                            int32_t* __restrict__ src, int32_t* __restrict__ dst,
                                            int32_t* __restrict__ cnt, uint32_t* __restrict__ bad) {
  __shared__ int32_t front[kMdrwMaxFs];
  __shared__ uint32_t wts[kMdrwMaxFs];
  const int lane = lane_id();
  const uint32_t c3 = (kDomMdrw << 28) | request;
  for (int32_t s = blockIdx.x; s < S; s += gridDim.x) {
    if (bad) {
      bool out = false;
      for (int q = lane; q < fs; q += kWave) {
        const int32_t v = seeds[(int64_t)s * fs + q];
        out |= v < 0 || (int64_t)v >= G.V;
      }
      if (ballot(out)) {
        if (lane == 0) {
          cnt[s] = 0;
          atomicOr(bad, 1u);
        }
        continue;
      }
    }
    uint64_t T = 0;
    for (int q0 = 0; q0 < fs; q0 += kWave) {
      const int q = q0 + lane;
      uint32_t w = 0;
      if (q < fs) {
        const int32_t v = seeds[(int64_t)s * fs + q];
        front[q] = v;
        w = VertexBias(v, &hv);
        wts[q] = w;
      }
      T += readlane_u64(wave_incl_scan_u64((uint64_t)w), kWave - 1);
    }
    wave_fence();
    int32_t c = 0;
    for (int32_t t = 0; t < depth; ++t) {
      const U4 o = philox4x32_10((uint32_t)t, 0u, (uint32_t)s, c3, k0, k1);
      int32_t q = 0;
      if (T == 0) {
        q = (int32_t)mulhi64(philox_t64(o), (uint64_t)fs);
      } else {
        const uint64_t x = mulhi64(philox_t64(o), T);
        uint64_t acc = 0;
        for (int q0 = 0; q0 < fs; q0 += kWave) {
          const int qi = q0 + lane;
          const uint64_t w = qi < fs ? wts[qi] : 0u;
          const uint64_t sc = wave_incl_scan_u64(w);
          const uint64_t tot = readlane_u64(sc, kWave - 1);
          if (acc + tot > x) {
            q = q0 + __builtin_ctzll(ballot(acc + sc > x));
            break;
          }
          acc += tot;
        }
      }
      const int32_t u = front[q];
      const int32_t d = G.deg[u];
      if (d == 0) continue;
      const U4 o2 = philox4x32_10((uint32_t)t, 1u, (uint32_t)s, c3, k0, k1);
      const int32_t nb = G.adj[G.off[u] + (int64_t)mulhi64(philox_t64(o2), (uint64_t)d)];
      if (lane == 0) {
        src[(int64_t)s * depth + c] = u;
        dst[(int64_t)s * depth + c] = nb;
        front[q] = nb;
        wts[q] = VertexBias(nb, &hv);
      }
      T = T - (uint64_t)VertexBias(u, &hv) + (uint64_t)VertexBias(nb, &hv);
      ++c;
      wave_fence();
    }
    if (lane == 0) cnt[s] = c;
  }
}

// Per-sample edge offsets from the per-sample counts: one workgroup of 1024
// threads, each summing a contiguous run of samples, then a block scan.
__global__ __launch_bounds__(1024) void k_mdrw_offsets(const int32_t* __restrict__ cnt, int32_t S,
                                                       int64_t* __restrict__ offsets) {
  __shared__ uint64_t wsum[16];
  const int t = threadIdx.x, lane = lane_id(), wid = t / kWave;
  const int64_t per = ((int64_t)S + 1023) / 1024;
  const int64_t b = (int64_t)t * per, e = min((int64_t)S, b + per);
  uint64_t loc = 0;
  for (int64_t i = b; i < e; ++i) loc += (uint64_t)cnt[i];
  const uint64_t inc = wave_incl_scan_u64(loc);
  if (lane == kWave - 1) wsum[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    const uint64_t x = lane < 16 ? wsum[lane] : 0;
    const uint64_t xi = wave_incl_scan_u64(x);
    if (lane < 16) wsum[lane] = xi - x;
  }
  __syncthreads();
  uint64_t run = wsum[wid] + inc - loc;
  for (int64_t i = b; i < e; ++i) {
    offsets[i] = (int64_t)run;
    run += (uint64_t)cnt[i];
  }
  if (t == 1023) offsets[S] = (int64_t)run;  // the last thread's prefix + its run = the total
}

// Moves each sample's edges from its depth-strided slots to its offset: one
// wave per sample.
__global__ __launch_bounds__(256) void k_mdrw_compact(const int32_t* __restrict__ cnt, const int64_t* __restrict__ offsets,
                                                      const int32_t* __restrict__ ssrc, const int32_t* __restrict__ sdst,
                                                      int32_t S, int32_t depth, int32_t* __restrict__ src,
                                                      int32_t* __restrict__ dst) {
  const int lane = lane_id();
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t s = (int64_t)blockIdx.x * 4 + threadIdx.x / kWave; s < S; s += nw) {
    const int32_t c = cnt[s];
    const int64_t o = offsets[s];
    for (int32_t i = lane; i < c; i += kWave) {
      src[o + i] = ssrc[s * depth + i];
      dst[o + i] = sdst[s * depth + i];
    }
  }
}

__global__ void k_mdrw_clear(uint32_t* __restrict__ p) {
  if (threadIdx.x == 0) *p = 0u;
}

// ------------------------------------------------------ allocation-free MDRW
// Device scratch of csaw_mdrw_noalloc: a flag word, then per-sample counts and
// the depth-strided edge slots.  Each GPU holds its own copy for the life of
// the process; host code reaches it only through the address that
// hipGetSymbolAddress returned on that device (mdrw_scratch_ptr).
constexpr int64_t kMdrwScratchBytes = 32ll << 20;
static __device__ __align__(256) unsigned char scratch[kMdrwScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int mdrw_scratch_setup() {
  int dev = 0;
  CSAW_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_last_error("device id beyond the scratch table");
    return CSAW_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return CSAW_OK;
  void* p = nullptr;
  CSAW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  return CSAW_OK;
}

// The scratch bytes one call needs (flag, counts, two slot arrays).
static int64_t mdrw_scratch_need(int64_t samples, int64_t depth) {
  const auto up = [](int64_t b) { return (b + 255) & ~(int64_t)255; };
  return 256 + up(4 * samples) + 2 * up(4 * samples * depth);
}

}  // namespace csaw

using namespace csaw;

extern "C" int csaw_mdrw(csaw_graph* g, uint64_t seed, uint32_t request_id, int32_t samples, int32_t frontier_size,
                         int32_t depth, const int32_t* seeds, int32_t* src, int32_t* dst, int64_t* offsets,
                         csaw_stats* stats) {
  // request ids share counter word 3 with the domain tag in bits 28-31
  if (!g || samples < 0 || frontier_size < 1 || frontier_size > kMdrwMaxFs || depth < 0 || !offsets ||
      request_id >= (1u << 28) ||
      (samples > 0 && (!seeds || (depth > 0 && (!src || !dst)))))
    return CSAW_ERR_INVALID_ARG;
  const int64_t ns = (int64_t)samples * frontier_size;
  for (int64_t i = 0; i < ns; ++i)
    if (seeds[i] < 0 || seeds[i] >= g->V) return CSAW_ERR_INVALID_ARG;
  const int64_t slots = std::max<int64_t>((int64_t)samples * depth, 1);
  int32_t *d_seeds = nullptr, *d_src = nullptr, *d_dst = nullptr, *d_cnt = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(d_seeds);
    (void)hipFree(d_src);
    (void)hipFree(d_dst);
    (void)hipFree(d_cnt);
  };
  auto body = [&]() -> int {
    CSAW_HIP(hipMalloc(&d_seeds, sizeof(int32_t) * std::max<int64_t>(ns, 1)));
    CSAW_HIP(hipMalloc(&d_src, sizeof(int32_t) * slots));
    CSAW_HIP(hipMalloc(&d_dst, sizeof(int32_t) * slots));
    CSAW_HIP(hipMalloc(&d_cnt, sizeof(int32_t) * std::max<int32_t>(samples, 1)));
    if (ns > 0) CSAW_HIP(hipMemcpy(d_seeds, seeds, sizeof(int32_t) * ns, hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    CSAW_HIP(hipEventCreate(&e0));
    CSAW_HIP(hipEventCreate(&e1));
    CSAW_HIP(hipEventRecord(e0, g->stream));
    if (samples > 0) {
      const int grid = std::min<int>(samples, g->num_cu * 16);
      k_mdrw<<<grid, 64, 0, g->stream>>>(g->view(), g->hook_view(), (uint32_t)seed, (uint32_t)(seed >> 32), request_id,
                                         samples, frontier_size, depth, d_seeds, d_src, d_dst, d_cnt, nullptr);
      CSAW_LAUNCH_CHECK();
    }
    CSAW_HIP(hipEventRecord(e1, g->stream));
    CSAW_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CSAW_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::vector<int32_t> cnt(samples), hs(slots), hd(slots);
    if (samples > 0) {
      CSAW_HIP(hipMemcpy(cnt.data(), d_cnt, sizeof(int32_t) * samples, hipMemcpyDeviceToHost));
      CSAW_HIP(hipMemcpy(hs.data(), d_src, sizeof(int32_t) * slots, hipMemcpyDeviceToHost));
      CSAW_HIP(hipMemcpy(hd.data(), d_dst, sizeof(int32_t) * slots, hipMemcpyDeviceToHost));
    }
    int64_t at = 0;
    for (int32_t s = 0; s < samples; ++s) {
      offsets[s] = at;
      for (int32_t i = 0; i < cnt[s]; ++i, ++at) {
        src[at] = hs[(int64_t)s * depth + i];
        dst[at] = hd[(int64_t)s * depth + i];
      }
    }
    offsets[samples] = at;
    if (stats) {
      memset(stats, 0, sizeof(*stats));
      stats->edges = at;
      stats->kernel_ms = ms;
    }
    return CSAW_OK;
  };
  int st = body();
  cleanup();
  return st;
}

extern "C" int csaw_mdrw_noalloc(csaw_graph* g, uint64_t seed, uint32_t request_id, int32_t samples,
                                 int32_t frontier_size, int32_t depth, const int32_t* d_seeds, int32_t* d_src,
                                 int32_t* d_dst, int64_t* d_offsets, void* stream, int sync) {
  if (!g || samples < 0 || frontier_size < 1 || frontier_size > kMdrwMaxFs || depth < 0 || !d_offsets ||
      request_id >= (1u << 28) || (samples > 0 && (!d_seeds || (depth > 0 && (!d_src || !d_dst)))))
    return CSAW_ERR_INVALID_ARG;
  const int64_t need = mdrw_scratch_need(samples, depth);
  if (need > kMdrwScratchBytes) {
    set_last_error("csaw_mdrw_noalloc: samples x depth = " + std::to_string((int64_t)samples * depth) +
                   " needs " + std::to_string(need) + " bytes of scratch; the maximum is " +
                   std::to_string(kMdrwScratchBytes) + " (about samples * (8 * depth + 4))");
    return CSAW_ERR_CAPACITY;
  }
  int dev = 0;
  CSAW_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return CSAW_ERR_UNSUPPORTED;
  unsigned char* base = g_scratch_ptr[dev];
  if (!base) {
    CSAW_TRY(mdrw_scratch_setup());
    base = g_scratch_ptr[dev];
  }
  const auto up = [](int64_t b) { return (b + 255) & ~(int64_t)255; };
  uint32_t* bad = reinterpret_cast<uint32_t*>(base);
  int32_t* cnt = reinterpret_cast<int32_t*>(base + 256);
  int32_t* ssrc = reinterpret_cast<int32_t*>(base + 256 + up(4 * (int64_t)samples));
  int32_t* sdst = reinterpret_cast<int32_t*>(base + 256 + up(4 * (int64_t)samples) + up(4 * (int64_t)samples * depth));
  hipStream_t st = stream ? (hipStream_t)stream : g->stream;
  k_mdrw_clear<<<1, 64, 0, st>>>(bad);
  CSAW_LAUNCH_CHECK();
  if (samples > 0) {
    const int grid = std::min<int>(samples, g->num_cu * 16);
    k_mdrw<<<grid, 64, 0, st>>>(g->view(), g->hook_view(), (uint32_t)seed, (uint32_t)(seed >> 32), request_id, samples,
                                frontier_size, depth, d_seeds, ssrc, sdst, cnt, bad);
    CSAW_LAUNCH_CHECK();
  }
  k_mdrw_offsets<<<1, 1024, 0, st>>>(cnt, samples, d_offsets);
  CSAW_LAUNCH_CHECK();
  if (samples > 0 && depth > 0) {
    const int grid = (int)std::min<int64_t>(((int64_t)samples + 3) / 4, (int64_t)g->num_cu * 8);
    k_mdrw_compact<<<grid, 256, 0, st>>>(cnt, d_offsets, ssrc, sdst, samples, depth, d_src, d_dst);
    CSAW_LAUNCH_CHECK();
  }
  if (sync) {
    uint32_t flag = 0;
    CSAW_HIP(hipMemcpyAsync(&flag, bad, sizeof(flag), hipMemcpyDeviceToHost, st));
    CSAW_HIP(hipStreamSynchronize(st));
    if (flag) {
      set_last_error("csaw_mdrw_noalloc: a seed is outside [0, V) (its sample emitted nothing)");
      return CSAW_ERR_INVALID_ARG;
    }
  }
  return CSAW_OK;
}
