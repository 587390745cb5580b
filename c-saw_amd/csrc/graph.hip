// Graph construction on the device: upload + validation, GPU RMAT generator,
// degree array, FP32 sequential CDF cache and the integer hub CDF cache.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "cdf_index.hpp"
#include "graph.hpp"

namespace csaw {

namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

int device_init(int* num_cu) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    set_last_error("no HIP device visible");
    return CSAW_ERR_NO_DEVICE;
  }
  int dev = 0;
  CSAW_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  CSAW_HIP(hipGetDeviceProperties(&prop, dev));
  *num_cu = prop.multiProcessorCount;
  return CSAW_OK;
}

// ------------------------------------------------------------- kernels
__global__ void k_degree(const int64_t* __restrict__ off, int64_t V, int32_t* __restrict__ deg,
                         unsigned long long* __restrict__ maxd, uint32_t* __restrict__ bad) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
    int64_t d = off[v + 1] - off[v];
    if (d < 0 || d > 0x7fffffff) {
      atomicOr(bad, 1u);
      d = 0;
    }
    deg[v] = (int32_t)d;
    atomicMax(maxd, (unsigned long long)d);
  }
}

__global__ void k_check_ids(const int32_t* __restrict__ adj, int64_t E, int64_t V, uint32_t* __restrict__ bad) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < E; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t u = adj[i];
    if (u < 0 || u >= V) atomicOr(bad, 2u);
  }
}

__global__ void k_widen_offsets(const int32_t* __restrict__ in, int64_t n, int64_t* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

// RMAT pair generation: level l of edge e uses output word (l & 3) of
// philox(key; e lo, e hi, l >> 2, kDomRmat << 28); thresholds a, a+b,
// a+b+c as u32 fractions; bit (scale-1-l) of (src, dst).
__global__ void k_rmat_pairs(int scale, int64_t M, uint32_t ta, uint32_t tab, uint32_t tabc, uint32_t k0,
                             uint32_t k1, uint64_t* __restrict__ keys) {
  const uint64_t sentinel = (scale == 32) ? ~0ull : ((1ull << (2 * scale)) - 1);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < M; e += (int64_t)gridDim.x * blockDim.x) {
    uint32_t src = 0, dst = 0;
    U4 o{0, 0, 0, 0};
    for (int l = 0; l < scale; ++l) {
      if ((l & 3) == 0) o = philox4x32_10((uint32_t)e, (uint32_t)(e >> 32), (uint32_t)(l >> 2), kDomRmat << 28, k0, k1);
      const uint32_t r = (l & 3) == 0 ? o.x : (l & 3) == 1 ? o.y : (l & 3) == 2 ? o.z : o.w;
      const uint32_t bit = 1u << (scale - 1 - l);
      if (r < ta) {
      } else if (r < tab) {
        dst |= bit;
      } else if (r < tabc) {
        src |= bit;
      } else {
        src |= bit;
        dst |= bit;
      }
    }
    uint64_t a = ((uint64_t)src << scale) | dst, b = ((uint64_t)dst << scale) | src;
    if (src == dst) a = b = sentinel;
    keys[2 * e] = a;
    keys[2 * e + 1] = b;
  }
}

// unique sorted keys -> adjacency ids + offsets (off[u] = first index of u)
__global__ void k_keys_to_csr(const uint64_t* __restrict__ keys, int64_t E, int scale, int64_t V,
                              int32_t* __restrict__ adj, int64_t* __restrict__ off) {
  const uint64_t mask = (1ull << scale) - 1;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= E; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t u = i < E ? (int64_t)(keys[i] >> scale) : V;
    const int64_t p = i > 0 ? (int64_t)(keys[i - 1] >> scale) : -1;
    if (i < E) adj[i] = (int32_t)(keys[i] & mask);
    for (int64_t v = p + 1; v <= u; ++v) off[v] = i;
  }
}

__global__ void k_weights(uint32_t k0, uint32_t k1, int64_t E, float* __restrict__ w) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
    U4 o = philox4x32_10((uint32_t)e, (uint32_t)(e >> 32), 0u, kDomWeights << 28, k0, k1);
    w[e] = (float)((o.x >> 8) + 1u) * 5.9604644775390625e-08f;
  }
}

// Sequential FP32 CDF per vertex, strictly left to right (one thread per
// vertex; never a parallel scan, whose rounding would differ).  Vertices of
// degree above skip_above are left to k_fp32_cdf_hubs.
__global__ void k_fp32_cdf(const int64_t* __restrict__ off, const float* __restrict__ w, int64_t V,
                           int64_t skip_above, float* __restrict__ cdf) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
    float acc = 0.0f;
    const int64_t b = off[v], e = off[v + 1];
    if (e - b > skip_above) continue;
    for (int64_t i = b; i < e; ++i) {
      acc = __fadd_rn(acc, w[i]);
      cdf[i] = acc;
    }
  }
}

// The same left-to-right fold for the hubs, one wave per hub: the wave loads
// 64 weights at a time (the next 64 in flight meanwhile) and every lane
// applies the 64 additions in order to the same running sum, keeping the one
// at its own position.
__global__ __launch_bounds__(256) void k_fp32_cdf_hubs(const int64_t* __restrict__ off, const float* __restrict__ w,
                                                       const int32_t* __restrict__ hubs, int32_t n_hubs,
                                                       float* __restrict__ cdf) {
  const int lane = lane_id();
  for (int64_t h = (int64_t)blockIdx.x * 4 + threadIdx.x / kWave; h < n_hubs; h += (int64_t)gridDim.x * 4) {
    const int32_t v = hubs[h];
    const int64_t b = off[v], e = off[v + 1];
    float acc = 0.0f;
    float x = b + lane < e ? w[b + lane] : 0.0f;
    for (int64_t base = b; base < e; base += kWave) {
      const float nx = base + kWave + lane < e ? w[base + kWave + lane] : 0.0f;
      const int cnt = (int)min((int64_t)kWave, e - base);
      float mine = 0.0f;
      for (int j = 0; j < cnt; ++j) {
        acc = __fadd_rn(acc, __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j)));
        if (lane == j) mine = acc;
      }
      if (lane < cnt) cdf[base + lane] = mine;
      x = nx;
    }
  }
}

__global__ void k_seeds(uint32_t k0, uint32_t k1, uint32_t request, int64_t V, int64_t n, int32_t* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    U4 o = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), request, kDomSeeds << 28, k0, k1);
    out[i] = (int32_t)mulhi64(philox_t64(o), (uint64_t)V);
  }
}

// Hub cache, split across workgroups: each workgroup scans one segment of
// one hub's neighbour list (kHubSeg entries) and writes the segment-local
// inclusive CDF plus the segment total; a second pass adds segment offsets.
constexpr int kHubThreads = 256;

template <int KIND>  // 0 degree, 1 user hook
__global__ __launch_bounds__(kHubThreads) void k_hub_segments(GraphDev G, csaw_graph_view hv, const int32_t* __restrict__ seg_hub,
                                                              const int32_t* __restrict__ seg_idx, int64_t n_seg,
                                                              uint64_t* __restrict__ pool, uint64_t* __restrict__ seg_tot) {
  __shared__ uint64_t wsum[kHubThreads / kWave];
  __shared__ uint64_t carry_s;
  for (int64_t s = blockIdx.x; s < n_seg; s += gridDim.x) {
    const int32_t h = seg_hub[s];
    const int32_t v = G.hub_vid[h];
    const int64_t o = G.off[v];
    const int64_t d = G.off[v + 1] - o;
    const int64_t b = (int64_t)seg_idx[s] * kHubSeg;
    const int64_t e = b + kHubSeg < d ? b + kHubSeg : d;
    const int64_t pbase = G.hub_pool[h];
    if (threadIdx.x == 0) carry_s = 0;
    // the segment's neighbour ids, then their biases, all requested before
    // any is used (16 rows of 256 entries): the list may be read in place
    // over the bus, where every dependent round trip costs microseconds
    constexpr int kRows = kHubSeg / kHubThreads;
    int32_t u[kRows];
    uint64_t w[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int64_t i = b + r * kHubThreads + threadIdx.x;
      u[r] = i < e ? G.adj[o + i] : -1;
    }
#pragma unroll
    for (int r = 0; r < kRows; ++r)
      w[r] = u[r] < 0 ? 0ull : (KIND == 0 ? (uint64_t)G.deg[u[r]] : (uint64_t)EdgeBias(u[r], &hv));
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int64_t i = b + r * kHubThreads + threadIdx.x;
      const uint64_t sc = wave_incl_scan_u64(w[r]);
      const int wid = threadIdx.x / kWave;
      if (lane_id() == kWave - 1) wsum[wid] = sc;
      __syncthreads();
      uint64_t pre = carry_s;
      for (int k = 0; k < wid; ++k) pre += wsum[k];
      if (i < e) pool[pbase + i] = pre + sc;
      __syncthreads();
      if (threadIdx.x == kHubThreads - 1) carry_s = pre + sc;
      __syncthreads();
    }
    if (threadIdx.x == 0) seg_tot[s] = carry_s;
    __syncthreads();
  }
}

// seg_first[h]..seg_first[h+1]: segments of hub h; exclusive offsets per segment
__global__ void k_hub_seg_offsets(const int64_t* __restrict__ seg_first, int32_t n_hubs, uint64_t* __restrict__ seg_tot) {
  for (int32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < n_hubs; h += gridDim.x * blockDim.x) {
    uint64_t acc = 0;
    for (int64_t s = seg_first[h]; s < seg_first[h + 1]; ++s) {
      uint64_t t = seg_tot[s];
      seg_tot[s] = acc;
      acc += t;
    }
  }
}

__global__ __launch_bounds__(kHubThreads) void k_hub_add(GraphDev G, const int32_t* __restrict__ seg_hub,
                                                        const int32_t* __restrict__ seg_idx, int64_t n_seg,
                                                        const uint64_t* __restrict__ seg_off, uint64_t* __restrict__ pool) {
  for (int64_t s = blockIdx.x; s < n_seg; s += gridDim.x) {
    const int32_t h = seg_hub[s];
    const int32_t v = G.hub_vid[h];
    const int64_t d = G.off[v + 1] - G.off[v];
    const int64_t b = (int64_t)seg_idx[s] * kHubSeg;
    const int64_t e = b + kHubSeg < d ? b + kHubSeg : d;
    const uint64_t add = seg_off[s];
    if (add == 0) continue;
    for (int64_t i = b + threadIdx.x; i < e; i += kHubThreads) pool[G.hub_pool[h] + i] += add;
  }
}

static int grid_for(int64_t n, int threads, int num_cu) {
  int64_t g = (n + threads - 1) / threads;
  int64_t cap = (int64_t)num_cu * 32;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

int graph_alloc_common(csaw_graph* g) {
  CSAW_TRY(device_init(&g->num_cu));
  CSAW_HIP(hipGetDevice(&g->device));
  CSAW_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  const char* keys = getenv("CSAW_INT_CDF_KEYS");
  g->wide_int_cdf = keys && strcmp(keys, "64") == 0;
  CSAW_TRY(mdrw_scratch_setup());
  return CSAW_OK;
}

// Degree array + validation.
static int finish_graph(csaw_graph* g) {
  CSAW_HIP(hipMalloc(&g->deg, sizeof(int32_t) * (g->V > 0 ? g->V : 1)));
  unsigned long long* dmax;
  uint32_t* bad;
  CSAW_HIP(hipMalloc(&dmax, sizeof(unsigned long long) + sizeof(uint32_t)));
  bad = (uint32_t*)(dmax + 1);
  CSAW_HIP(hipMemsetAsync(dmax, 0, sizeof(unsigned long long) + sizeof(uint32_t), g->stream));
  if (g->V > 0) {
    k_degree<<<grid_for(g->V, 256, g->num_cu), 256, 0, g->stream>>>(g->off, g->V, g->deg, dmax, bad);
    CSAW_LAUNCH_CHECK();
  }
  if (g->E > 0) {
    k_check_ids<<<grid_for(g->E, 256, g->num_cu), 256, 0, g->stream>>>(g->adj, g->E, g->V, bad);
    CSAW_LAUNCH_CHECK();
  }
  unsigned long long hm = 0;
  uint32_t hb = 0;
  CSAW_HIP(hipMemcpyAsync(&hm, dmax, sizeof(hm), hipMemcpyDeviceToHost, g->stream));
  CSAW_HIP(hipMemcpyAsync(&hb, bad, sizeof(hb), hipMemcpyDeviceToHost, g->stream));
  CSAW_HIP(hipStreamSynchronize(g->stream));
  CSAW_HIP(hipFree(dmax));
  g->max_deg = (int64_t)hm;
  if (hb) {
    set_last_error(hb & 1 ? "offsets not monotone or degree >= 2^31" : "neighbour id out of range");
    return CSAW_ERR_INVALID_GRAPH;
  }
  return CSAW_OK;
}

// Vertices of degree above kHubMin, sorted by id, with their pool offsets.
static int ensure_hub_list(csaw_graph* g) {
  if (g->n_hubs >= 0) return CSAW_OK;
  std::vector<int32_t> deg(g->V);
  if (g->V > 0) CSAW_HIP(hipMemcpy(deg.data(), g->deg, sizeof(int32_t) * g->V, hipMemcpyDeviceToHost));
  std::vector<int32_t> vid;
  std::vector<int64_t> pool{0};
  for (int64_t v = 0; v < g->V; ++v)
    if (deg[v] > kHubMin) {
      vid.push_back((int32_t)v);
      pool.push_back(pool.back() + deg[v]);
    }
  CSAW_HIP(hipMalloc(&g->hub_vid, sizeof(int32_t) * (vid.size() + 1)));
  CSAW_HIP(hipMalloc(&g->hub_pool, sizeof(int64_t) * pool.size()));
  if (!vid.empty())
    CSAW_HIP(hipMemcpy(g->hub_vid, vid.data(), sizeof(int32_t) * vid.size(), hipMemcpyHostToDevice));
  CSAW_HIP(hipMemcpy(g->hub_pool, pool.data(), sizeof(int64_t) * pool.size(), hipMemcpyHostToDevice));
  g->n_hubs = (int32_t)vid.size();
  g->hub_entries = pool.back();
  return CSAW_OK;
}

int graph_prepare_bias(csaw_graph* g, int32_t bias) {
  if (bias == CSAW_BIAS_WEIGHT) {
    if (!g->w) {
      set_last_error("CSAW_BIAS_WEIGHT needs edge weights");
      return CSAW_ERR_INVALID_ARG;
    }
    if (!g->cdf) {
      // padded to 16 entries: the search index reads aligned 16-entry blocks
      const int64_t n = pad16(g->E > 0 ? g->E : 1);
      CSAW_HIP(hipMalloc(&g->cdf, sizeof(float) * n));
      CSAW_HIP(hipMemsetAsync(g->cdf, 0, sizeof(float) * n, g->stream));
      if (g->V > 0) {
        CSAW_TRY(ensure_hub_list(g));
        k_fp32_cdf<<<grid_for(g->V, 256, g->num_cu), 256, 0, g->stream>>>(g->off, g->w, g->V, (int64_t)kHubMin,
                                                                           g->cdf);
        CSAW_LAUNCH_CHECK();
        if (g->n_hubs > 0) {
          const int grid = (int)std::min<int64_t>((g->n_hubs + 3) / 4, (int64_t)g->num_cu * 8);
          k_fp32_cdf_hubs<<<grid, 256, 0, g->stream>>>(g->off, g->w, g->hub_vid, g->n_hubs, g->cdf);
          CSAW_LAUNCH_CHECK();
        }
      }
      CSAW_HIP(hipStreamSynchronize(g->stream));
    }
    return CSAW_OK;
  }
  if (bias != CSAW_BIAS_DEGREE && bias != CSAW_BIAS_USER) return CSAW_OK;
  const int kind = bias == CSAW_BIAS_DEGREE ? 0 : 1;
  CSAW_TRY(ensure_hub_list(g));
  if (g->hub_cdf[kind] || g->n_hubs == 0) return CSAW_OK;
  std::vector<int64_t> pool(g->n_hubs + 1);
  CSAW_HIP(hipMemcpy(pool.data(), g->hub_pool, sizeof(int64_t) * pool.size(), hipMemcpyDeviceToHost));
  CSAW_HIP(hipMalloc(&g->hub_cdf[kind], sizeof(uint64_t) * g->hub_entries));
  CSAW_TRY(hub_cdf_build(g->view(), g->hook_view(), kind, pool, g->hub_cdf[kind], g->stream, g->num_cu));
  CSAW_HIP(hipStreamSynchronize(g->stream));
  return CSAW_OK;
}

// Builds the inclusive integer CDFs of the hubs G.hub_vid[0..n) into out
// (pool offsets host_pool, n+1 entries), each hub split across workgroups.
int hub_cdf_build(const GraphDev& G, const csaw_graph_view& hv, int kind, const std::vector<int64_t>& pool,
                  uint64_t* out, hipStream_t stream, int num_cu) {
  const int32_t n_hubs = (int32_t)pool.size() - 1;
  if (n_hubs <= 0) return CSAW_OK;
  std::vector<int32_t> seg_hub, seg_idx;
  std::vector<int64_t> seg_first{0};
  for (int32_t h = 0; h < n_hubs; ++h) {
    int64_t d = pool[h + 1] - pool[h];
    int64_t ns = (d + kHubSeg - 1) / kHubSeg;
    for (int64_t s = 0; s < ns; ++s) {
      seg_hub.push_back(h);
      seg_idx.push_back((int32_t)s);
    }
    seg_first.push_back((int64_t)seg_hub.size());
  }
  const int64_t n_seg = (int64_t)seg_hub.size();
  int32_t *d_seg_hub, *d_seg_idx;
  int64_t* d_seg_first;
  uint64_t* d_seg_tot;
  CSAW_HIP(hipMalloc(&d_seg_hub, sizeof(int32_t) * n_seg));
  CSAW_HIP(hipMalloc(&d_seg_idx, sizeof(int32_t) * n_seg));
  CSAW_HIP(hipMalloc(&d_seg_first, sizeof(int64_t) * seg_first.size()));
  CSAW_HIP(hipMalloc(&d_seg_tot, sizeof(uint64_t) * n_seg));
  CSAW_HIP(hipMemcpyAsync(d_seg_hub, seg_hub.data(), sizeof(int32_t) * n_seg, hipMemcpyHostToDevice, stream));
  CSAW_HIP(hipMemcpyAsync(d_seg_idx, seg_idx.data(), sizeof(int32_t) * n_seg, hipMemcpyHostToDevice, stream));
  CSAW_HIP(hipMemcpyAsync(d_seg_first, seg_first.data(), sizeof(int64_t) * seg_first.size(), hipMemcpyHostToDevice,
                          stream));
  CSAW_TRY(hub_cdf_segments(G, hv, kind, n_hubs, d_seg_hub, d_seg_idx, n_seg, d_seg_first, d_seg_tot, out, stream,
                            num_cu));
  CSAW_HIP(hipStreamSynchronize(stream));
  CSAW_HIP(hipFree(d_seg_hub));
  CSAW_HIP(hipFree(d_seg_idx));
  CSAW_HIP(hipFree(d_seg_first));
  CSAW_HIP(hipFree(d_seg_tot));
  return CSAW_OK;
}

int hub_cdf_segments(const GraphDev& G, const csaw_graph_view& hv, int kind, int32_t n_hubs, const int32_t* seg_hub,
                     const int32_t* seg_idx, int64_t n_seg, const int64_t* seg_first, uint64_t* seg_tot,
                     uint64_t* out, hipStream_t stream, int num_cu) {
  if (n_hubs <= 0 || n_seg <= 0) return CSAW_OK;
  const int grid = (int)std::min<int64_t>(n_seg, (int64_t)num_cu * 8);
  if (kind == 0)
    k_hub_segments<0><<<grid, kHubThreads, 0, stream>>>(G, hv, seg_hub, seg_idx, n_seg, out, seg_tot);
  else
    k_hub_segments<1><<<grid, kHubThreads, 0, stream>>>(G, hv, seg_hub, seg_idx, n_seg, out, seg_tot);
  CSAW_LAUNCH_CHECK();
  k_hub_seg_offsets<<<grid_for(n_hubs, 64, num_cu), 64, 0, stream>>>(seg_first, n_hubs, seg_tot);
  CSAW_LAUNCH_CHECK();
  k_hub_add<<<grid, kHubThreads, 0, stream>>>(G, seg_hub, seg_idx, n_seg, seg_tot, out);
  CSAW_LAUNCH_CHECK();
  return CSAW_OK;
}

// Full integer CDF cache (the reference's probability cache Cp,
// functions.cuh:190-208 and 329-341, there compiled in but disabled by
// cache=0 at main.cu:26): one inclusive uint64 CDF per CSR entry, so a draw
// costs a binary search instead of a pass over the neighbour list.
template <int KIND>
__global__ __launch_bounds__(256) void k_int_cdf_small(GraphDev G, csaw_graph_view hv, uint64_t* __restrict__ C) {
  const int64_t nw = (int64_t)gridDim.x * 4;
  const int lane = lane_id();
  for (int64_t v = blockIdx.x * 4 + threadIdx.x / kWave; v < G.V; v += nw) {
    const int32_t d = G.deg[v];
    if (d == 0 || d > kHubMin) continue;
    const int64_t o = G.off[v];
    uint64_t acc = 0;
    for (int i0 = 0; i0 < d; i0 += kWave) {
      const int i = i0 + lane;
      uint64_t w = 0;
      if (i < d) {
        const int32_t u = G.adj[o + i];
        w = KIND == 0 ? (uint64_t)G.deg[u] : (uint64_t)EdgeBias(u, &hv);
      }
      const uint64_t s = wave_incl_scan_u64(w);
      if (i < d) C[o + i] = acc + s;
      acc += readlane_u64(s, kWave - 1);
    }
  }
}

__global__ void k_int_cdf_hubs(GraphDev G, const uint64_t* __restrict__ pool, uint64_t* __restrict__ C) {
  for (int32_t h = blockIdx.x; h < G.n_hubs; h += gridDim.x) {
    const int32_t v = G.hub_vid[h];
    const int64_t o = G.off[v], d = G.off[v + 1] - o, p = G.hub_pool[h];
    for (int64_t i = threadIdx.x; i < d; i += blockDim.x) C[o + i] = pool[p + i];
  }
}

__global__ void k_max_u64(const uint64_t* __restrict__ C, int64_t n, unsigned long long* __restrict__ out) {
  unsigned long long m = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = C[i] > m ? C[i] : m;
  for (int off = 32; off >= 1; off >>= 1) {
    const unsigned long long o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  if (lane_id() == 0 && m) atomicMax(out, m);
}

__global__ void k_narrow_u32(const uint64_t* __restrict__ C, int64_t n, uint32_t* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (uint32_t)C[i];
}

int graph_prepare_int_cache(csaw_graph* g, int32_t bias) {
  if (bias != CSAW_BIAS_DEGREE && bias != CSAW_BIAS_USER) return CSAW_OK;
  const int kind = bias == CSAW_BIAS_DEGREE ? 0 : 1;
  if (g->int_cdf[kind] || g->int_cdf32[kind]) return CSAW_OK;
  CSAW_TRY(graph_prepare_bias(g, bias));
  const int64_t n = pad16(g->E > 0 ? g->E : 1);
  CSAW_HIP(hipMalloc(&g->int_cdf[kind], sizeof(uint64_t) * n));
  CSAW_HIP(hipMemsetAsync(g->int_cdf[kind], 0, sizeof(uint64_t) * n, g->stream));
  GraphDev G = g->view();
  csaw_graph_view hv = g->hook_view();
  if (g->V > 0) {
    const int grid = grid_for((g->V + 3) / 4, 1, g->num_cu);
    if (kind == 0) k_int_cdf_small<0><<<grid, 256, 0, g->stream>>>(G, hv, g->int_cdf[kind]);
    else k_int_cdf_small<1><<<grid, 256, 0, g->stream>>>(G, hv, g->int_cdf[kind]);
    CSAW_LAUNCH_CHECK();
  }
  if (g->n_hubs > 0) {
    k_int_cdf_hubs<<<std::min(g->n_hubs, g->num_cu * 8), 256, 0, g->stream>>>(G, g->hub_cdf[kind], g->int_cdf[kind]);
    CSAW_LAUNCH_CHECK();
  }
  // Every per-vertex total below 2^32 (the largest entry is the largest
  // total): keep the cache as uint32, half the bytes per search.
  unsigned long long* d_max = nullptr;
  CSAW_HIP(hipMalloc(&d_max, sizeof(unsigned long long)));
  CSAW_HIP(hipMemsetAsync(d_max, 0, sizeof(unsigned long long), g->stream));
  if (g->E > 0) {
    k_max_u64<<<grid_for(g->E, 256, g->num_cu), 256, 0, g->stream>>>(g->int_cdf[kind], g->E, d_max);
    CSAW_LAUNCH_CHECK();
  }
  unsigned long long mx = 0;
  CSAW_HIP(hipMemcpyAsync(&mx, d_max, sizeof(mx), hipMemcpyDeviceToHost, g->stream));
  CSAW_HIP(hipStreamSynchronize(g->stream));
  CSAW_HIP(hipFree(d_max));
  if (mx < (1ull << 32) && !g->wide_int_cdf) {
    CSAW_HIP(hipMalloc(&g->int_cdf32[kind], sizeof(uint32_t) * n));
    k_narrow_u32<<<grid_for(n, 256, g->num_cu), 256, 0, g->stream>>>(g->int_cdf[kind], n, g->int_cdf32[kind]);
    CSAW_LAUNCH_CHECK();
    CSAW_HIP(hipStreamSynchronize(g->stream));
    CSAW_HIP(hipFree(g->int_cdf[kind]));
    g->int_cdf[kind] = nullptr;
  }
  return CSAW_OK;
}

// Level l of the search index: K[b] = C[min(16^l (b+1) - 1, E-1)] for b < n,
// zero for the padding b in [n, npad).
template <typename T>
__global__ void k_index_level(const T* __restrict__ C, int64_t E, int sh, int64_t n, int64_t npad, T* __restrict__ K) {
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < npad; b += (int64_t)gridDim.x * blockDim.x) {
    T v = 0;
    if (b < n) {
      const int64_t e = ((b + 1) << sh) - 1;
      v = C[e < E ? e : E - 1];
    }
    K[b] = v;
  }
}

template <typename T>
static int build_index(const T* C, int64_t E, T** out, hipStream_t stream, int num_cu) {
  const int64_t tot = cdf_index_size(E);
  CSAW_HIP(hipMalloc(out, sizeof(T) * (tot > 0 ? tot : 16)));
  int64_t at = 0;
  for (int l = 1; l <= cdf_index_levels(E); ++l) {
    const int sh = 4 * l;
    const int64_t n = (E + ((int64_t)1 << sh) - 1) >> sh, npad = cdf_index_level_size(E, l);
    k_index_level<T><<<grid_for(npad, 256, num_cu), 256, 0, stream>>>(C, E, sh, n, npad, *out + at);
    CSAW_LAUNCH_CHECK();
    at += npad;
  }
  CSAW_HIP(hipStreamSynchronize(stream));
  return CSAW_OK;
}

template <typename T>
__global__ void k_vertex_records(const int64_t* __restrict__ off, const T* __restrict__ C, int64_t V,
                                 VertexRecord* __restrict__ out) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = off[v];
    const int64_t d = off[v + 1] - o;
    VertexRecord r;
    r.off = o;
    uint32_t t = 0;
    if (d > 0) memcpy(&t, &C[o + d - 1], 4);
    r.total = t;
    r.deg = (int32_t)d;
    out[v] = r;
  }
}

int graph_prepare_vertex_records(csaw_graph* g, int32_t bias) {
  const int k = bias == CSAW_BIAS_DEGREE ? 0 : (bias == CSAW_BIAS_USER ? 1 : (bias == CSAW_BIAS_WEIGHT ? 2 : -1));
  if (k < 0 || g->vrec[k] || g->V == 0) return CSAW_OK;
  if (k == 2 && !g->cdf) return CSAW_OK;
  if (k < 2 && !g->int_cdf32[k]) return CSAW_OK;
  CSAW_HIP(hipMalloc(&g->vrec[k], sizeof(VertexRecord) * g->V));
  const int grid = grid_for(g->V, 256, g->num_cu);
  if (k == 2)
    k_vertex_records<float><<<grid, 256, 0, g->stream>>>(g->off, g->cdf, g->V, g->vrec[k]);
  else
    k_vertex_records<uint32_t><<<grid, 256, 0, g->stream>>>(g->off, g->int_cdf32[k], g->V, g->vrec[k]);
  CSAW_LAUNCH_CHECK();
  CSAW_HIP(hipStreamSynchronize(g->stream));
  return CSAW_OK;
}

int graph_prepare_index(csaw_graph* g, int32_t bias) {
  if (bias == CSAW_BIAS_WEIGHT) {
    CSAW_TRY(graph_prepare_bias(g, bias));
    if (!g->cdf_idx) CSAW_TRY(build_index<float>(g->cdf, g->E, &g->cdf_idx, g->stream, g->num_cu));
    return CSAW_OK;
  }
  if (bias != CSAW_BIAS_DEGREE && bias != CSAW_BIAS_USER) return CSAW_OK;
  const int kind = bias == CSAW_BIAS_DEGREE ? 0 : 1;
  CSAW_TRY(graph_prepare_int_cache(g, bias));
  if (g->int_cdf32[kind]) {
    if (!g->int_idx32[kind])
      CSAW_TRY(build_index<uint32_t>(g->int_cdf32[kind], g->E, &g->int_idx32[kind], g->stream, g->num_cu));
  } else if (!g->int_idx[kind]) {
    CSAW_TRY(build_index<uint64_t>(g->int_cdf[kind], g->E, &g->int_idx[kind], g->stream, g->num_cu));
  }
  return CSAW_OK;
}

}  // namespace csaw

using namespace csaw;

csaw::GraphDev csaw_graph::view(bool use_cache) const {
  GraphDev G;
  G.off = off;
  G.adj = adj;
  G.deg = deg;
  G.w = w;
  G.cdf = cdf;
  G.hub_vid = hub_vid;
  G.hub_pool = hub_pool;
  G.hub_cdf[0] = hub_cdf[0];
  G.hub_cdf[1] = hub_cdf[1];
  G.int_cdf[0] = use_cache ? int_cdf[0] : nullptr;
  G.int_cdf[1] = use_cache ? int_cdf[1] : nullptr;
  // an integer cache is searched through its index: both or neither
  G.int_idx[0] = use_cache ? int_idx[0] : nullptr;
  G.int_idx[1] = use_cache ? int_idx[1] : nullptr;
  for (int k = 0; k < 2; ++k) {
    G.int_cdf32[k] = use_cache ? int_cdf32[k] : nullptr;
    G.int_idx32[k] = use_cache ? int_idx32[k] : nullptr;
  }
  G.cdf_idx = use_cache ? cdf_idx : nullptr;
  for (int k = 0; k < 3; ++k) G.vrec[k] = use_cache ? vrec[k] : nullptr;
  G.n_hubs = n_hubs < 0 ? 0 : n_hubs;
  G.V = V;
  G.E = E;
  return G;
}

csaw_graph_view csaw_graph::hook_view() const {
  csaw_graph_view h;
  h.beg_pos = off;
  h.adj_list = adj;
  h.degree_list = deg;
  h.weight_list = w;
  h.vert_count = V;
  h.edge_count = E;
  return h;
}

static void graph_free(csaw_graph* g) {
  if (!g) return;
  (void)hipFree(g->off);
  (void)hipFree(g->adj);
  (void)hipFree(g->deg);
  (void)hipFree(g->w);
  (void)hipFree(g->cdf);
  (void)hipFree(g->hub_vid);
  (void)hipFree(g->hub_pool);
  (void)hipFree(g->hub_cdf[0]);
  (void)hipFree(g->hub_cdf[1]);
  (void)hipFree(g->int_cdf[0]);
  (void)hipFree(g->int_cdf[1]);
  (void)hipFree(g->int_idx[0]);
  (void)hipFree(g->int_idx[1]);
  for (int k = 0; k < 2; ++k) {
    (void)hipFree(g->int_cdf32[k]);
    (void)hipFree(g->int_idx32[k]);
  }
  (void)hipFree(g->cdf_idx);
  for (int k = 0; k < 3; ++k) (void)hipFree(g->vrec[k]);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

extern "C" int csaw_graph_destroy(csaw_graph* g) {
  graph_free(g);
  return CSAW_OK;
}

extern "C" int csaw_graph_create(const void* offsets, int offsets_are_i64, const int32_t* indices,
                                 const float* weights, int64_t V, int64_t E, csaw_graph** out) {
  if (!out || !offsets || (E > 0 && !indices) || V < 0 || E < 0 || V > 0x7fffffffLL) {
    set_last_error("csaw_graph_create: invalid argument");
    return CSAW_ERR_INVALID_ARG;
  }
  *out = nullptr;
  // host-side checks of the offsets (the reference asserts nothing, graph.hpp:37)
  for (int64_t v = 0; v <= V; ++v) {
    int64_t x = offsets_are_i64 ? ((const int64_t*)offsets)[v] : ((const int32_t*)offsets)[v];
    int64_t px = v ? (offsets_are_i64 ? ((const int64_t*)offsets)[v - 1] : ((const int32_t*)offsets)[v - 1]) : 0;
    if ((v == 0 && x != 0) || x < px || (v == V && x != E)) {
      set_last_error("csaw_graph_create: offsets must start at 0, be monotone and end at E");
      return CSAW_ERR_INVALID_GRAPH;
    }
  }
  csaw_graph* g = new csaw_graph();
  int st = graph_alloc_common(g);
  if (st != CSAW_OK) { graph_free(g); return st; }
  g->V = V;
  g->E = E;
  auto body = [&]() -> int {
    CSAW_HIP(hipMalloc(&g->off, sizeof(int64_t) * (V + 1)));
    CSAW_HIP(hipMalloc(&g->adj, sizeof(int32_t) * (E > 0 ? E : 1)));
    if (offsets_are_i64) {
      CSAW_HIP(hipMemcpy(g->off, offsets, sizeof(int64_t) * (V + 1), hipMemcpyHostToDevice));
    } else {
      int32_t* tmp;
      CSAW_HIP(hipMalloc(&tmp, sizeof(int32_t) * (V + 1)));
      CSAW_HIP(hipMemcpy(tmp, offsets, sizeof(int32_t) * (V + 1), hipMemcpyHostToDevice));
      k_widen_offsets<<<grid_for(V + 1, 256, g->num_cu), 256, 0, g->stream>>>(tmp, V + 1, g->off);
      CSAW_LAUNCH_CHECK();
      CSAW_HIP(hipStreamSynchronize(g->stream));
      CSAW_HIP(hipFree(tmp));
    }
    if (E > 0) CSAW_HIP(hipMemcpy(g->adj, indices, sizeof(int32_t) * E, hipMemcpyHostToDevice));
    if (weights) {
      CSAW_HIP(hipMalloc(&g->w, sizeof(float) * (E > 0 ? E : 1)));
      if (E > 0) CSAW_HIP(hipMemcpy(g->w, weights, sizeof(float) * E, hipMemcpyHostToDevice));
    }
    return finish_graph(g);
  };
  st = body();
  if (st != CSAW_OK) { graph_free(g); return st; }
  *out = g;
  return CSAW_OK;
}

static uint32_t thr32(double x) {
  if (x <= 0.0) return 0;
  if (x >= 1.0) return 0xFFFFFFFFu;
  return (uint32_t)(x * 4294967296.0);
}

namespace csaw {
// GPU RMAT: Philox pairs, both directions, radix sort, unique, CSR.
int rmat_generate_device(int scale, int edge_factor, double a, double b, double c, uint64_t key, hipStream_t stream,
                         int num_cu, int64_t* V_out, int64_t* E_out, int64_t** d_off, int32_t** d_adj) {
  const int64_t V = (int64_t)1 << scale;
  const int64_t M = (int64_t)edge_factor << scale;
  const size_t n = (size_t)(2 * M);
  uint64_t *k0, *k1;
  CSAW_HIP(hipMalloc(&k0, sizeof(uint64_t) * n));
  CSAW_HIP(hipMalloc(&k1, sizeof(uint64_t) * n));
  k_rmat_pairs<<<grid_for(M, 256, num_cu), 256, 0, stream>>>(scale, M, thr32(a), thr32(a + b), thr32(a + b + c),
                                                             (uint32_t)key, (uint32_t)(key >> 32), k0);
  CSAW_LAUNCH_CHECK();
  rocprim::double_buffer<uint64_t> db(k0, k1);
  size_t tmp_bytes = 0;
  CSAW_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, db, n, 0, 2 * scale, stream));
  void* tmp;
  CSAW_HIP(hipMalloc(&tmp, tmp_bytes));
  CSAW_HIP(rocprim::radix_sort_keys(tmp, tmp_bytes, db, n, 0, 2 * scale, stream));
  CSAW_HIP(hipStreamSynchronize(stream));
  CSAW_HIP(hipFree(tmp));
  uint64_t* sorted = db.current();
  uint64_t* uniq = db.alternate();
  uint32_t* d_cnt;
  CSAW_HIP(hipMalloc(&d_cnt, sizeof(uint32_t)));
  tmp_bytes = 0;
  CSAW_HIP(rocprim::unique(nullptr, tmp_bytes, sorted, uniq, d_cnt, n, rocprim::equal_to<uint64_t>(), stream));
  CSAW_HIP(hipMalloc(&tmp, tmp_bytes));
  CSAW_HIP(rocprim::unique(tmp, tmp_bytes, sorted, uniq, d_cnt, n, rocprim::equal_to<uint64_t>(), stream));
  uint32_t cnt = 0;
  CSAW_HIP(hipMemcpyAsync(&cnt, d_cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  CSAW_HIP(hipStreamSynchronize(stream));
  CSAW_HIP(hipFree(tmp));
  CSAW_HIP(hipFree(d_cnt));
  int64_t E = cnt;
  if (E > 0) {
    uint64_t last = 0;
    CSAW_HIP(hipMemcpy(&last, uniq + E - 1, sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (last == ((1ull << (2 * scale)) - 1)) --E;  // self-loop sentinel
  }
  CSAW_HIP(hipMalloc(d_off, sizeof(int64_t) * (V + 1)));
  CSAW_HIP(hipMalloc(d_adj, sizeof(int32_t) * (E > 0 ? E : 1)));
  k_keys_to_csr<<<grid_for(E + 1, 256, num_cu), 256, 0, stream>>>(uniq, E, scale, V, *d_adj, *d_off);
  CSAW_LAUNCH_CHECK();
  CSAW_HIP(hipStreamSynchronize(stream));
  CSAW_HIP(hipFree(k0));
  CSAW_HIP(hipFree(k1));
  *V_out = V;
  *E_out = E;
  return CSAW_OK;
}
}  // namespace csaw

extern "C" int csaw_graph_create_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t key,
                                      int with_weights, csaw_graph** out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) {
    set_last_error("csaw_graph_create_rmat: invalid argument");
    return CSAW_ERR_INVALID_ARG;
  }
  *out = nullptr;
  csaw_graph* g = new csaw_graph();
  int st = graph_alloc_common(g);
  if (st != CSAW_OK) { graph_free(g); return st; }
  auto body = [&]() -> int {
    CSAW_TRY(rmat_generate_device(scale, edge_factor, a, b, c, key, g->stream, g->num_cu, &g->V, &g->E, &g->off,
                                  &g->adj));
    const int64_t E = g->E;
    if (with_weights) {
      CSAW_HIP(hipMalloc(&g->w, sizeof(float) * (E > 0 ? E : 1)));
      if (E > 0) {
        k_weights<<<grid_for(E, 256, g->num_cu), 256, 0, g->stream>>>((uint32_t)key, (uint32_t)(key >> 32), E, g->w);
        CSAW_LAUNCH_CHECK();
      }
    }
    return finish_graph(g);
  };
  st = body();
  if (st != CSAW_OK) { graph_free(g); return st; }
  *out = g;
  return CSAW_OK;
}

// The reference's on-disk CSR (graph.hpp:18-74): beg.bin holds V+1 int64
// offsets, csr.bin E int64 neighbour ids; both are narrowed on load (ids to
// int32).  The optional weight file holds E float32 values.
static int read_file(const char* path, std::vector<char>& buf) {
  FILE* f = fopen(path, "rb");
  if (!f) {
    set_last_error(std::string("cannot open ") + path);
    return CSAW_ERR_IO;
  }
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  buf.resize(sz > 0 ? (size_t)sz : 0);
  size_t got = sz > 0 ? fread(buf.data(), 1, (size_t)sz, f) : 0;
  fclose(f);
  if ((long)got != sz) {
    set_last_error(std::string("short read of ") + path);
    return CSAW_ERR_IO;
  }
  return CSAW_OK;
}

extern "C" int csaw_graph_load_bin(const char* beg_file, const char* csr_file, const char* weight_file,
                                   csaw_graph** out) {
  if (!beg_file || !csr_file || !out) return CSAW_ERR_INVALID_ARG;
  std::vector<char> beg, csr, wf;
  CSAW_TRY(read_file(beg_file, beg));
  CSAW_TRY(read_file(csr_file, csr));
  if (beg.size() < 8 || beg.size() % 8 || csr.size() % 8) {
    set_last_error("beg.bin / csr.bin sizes are not multiples of 8 bytes");
    return CSAW_ERR_INVALID_GRAPH;
  }
  const int64_t V = (int64_t)(beg.size() / 8) - 1;
  const int64_t* off = (const int64_t*)beg.data();
  const int64_t E = off[V];
  if (E < 0 || (int64_t)(csr.size() / 8) < E) {
    set_last_error("csr.bin holds fewer ids than beg.bin's last offset");
    return CSAW_ERR_INVALID_GRAPH;
  }
  std::vector<int32_t> adj(E);
  const int64_t* ids = (const int64_t*)csr.data();
  for (int64_t i = 0; i < E; ++i) {
    if (ids[i] < 0 || ids[i] >= V) {
      set_last_error("csr.bin id out of range");
      return CSAW_ERR_INVALID_GRAPH;
    }
    adj[i] = (int32_t)ids[i];
  }
  const float* w = nullptr;
  if (weight_file) {
    CSAW_TRY(read_file(weight_file, wf));
    if ((int64_t)(wf.size() / 4) < E) {
      set_last_error("weight file holds fewer than E float32 values");
      return CSAW_ERR_INVALID_GRAPH;
    }
    w = (const float*)wf.data();
  }
  return csaw_graph_create(off, 1, adj.data(), w, V, E, out);
}

extern "C" int csaw_graph_info(const csaw_graph* g, int64_t* V, int64_t* E, int64_t* max_degree) {
  if (!g) return CSAW_ERR_INVALID_ARG;
  if (V) *V = g->V;
  if (E) *E = g->E;
  if (max_degree) *max_degree = g->max_deg;
  return CSAW_OK;
}

extern "C" int csaw_graph_copy_csr(const csaw_graph* g, int64_t* offsets, int32_t* indices, float* weights) {
  if (!g) return CSAW_ERR_INVALID_ARG;
  if (offsets) CSAW_HIP(hipMemcpy(offsets, g->off, sizeof(int64_t) * (g->V + 1), hipMemcpyDeviceToHost));
  if (indices && g->E > 0) CSAW_HIP(hipMemcpy(indices, g->adj, sizeof(int32_t) * g->E, hipMemcpyDeviceToHost));
  if (weights) {
    if (!g->w) return CSAW_ERR_INVALID_ARG;
    if (g->E > 0) CSAW_HIP(hipMemcpy(weights, g->w, sizeof(float) * g->E, hipMemcpyDeviceToHost));
  }
  return CSAW_OK;
}

extern "C" int csaw_seeds(const csaw_graph* g, uint64_t key, uint32_t request, int64_t n, int32_t* out) {
  if (!g || n < 0 || (n > 0 && !out) || g->V < 1) return CSAW_ERR_INVALID_ARG;
  if (n == 0) return CSAW_OK;
  int32_t* d;
  CSAW_HIP(hipMalloc(&d, sizeof(int32_t) * n));
  k_seeds<<<grid_for(n, 256, g->num_cu), 256, 0, g->stream>>>((uint32_t)key, (uint32_t)(key >> 32), request, g->V, n, d);
  CSAW_LAUNCH_CHECK();
  CSAW_HIP(hipMemcpyAsync(out, d, sizeof(int32_t) * n, hipMemcpyDeviceToHost, g->stream));
  CSAW_HIP(hipStreamSynchronize(g->stream));
  CSAW_HIP(hipFree(d));
  return CSAW_OK;
}

extern "C" const char* csaw_last_error(void) { return csaw::last_error(); }

// Host-side seeds for graphs that only live in host memory.
extern "C" int csaw_seeds_v(int64_t V, uint64_t key, uint32_t request, int64_t n, int32_t* out) {
  if (V < 1 || n < 0 || (n > 0 && !out)) return CSAW_ERR_INVALID_ARG;
  for (int64_t i = 0; i < n; ++i) {
    const U4 o = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), request, kDomSeeds << 28, (uint32_t)key,
                               (uint32_t)(key >> 32));
    out[i] = (int32_t)mulhi64(philox_t64(o), (uint64_t)V);
  }
  return CSAW_OK;
}
