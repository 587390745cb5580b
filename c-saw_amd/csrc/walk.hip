// Biased random walks (FrontierSize 1, NeighborSize 1, Depth = walk length;
// the reference's README.md:34 "Depth/Length" path through check(),
// non-stream/main.cu:21-93, with one draw per step).
//
// One lane per walker.  Each step: degree and offset of the current vertex,
// one ITS draw philox(key; step, walker lo, walker hi, DOM_WALK<<28 | request),
// then the neighbour.  Steps are staged in LDS in tiles of kWalkTile and
// written as 64-byte row segments of the [walkers x (len+1)] matrix.
#include <algorithm>

#include "graph.hpp"
#include "sample_kernels.hpp"
#include "select.hpp"

namespace csaw {

constexpr int kWalkTile = 16;
constexpr int kWalkThreads = 256;

// FP32 pick given the vertex's CDF total T = C[o + d - 1].
__device__ __forceinline__ int32_t walk_pick_f32(const GraphDev& G, int64_t o, int32_t d, float T, const U4& r) {
  if (T > 0.0f) {
    const float x = __fmul_rn(philox_u24(r), T);
    if (G.cdf_idx) return cdf_index_search<float>(G.cdf, G.cdf_idx, G.E, o, d, x);
    return (int32_t)upper_bound_dev<float>(G.cdf + o, d, x);
  }
  return (int32_t)mulhi64(philox_t64(r), (uint64_t)d);
}

// Integer pick over a uint32 CDF cache given its total T = C[o + d - 1].
__device__ __forceinline__ int32_t walk_pick_i32(const GraphDev& G, int ck, int64_t o, int32_t d, uint32_t T,
                                                 const U4& r) {
  const uint64_t t = philox_t64(r);
  if (T > 0) return cdf_index_search<uint32_t>(G.int_cdf32[ck], G.int_idx32[ck], G.E, o, d, (uint32_t)mulhi64(t, T));
  return (int32_t)mulhi64(t, (uint64_t)d);
}

template <int KIND>
__device__ __forceinline__ int32_t walk_pick(const GraphDev& G, const csaw_graph_view& hv, int32_t bias, int32_t v,
                                             int64_t o, int32_t d, const U4& r) {
  const uint64_t t = philox_t64(r);
  if (KIND == kF32) return walk_pick_f32(G, o, d, G.cdf[o + d - 1], r);
  if (KIND == kInt) {
    const int ck = bias == CSAW_BIAS_USER ? 1 : 0;
    if (G.int_cdf32[ck] && G.int_idx32[ck]) {
      const uint64_t T = G.int_cdf32[ck][o + d - 1];
      if (T > 0)
        return cdf_index_search<uint32_t>(G.int_cdf32[ck], G.int_idx32[ck], G.E, o, d, (uint32_t)mulhi64(t, T));
      return (int32_t)mulhi64(t, (uint64_t)d);
    }
    if (G.int_cdf[ck] && G.int_idx[ck]) {
      const uint64_t T = G.int_cdf[ck][o + d - 1];
      if (T > 0) return cdf_index_search<uint64_t>(G.int_cdf[ck], G.int_idx[ck], G.E, o, d, mulhi64(t, T));
      return (int32_t)mulhi64(t, (uint64_t)d);
    }
    if (d > kHubMin) {
      const int32_t h = hub_find(G.hub_vid, G.n_hubs, v);
      const uint64_t* C = G.hub_cdf[bias == CSAW_BIAS_USER ? 1 : 0] + G.hub_pool[h];
      const uint64_t T = C[d - 1];
      if (T > 0) return (int32_t)upper_bound_dev<uint64_t>(C, d, mulhi64(t, T));
      return (int32_t)mulhi64(t, (uint64_t)d);
    }
    uint64_t T = 0;
    for (int32_t i = 0; i < d; ++i) T += int_bias(G, hv, bias, G.adj[o + i]);
    if (T == 0) return (int32_t)mulhi64(t, (uint64_t)d);
    const uint64_t x = mulhi64(t, T);
    uint64_t acc = 0;
    int32_t i = 0;
    for (; i < d - 1; ++i) {
      acc += int_bias(G, hv, bias, G.adj[o + i]);
      if (acc > x) break;
    }
    return i;
  }
  return (int32_t)mulhi64(t, (uint64_t)d);
}

template <int KIND>
__global__ __launch_bounds__(kWalkThreads) void k_walk(GraphDev G, csaw_graph_view hv, int32_t bias, uint32_t k0,
                                                       uint32_t k1, uint32_t request, const int32_t* __restrict__ seeds,
                                                       int64_t n, int32_t len, int32_t* __restrict__ W,
                                                       unsigned long long* __restrict__ steps_out) {
  __shared__ int32_t tile[kWalkTile][kWalkThreads + 1];
  const uint32_t c3 = (kDomWalk << 28) | request;
  // vertex records of the bias's cache (offset, degree, CDF total in one load)
  const int ck = bias == CSAW_BIAS_USER ? 1 : 0;
  const VertexRecord* __restrict__ vt =
      KIND == kF32 ? G.vrec[2] : (KIND == kInt && G.int_cdf32[ck] && G.int_idx32[ck] ? G.vrec[ck] : nullptr);
  unsigned long long moved = 0, bytes = 0;
  for (int64_t w0 = (int64_t)blockIdx.x * kWalkThreads; w0 < n; w0 += (int64_t)gridDim.x * kWalkThreads) {
    const int64_t w = w0 + threadIdx.x;
    const bool act = w < n;
    int32_t v = act ? seeds[w] : -1;
    int32_t* row = W + w * (int64_t)(len + 1);
    if (act) row[0] = v;
    for (int32_t t0 = 0; t0 < len; t0 += kWalkTile) {
      const int32_t tn = min(kWalkTile, len - t0);
      for (int32_t k = 0; k < tn; ++k) {
        const int32_t t = t0 + k;
        if (act && v >= 0) {
          int64_t o;
          int32_t d;
          uint32_t T = 0;
          if (KIND != kUniform && vt) {
            const VertexRecord rec = vt[v];  // offset, degree and CDF total in one 16-byte load
            o = rec.off;
            d = rec.deg;
            T = rec.total;
          } else {
            o = G.off[v];
            d = (int32_t)(G.off[v + 1] - o);  // same line as off[v]
          }
          if (d == 0) {
            v = -1;
          } else {
            const U4 r = philox4x32_10((uint32_t)t, (uint32_t)w, (uint32_t)(w >> 32), c3, k0, k1);
            const int32_t pick = KIND != kUniform && vt
                                     ? (KIND == kF32 ? walk_pick_f32(G, o, d, __uint_as_float(T), r)
                                                     : walk_pick_i32(G, ck, o, d, T, r))
                                     : walk_pick<KIND>(G, hv, bias, v, o, d, r);
            v = G.adj[o + pick];
            ++moved;
            // algorithmic bytes of the step: offset 8 + degree 4 + picked id 4 +
            // walk entry 4, plus the CDF total and ceil(log2 d) probes (FP32 /
            // cached integer) or two passes over adj + degree (integer, built)
            const int l = ceil_log2_u32((uint32_t)d);
            const bool cached = KIND == kInt && (G.int_cdf[bias == CSAW_BIAS_USER ? 1 : 0] ||
                                                 G.int_cdf32[bias == CSAW_BIAS_USER ? 1 : 0]);
            bytes += 20 + (KIND == kF32 ? 4 + 4 * l
                                        : (KIND == kInt ? (cached || d > kHubMin ? 8 + 8 * l : 16 * (unsigned long long)d)
                                                        : 0));
          }
        }
        tile[k][threadIdx.x] = v;
      }
      if (act)
        for (int32_t k = 0; k < tn; ++k) row[1 + t0 + k] = tile[k][threadIdx.x];
    }
  }
  // wave-aggregated step count
  unsigned long long s = moved, b = bytes;
  for (int off = 32; off >= 1; off >>= 1) {
    s += __shfl_xor(s, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  if (lane_id() == 0 && s) {
    atomicAdd(steps_out, s);
    atomicAdd(steps_out + 1, b);
  }
}

int launch_walk(hipStream_t st, int num_cu, int per_cu, const GraphDev& G, const csaw_graph_view& hv, int32_t bias,
                uint64_t key, uint32_t request, const int32_t* seeds, int64_t n, int32_t len, int32_t* W,
                unsigned long long* steps) {
  const int64_t need = (n + kWalkThreads - 1) / kWalkThreads;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)num_cu * per_cu));
  const int kind = bias_kind(bias);
  if (kind == kUniform)
    k_walk<kUniform><<<grid, kWalkThreads, 0, st>>>(G, hv, bias, (uint32_t)key, (uint32_t)(key >> 32), request, seeds, n,
                                                    len, W, steps);
  else if (kind == kInt)
    k_walk<kInt><<<grid, kWalkThreads, 0, st>>>(G, hv, bias, (uint32_t)key, (uint32_t)(key >> 32), request, seeds, n, len,
                                                W, steps);
  else
    k_walk<kF32><<<grid, kWalkThreads, 0, st>>>(G, hv, bias, (uint32_t)key, (uint32_t)(key >> 32), request, seeds, n, len,
                                                W, steps);
  CSAW_LAUNCH_CHECK();
  return CSAW_OK;
}

}  // namespace csaw

using namespace csaw;

struct csaw_walk_plan {
  csaw_graph* g = nullptr;
  int32_t bias = 0;
  uint64_t key = 0;
  uint32_t request = 0;
  int64_t n = 0;
  int32_t len = 0;
  int32_t* d_seeds = nullptr;
  int32_t* d_W = nullptr;
  unsigned long long* d_steps = nullptr;
  hipStream_t stream = nullptr;
  // workgroups per CU: 1,048,576 walkers x 80 on RMAT-22 take 4.3 ms at 4 (4,096
  // tiles of 256 walkers spread evenly, four per workgroup), 4.7 ms at 8, 5.5 at 2
  int per_cu = 4;
  double last_ms = 0.0;
  int64_t last_steps = 0;
  double last_bytes = 0.0;
};

extern "C" int csaw_walk_plan_create(csaw_graph* g, int32_t bias, uint64_t seed, uint32_t request_id,
                                     const int32_t* seeds, int64_t n_walkers, int32_t len, csaw_walk_plan** out) {
  // request ids share counter word 3 with the domain tag in bits 28-31
  if (!g || !out || n_walkers < 0 || len < 0 || (n_walkers > 0 && !seeds) || bias < 0 || bias > CSAW_BIAS_USER ||
      request_id >= (1u << 28))
    return CSAW_ERR_INVALID_ARG;
  *out = nullptr;
  for (int64_t i = 0; i < n_walkers; ++i)
    if (seeds[i] < 0 || seeds[i] >= g->V) return CSAW_ERR_INVALID_ARG;
  // walks search the CDF cache (FP32 CDF or integer cache) through its index
  CSAW_TRY(graph_prepare_index(g, bias));
  CSAW_TRY(graph_prepare_vertex_records(g, bias));
  csaw_walk_plan* p = new csaw_walk_plan();
  p->g = g;
  p->bias = bias;
  p->key = seed;
  p->request = request_id;
  p->n = n_walkers;
  p->len = len;
  auto body = [&]() -> int {
    CSAW_HIP(hipMalloc(&p->d_seeds, sizeof(int32_t) * std::max<int64_t>(n_walkers, 1)));
    CSAW_HIP(hipMalloc(&p->d_W, sizeof(int32_t) * std::max<int64_t>(n_walkers * (len + 1), 1)));
    CSAW_HIP(hipMalloc(&p->d_steps, 2 * sizeof(unsigned long long)));
    if (n_walkers > 0)
      CSAW_HIP(hipMemcpy(p->d_seeds, seeds, sizeof(int32_t) * n_walkers, hipMemcpyHostToDevice));
    CSAW_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    return CSAW_OK;
  };
  int st = body();
  if (st != CSAW_OK) {
    (void)hipFree(p->d_seeds);
    (void)hipFree(p->d_W);
    (void)hipFree(p->d_steps);
    delete p;
    return st;
  }
  *out = p;
  return CSAW_OK;
}

extern "C" int csaw_walk_plan_run(csaw_walk_plan* p, void* stream, int sync) {
  if (!p) return CSAW_ERR_INVALID_ARG;
  hipStream_t st = stream ? (hipStream_t)stream : p->stream;
  hipEvent_t e0, e1;
  CSAW_HIP(hipEventCreate(&e0));
  CSAW_HIP(hipEventCreate(&e1));
  CSAW_HIP(hipMemsetAsync(p->d_steps, 0, 2 * sizeof(unsigned long long), st));
  CSAW_HIP(hipEventRecord(e0, st));
  CSAW_TRY(launch_walk(st, p->g->num_cu, p->per_cu, p->g->view(true), p->g->hook_view(), p->bias, p->key, p->request,
                       p->d_seeds, p->n, p->len, p->d_W, p->d_steps));
  CSAW_HIP(hipEventRecord(e1, st));
  if (sync) {
    CSAW_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CSAW_HIP(hipEventElapsedTime(&ms, e0, e1));
    p->last_ms = ms;
    unsigned long long s[2] = {0, 0};
    CSAW_HIP(hipMemcpy(s, p->d_steps, sizeof(s), hipMemcpyDeviceToHost));
    p->last_steps = (int64_t)s[0];
    p->last_bytes = (double)s[1];
  }
  CSAW_HIP(hipEventDestroy(e0));
  CSAW_HIP(hipEventDestroy(e1));
  return CSAW_OK;
}

extern "C" int csaw_walk_plan_result(const csaw_walk_plan* p, int32_t* walks, int64_t* steps, double* kernel_ms,
                                     double* algo_bytes) {
  if (!p) return CSAW_ERR_INVALID_ARG;
  if (walks && p->n > 0)
    CSAW_HIP(hipMemcpy(walks, p->d_W, sizeof(int32_t) * p->n * (p->len + 1), hipMemcpyDeviceToHost));
  if (steps) *steps = p->last_steps;
  if (kernel_ms) *kernel_ms = p->last_ms;
  if (algo_bytes) *algo_bytes = p->last_bytes;
  return CSAW_OK;
}

extern "C" int csaw_walk_plan_destroy(csaw_walk_plan* p) {
  if (!p) return CSAW_OK;
  (void)hipFree(p->d_seeds);
  (void)hipFree(p->d_W);
  (void)hipFree(p->d_steps);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return CSAW_OK;
}

extern "C" int csaw_random_walk(csaw_graph* g, int32_t bias, uint64_t seed, uint32_t request_id, const int32_t* seeds,
                                int64_t n_walkers, int32_t len, int32_t* walks, csaw_stats* stats) {
  csaw_walk_plan* p = nullptr;
  CSAW_TRY(csaw_walk_plan_create(g, bias, seed, request_id, seeds, n_walkers, len, &p));
  int st = csaw_walk_plan_run(p, nullptr, 1);
  if (st == CSAW_OK) st = csaw_walk_plan_result(p, walks, nullptr, nullptr, nullptr);
  if (st == CSAW_OK && stats) {
    memset(stats, 0, sizeof(*stats));
    stats->edges = p->last_steps;
    stats->kernel_ms = p->last_ms;
  }
  csaw_walk_plan_destroy(p);
  return st;
}
