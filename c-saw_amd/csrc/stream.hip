// Streaming sampler for graphs larger than the device budget: the reference's
// streaming path (streaming/streaming_sampling.cu:412-425 block_augument,
// 427-728 Sampler with the per-partition queues of check(), 199-384),
// re-designed for a real out-of-core run:
//  - the CSR stays in pinned host memory; the offsets and degrees are
//    device-resident (every kernel needs d(v); degree bias needs d(u) of any
//    neighbour);
//  - vertex ranges with equal edge counts form the partitions; per hop the
//    frontier is binned by partition and each partition's adjacency slice is
//    copied on a second stream into one of two device buffers while the
//    previous partition expands (double buffering);
//  - for integer biases, every distinct frontier vertex of the partition gets
//    its inclusive CDF built once per (hop, partition), at its own position
//    in a slice-shaped array (one wave per vertex; vertices above kHubSeg
//    neighbours split across workgroups), so every draw is a search of that
//    array instead of a pass over the neighbour list per frontier entry;
//  - output positions come from the same global prefix sums as the
//    in-memory run, so the edge list is bit-identical to it.
// In batched mode each partition's entries are binned into chunks by degree
// class and every chunk runs the kernel variant the schedule picks (the
// fixed heuristic, a forced variant, a random one or the learned cost
// model), exactly as in the in-memory batched run; in plain mode one wave
// expands each listed entry.
// Every device allocation is charged against budget_bytes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "graph.hpp"
#include "pack.hpp"
#include "plan.hpp"
#include "sample_kernels.hpp"

struct csaw_host_graph {
  int64_t V = 0, E = 0, max_deg = 0;
  int64_t* off = nullptr;  // pinned host, V+1
  int32_t* adj = nullptr;  // pinned host, E
  // packed copy of adj (csaw_host_graph_pack): blocks of 64 consecutive
  // entries, block b at byte boff[b] of pdata
  uint8_t* pdata = nullptr;     // pinned
  uint32_t* boff32 = nullptr;   // pinned, low 32 bits of boff (nblk + 1)
  std::vector<uint64_t> boff;   // nblk + 1
  int64_t nblk = 0;
};

namespace csaw {

__global__ void k_stream_degree(const int64_t* __restrict__ off, int64_t V, int32_t* __restrict__ deg) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
    deg[v] = (int32_t)(off[v + 1] - off[v]);
}

// Host bytes one frontier entry would read if its partition's adjacency
// were read in place over the bus instead of copied: a 64-byte request per
// picked neighbour, plus, for the first entry of each distinct vertex, the
// neighbour list an integer CDF is built from.
constexpr int64_t kBusRead = 64;
constexpr double kDirectFactor = 2.0;  // default: in place when 2x that estimate < the slice

// partition id per frontier entry, per-partition counts and direct-read byte
// estimates; for integer biases the vertices whose entries draw (d > 0,
// F > 0, not a take-all copy) are marked in fbits (up to kHubSeg neighbours)
// or hubbits (more)
__global__ __launch_bounds__(256) void k_part_bin(GraphDev G, const ReqDev* __restrict__ R, HopDev H,
                                                  const int32_t* __restrict__ vb, int P, uint8_t* __restrict__ pid,
                                                  uint32_t* __restrict__ cnt, unsigned long long* __restrict__ pbytes,
                                                  uint32_t* __restrict__ fbits, uint32_t* __restrict__ hubbits) {
  __shared__ uint32_t lcnt[256];
  __shared__ unsigned long long lbytes[256];
  for (int i = threadIdx.x; i < P; i += blockDim.x) {
    lcnt[i] = 0;
    lbytes[i] = 0;
  }
  __syncthreads();
  const int64_t n = *H.n_ptr;
  const ReqDev& r = R[0];
  const int32_t F = r.fanout[H.hop];
  const bool int_bias = r.bias == CSAW_BIAS_DEGREE || r.bias == CSAW_BIAS_USER;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t v = H.fv[e];
    int lo = 0, hi = P;  // last p with vb[p] <= v
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (vb[mid] <= v) lo = mid; else hi = mid;
    }
    pid[e] = (uint8_t)lo;
    atomicAdd(&lcnt[lo], 1u);
    const int32_t d = G.deg[v];
    const bool picks = d > 0 && F > 0 && !(r.dedup && d <= F);
    unsigned long long eb = (unsigned long long)(kBusRead * entry_k(d, F, r.dedup));
    if (int_bias && picks) {
      // the CDF build reads the list once per distinct vertex: counted by
      // the entry that sets the vertex's mark
      uint32_t* bits = d > kHubSeg ? hubbits : fbits;
      const uint32_t bit = 1u << (v & 31);
      if (!(bits[v >> 5] & bit) && !(atomicOr(&bits[v >> 5], bit) & bit)) eb += 4ull * (unsigned long long)d;
    }
    atomicAdd(&lbytes[lo], eb);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P; i += blockDim.x) {
    if (lcnt[i]) atomicAdd(&cnt[i], lcnt[i]);
    if (lbytes[i]) atomicAdd(&pbytes[i], lbytes[i]);
  }
}

// The vertices marked in fbits within [v0, v1), compacted (in no particular
// order) into list[0 .. *n).
__global__ __launch_bounds__(256) void k_mark_list(const uint32_t* __restrict__ fbits, int32_t v0, int32_t v1,
                                                   int32_t* __restrict__ list, uint32_t* __restrict__ n) {
  const int64_t w0 = v0 >> 5, w1 = ((int64_t)v1 + 31) >> 5;
  for (int64_t wb = w0 + blockIdx.x * (int64_t)blockDim.x; wb < w1; wb += (int64_t)gridDim.x * blockDim.x) {
    const int64_t wi = wb + threadIdx.x;
    uint32_t m = 0;
    if (wi < w1) {
      m = fbits[wi];
      const int64_t lo = (int64_t)v0 - wi * 32, hi = (int64_t)v1 - wi * 32;  // keep bits lo .. hi-1
      if (lo > 0) m &= lo >= 32 ? 0u : ~0u << lo;
      if (hi < 32) m &= hi <= 0 ? 0u : ~0u >> (32 - hi);
    }
    const uint32_t c = (uint32_t)__popc(m);
    const uint32_t incl = wave_incl_scan_u32(c);
    uint32_t base = 0;
    if (lane_id() == kWave - 1 && incl) base = atomicAdd(n, incl);
    base = readlane_u(base, kWave - 1) + incl - c;
    while (m) {
      const int b = __builtin_ctz(m);
      m &= m - 1;
      list[base++] = (int32_t)(wi * 32 + b);
    }
  }
}

// Inclusive integer CDFs of the listed vertices (up to kHubSeg neighbours),
// one wave per vertex, written at their slice positions: C[o - a + i] for
// entry i of vertex v (o = off[v]).
template <int KIND>  // 0 degree, 1 user hook
__global__ __launch_bounds__(256) void k_frontier_cdf(GraphDev G, csaw_graph_view hv, const int32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ n_ptr, int64_t a,
                                                      uint64_t* __restrict__ C) {
  const int lane = lane_id();
  const int64_t n = *n_ptr;
  for (int64_t k = blockIdx.x * 4 + threadIdx.x / kWave; k < n; k += (int64_t)gridDim.x * 4) {
    const int32_t v = list[k];
    const int64_t o = G.off[v];
    const int32_t d = (int32_t)(G.off[v + 1] - o);
    uint64_t acc = 0;
    // eight 64-entry rows at a time: their ids, then their biases, all
    // requested before the scans (the list may be read over the bus)
    constexpr int kRows = 8;
    for (int i0 = 0; i0 < d; i0 += kRows * kWave) {
      int32_t u[kRows];
      uint64_t w[kRows];
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        const int i = i0 + r * kWave + lane;
        u[r] = i < d ? G.adj[o + i] : -1;
      }
#pragma unroll
      for (int r = 0; r < kRows; ++r)
        w[r] = u[r] < 0 ? 0ull : (KIND == 0 ? (uint64_t)G.deg[u[r]] : (uint64_t)EdgeBias(u[r], &hv));
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        const int i = i0 + r * kWave + lane;
        const uint64_t sc = wave_incl_scan_u64(w[r]);
        if (i < d) C[o - a + i] = acc + sc;
        acc += readlane_u64(sc, kWave - 1);
      }
    }
  }
}

// Entries grouped by partition into list (part_off[p] + slot): each
// workgroup takes 1024 consecutive entries, ranks them per partition in LDS
// and reserves each partition's run with one global atomic.
__global__ __launch_bounds__(256) void k_part_scatter(HopDev H, const uint8_t* __restrict__ pid, int P,
                                                      const uint32_t* __restrict__ part_off,
                                                      uint32_t* __restrict__ cursor, int32_t* __restrict__ list) {
  __shared__ uint32_t lcnt[256];
  __shared__ uint32_t lbase[256];
  const int64_t n = *H.n_ptr;
  for (int64_t b0 = blockIdx.x * 1024ll; b0 < n; b0 += (int64_t)gridDim.x * 1024) {
    for (int i = threadIdx.x; i < P; i += blockDim.x) lcnt[i] = 0;
    __syncthreads();
    int pp[4];
    uint32_t slot[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t e = b0 + k * 256 + threadIdx.x;
      pp[k] = e < n ? (int)pid[e] : -1;
      if (pp[k] >= 0) slot[k] = atomicAdd(&lcnt[pp[k]], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < P; i += blockDim.x)
      if (lcnt[i]) lbase[i] = part_off[i] + atomicAdd(&cursor[i], lcnt[i]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (pp[k] >= 0) list[lbase[pp[k]] + slot[k]] = (int32_t)(b0 + k * 256 + threadIdx.x);
    __syncthreads();
  }
}

constexpr int64_t kChunkedMin = 16384;  // frontier entries below which a partition is not binned
constexpr int64_t kPieceBlocks = 1 << 17;  // packed blocks per staged piece (8M entries)

// One wave per block of a staged piece: blocks pb0 + k (k < nb) at bytes
// (off[k] - off[0]) of data; entries outside [a, e) are not written; the
// slice holds entry i at slice[i - a].
__global__ __launch_bounds__(256) void k_unpack(const uint8_t* __restrict__ data, const uint32_t* __restrict__ off,
                                                int64_t pb0, int64_t nb, int64_t E, int64_t a, int64_t e,
                                                int32_t* __restrict__ slice) {
  const int lane = lane_id();
  for (int64_t k = blockIdx.x * 4 + threadIdx.x / kWave; k < nb; k += (int64_t)gridDim.x * 4) {
    const uint8_t* blk = data + (uint32_t)(off[k] - off[0]);
    const int64_t first = (pb0 + k) * kPackBlock;
    const int n = (int)min((int64_t)kPackBlock, E - first);
    const int w = blk[0];
    uint32_t v = 0;
    if (lane == 0) {
      v = (uint32_t)blk[1] | ((uint32_t)blk[2] << 8) | ((uint32_t)blk[3] << 16) | ((uint32_t)blk[4] << 24);
    } else if (lane < n) {
      const uint8_t* q = blk + 5 + (lane - 1) * w;
      uint32_t z = 0;
      for (int i = 0; i < w; ++i) z |= (uint32_t)q[i] << (8 * i);
      v = (z >> 1) ^ (0u - (z & 1u));  // zigzag decode
    }
    const uint32_t x = wave_incl_scan_u32(v);  // ids: first + running differences (mod 2^32)
    const int64_t i = first + lane;
    if (lane < n && i >= a && i < e) slice[i - a] = (int32_t)x;
  }
}

struct Budget {
  int64_t limit, used = 0, peak = 0;
  std::vector<void*> ptrs;
  explicit Budget(int64_t l) : limit(l) {}
  template <typename T>
  int alloc(T** p, int64_t count) {
    const int64_t bytes = std::max<int64_t>(count, 1) * (int64_t)sizeof(T);
    if (used + bytes > limit) {
      set_last_error("device budget exceeded: need " + std::to_string(used + bytes) + " of " + std::to_string(limit));
      return CSAW_ERR_BUDGET_TOO_SMALL;
    }
    CSAW_HIP(hipMalloc((void**)p, bytes));
    ptrs.push_back((void*)*p);
    used += bytes;
    peak = std::max(peak, used);
    return CSAW_OK;
  }
  ~Budget() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

static int grid_of(int64_t n, int num_cu) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)num_cu * 8));
}

}  // namespace csaw

using namespace csaw;

extern "C" int csaw_host_graph_create_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t key,
                                           csaw_host_graph** out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1) return CSAW_ERR_INVALID_ARG;
  *out = nullptr;
  int num_cu = 0;
  CSAW_TRY(device_init(&num_cu));
  hipStream_t st;
  CSAW_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  int64_t V = 0, E = 0;
  int64_t* d_off = nullptr;
  int32_t* d_adj = nullptr;
  int rc = rmat_generate_device(scale, edge_factor, a, b, c, key, st, num_cu, &V, &E, &d_off, &d_adj);
  csaw_host_graph* g = nullptr;
  if (rc == CSAW_OK) {
    g = new csaw_host_graph();
    g->V = V;
    g->E = E;
    auto body = [&]() -> int {
      CSAW_HIP(hipHostMalloc((void**)&g->off, sizeof(int64_t) * (V + 1), hipHostMallocDefault));
      CSAW_HIP(hipHostMalloc((void**)&g->adj, sizeof(int32_t) * std::max<int64_t>(E, 1), hipHostMallocDefault));
      CSAW_HIP(hipMemcpy(g->off, d_off, sizeof(int64_t) * (V + 1), hipMemcpyDeviceToHost));
      if (E > 0) CSAW_HIP(hipMemcpy(g->adj, d_adj, sizeof(int32_t) * E, hipMemcpyDeviceToHost));
      return CSAW_OK;
    };
    rc = body();
  }
  (void)hipFree(d_off);
  (void)hipFree(d_adj);
  (void)hipStreamDestroy(st);
  if (rc != CSAW_OK) {
    if (g) {
      (void)hipHostFree(g->off);
      (void)hipHostFree(g->adj);
      delete g;
    }
    return rc;
  }
  for (int64_t v = 0; v < V; ++v) g->max_deg = std::max(g->max_deg, g->off[v + 1] - g->off[v]);
  *out = g;
  return CSAW_OK;
}

extern "C" int csaw_host_graph_from_csr(const int64_t* off, const int32_t* adj, int64_t V, int64_t E,
                                        csaw_host_graph** out) {
  if (!off || (E > 0 && !adj) || !out || V < 0 || V > 0x7fffffffLL || off[0] != 0 || off[V] != E)
    return CSAW_ERR_INVALID_ARG;
  for (int64_t v = 0; v < V; ++v)
    if (off[v + 1] < off[v]) return CSAW_ERR_INVALID_GRAPH;
  for (int64_t i = 0; i < E; ++i)
    if (adj[i] < 0 || adj[i] >= V) return CSAW_ERR_INVALID_GRAPH;
  csaw_host_graph* g = new csaw_host_graph();
  g->V = V;
  g->E = E;
  if (hipHostMalloc((void**)&g->off, sizeof(int64_t) * (V + 1), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&g->adj, sizeof(int32_t) * std::max<int64_t>(E, 1), hipHostMallocDefault) != hipSuccess) {
    (void)hipHostFree(g->off);
    delete g;
    return CSAW_ERR_OOM;
  }
  memcpy(g->off, off, sizeof(int64_t) * (V + 1));
  if (E > 0) memcpy(g->adj, adj, sizeof(int32_t) * E);
  for (int64_t v = 0; v < V; ++v) g->max_deg = std::max(g->max_deg, off[v + 1] - off[v]);
  *out = g;
  return CSAW_OK;
}

extern "C" int csaw_host_graph_info(const csaw_host_graph* g, int64_t* V, int64_t* E, int64_t* max_degree,
                                    const int64_t** offsets, const int32_t** indices) {
  if (!g) return CSAW_ERR_INVALID_ARG;
  if (V) *V = g->V;
  if (E) *E = g->E;
  if (max_degree) *max_degree = g->max_deg;
  if (offsets) *offsets = g->off;
  if (indices) *indices = g->adj;
  return CSAW_OK;
}

extern "C" int csaw_host_graph_destroy(csaw_host_graph* g) {
  if (!g) return CSAW_OK;
  (void)hipHostFree(g->off);
  (void)hipHostFree(g->adj);
  if (g->pdata) (void)hipHostFree(g->pdata);
  if (g->boff32) (void)hipHostFree(g->boff32);
  delete g;
  return CSAW_OK;
}

extern "C" int csaw_host_graph_pack(csaw_host_graph* g, int32_t threads, int64_t* packed_bytes) {
  if (!g || threads < 0) return CSAW_ERR_INVALID_ARG;
  if (!g->pdata) {
    const int64_t nblk = (g->E + kPackBlock - 1) / kPackBlock;
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(threads > 0 ? threads : 16, std::max<int64_t>(nblk, 1)));
    std::vector<uint64_t> boff(nblk + 1, 0);
    auto for_blocks = [&](auto fn) {
      std::vector<std::thread> ts;
      for (int t = 0; t < nt; ++t)
        ts.emplace_back([&, t]() {
          for (int64_t b = nblk * t / nt; b < nblk * (t + 1) / nt; ++b) fn(b);
        });
      for (auto& th : ts) th.join();
    };
    for_blocks([&](int64_t b) {
      const int64_t i0 = b * kPackBlock;
      boff[b + 1] = (uint64_t)pack_block(g->adj + i0, (int)std::min<int64_t>(kPackBlock, g->E - i0), nullptr);
    });
    for (int64_t b = 0; b < nblk; ++b) boff[b + 1] += boff[b];
    uint8_t* pdata = nullptr;
    uint32_t* b32 = nullptr;
    if (hipHostMalloc((void**)&pdata, std::max<uint64_t>(boff[nblk], 1), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b32, sizeof(uint32_t) * (nblk + 1), hipHostMallocDefault) != hipSuccess) {
      if (pdata) (void)hipHostFree(pdata);
      set_last_error("cannot pin memory for the packed adjacency");
      return CSAW_ERR_OOM;
    }
    for_blocks([&](int64_t b) {
      const int64_t i0 = b * kPackBlock;
      pack_block(g->adj + i0, (int)std::min<int64_t>(kPackBlock, g->E - i0), pdata + boff[b]);
      b32[b] = (uint32_t)boff[b];
    });
    b32[nblk] = (uint32_t)boff[nblk];
    g->pdata = pdata;
    g->boff32 = b32;
    g->boff = std::move(boff);
    g->nblk = nblk;
  }
  if (packed_bytes) *packed_bytes = (int64_t)g->boff[g->nblk] + 4 * (g->nblk + 1);
  return CSAW_OK;
}

extern "C" int csaw_stream_sample_opts(csaw_host_graph* hg, const csaw_request* req, int64_t budget_bytes,
                                       int32_t n_partitions, const csaw_stream_opts* opts, csaw_edges** out,
                                       csaw_stream_stats* stats) {
  if (!hg || !req || !out || n_partitions < 1 || n_partitions > 255 || budget_bytes <= 0) return CSAW_ERR_INVALID_ARG;
  *out = nullptr;
  csaw_stream_opts o;
  memset(&o, 0, sizeof(o));
  o.mode = CSAW_MODE_BATCHED;
  o.sched = CSAW_SCHED_HEURISTIC;
  if (opts) o = *opts;
  if (o.mode != CSAW_MODE_PLAIN && o.mode != CSAW_MODE_BATCHED) return CSAW_ERR_INVALID_ARG;
  if (o.slice_policy < CSAW_SLICE_AUTO || o.slice_policy > CSAW_SLICE_DIRECT) return CSAW_ERR_INVALID_ARG;
  if (o.sched < CSAW_SCHED_HEURISTIC || o.sched > CSAW_SCHED_RANDOM) return CSAW_ERR_INVALID_ARG;
  if (req->dedup < 0 || req->dedup > 2) return CSAW_ERR_INVALID_ARG;
  CSAW_TRY(validate_tuning(o.tuning));
  if (o.sched == CSAW_SCHED_FORCED && (o.variant < 0 || o.variant >= CSAW_NUM_VARIANTS)) return CSAW_ERR_INVALID_ARG;
  if (o.mode == CSAW_MODE_BATCHED && o.sched == CSAW_SCHED_MLP && !o.mlp) {
    set_last_error("CSAW_SCHED_MLP needs a trained csaw_mlp");
    return CSAW_ERR_INVALID_ARG;
  }
  const bool batched = o.mode == CSAW_MODE_BATCHED;
  const csaw_request& r = *req;
  if (r.samples < 0 || r.depth < 0 || r.depth > CSAW_MAX_HOPS || (r.samples > 0 && !r.seeds)) return CSAW_ERR_INVALID_ARG;
  if (r.bias == CSAW_BIAS_WEIGHT) {
    set_last_error("streaming supports uniform, degree and user biases");
    return CSAW_ERR_UNSUPPORTED;
  }
  for (int h = 0; h < r.depth; ++h)
    if (r.fanouts[h] < 0 || r.fanouts[h] > CSAW_MAX_FANOUT) return CSAW_ERR_INVALID_ARG;
  for (int32_t s = 0; s < r.samples; ++s)
    if (r.seeds[s] < 0 || r.seeds[s] >= hg->V) return CSAW_ERR_INVALID_ARG;
  int num_cu = 0;
  CSAW_TRY(device_init(&num_cu));
  const int64_t V = hg->V, E = hg->E;
  const int P = n_partitions;
  const int32_t S = r.samples;
  // equal-edge vertex ranges
  std::vector<int32_t> vb(P + 1);
  vb[0] = 0;
  vb[P] = (int32_t)V;
  for (int p = 1; p < P; ++p) {
    const int64_t target = E * p / P;
    vb[p] = (int32_t)(std::lower_bound(hg->off, hg->off + V + 1, target) - hg->off);
    vb[p] = std::max(vb[p], vb[p - 1]);
  }
  int64_t slice_max = 1, hubs_max = 1;
  for (int p = 0; p < P; ++p) {
    slice_max = std::max(slice_max, hg->off[vb[p + 1]] - hg->off[vb[p]]);
    int64_t nh = 0;
    for (int32_t v = vb[p]; v < vb[p + 1]; ++v) nh += (hg->off[v + 1] - hg->off[v]) > kHubSeg;
    hubs_max = std::max(hubs_max, nh);
  }
  // per-hop capacities
  std::vector<int64_t> cap(r.depth), nmax(r.depth);
  int64_t ub = S, mx = std::max<int32_t>(S, 1);
  for (int h = 0; h < r.depth; ++h) {
    nmax[h] = ub;
    ub = ub * r.fanouts[h];
    cap[h] = ub;
    mx = std::max(mx, nmax[h]);
    if (ub > ((int64_t)1 << 32)) return CSAW_ERR_CAPACITY;
  }
  const bool int_bias = r.bias == CSAW_BIAS_DEGREE || r.bias == CSAW_BIAS_USER;
  const int kind = bias_kind(r.bias);

  Budget B(budget_bytes);
  hipStream_t cs = nullptr, xs = nullptr;
  hipEvent_t ready[2] = {nullptr, nullptr}, freed[2] = {nullptr, nullptr}, t0 = nullptr, t1 = nullptr;
  std::vector<hipEvent_t> evs;  // per processed partition: bin begin/end, expand begin/end
  std::vector<hipEvent_t> cev;  // per slice copy: begin/end on the copy stream
  std::vector<hipEvent_t> hev;  // per hop: start on the compute stream
  uint32_t* h_qlen = nullptr;   // pinned: queue lengths per (hop, partition)
  std::vector<int32_t> h_src, h_dst, h_smp;
  std::vector<int64_t> h_ho;
  double h2d_bytes = 0, h2d_raw_bytes = 0, h2d_ms = 0, expand_ms = 0, copy_ms = 0, hop_ms[CSAW_MAX_HOPS] = {};
  const bool packed = hg->pdata != nullptr && !o.raw_slices;
  hipStream_t us = nullptr;  // unpack stream (packed slices)
  hipEvent_t staged[2] = {nullptr, nullptr}, consumed[2] = {nullptr, nullptr};
  int64_t piece_tick = 0;
  int64_t n_chunks_all = 0, var_chunks[CSAW_NUM_VARIANTS] = {};
  int32_t direct_parts = 0, reused_slices = 0;
  int64_t log_rows = 0;
  auto body = [&]() -> int {
    CSAW_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    CSAW_HIP(hipStreamCreateWithFlags(&xs, hipStreamNonBlocking));
    if (packed) {
      CSAW_HIP(hipStreamCreateWithFlags(&us, hipStreamNonBlocking));
      for (int i = 0; i < 2; ++i) {
        CSAW_HIP(hipEventCreateWithFlags(&staged[i], hipEventDisableTiming));
        CSAW_HIP(hipEventCreateWithFlags(&consumed[i], hipEventDisableTiming));
        CSAW_HIP(hipEventRecord(consumed[i], us));  // both staging buffers start free
      }
    }
    for (int i = 0; i < 2; ++i) {
      CSAW_HIP(hipEventCreateWithFlags(&ready[i], hipEventDisableTiming));
      CSAW_HIP(hipEventCreateWithFlags(&freed[i], hipEventDisableTiming));
    }
    CSAW_HIP(hipEventCreate(&t0));
    CSAW_HIP(hipEventCreate(&t1));
    // resident: offsets, degrees
    int64_t* d_off;
    int32_t* d_deg;
    CSAW_TRY(B.alloc(&d_off, V + 1));
    CSAW_TRY(B.alloc(&d_deg, V));
    CSAW_HIP(hipMemcpyAsync(d_off, hg->off, sizeof(int64_t) * (V + 1), hipMemcpyHostToDevice, cs));
    k_stream_degree<<<grid_of(V, num_cu), 256, 0, cs>>>(d_off, V, d_deg);
    CSAW_LAUNCH_CHECK();
    int32_t* slice[2];
    CSAW_TRY(B.alloc(&slice[0], slice_max));
    CSAW_TRY(B.alloc(&slice[1], slice_max));
    uint8_t* sdata[2] = {nullptr, nullptr};
    uint32_t* soff[2] = {nullptr, nullptr};
    if (packed) {
      for (int i = 0; i < 2; ++i) {
        CSAW_TRY(B.alloc(&sdata[i], kPieceBlocks * kPackMaxBlockBytes));
        CSAW_TRY(B.alloc(&soff[i], kPieceBlocks + 1));
      }
    }
    int held[2] = {-1, -1};  // partition whose slice each buffer holds
    int64_t last_use[2] = {0, 0}, use_tick = 0;
    // the pinned host adjacency as the device addresses it, for partitions
    // read in place
    int32_t* adj_bus = nullptr;
    if (o.slice_policy != CSAW_SLICE_COPY && E > 0 &&
        hipHostGetDevicePointer((void**)&adj_bus, hg->adj, 0) != hipSuccess) {
      adj_bus = nullptr;
      (void)hipGetLastError();
    }
    if (o.slice_policy == CSAW_SLICE_DIRECT && E > 0 && !adj_bus) {
      set_last_error("the host adjacency cannot be mapped for in-place reads");
      return CSAW_ERR_UNSUPPORTED;
    }
    // integer biases: the partition's frontier CDFs at slice positions, the
    // marks, and the hub segment lists
    uint64_t* pool = nullptr;
    int32_t *d_hub_vid = nullptr, *d_seg_hub = nullptr, *d_seg_idx = nullptr;
    int64_t *d_hub_pos = nullptr, *d_seg_first = nullptr;
    uint64_t* d_seg_tot = nullptr;
    uint32_t *d_hubbits = nullptr, *d_fbits = nullptr, *d_fcount = nullptr;
    int32_t* d_fvlist = nullptr;
    const int64_t seg_max = slice_max / kHubSeg + hubs_max + 1;
    if (int_bias) {
      CSAW_TRY(B.alloc(&pool, slice_max));
      CSAW_TRY(B.alloc(&d_hubbits, (V + 31) / 32));
      CSAW_TRY(B.alloc(&d_fbits, (V + 31) / 32));
      CSAW_TRY(B.alloc(&d_fcount, 1));
      CSAW_TRY(B.alloc(&d_fvlist, mx));  // distinct vertices of one partition's frontier
      CSAW_TRY(B.alloc(&d_hub_vid, hubs_max));
      CSAW_TRY(B.alloc(&d_hub_pos, hubs_max));
      CSAW_TRY(B.alloc(&d_seg_first, hubs_max + 1));
      CSAW_TRY(B.alloc(&d_seg_hub, seg_max));
      CSAW_TRY(B.alloc(&d_seg_idx, seg_max));
      CSAW_TRY(B.alloc(&d_seg_tot, seg_max));
    }
    // request / frontier / outputs
    ReqDev rd;
    memset(&rd, 0, sizeof(rd));
    for (int h = 0; h < r.depth; ++h) rd.fanout[h] = r.fanouts[h];
    rd.depth = r.depth;
    rd.bias = r.bias;
    rd.dedup = r.dedup == 2 ? 2 : (r.dedup ? 1 : 0);
    rd.samples = S;
    rd.key_lo = (uint32_t)r.seed;
    rd.key_hi = (uint32_t)(r.seed >> 32);
    rd.request_id = r.request_id;
    ReqDev* d_req;
    int32_t *d_gs2req, *d_fv0, *d_fs0;
    int64_t* d_fo0;
    CSAW_TRY(B.alloc(&d_req, 1));
    CSAW_TRY(B.alloc(&d_gs2req, S));
    CSAW_TRY(B.alloc(&d_fv0, S));
    CSAW_TRY(B.alloc(&d_fs0, S));
    CSAW_TRY(B.alloc(&d_fo0, S + 1));
    std::vector<int32_t> zeros(S, 0), iota_s(S);
    std::vector<int64_t> fo0(S + 1);
    std::iota(iota_s.begin(), iota_s.end(), 0);
    std::iota(fo0.begin(), fo0.end(), (int64_t)0);
    CSAW_HIP(hipMemcpyAsync(d_req, &rd, sizeof(rd), hipMemcpyHostToDevice, cs));
    if (S > 0) {
      CSAW_HIP(hipMemcpyAsync(d_gs2req, zeros.data(), sizeof(int32_t) * S, hipMemcpyHostToDevice, cs));
      CSAW_HIP(hipMemcpyAsync(d_fv0, r.seeds, sizeof(int32_t) * S, hipMemcpyHostToDevice, cs));
      CSAW_HIP(hipMemcpyAsync(d_fs0, iota_s.data(), sizeof(int32_t) * S, hipMemcpyHostToDevice, cs));
    }
    CSAW_HIP(hipMemcpyAsync(d_fo0, fo0.data(), sizeof(int64_t) * (S + 1), hipMemcpyHostToDevice, cs));
    std::vector<int64_t*> d_eo(r.depth);
    std::vector<int32_t*> d_src(r.depth), d_dst(r.depth), d_smp(r.depth);
    for (int h = 0; h < r.depth; ++h) {
      CSAW_TRY(B.alloc(&d_eo[h], S + 1));
      CSAW_TRY(B.alloc(&d_src[h], cap[h]));
      CSAW_TRY(B.alloc(&d_dst[h], cap[h]));
      CSAW_TRY(B.alloc(&d_smp[h], cap[h]));
    }
    const int64_t nblk = (mx + kBlockEntries - 1) / kBlockEntries;
    uint32_t *d_local, *d_blk, *d_cnt, *d_poff, *d_cursor;
    uint64_t* d_grp;
    int64_t* d_tot;
    uint8_t* d_pid;
    int32_t *d_list, *d_vb;
    CSAW_TRY(B.alloc(&d_local, mx));
    CSAW_TRY(B.alloc(&d_blk, nblk));
    CSAW_TRY(B.alloc(&d_grp, (nblk + 1023) / 1024));
    CSAW_TRY(B.alloc(&d_tot, std::max(r.depth, 1)));
    CSAW_TRY(B.alloc(&d_pid, mx));
    CSAW_TRY(B.alloc(&d_list, mx));
    CSAW_TRY(B.alloc(&d_cnt, 3 * P));
    unsigned long long* d_pbytes;
    CSAW_TRY(B.alloc(&d_pbytes, P));
    d_poff = d_cnt + P;
    d_cursor = d_cnt + 2 * P;
    CSAW_TRY(B.alloc(&d_vb, P + 1));
    CSAW_HIP(hipMemcpyAsync(d_vb, vb.data(), sizeof(int32_t) * (P + 1), hipMemcpyHostToDevice, cs));
    LaunchCfg L;
    L.stream = cs;
    L.num_cu = num_cu;
    if (o.tuning) apply_tuning(L, o.tuning);
    // batched mode: chunk descriptors of one partition, variant queues, counters
    int32_t* d_order = nullptr;
    Chunk* d_chunks = nullptr;
    uint32_t *d_queues = nullptr, *d_ctr = nullptr;
    uint16_t* d_mlp = nullptr;
    ChunkLog* d_log = nullptr;
    uint32_t chunk_cap = 0;
    const int32_t expand_waves = L.num_cu * L.expand_per_cu * 4;
    if (batched) {
      chunk_cap = (uint32_t)std::min<int64_t>(nblk * kChunksPerBlock + expand_waves, 0xFFFFFFF0LL);
      CSAW_TRY(B.alloc(&d_order, mx));
      CSAW_TRY(B.alloc(&d_chunks, chunk_cap));
      CSAW_TRY(B.alloc(&d_queues, (int64_t)kNumQueues * chunk_cap));
      CSAW_TRY(B.alloc(&d_ctr, kCtrCount));
      CSAW_TRY(launch_fill(L, d_ctr, 0u, kCtrCount));
      if (o.log_chunks) CSAW_TRY(B.alloc(&d_log, chunk_cap));
      if (o.sched == CSAW_SCHED_MLP) {
        CSAW_TRY(B.alloc(&d_mlp, kMlpParams));
        CSAW_TRY(mlp_upload_bf16(o.mlp, d_mlp));
      }
      CSAW_HIP(hipHostMalloc((void**)&h_qlen, sizeof(uint32_t) * kNumQueues * std::max(r.depth, 1) * P,
                             hipHostMallocDefault));
      memset(h_qlen, 0, sizeof(uint32_t) * kNumQueues * std::max(r.depth, 1) * P);
    }

    GraphDev G;
    memset(&G, 0, sizeof(G));
    G.off = d_off;
    G.deg = d_deg;
    G.V = V;
    G.E = E;
    csaw_graph_view hv;
    hv.beg_pos = d_off;
    hv.adj_list = nullptr;
    hv.degree_list = d_deg;
    hv.weight_list = nullptr;
    hv.vert_count = V;
    hv.edge_count = E;
    std::vector<uint32_t> hubbits_h;
    bool kinds_present[3] = {false, false, false};
    kinds_present[kind] = true;
    auto event = [&](hipEvent_t* e) -> int {
      CSAW_HIP(hipEventCreate(e));
      evs.push_back(*e);
      return CSAW_OK;
    };
    CSAW_HIP(hipEventRecord(t0, cs));
    for (int h = 0; h < r.depth; ++h) {
      hev.push_back(nullptr);
      CSAW_HIP(hipEventCreate(&hev.back()));
      CSAW_HIP(hipEventRecord(hev.back(), cs));
      HopDev H;
      memset(&H, 0, sizeof(H));
      H.fv = h == 0 ? d_fv0 : d_dst[h - 1];
      H.fs = h == 0 ? d_fs0 : d_smp[h - 1];
      H.fo = h == 0 ? d_fo0 : d_eo[h - 1];
      H.n_ptr = H.fo + S;
      H.hop = h;
      H.gb = S;
      H.local_excl = d_local;
      H.blk_excl = d_blk;
      H.grp_excl = d_grp;
      H.src = d_src[h];
      H.dst = d_dst[h];
      H.smp = d_smp[h];
      H.eo = d_eo[h];
      if (batched) {
        H.order = d_order;
        H.chunks = d_chunks;
        H.n_chunks = d_ctr + kCtrChunks;
        H.chunk_cap = chunk_cap;
        H.flags = d_ctr + kCtrFlags;
        H.expand_waves = expand_waves;
        H.chunk_min = L.chunk_min;
        H.group_min = L.group_min;
      }
      CSAW_TRY(launch_count(L, G, d_req, d_gs2req, H, false, nmax[h]));
      CSAW_TRY(launch_scans(L, H, d_tot + h, nmax[h]));
      // bin the frontier by partition
      CSAW_TRY(launch_fill(L, d_cnt, 0u, 3 * P));
      CSAW_TRY(launch_fill(L, reinterpret_cast<uint32_t*>(d_pbytes), 0u, 2 * P));
      if (int_bias) {
        CSAW_TRY(launch_fill(L, d_hubbits, 0u, (V + 31) / 32));
        CSAW_TRY(launch_fill(L, d_fbits, 0u, (V + 31) / 32));
      }
      k_part_bin<<<grid_of(nmax[h], num_cu), 256, 0, cs>>>(G, d_req, H, d_vb, P, d_pid, d_cnt, d_pbytes, d_fbits,
                                                           d_hubbits);
      CSAW_LAUNCH_CHECK();
      std::vector<uint32_t> cnt(P), poff(P + 1, 0);
      std::vector<unsigned long long> pbytes(P);
      CSAW_HIP(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(uint32_t) * P, hipMemcpyDeviceToHost, cs));
      CSAW_HIP(hipMemcpyAsync(pbytes.data(), d_pbytes, sizeof(unsigned long long) * P, hipMemcpyDeviceToHost, cs));
      CSAW_HIP(hipStreamSynchronize(cs));
      for (int p = 0; p < P; ++p) poff[p + 1] = poff[p] + cnt[p];
      CSAW_HIP(hipMemcpyAsync(d_poff, poff.data(), sizeof(uint32_t) * P, hipMemcpyHostToDevice, cs));
      k_part_scatter<<<grid_of((nmax[h] + 3) / 4, num_cu), 256, 0, cs>>>(H, d_pid, P, d_poff, d_cursor, d_list);
      CSAW_LAUNCH_CHECK();
      // unique frontier hubs per partition (sorted by id)
      std::vector<std::vector<int32_t>> hubs(P);
      if (int_bias) {
        hubbits_h.resize((V + 31) / 32);
        CSAW_HIP(hipMemcpyAsync(hubbits_h.data(), d_hubbits, sizeof(uint32_t) * hubbits_h.size(),
                                hipMemcpyDeviceToHost, cs));
        CSAW_HIP(hipStreamSynchronize(cs));
        int p = 0;
        for (size_t wi = 0; wi < hubbits_h.size(); ++wi) {
          uint32_t m = hubbits_h[wi];
          while (m) {
            const int32_t v = (int32_t)(wi * 32 + __builtin_ctz(m));
            m &= m - 1;
            while (v >= vb[p + 1]) ++p;
            hubs[p].push_back(v);
          }
        }
      }
      // Partitions with frontier entries, ascending on even hops and
      // descending on odd ones, so the two slices left on the device by one
      // hop are the first two the next hop needs.  Each is either read in
      // place over the bus (few bytes of it needed) or copied on the copy
      // stream into one of two buffers, one partition ahead (double buffering).
      std::vector<int> todo;
      for (int q = 0; q < P; ++q) {
        const int p = (h & 1) ? P - 1 - q : q;
        if (cnt[p] > 0) todo.push_back(p);
      }
      std::vector<char> direct(P, 0);
      for (int p = 0; p < P; ++p) {
        const double slice_b = (double)(hg->off[vb[p + 1]] - hg->off[vb[p]]) * sizeof(int32_t);
        const bool held_p = held[0] == p || held[1] == p;
        if (o.slice_policy == CSAW_SLICE_DIRECT) direct[p] = 1;
        else if (o.slice_policy == CSAW_SLICE_AUTO && adj_bus && !held_p)
          direct[p] = (double)pbytes[p] * (o.direct_factor > 0.0f ? o.direct_factor : kDirectFactor) < slice_b;
      }
      // *out = a buffer, other than avoid, holding partition p (the copy is
      // issued when neither buffer holds it)
      // raw: copy unpacked even when the graph is packed (unpacking needs
      // CUs, which a hop read in place is using)
      auto stage = [&](int p, int avoid, int* out, bool raw = false) -> int {
        if ((held[0] == p && avoid != 0) || (held[1] == p && avoid != 1)) {
          *out = held[0] == p && avoid != 0 ? 0 : 1;
          reused_slices += 1;
          return CSAW_OK;
        }
        const int b = avoid >= 0 ? 1 - avoid
                                 : (held[0] < 0 ? 0 : (held[1] < 0 ? 1 : (last_use[0] <= last_use[1] ? 0 : 1)));
        const int64_t a = hg->off[vb[p]], e = hg->off[vb[p + 1]];
        auto mark = [&]() -> int {  // copy-stream timing
          cev.push_back(nullptr);
          CSAW_HIP(hipEventCreate(&cev.back()));
          CSAW_HIP(hipEventRecord(cev.back(), xs));
          return CSAW_OK;
        };
        if (!packed || raw) {
          CSAW_HIP(hipStreamWaitEvent(xs, freed[b], 0));
          CSAW_TRY(mark());
          if (e > a)
            CSAW_HIP(hipMemcpyAsync(slice[b], hg->adj + a, sizeof(int32_t) * (e - a), hipMemcpyHostToDevice, xs));
          CSAW_TRY(mark());
          CSAW_HIP(hipEventRecord(ready[b], xs));
          h2d_bytes += (double)(e - a) * sizeof(int32_t);
        } else {
          // packed blocks covering [a, e), in pieces through two staging
          // buffers: piece j+1 crosses the bus while piece j is unpacked
          // into the slice buffer (on the unpack stream, after the buffer's
          // previous partition has expanded)
          CSAW_HIP(hipStreamWaitEvent(us, freed[b], 0));
          const int64_t b0 = a / kPackBlock, b1 = (e + kPackBlock - 1) / kPackBlock;
          for (int64_t pb = b0; pb < b1; pb += kPieceBlocks) {
            const int64_t nb = std::min<int64_t>(kPieceBlocks, b1 - pb);
            const int sb = (int)(piece_tick++ & 1);
            const uint64_t d0 = hg->boff[pb], d1 = hg->boff[pb + nb];
            CSAW_HIP(hipStreamWaitEvent(xs, consumed[sb], 0));
            CSAW_TRY(mark());
            CSAW_HIP(hipMemcpyAsync(soff[sb], hg->boff32 + pb, sizeof(uint32_t) * (nb + 1), hipMemcpyHostToDevice, xs));
            CSAW_HIP(hipMemcpyAsync(sdata[sb], hg->pdata + d0, d1 - d0, hipMemcpyHostToDevice, xs));
            CSAW_TRY(mark());
            CSAW_HIP(hipEventRecord(staged[sb], xs));
            CSAW_HIP(hipStreamWaitEvent(us, staged[sb], 0));
            const int g = (int)std::max<int64_t>(1, std::min<int64_t>((nb + 3) / 4, (int64_t)num_cu * 8));
            k_unpack<<<g, 256, 0, us>>>(sdata[sb], soff[sb], pb, nb, E, a, e, slice[b]);
            CSAW_LAUNCH_CHECK();
            CSAW_HIP(hipEventRecord(consumed[sb], us));
            h2d_bytes += (double)(d1 - d0) + 4.0 * (double)(nb + 1);
          }
          CSAW_HIP(hipEventRecord(ready[b], us));
        }
        h2d_raw_bytes += (double)(e - a) * sizeof(int32_t);
        held[b] = p;
        *out = b;
        return CSAW_OK;
      };
      // the freed events start signalled
      for (int i = 0; i < 2; ++i) CSAW_HIP(hipEventRecord(freed[i], cs));
      std::vector<int> bufof(todo.size(), -1);
      auto next_copy = [&](size_t from) -> size_t {
        while (from < todo.size() && direct[todo[from]]) ++from;
        return from;
      };
      size_t staged = next_copy(0);
      if (staged < todo.size()) {
        CSAW_TRY(stage(todo[staged], -1, &bufof[staged]));
      } else if (h + 1 < r.depth && P > 1 && o.slice_policy != CSAW_SLICE_DIRECT) {
        // a hop read entirely in place leaves both buffers idle: fill them
        // with the next hop's first two partitions meanwhile (if that hop
        // reads them in place instead, the copies were only wasted bus time)
        const int q0 = ((h + 1) & 1) ? P - 1 : 0, q1 = ((h + 1) & 1) ? P - 2 : 1;
        int b0 = -1, b1 = -1;
        const int32_t keep = reused_slices;  // counted when the next hop finds them
        CSAW_TRY(stage(q0, -1, &b0, true));
        CSAW_TRY(stage(q1, b0, &b1, true));
        reused_slices = keep;
      }
      for (size_t t = 0; t < todo.size(); ++t) {
        const int p = todo[t];
        const int buf = direct[p] ? -1 : bufof[t];
        if (direct[p]) direct_parts += 1;
        if (buf >= 0) {
          // stage the next copied partition into the other buffer
          const size_t t2 = next_copy(t + 1);
          if (t2 < todo.size() && t2 > staged) {
            staged = t2;
            CSAW_TRY(stage(todo[t2], buf, &bufof[t2]));
          }
          last_use[buf] = ++use_tick;
        }
        // small partitions are expanded one wave per entry whatever the
        // schedule: binning and choosing cost more than they save there
        const bool chunked = batched && (int64_t)cnt[p] >= (o.bin_min_entries > 0 ? o.bin_min_entries : kChunkedMin);
        hipEvent_t eb0, eb1, ex0, ex1;
        CSAW_TRY(event(&eb0));
        CSAW_TRY(event(&eb1));
        CSAW_TRY(event(&ex0));
        CSAW_TRY(event(&ex1));
        CSAW_HIP(hipEventRecord(eb0, cs));
        if (chunked) {
          // chunks of this partition's entries and their variants (needs only
          // degrees, so it runs while the slice is still in flight)
          CSAW_TRY(launch_fill(L, d_ctr, 0u, kCtrCount - 1));  // keeps the flags word
          CSAW_TRY(launch_bin_list(L, G, d_req, d_gs2req, H, d_list + poff[p], (int64_t)cnt[p]));
          ChooseArgs A;
          memset(&A, 0, sizeof(A));
          A.sched = o.sched;
          A.variant = o.variant;
          A.explore_key = o.explore_key;
          A.qlen = d_ctr + kCtrQlen;
          A.queues = d_queues;
          A.queue_cap = chunk_cap;
          A.chosen = nullptr;
          A.int_cache = int_bias ? 1 : 0;  // every drawing vertex has its CDF
          A.in_place = direct[p] ? 1 : 0;
          A.mlp.w = d_mlp;
          A.mlp_w = d_mlp;
          const int64_t nblk_p = ((int64_t)cnt[p] + kBlockEntries - 1) / kBlockEntries;
          CSAW_TRY(launch_choose(L, A, H, std::min<int64_t>(nblk_p * kChunksPerBlock + expand_waves, chunk_cap)));
          CSAW_HIP(hipMemcpyAsync(h_qlen + ((size_t)h * P + p) * kNumQueues, d_ctr + kCtrQlen,
                                  sizeof(uint32_t) * kNumQueues, hipMemcpyDeviceToHost, cs));
        }
        CSAW_HIP(hipEventRecord(eb1, cs));
        if (buf >= 0) CSAW_HIP(hipStreamWaitEvent(cs, ready[buf], 0));
        CSAW_HIP(hipEventRecord(ex0, cs));
        GraphDev Gp = G;
        // global offsets index the slice (or the host adjacency, read in place)
        Gp.adj = buf >= 0 ? slice[buf] - hg->off[vb[p]] : adj_bus;
        csaw_graph_view hvp = hv;
        hvp.adj_list = Gp.adj;
        Gp.n_hubs = 0;
        std::vector<int32_t> seg_hub, seg_idx;
        std::vector<int64_t> seg_first, hub_pos;
        if (int_bias) {
          // CDFs of the partition's distinct frontier vertices at their slice
          // positions; every draw of this partition then searches them
          const int ck = r.bias == CSAW_BIAS_USER ? 1 : 0;
          const int64_t a = hg->off[vb[p]];
          const int64_t words = ((int64_t)vb[p + 1] + 31) / 32 - vb[p] / 32;
          CSAW_TRY(launch_fill(L, d_fcount, 0u, 1));
          k_mark_list<<<grid_of(words, num_cu), 256, 0, cs>>>(d_fbits, vb[p], vb[p + 1], d_fvlist, d_fcount);
          CSAW_LAUNCH_CHECK();
          const int gcdf = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)cnt[p] + 3) / 4, (int64_t)num_cu * 8));
          if (ck == 0) k_frontier_cdf<0><<<gcdf, 256, 0, cs>>>(Gp, hvp, d_fvlist, d_fcount, a, pool);
          else k_frontier_cdf<1><<<gcdf, 256, 0, cs>>>(Gp, hvp, d_fvlist, d_fcount, a, pool);
          CSAW_LAUNCH_CHECK();
          if (!hubs[p].empty()) {
            seg_first.push_back(0);
            for (size_t i = 0; i < hubs[p].size(); ++i) {
              const int32_t v = hubs[p][i];
              const int64_t d = hg->off[v + 1] - hg->off[v];
              hub_pos.push_back(hg->off[v] - a);
              for (int64_t sg = 0; sg * kHubSeg < d; ++sg) {
                seg_hub.push_back((int32_t)i);
                seg_idx.push_back((int32_t)sg);
              }
              seg_first.push_back((int64_t)seg_hub.size());
            }
            const int32_t nh = (int32_t)hubs[p].size();
            const int64_t ns = (int64_t)seg_hub.size();
            CSAW_HIP(hipMemcpyAsync(d_hub_vid, hubs[p].data(), sizeof(int32_t) * nh, hipMemcpyHostToDevice, cs));
            CSAW_HIP(hipMemcpyAsync(d_hub_pos, hub_pos.data(), sizeof(int64_t) * nh, hipMemcpyHostToDevice, cs));
            CSAW_HIP(hipMemcpyAsync(d_seg_first, seg_first.data(), sizeof(int64_t) * (nh + 1), hipMemcpyHostToDevice, cs));
            CSAW_HIP(hipMemcpyAsync(d_seg_hub, seg_hub.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, cs));
            CSAW_HIP(hipMemcpyAsync(d_seg_idx, seg_idx.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, cs));
            GraphDev Gh = Gp;
            Gh.hub_vid = d_hub_vid;
            Gh.hub_pool = d_hub_pos;
            Gh.n_hubs = nh;
            CSAW_TRY(hub_cdf_segments(Gh, hvp, ck, nh, d_seg_hub, d_seg_idx, ns, d_seg_first, d_seg_tot, pool, cs,
                                      num_cu));
          }
          Gp.int_cdf[ck] = pool - a;  // global offsets index the slice-shaped array
        }
        if (chunked) {
          ExpandArgs X;
          X.queues = d_queues;
          X.queue_cap = chunk_cap;
          X.qlen = d_ctr + kCtrQlen;
          X.qhead = d_ctr + kCtrQhead;
          X.log = d_log;
          X.int_cache = int_bias ? 1 : 0;
          X.remap = r.dedup == 2 ? 1 : 0;
          CSAW_TRY(launch_expand_chunks(L, Gp, hvp, d_req, d_gs2req, H, X, kinds_present, nullptr));
        } else {
          CSAW_TRY(launch_expand_list(L, kind, Gp, hvp, d_req, d_gs2req, H, d_list + poff[p], (int64_t)cnt[p]));
        }
        CSAW_HIP(hipEventRecord(ex1, cs));
        if (buf >= 0) CSAW_HIP(hipEventRecord(freed[buf], cs));
        if (!seg_hub.empty()) CSAW_HIP(hipStreamSynchronize(cs));  // hub list buffers are reused
        if (d_log && chunked) {
          // this partition's chunks and their timings, before the next reuses them
          CSAW_HIP(hipStreamSynchronize(cs));
          uint32_t nch = 0;
          CSAW_HIP(hipMemcpy(&nch, d_ctr + kCtrChunks, sizeof(uint32_t), hipMemcpyDeviceToHost));
          nch = std::min(nch, chunk_cap);
          std::vector<Chunk> ch(nch);
          std::vector<ChunkLog> lg(nch);
          if (nch) {
            CSAW_HIP(hipMemcpy(ch.data(), d_chunks, sizeof(Chunk) * nch, hipMemcpyDeviceToHost));
            CSAW_HIP(hipMemcpy(lg.data(), d_log, sizeof(ChunkLog) * nch, hipMemcpyDeviceToHost));
          }
          for (uint32_t i = 0; i < nch; ++i) {
            const Chunk& c = ch[i];
            if (c.bucket == kBucketCopy) continue;
            if (log_rows < o.log_cap) {
              if (o.log_features)
                chunk_features(c.len, c.sum_d, c.max_d, c.sum_f, c.max_f, bucket_kind(c.bucket), c.hop, c.bucket, nch,
                               o.log_features + log_rows * kMlpIn, direct[p]);
              if (o.log_variant) o.log_variant[log_rows] = lg[i].variant;
              if (o.log_log2_ticks) o.log_log2_ticks[log_rows] = std::log2((float)std::max<uint32_t>(lg[i].ticks, 1u));
              if (o.log_keys) o.log_keys[log_rows] = ((uint64_t)h << 40) | ((uint64_t)p << 32) | c.key;
            }
            ++log_rows;
          }
        }
      }
    }
    CSAW_HIP(hipEventRecord(t1, cs));
    CSAW_HIP(hipEventSynchronize(t1));
    float ms = 0.f;
    CSAW_HIP(hipEventElapsedTime(&ms, t0, t1));
    h2d_ms = ms;
    for (size_t i = 0; i + 1 < evs.size(); i += 2) {
      float m = 0.f;
      CSAW_HIP(hipEventElapsedTime(&m, evs[i], evs[i + 1]));
      expand_ms += m;
    }
    CSAW_HIP(hipStreamSynchronize(xs));
    if (us) CSAW_HIP(hipStreamSynchronize(us));
    for (size_t i = 0; i + 1 < cev.size(); i += 2) {
      float m = 0.f;
      CSAW_HIP(hipEventElapsedTime(&m, cev[i], cev[i + 1]));
      copy_ms += m;
    }
    for (size_t i = 0; i < hev.size(); ++i) {
      float m = 0.f;
      CSAW_HIP(hipEventElapsedTime(&m, hev[i], i + 1 < hev.size() ? hev[i + 1] : t1));
      hop_ms[i] = m;
    }
    if (batched) {
      uint32_t flags = 0;
      CSAW_HIP(hipMemcpy(&flags, d_ctr + kCtrFlags, sizeof(uint32_t), hipMemcpyDeviceToHost));
      if (flags & 1u) {
        set_last_error("chunk descriptor capacity exceeded");
        return CSAW_ERR_CAPACITY;
      }
      for (int64_t i = 0; i < (int64_t)std::max(r.depth, 1) * P; ++i)
        for (int q = 0; q < kNumQueues; ++q) {
          const uint32_t c = h_qlen[i * kNumQueues + q];
          n_chunks_all += c;
          if (q != kQueueCopy) var_chunks[q / 3] += c;
        }
    }
    // copy out
    std::vector<int64_t> lo(r.depth), hi(r.depth);
    int64_t n_all = 0;
    for (int h = 0; h < r.depth; ++h) {
      CSAW_HIP(hipMemcpy(&lo[h], d_eo[h], sizeof(int64_t), hipMemcpyDeviceToHost));
      CSAW_HIP(hipMemcpy(&hi[h], d_eo[h] + S, sizeof(int64_t), hipMemcpyDeviceToHost));
      n_all += hi[h] - lo[h];
    }
    h_src.resize(n_all);
    h_dst.resize(n_all);
    h_smp.resize(n_all);
    h_ho.resize((size_t)r.depth * (S + 1));
    int64_t at = 0;
    for (int h = 0; h < r.depth; ++h) {
      const int64_t n = hi[h] - lo[h];
      if (n > 0) {
        CSAW_HIP(hipMemcpy(h_src.data() + at, d_src[h] + lo[h], sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        CSAW_HIP(hipMemcpy(h_dst.data() + at, d_dst[h] + lo[h], sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        CSAW_HIP(hipMemcpy(h_smp.data() + at, d_smp[h] + lo[h], sizeof(int32_t) * n, hipMemcpyDeviceToHost));
      }
      std::vector<int64_t> eo(S + 1);
      CSAW_HIP(hipMemcpy(eo.data(), d_eo[h], sizeof(int64_t) * (S + 1), hipMemcpyDeviceToHost));
      for (int32_t s = 0; s <= S; ++s) h_ho[(size_t)h * (S + 1) + s] = at + eo[s] - lo[h];
      at += n;
    }
    return CSAW_OK;
  };
  int st = body();
  for (int i = 0; i < 2; ++i) {
    if (ready[i]) (void)hipEventDestroy(ready[i]);
    if (freed[i]) (void)hipEventDestroy(freed[i]);
  }
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  for (hipEvent_t e : evs) (void)hipEventDestroy(e);
  for (hipEvent_t e : cev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : hev)
    if (e) (void)hipEventDestroy(e);
  if (cs) (void)hipStreamSynchronize(cs);
  if (xs) (void)hipStreamSynchronize(xs);
  if (us) (void)hipStreamSynchronize(us);
  for (int i = 0; i < 2; ++i) {
    if (staged[i]) (void)hipEventDestroy(staged[i]);
    if (consumed[i]) (void)hipEventDestroy(consumed[i]);
  }
  if (cs) (void)hipStreamDestroy(cs);
  if (xs) (void)hipStreamDestroy(xs);
  if (us) (void)hipStreamDestroy(us);
  if (h_qlen) (void)hipHostFree(h_qlen);
  if (st != CSAW_OK) return st;
  if (o.log_n) *o.log_n = log_rows;
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->edges = (int64_t)h_src.size();
    stats->device_ms = h2d_ms;
    stats->h2d_bytes = h2d_bytes;
    stats->partitions = P;
    stats->peak_device_bytes = B.peak;
    stats->slice_bytes = slice_max * (int64_t)sizeof(int32_t);
    stats->expand_ms = expand_ms;
    stats->copy_ms = copy_ms;
    stats->h2d_raw_bytes = h2d_raw_bytes;
    for (int h = 0; h < CSAW_MAX_HOPS; ++h) stats->hop_ms[h] = hop_ms[h];
    stats->direct_partitions = direct_parts;
    stats->reused_slices = reused_slices;
    stats->chunks = n_chunks_all;
    for (int v = 0; v < CSAW_NUM_VARIANTS; ++v) stats->variant_chunks[v] = var_chunks[v];
  }
  return csaw_edges_from_host(r.depth, S, (int64_t)h_src.size(), h_src.data(), h_dst.data(), h_smp.data(),
                              h_ho.data(), out);
}

extern "C" int csaw_stream_sample(csaw_host_graph* hg, const csaw_request* req, int64_t budget_bytes,
                                  int32_t n_partitions, csaw_edges** out, csaw_stream_stats* stats) {
  return csaw_stream_sample_opts(hg, req, budget_bytes, n_partitions, nullptr, out, stats);
}
