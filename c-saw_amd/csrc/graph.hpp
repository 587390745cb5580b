// Device-resident CSR graph (replaces gpu_graph, non-stream/gpu_graph.cuh:10-65,
// and the host loader graph<>, non-stream/graph.hpp:4-111).
#pragma once
#include <vector>

#include "common.hpp"

struct csaw_graph {
  int64_t V = 0, E = 0, max_deg = 0;
  int64_t* off = nullptr;  // V+1
  int32_t* adj = nullptr;  // E
  int32_t* deg = nullptr;  // V
  float* w = nullptr;      // E (optional)
  float* cdf = nullptr;    // E (optional, built on demand)
  // hub cache: vertices with degree > kHubMin, sorted by id
  int32_t n_hubs = -1;     // -1: not computed
  int32_t* hub_vid = nullptr;
  int64_t* hub_pool = nullptr;  // n_hubs+1 pool offsets
  uint64_t* hub_cdf[2] = {nullptr, nullptr};  // [DEGREE, USER]
  uint64_t* int_cdf[2] = {nullptr, nullptr};  // full per-entry CDF cache [DEGREE, USER]
  // 16-ary search indexes over cdf and int_cdf (CDF arrays padded to 16)
  float* cdf_idx = nullptr;
  uint64_t* int_idx[2] = {nullptr, nullptr};
  // the integer cache as uint32 when every total fits (then int_cdf is null)
  uint32_t* int_cdf32[2] = {nullptr, nullptr};
  uint32_t* int_idx32[2] = {nullptr, nullptr};
  int64_t hub_entries = 0;
  // vertex records (offset, CDF total, degree) over the cached CDFs:
  // [DEGREE, USER] from the uint32 integer caches, [WEIGHT] from the FP32 CDF
  csaw::VertexRecord* vrec[3] = {nullptr, nullptr, nullptr};
  hipStream_t stream = nullptr;
  int device = 0;
  int num_cu = 0;
  // keep integer CDF caches as uint64 even when every total fits 32 bits
  // (environment CSAW_INT_CDF_KEYS=64 at graph creation; covers the 64-bit
  // search paths on graphs whose totals are small)
  bool wide_int_cdf = false;

  csaw::GraphDev view(bool use_cache = false) const;
  csaw_graph_view hook_view() const;
};

namespace csaw {
// Builds the vertex records of a bias's cached CDF (after graph_prepare_index;
// nothing for uniform, or for an integer cache kept as uint64).
int graph_prepare_vertex_records(csaw_graph* g, int32_t bias);
// Builds lazily what a bias kind needs: FP32 CDF (CSAW_BIAS_WEIGHT) or the
// integer hub cache (CSAW_BIAS_DEGREE / CSAW_BIAS_USER).
int graph_prepare_bias(csaw_graph* g, int32_t bias);
int graph_prepare_int_cache(csaw_graph* g, int32_t bias);
// Builds the search index of the CDF a bias uses from the cache (FP32 CDF for
// CSAW_BIAS_WEIGHT, the integer CDF cache otherwise), building that CDF first.
int graph_prepare_index(csaw_graph* g, int32_t bias);
int hub_cdf_build(const GraphDev& G, const csaw_graph_view& hv, int kind, const std::vector<int64_t>& pool,
                  uint64_t* out, hipStream_t stream, int num_cu);
// Hub CDFs from a prepared segment list (no allocation, no synchronisation):
// segment s covers entries [seg_idx[s] * kHubSeg, +kHubSeg) of hub
// seg_hub[s] (vertex G.hub_vid[h], CDF written at out[G.hub_pool[h] + i]);
// seg_first[h] .. seg_first[h+1] are hub h's segments; seg_tot is scratch.
constexpr int kHubSeg = 4096;
int hub_cdf_segments(const GraphDev& G, const csaw_graph_view& hv, int kind, int32_t n_hubs, const int32_t* seg_hub,
                     const int32_t* seg_idx, int64_t n_seg, const int64_t* seg_first, uint64_t* seg_tot,
                     uint64_t* out, hipStream_t stream, int num_cu);
int rmat_generate_device(int scale, int edge_factor, double a, double b, double c, uint64_t key, hipStream_t stream,
                         int num_cu, int64_t* V, int64_t* E, int64_t** d_off, int32_t** d_adj);
int graph_alloc_common(csaw_graph* g);
// Looks up this device's address of the allocation-free MDRW scratch (once
// per device; called with the device current when a graph is created).
int mdrw_scratch_setup();
int device_init(int* num_cu);
}  // namespace csaw
