// csaw_api.h -- device-side plug-in hooks (the drop-in for non-stream/api.cuh:4-27).
//
// As in the reference, hooks are compiled in: edit the three functions
// below and rebuild the library.  They are used when a request selects
// CSAW_BIAS_USER (EdgeBias) and by the multi-dimensional random walk
// (VertexBias); Update is applied to every selected neighbour of every
// bias kind.  Defaults match the reference file: VertexBias and EdgeBias
// return the degree (api.cuh:8, 17 -- with the degree taken from the
// offsets, off[v+1]-off[v], as streaming/streaming_sampling.cu:251-253 does),
// Update returns the selected vertex (api.cuh:26).
//
// Contract: biases are non-negative 32-bit integers (the CDF is an exact
// uint64 prefix sum); Update must return a vertex id in [0, V).  The
// CPU oracle implements the same defaults; a changed hook needs the same
// change there for the parity tests to apply.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// What the hooks see of the graph (gpu_graph, gpu_graph.cuh:10-21).
struct csaw_graph_view {
  const int64_t* beg_pos;     // V+1 offsets
  const int32_t* adj_list;    // E neighbour ids
  const int32_t* degree_list; // V degrees (per vertex)
  const float* weight_list;   // E weights or nullptr
  int64_t vert_count;
  int64_t edge_count;
};

__device__ __forceinline__ uint32_t VertexBias(int32_t vertexID, const csaw_graph_view* graph) {
  // For MDRW (api.cuh:7-8)
  return (uint32_t)graph->degree_list[vertexID];
}

__device__ __forceinline__ uint32_t EdgeBias(int32_t vertexID, const csaw_graph_view* graph) {
  // For biased neighbour sampling / layer sampling (api.cuh:16-17)
  return (uint32_t)graph->degree_list[vertexID];
}

__device__ __forceinline__ int32_t Update(const csaw_graph_view* G, int32_t selected, int32_t source) {
  (void)G;
  (void)source;
  return selected;
}
