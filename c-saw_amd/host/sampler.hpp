// C++ host interface mirroring the reference's sampler entry
// (non-stream/sampler.cuh:1-8): the same struct arguments and the same
// Sampler() signature and argument meaning, implemented over the C ABI.
//
//   beg, csr         paths of beg.bin (V+1 int64 offsets) and csr.bin (E int64
//                    ids), or "rmat:<scale>:<edge_factor>" for a generated graph
//   n_blocks         grid override of the expansion kernel (0: automatic)
//   n_threads        accepted for compatibility (workgroups are 256 lanes)
//   n_subgraph       number of samples
//   FrontierSize     1: neighbour sampling / random walk; > 1: multi-dimensional walk
//   NeighborSize     fanout per hop
//   Depth            hops (walk length for NeighborSize 1 with uniform bias)
//   rank             request id (keeps per-rank streams independent)
//
// Unlike the reference (which never sets sampled_edges, main.cu:305, and
// exits on errors), sampled_edges is filled, time is the device time in
// seconds, and errors are reported with sampled_edges = -status.
#pragma once
#include "../../include/csaw.h"

struct arguments {
  int sampled_edges;
  double time;
};

struct arguments Sampler(char beg[100], char csr[100], int n_blocks, int n_threads, int n_subgraph, int FrontierSize,
                         int NeighborSize, int Depth, struct arguments args, int rank);
