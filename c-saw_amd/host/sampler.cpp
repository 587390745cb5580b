// Sampler() over the C ABI (see sampler.hpp).  Bias policy: the compiled-in
// user hooks (CSAW_BIAS_USER), exactly like the reference, whose selection
// always went through api.cuh's EdgeBias.
#include "sampler.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static int open_graph(const char* beg, const char* csr, csaw_graph** g) {
  if (strncmp(beg, "rmat:", 5) == 0) {
    int scale = 0, ef = 0;
    if (sscanf(beg + 5, "%d:%d", &scale, &ef) != 2) return CSAW_ERR_INVALID_ARG;
    return csaw_graph_create_rmat(scale, ef, 0.57, 0.19, 0.19, 42, 0, g);
  }
  return csaw_graph_load_bin(beg, csr, nullptr, g);
}

struct arguments Sampler(char beg[100], char csr[100], int n_blocks, int n_threads, int n_subgraph, int FrontierSize,
                         int NeighborSize, int Depth, struct arguments args, int rank) {
  (void)n_threads;
  args.sampled_edges = 0;
  args.time = 0.0;
  csaw_graph* g = nullptr;
  int st = open_graph(beg, csr, &g);
  if (st != CSAW_OK) {
    fprintf(stderr, "Sampler: %s\n", csaw_last_error());
    args.sampled_edges = -st;
    return args;
  }
  int64_t V = 0;
  csaw_graph_info(g, &V, nullptr, nullptr);
  const int fs = FrontierSize < 1 ? 1 : FrontierSize;
  std::vector<int32_t> seeds((size_t)n_subgraph * fs);
  st = csaw_seeds(g, 42, (uint32_t)rank, (int64_t)seeds.size(), seeds.data());
  csaw_params p;
  memset(&p, 0, sizeof(p));
  p.samples = n_subgraph;
  p.frontier_size = fs;
  p.neighbor_size = NeighborSize;
  p.depth = Depth;
  p.bias = CSAW_BIAS_USER;
  p.dedup = 1;
  p.seed = 42;
  p.request_id = (uint32_t)rank;
  p.n_blocks = n_blocks;
  csaw_edges* e = nullptr;
  csaw_stats s;
  memset(&s, 0, sizeof(s));
  if (st == CSAW_OK) st = csaw_sample(g, &p, seeds.data(), &e, &s);
  if (st != CSAW_OK) {
    fprintf(stderr, "Sampler: %s\n", csaw_last_error());
    args.sampled_edges = -st;
  } else {
    args.sampled_edges = (int)s.edges;
    args.time = s.kernel_ms * 1e-3;
  }
  csaw_edges_free(e);
  csaw_graph_destroy(g);
  return args;
}
