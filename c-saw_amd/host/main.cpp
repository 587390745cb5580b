// Command line mirroring the reference's driver (non-stream/mpi_main.cpp:17-49):
//   csaw_cli <dataset name> <beg file> <csr file> <ThreadBlocks> <Threads>
//            <# of samples> <FrontierSize> <NeighborSize> <Depth/Length> <#GPUs>
// prints "Edges sampled, dataset name, min-time, max-time" (README.md:38-41).
// One process drives one GPU; #GPUs other than 1 is refused (the reference
// divided the samples by it and reduced times over MPI ranks, mpi_main.cpp:38-42).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "sampler.hpp"

int main(int argc, char* argv[]) {
  if (argc != 11) {
    printf("Input: ./exe <dataset name> <beg file> <csr file> <ThreadBlocks> <Threads> <# of samples> "
           "<FrontierSize> <NeighborSize> <Depth/Length> <#GPUs>\n");
    return 1;
  }
  const int n_blocks = atoi(argv[4]);
  const int n_threads = atoi(argv[5]);
  const int samples = atoi(argv[6]);
  const int frontier = atoi(argv[7]);
  const int neighbors = atoi(argv[8]);
  const int depth = atoi(argv[9]);
  const int gpus = atoi(argv[10]);
  if (gpus != 1) {
    fprintf(stderr, "this build drives one GPU per process (#GPUs must be 1)\n");
    return 2;
  }
  char beg[100], csr[100];
  snprintf(beg, sizeof(beg), "%s", argv[2]);
  snprintf(csr, sizeof(csr), "%s", argv[3]);
  struct arguments args = {0, 0.0};
  args = Sampler(beg, csr, n_blocks, n_threads, samples, frontier, neighbors, depth, args, 0);
  if (args.sampled_edges < 0) return 3;
  printf("%d,%s,%f,%f\n", args.sampled_edges, argv[1], args.time, args.time);
  return 0;
}
