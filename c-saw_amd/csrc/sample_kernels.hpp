// Launch interfaces of the per-hop sampling kernels.
#pragma once
#include <algorithm>

#include "common.hpp"
#include "mlp.hpp"

namespace csaw {

constexpr int kNumQueues = 19;  // variant * 3 + kind, the copy-all queue, 3 remap queues
constexpr int kQueueCopy = 15;
constexpr int kQueueRemap0 = 16; // + kind: region-remap entries (one wave per entry)

enum KernKind : int {
  kKernCount = 0,
  kKernScan,
  kKernChoose,
  kKernCopy,
  kKernThread,
  kKernWave,
  kKernBlock,
  kKernPlain,
  kKernWalk,
  kKernMdrw,
  kKernFill,
  kNumKernKinds
};
static const char* const kKernNames[kNumKernKinds] = {"count", "scan", "choose", "expand_copy", "expand_thread",
                                                      "expand_wave", "expand_block", "expand_plain", "walk", "mdrw",
                                                      "fill"};

struct ChunkLog {
  uint32_t key;
  uint32_t chunk;
  uint32_t ticks;   // wall_clock64 ticks (100 MHz constant clock)
  uint8_t hop;
  uint8_t variant;
  uint8_t pad[2];
};

struct LaunchCfg {
  hipStream_t stream = nullptr;
  int num_cu = 256;
  int count_per_cu = 8;
  int expand_per_cu = 8;   // WAVE / plain / copy workgroups per CU
  int thread_per_cu = 4;   // THREAD workgroups per CU
  int block_per_cu = 2;    // BLOCK workgroups per CU
  int grid_override = 0;
  int chunk_min = kChunkMin;  // tile shapes: smallest chunk piece,
  int group_min = kGroupMin;  // smallest sub-wave group
};

struct ExpandArgs {
  const uint32_t* queues;
  int64_t queue_cap;
  const uint32_t* qlen;
  uint32_t* qhead;
  ChunkLog* log;
  int32_t int_cache;  // integer biases search the CDF cache (no per-entry build)
  int32_t remap;      // some request uses the region-remap mode (its queues may be non-empty)
};

struct ChooseArgs {
  int32_t sched;
  int32_t variant;
  uint64_t explore_key;
  uint32_t* qlen;
  uint32_t* queues;
  int64_t queue_cap;
  uint8_t* chosen;   // per chunk variant (may be null)
  int32_t int_cache; // integer CDF cache available (THREAD valid for any degree)
  int32_t in_place;  // the chunks' neighbour lists are read over the bus
  MlpDev mlp;
  const void* mlp_w; // non-null when a model is loaded
};

// Per-kernel-kind timing with hipEvents on the launch stream.
struct TimingHook {
  virtual void begin(int which, hipStream_t s) = 0;
  virtual void end(int which, hipStream_t s) = 0;
  virtual ~TimingHook() {}
};

__device__ int mlp_choose(const ChooseArgs& A, const float* wl, const Chunk& c, int kind, const HopDev& H);

int launch_count(const LaunchCfg& L, const GraphDev& G, const ReqDev* R, const int32_t* gs2req, const HopDev& H,
                 bool bin, int64_t max_entries);
// Bins the n entries list[0 .. n) into chunks (H.order by list position).
int launch_bin_list(const LaunchCfg& L, const GraphDev& G, const ReqDev* R, const int32_t* gs2req, const HopDev& H,
                    const int32_t* list, int64_t n);
int launch_scans(const LaunchCfg& L, const HopDev& H, int64_t* total, int64_t max_entries);
int launch_fill(const LaunchCfg& L, uint32_t* p, uint32_t v, int64_t n);
int launch_digest(hipStream_t st, int num_cu, const int32_t* src, const int32_t* dst, const int32_t* smp, int64_t lo,
                  int64_t hi, int64_t pos0, int64_t gbase, unsigned long long* out);
int launch_expand_plain(const LaunchCfg& L, int kind, const GraphDev& G, const csaw_graph_view& hv, const ReqDev* R,
                        const int32_t* gs2req, const HopDev& H, int64_t max_entries, int gs);
int launch_choose(const LaunchCfg& L, const ChooseArgs& A, const HopDev& H, int64_t max_chunks);
int launch_expand_list(const LaunchCfg& L, int kind, const GraphDev& G, const csaw_graph_view& hv, const ReqDev* R,
                       const int32_t* gs2req, const HopDev& H, const int32_t* list, int64_t n);
// The variant kernels of one hop; with n_streams > 0 kernel i goes to
// streams[i % n_streams] (heaviest kinds first), else all to L.stream.
int launch_expand_chunks(const LaunchCfg& L, const GraphDev& G, const csaw_graph_view& hv, const ReqDev* R,
                         const int32_t* gs2req, const HopDev& H, const ExpandArgs& A, const bool* kinds_present,
                         TimingHook* th, const hipStream_t* streams = nullptr, int n_streams = 0);

}  // namespace csaw
