// Host-side plan state (see plan.cpp).
#pragma once
#include <utility>
#include <vector>

#include "graph.hpp"
#include "sample_kernels.hpp"

namespace csaw {

// layout of the small counter block reset at every hop
enum : int {
  kCtrChunks = 0,
  kCtrQlen = 1,
  kCtrQhead = kCtrQlen + kNumQueues,
  kCtrFlags = kCtrQhead + kNumQueues,
  kCtrCount = kCtrFlags + 1
};

struct EventTimer : TimingHook {
  struct Rec {
    int kind;
    hipEvent_t a, b;
  };
  bool enabled = false;
  std::vector<Rec> pool;
  size_t used = 0;
  double total_ms[kNumKernKinds] = {};
  int64_t launches[kNumKernKinds] = {};
  double total_bytes[kNumKernKinds] = {};
  void begin(int which, hipStream_t s) override;
  void end(int which, hipStream_t s) override;
  void collect();
  ~EventTimer() override;
};

struct Unit {
  std::vector<int> reqs;        // plan request indices, sorted by depth (desc)
  std::vector<int64_t> gbase;   // global sample base per unit request
  int64_t G = 0;
  int depth_max = 0;
  bool kinds[3] = {false, false, false};
  bool remap = false;           // some request uses the region-remap mode
  std::vector<int32_t> gb;      // per hop: first inactive global sample
  std::vector<int64_t> cap;     // per hop: edge capacity
  std::vector<int64_t> nmax;    // per hop: frontier entries upper bound
  ReqDev* d_req = nullptr;
  int32_t* d_gs2req = nullptr;
  int32_t* d_fv0 = nullptr;
  int32_t* d_fs0 = nullptr;
  int64_t* d_fo0 = nullptr;
  std::vector<int64_t*> d_eo;
  std::vector<int32_t*> d_src, d_dst, d_smp;
};

struct HopLog {
  int hop = 0;
  std::vector<Chunk> chunks;
  std::vector<ChunkLog> log;
};

void apply_tuning(LaunchCfg& L, const csaw_tuning* t);
// CSAW_ERR_INVALID_ARG for shapes out of range (workgroups per CU 0..64,
// chunk_min a power of two in 8..1024, group_min 4..32; 0 keeps a default).
int validate_tuning(const csaw_tuning* t);
int mlp_upload_bf16(const csaw_mlp* m, uint16_t* dst);

}  // namespace csaw

struct csaw_plan {
  csaw_graph* g = nullptr;
  std::vector<csaw_request> reqs;
  std::vector<std::vector<int32_t>> seeds;
  std::vector<std::pair<int, int>> req_unit;  // (unit, index in unit)
  csaw_plan_opts opts;
  std::vector<csaw::Unit> units;
  uint32_t* local_excl = nullptr;
  int32_t* order = nullptr;
  uint32_t* blk_excl = nullptr;
  uint64_t* grp_excl = nullptr;
  csaw::Chunk* chunks = nullptr;
  uint32_t chunk_cap = 0;
  uint32_t* counters = nullptr;
  uint32_t* queues = nullptr;
  int64_t queue_cap = 0;
  int64_t* totals = nullptr;
  uint8_t* chosen = nullptr;
  csaw::ChunkLog* log = nullptr;
  uint16_t* d_mlp = nullptr;
  unsigned long long* bytes_acc = nullptr;  // algorithmic bytes per kernel kind (timing runs)
  csaw::LaunchCfg L;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipGraphExec_t gexec = nullptr;
  hipStream_t graph_stream = nullptr;
  // batched mode: side streams the expansion kernels of a hop fan out to
  // (joined back before the next hop), fork / join events
  static constexpr int kSide = 3;
  hipStream_t side[kSide] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kSide] = {};
  csaw::EventTimer timer;
  double last_ms = 0.0;
  int64_t runs = 0;
  std::vector<csaw::HopLog> hop_logs;
  std::vector<int> pending_log_hops;
};
