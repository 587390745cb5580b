// Shared host/device definitions: status plumbing, wave64 primitives and the
// device-side views of graphs, requests and hop state.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/csaw.h"
#include "csaw_api.h"
#include "philox.hpp"

namespace csaw {

// ---------------------------------------------------------------- errors
void set_last_error(const std::string& msg);
const char* last_error();

#define CSAW_HIP(call)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      ::csaw::set_last_error(std::string(#call) + ": " + hipGetErrorString(e_) + " at " +   \
                             __FILE__ + ":" + std::to_string(__LINE__));                    \
      return e_ == hipErrorOutOfMemory ? CSAW_ERR_OOM : CSAW_ERR_HIP;                        \
    }                                                                                       \
  } while (0)

#define CSAW_TRY(expr)               \
  do {                               \
    int s_ = (expr);                 \
    if (s_ != CSAW_OK) return s_;    \
  } while (0)

#define CSAW_LAUNCH_CHECK() CSAW_HIP(hipGetLastError())

// ------------------------------------------------------------ constants
constexpr int kWave = 64;
constexpr int kBlockEntries = 1024;  // frontier entries per count block
constexpr int kBlockShift = 10;
constexpr int kGroupShift = 20;      // 1024 blocks per scan group
constexpr int kReselect = 400;       // bounded reselect (functions.cuh:844)
constexpr int kThreadIntCap = 32;    // THREAD variant: integer CDF up to 32 entries (SML_MID)
constexpr int kThreadFanCap = 32;    // THREAD variant: fanout up to 32
constexpr int kWaveIntCap = 1024;    // WAVE variant: LDS integer CDF up to 1024 (MID_LRG)
constexpr int kBlockIntCap = 8192;   // BLOCK variant: LDS integer CDF up to 8192
constexpr int kHubMin = kWaveIntCap; // integer-bias CDFs above this live in the hub cache
constexpr int kWaveBitmapCap = 16384;// WAVE bitmap taken-set, bits
constexpr int kBlockBitmapCap = 65536;
constexpr int kNumBuckets = 128;     // 0 = copy-all; 1 + bias*32 + ceil(log2 d)
constexpr int kBucketCopy = 0;
// Chunk descriptors of a hop: a block's buckets (at most 1 + 3 kinds x 8
// degree classes + 3 region-remap buckets non-empty) split into pieces of
// cmax entries give at most 28 + 1024 / cmax per block, and cmax >= n /
// (expand waves), so the hop has at most 28 * blocks + waves chunks.
constexpr int kChunksPerBlock = 28;
constexpr int kChunkMin = 8;         // smallest chunk piece (buckets split for small hops)

// Bias kinds inside kernels: integer biases share one code path.
enum BiasKind : int { kUniform = 0, kInt = 1, kF32 = 2 };

// ----------------------------------------------------------- device views
// One 16-byte record per vertex for a cached CDF: offset, degree and the
// CDF total C[off + deg - 1] (FP32 bits or uint32) in one aligned load,
// instead of the offset pair and then the total.
struct alignas(16) VertexRecord {
  int64_t off;
  uint32_t total;  // 0 when deg == 0
  int32_t deg;
};

struct GraphDev {
  const int64_t* off;
  const int32_t* adj;
  const int32_t* deg;
  const float* w;         // weights (may be null)
  const float* cdf;       // sequential FP32 CDF per vertex (may be null)
  const int32_t* hub_vid;       // sorted ids of vertices with degree > kHubMin
  const int64_t* hub_pool;      // pool offset per hub (n_hubs + 1)
  const uint64_t* hub_cdf[2];   // [DEGREE, USER]: inclusive uint64 CDFs of hubs
  const uint64_t* int_cdf[2];   // full per-entry CDF cache (null when not used)
  const float* cdf_idx;         // 16-ary search index over cdf (null when not used)
  const uint64_t* int_idx[2];   // 16-ary search indexes over int_cdf
  const uint32_t* int_cdf32[2]; // the cache as uint32 when all totals fit (int_cdf null then)
  const uint32_t* int_idx32[2];
  // [DEGREE, USER, WEIGHT] vertex records over the uint32 integer caches and
  // the FP32 CDF (null when that cache is not in the view)
  const VertexRecord* vrec[3] = {nullptr, nullptr, nullptr};
  int32_t n_hubs;
  int64_t V, E;
};

struct ReqDev {
  int32_t fanout[CSAW_MAX_HOPS];
  int32_t depth, bias, dedup, samples;
  uint32_t key_lo, key_hi, request_id, pad;
  int64_t gbase;
};

struct Chunk {
  uint32_t start;  // index into the order array
  uint16_t len;
  uint8_t bucket;
  uint8_t hop;
  uint32_t sum_f;
  uint32_t max_d;
  float sum_d;
  uint32_t key;    // block * kNumBuckets + bucket (stable across runs)
  uint8_t max_f;
  uint8_t pad;
  uint16_t sum_l;  // sum of ceil(log2 d) over the chunk's entries
};

struct HopDev {
  const int32_t* fv;    // frontier vertices
  const int32_t* fs;    // frontier global sample ids
  const int64_t* fo;    // per-sample frontier offsets (global sample index)
  const int64_t* n_ptr; // number of active entries (= fo[gb])
  int32_t hop;
  int32_t gb;           // first inactive global sample
  uint32_t* local_excl; // per entry
  uint32_t* blk_excl;   // per block of 1024 entries
  uint64_t* grp_excl;   // per group of 1024 blocks
  int32_t* order;       // bucket-sorted entry ids (per block region)
  Chunk* chunks;
  uint32_t* n_chunks;
  uint32_t chunk_cap;
  int32_t* src;
  int32_t* dst;
  int32_t* smp;
  int64_t* eo;          // per-sample edge offsets of this hop
  uint32_t* flags;      // bit 0: chunk overflow
  unsigned long long* bytes;  // algorithmic bytes per kernel kind (null: not counted)
  int32_t expand_waves; // waves of one expand launch (sizes chunks; 0: whole buckets)
  int32_t chunk_min;    // smallest chunk piece (a power of two, kChunkMin .. kBlockEntries)
  int32_t group_min;    // smallest sub-wave group (kGroupMin .. 32)
  int32_t only_kind;    // 1 + a bias kind: expand only that kind's entries (0: all)
};

// Hops of a batched plan with at most this many frontier entries skip the
// binning and the variant kernels (whose fixed costs dominate a small hop):
// one whole-hop kernel per bias kind expands them in frontier order.
constexpr int64_t kSmallHopEntries = 65536;

// Entries per chunk for a hop of n frontier entries: whole 1024-entry buckets
// when the hop gives every expand wave several of them, smaller pieces (down
// to kChunkMin) when it is small, so that the work still spreads over the
// waves.
__host__ __device__ __forceinline__ uint32_t chunk_max_for(int64_t n, int32_t waves, int32_t cmin = kChunkMin) {
  if (waves <= 0) return (uint32_t)kBlockEntries;
  const int64_t per = (n + waves - 1) / waves;
  uint32_t c = (uint32_t)(cmin < kChunkMin ? kChunkMin : cmin);
  while (c < (uint32_t)kBlockEntries && (int64_t)c < per) c <<= 1;
  return c;
}

// Sub-wave group size for a largest fanout: the fanout itself (at least
// gmin, kGroupMin by default), so that e.g. fanout 5 runs 12 entries per
// wave round instead of 8; above 32 one group is the whole wave.  Any size
// from kGroupMin to 32 gives the same edges (the autotuner picks gmin).
constexpr int kGroupMin = 4;
__host__ __device__ __forceinline__ int group_size_for(int max_f, int gmin = kGroupMin) {
  if (gmin < kGroupMin) gmin = kGroupMin;
  if (gmin > 32) gmin = 32;
  return max_f <= gmin ? gmin : (max_f <= 32 ? max_f : kWave);
}

__host__ __device__ __forceinline__ int ceil_log2_u32(uint32_t d) {
  // smallest c with 2^c >= d (d >= 1)
#if defined(__HIP_DEVICE_COMPILE__)
  return d <= 1 ? 0 : 32 - __clz((int)(d - 1));
#else
  return d <= 1 ? 0 : 32 - __builtin_clz(d - 1);
#endif
}

// Number of edges an entry emits: depends on d only (canonical semantics).
__host__ __device__ __forceinline__ int32_t entry_k(int64_t d, int32_t F, int32_t dedup) {
  if (d <= 0 || F <= 0) return 0;
  if (dedup) return d < F ? (int32_t)d : F;
  return F;
}

__host__ __device__ __forceinline__ int bias_kind(int32_t bias) {
  return bias == CSAW_BIAS_UNIFORM ? kUniform : (bias == CSAW_BIAS_WEIGHT ? kF32 : kInt);
}

// Degree classes of a chunk (by ceil(log2 d)): <=4, <=16, <=32, <=128,
// <=1024, <=8192, <=65536, larger.  Coarse classes keep chunks large.
constexpr int kDegClasses = 8;
__host__ __device__ __forceinline__ int degree_class(int l) {
  return l <= 2 ? 0 : l <= 4 ? 1 : l <= 5 ? 2 : l <= 7 ? 3 : l <= 10 ? 4 : l <= 13 ? 5 : l <= 16 ? 6 : 7;
}
// upper bound of ceil(log2 d) in a class
__host__ __device__ __forceinline__ int class_max_log(int cls) {
  const int m[kDegClasses] = {2, 4, 5, 7, 10, 13, 16, 31};
  return m[cls & 7];
}

// bucket id: 0 copy-all, else 1 + kDegClasses*kind + degree class (kind 0..2)
// Entries of region-remap requests (dedup == 2) that select (d > F) get
// one bucket per bias kind, expanded by their own whole-wave kernel.
constexpr int kBucketRemap0 = 1 + 3 * kDegClasses;

__host__ __device__ __forceinline__ int entry_bucket(int64_t d, int32_t F, int32_t dedup, int32_t bias) {
  if (dedup && d <= F) return kBucketCopy;
  int kind = bias_kind(bias);
  if (dedup == 2) return kBucketRemap0 + kind;
  int l = ceil_log2_u32((uint32_t)(d > 0x7fffffff ? 0x7fffffff : d));
  // a uniform draw costs the same at any degree: degrees up to 1024 share
  // one class, so a block's uniform entries form few, full chunks
  if (kind == kUniform && l <= 10) l = 10;
  return 1 + kDegClasses * kind + degree_class(l);
}
__host__ __device__ __forceinline__ int bucket_kind(int bucket) {
  return bucket >= kBucketRemap0 ? bucket - kBucketRemap0 : (bucket - 1) / kDegClasses;
}
__host__ __device__ __forceinline__ int bucket_class(int bucket) {
  return bucket >= kBucketRemap0 ? kDegClasses - 1 : (bucket - 1) % kDegClasses;
}

// Algorithmic bytes of one expansion (each datum counted once per use, no
// cache-line amplification): frontier entry 8 + offset 8 + degree 4 = 20;
// per emitted edge the picked id 4 + the edge written (src, dst, sample) 12.
// Integer biases built on the fly read adj + degree of every neighbour (8d);
// a cached CDF costs its total plus ceil(log2 d) probes per draw.
//   copy-all:  20 + 16d          uniform: 20 + 16k
//   int build: 20 + 8d + 16k     int cached: 28 + k(8l + 16)
//   fp32:      24 + k(4l + 16)   (l = ceil(log2 d))
__host__ __device__ __forceinline__ uint64_t expansion_bytes(int kind, bool built, int64_t n, int64_t sum_d,
                                                             int64_t sum_k, int l, bool copy) {
  if (copy) return 20 * n + 16 * sum_d;
  if (kind == kUniform) return 20 * n + 16 * sum_k;
  if (kind == kF32) return 24 * n + sum_k * (4 * l + 16);
  if (built) return 20 * n + 8 * sum_d + 16 * sum_k;
  return 28 * n + sum_k * (8 * l + 16);
}

// ------------------------------------------------------- wave primitives
#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int32_t readlane(int32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ uint32_t readlane_u(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int l) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

// Inclusive wave64 scans on DPP: four row_shr steps (1, 2, 4, 8) give every
// lane the prefix within its row of 16 lanes; row_bcast:15 then adds row
// r's last lane to row r+1 (rows 1 and 3), and row_bcast:31 adds lane 31 to
// rows 2 and 3.  Lanes whose DPP source lies outside the row (or whose row
// the row mask disables) receive the `old` operand, 0.  Every lane of the
// wave must be active (callers run the scan in wave-uniform control flow).
namespace dpp {
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t mov(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t mov64(uint64_t x) {
  return (uint64_t)mov<CTRL, ROWS>((uint32_t)x) | ((uint64_t)mov<CTRL, ROWS>((uint32_t)(x >> 32)) << 32);
}
}  // namespace dpp

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x) {
  x += dpp::mov<dpp::kRowShr1, 0xf>(x);
  x += dpp::mov<dpp::kRowShr2, 0xf>(x);
  x += dpp::mov<dpp::kRowShr4, 0xf>(x);
  x += dpp::mov<dpp::kRowShr8, 0xf>(x);
  x += dpp::mov<dpp::kRowBcast15, 0xa>(x);
  x += dpp::mov<dpp::kRowBcast31, 0xc>(x);
  return x;
}

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t x) {
  x += dpp::mov64<dpp::kRowShr1, 0xf>(x);
  x += dpp::mov64<dpp::kRowShr2, 0xf>(x);
  x += dpp::mov64<dpp::kRowShr4, 0xf>(x);
  x += dpp::mov64<dpp::kRowShr8, 0xf>(x);
  x += dpp::mov64<dpp::kRowBcast15, 0xa>(x);
  x += dpp::mov64<dpp::kRowBcast31, 0xc>(x);
  return x;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// first index i in [0,n) with C[i] > x (C inclusive, C[n-1] > x assumed)
template <typename T, typename Ptr>
__device__ __forceinline__ int64_t upper_bound_dev(Ptr C, int64_t n, T x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (C[mid] > x) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// hub lookup: index of v in the sorted hub id list, or -1
__device__ __forceinline__ int32_t hub_find(const int32_t* vid, int32_t n, int32_t v) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    int32_t mid = (lo + hi) >> 1;
    if (vid[mid] < v) lo = mid + 1; else hi = mid;
  }
  return (lo < n && vid[lo] == v) ? lo : -1;
}

// Selection counter word 3.
__device__ __forceinline__ uint32_t select_c3(uint32_t request, int hop) {
  return (kDomSelect << 28) | (request << 8) | ((uint32_t)hop & 0xFF);
}
#endif

}  // namespace csaw
