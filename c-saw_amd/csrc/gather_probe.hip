// Random-read ceiling of the GPU: the rate at which independent random
// 4-byte reads from a large device array complete.  Every sampled edge reads
// one neighbour id at a random position of the adjacency array, and every
// biased draw one random block of a CDF, so this rate (not the streaming HBM
// bandwidth) bounds the sampling kernels.  Built as its own library
// (libcsaw_gather.so) for the benchmark; not part of the sampling ABI.
//
// Lane t of the grid issues `per_lane` reads at positions
// mulhi64(mix64(seed ^ (t * per_lane + r)), n), eight in flight at a time,
// and writes the sum of what it read (so no load is dead).  mix64 is the
// SplitMix64 finaliser (Steele, Lea and Flood, OOPSLA 2014).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_gather(const int32_t* __restrict__ x, uint64_t n, int64_t per_lane,
                                                uint64_t seed, int32_t* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t acc = 0;
  const uint64_t base = (uint64_t)t * (uint64_t)per_lane;
  for (int64_t r = 0; r < per_lane; r += 8) {
    int32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = x[__umul64hi(mix64(seed ^ (base + r + k)), n)];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc += (uint32_t)v[k];
  }
  out[t] = (int32_t)acc;
}

// y[i] = x[i] + 1: a streaming read and write of known size (scales the
// FETCH_SIZE / WRITE_SIZE counters of the profiled kernels).
__global__ __launch_bounds__(256) void k_copy(const int32_t* __restrict__ x, int32_t* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i] + 1;
}

__global__ void k_fill(int32_t* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = (int32_t)i;
}

}  // namespace

// array_bytes: size of the array read (rounded down to 4 bytes); reads:
// reads per launch (rounded up to a multiple of 8 per lane); wg_per_cu:
// workgroups of 256 lanes per CU; iters: timed launches after one warm-up.
// ms_out: median launch time (ms); reads_out: reads per launch.  Returns 0,
// or a hipError_t value.
extern "C" int csaw_gather_probe(int64_t array_bytes, int64_t reads, int32_t wg_per_cu, int32_t iters,
                                 double* ms_out, int64_t* reads_out) {
  if (array_bytes < 4 || reads < 1 || wg_per_cu < 1 || iters < 1 || !ms_out || !reads_out) return (int)hipErrorInvalidValue;
  int dev = 0;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return (int)e;
  const int64_t n = array_bytes / 4;
  const int grid = prop.multiProcessorCount * wg_per_cu;
  const int64_t lanes = (int64_t)grid * 256;
  const int64_t per_lane = std::max<int64_t>(8, ((reads + lanes - 1) / lanes + 7) / 8 * 8);
  int32_t *x = nullptr, *out = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  std::vector<float> ms;
  auto run = [&]() -> hipError_t {
    hipError_t r;
    if ((r = hipMalloc(&x, sizeof(int32_t) * n)) != hipSuccess) return r;
    if ((r = hipMalloc(&out, sizeof(int32_t) * lanes)) != hipSuccess) return r;
    k_fill<<<prop.multiProcessorCount * 8, 256>>>(x, n);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    if ((r = hipEventCreate(&a)) != hipSuccess) return r;
    if ((r = hipEventCreate(&b)) != hipSuccess) return r;
    for (int i = 0; i <= iters; ++i) {
      if ((r = hipEventRecord(a, nullptr)) != hipSuccess) return r;
      k_gather<<<grid, 256>>>(x, (uint64_t)n, per_lane, 0x9e3779b97f4a7c15ull * (uint64_t)(i + 1), out);
      if ((r = hipGetLastError()) != hipSuccess) return r;
      if ((r = hipEventRecord(b, nullptr)) != hipSuccess) return r;
      if ((r = hipEventSynchronize(b)) != hipSuccess) return r;
      float t = 0.f;
      if ((r = hipEventElapsedTime(&t, a, b)) != hipSuccess) return r;
      if (i > 0) ms.push_back(t);
    }
    return hipSuccess;
  };
  e = run();
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  (void)hipFree(x);
  (void)hipFree(out);
  if (e != hipSuccess) return (int)e;
  std::sort(ms.begin(), ms.end());
  *ms_out = ms[ms.size() / 2];
  *reads_out = lanes * per_lane;
  return 0;
}

// Copies `bytes` (rounded down to 4) `iters` times; ms_out: median launch time.
extern "C" int csaw_copy_probe(int64_t bytes, int32_t iters, double* ms_out) {
  if (bytes < 4 || iters < 1 || !ms_out) return (int)hipErrorInvalidValue;
  int dev = 0;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return (int)e;
  const int64_t n = bytes / 4;
  int32_t *x = nullptr, *y = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  std::vector<float> ms;
  auto run = [&]() -> hipError_t {
    hipError_t r;
    if ((r = hipMalloc(&x, sizeof(int32_t) * n)) != hipSuccess) return r;
    if ((r = hipMalloc(&y, sizeof(int32_t) * n)) != hipSuccess) return r;
    k_fill<<<prop.multiProcessorCount * 8, 256>>>(x, n);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    if ((r = hipEventCreate(&a)) != hipSuccess) return r;
    if ((r = hipEventCreate(&b)) != hipSuccess) return r;
    for (int i = 0; i < iters; ++i) {
      if ((r = hipEventRecord(a, nullptr)) != hipSuccess) return r;
      k_copy<<<prop.multiProcessorCount * 8, 256>>>(x, y, n);
      if ((r = hipGetLastError()) != hipSuccess) return r;
      if ((r = hipEventRecord(b, nullptr)) != hipSuccess) return r;
      if ((r = hipEventSynchronize(b)) != hipSuccess) return r;
      float t = 0.f;
      if ((r = hipEventElapsedTime(&t, a, b)) != hipSuccess) return r;
      ms.push_back(t);
    }
    return hipSuccess;
  };
  e = run();
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  (void)hipFree(x);
  (void)hipFree(y);
  if (e != hipSuccess) return (int)e;
  std::sort(ms.begin(), ms.end());
  *ms_out = ms[ms.size() / 2];
  return 0;
}
