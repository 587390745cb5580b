// Cost model shared by host and device: chunk features, BF16 <-> FP32, and
// the 16 -> 64 -> 64 -> CSAW_NUM_VARIANTS MLP layout (BF16 weights, FP32
// accumulate, ReLU).  Output o_v is the predicted log2 of the chunk's device
// time (wall-clock ticks) under variant v.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "../../include/csaw.h"

namespace csaw {

constexpr int kMlpIn = 16;
constexpr int kMlpHid = 64;
constexpr int kMlpOut = CSAW_NUM_VARIANTS;
// parameter block layout (floats in the master copy, bf16 for weights in the working copy)
constexpr int kMlpW1 = 0;                            // [kMlpHid][kMlpIn]
constexpr int kMlpB1 = kMlpW1 + kMlpHid * kMlpIn;    // [kMlpHid]
constexpr int kMlpW2 = kMlpB1 + kMlpHid;             // [kMlpHid][kMlpHid]
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid;   // [kMlpHid]
constexpr int kMlpW3 = kMlpB2 + kMlpHid;             // [kMlpOut][kMlpHid]
constexpr int kMlpB3 = kMlpW3 + kMlpOut * kMlpHid;   // [kMlpOut]
constexpr int kMlpParams = kMlpB3 + kMlpOut;

struct MlpDev {
  const uint16_t* w;  // kMlpParams bf16 values (biases included, also bf16)
};

__host__ __device__ __forceinline__ float bf16_to_f32(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

__host__ __device__ __forceinline__ float flog2p1(float x) { return log2f(x + 1.0f); }

// Features of one chunk.  kind: 0 uniform, 1 integer, 2 fp32.  in_place:
// the chunk's neighbour lists are read from host memory over the bus (a
// streamed partition read in place); it adds 1 to the constant input f[0],
// so that models trained without such chunks see f[0] = 1 as before.
__host__ __device__ __forceinline__ void chunk_features(uint32_t len, float sum_d, uint32_t max_d, uint32_t sum_f,
                                                       uint32_t max_f, int kind, int hop, int bucket,
                                                       uint32_t n_chunks, float* f, int in_place = 0) {
  const float n = (float)(len > 0 ? len : 1);
  const float mean_d = sum_d / n;
  const float mean_f = (float)sum_f / n;
  f[0] = in_place ? 2.0f : 1.0f;
  f[1] = flog2p1(n) / 10.0f;
  f[2] = flog2p1(mean_d) / 20.0f;
  f[3] = flog2p1((float)max_d) / 20.0f;
  f[4] = mean_f / 64.0f;
  f[5] = (float)max_f / 64.0f;
  f[6] = kind == 0 ? 1.0f : 0.0f;
  f[7] = kind == 1 ? 1.0f : 0.0f;
  f[8] = kind == 2 ? 1.0f : 0.0f;
  f[9] = (float)hop / 8.0f;
  f[10] = mean_d > 0.0f ? fminf(mean_f / mean_d, 1.0f) : 1.0f;
  f[11] = max_d <= 32 ? 1.0f : 0.0f;
  f[12] = max_d <= 1024 ? 1.0f : 0.0f;
  f[13] = max_d <= 8192 ? 1.0f : 0.0f;
  f[14] = flog2p1((float)n_chunks) / 24.0f;
  f[15] = (float)((bucket - 1) % 8) / 8.0f;  // degree class
}

// Forward pass over weights already widened to FP32 (e.g. staged in LDS).
__host__ __device__ __forceinline__ void mlp_forward1_f32(const float* __restrict__ w, const float* x, float* out) {
  float h1[kMlpHid], h2[kMlpHid];
  for (int j = 0; j < kMlpHid; ++j) {
    float s = w[kMlpB1 + j];
    for (int i = 0; i < kMlpIn; ++i) s += w[kMlpW1 + j * kMlpIn + i] * x[i];
    h1[j] = s > 0.0f ? s : 0.0f;
  }
  for (int j = 0; j < kMlpHid; ++j) {
    float s = w[kMlpB2 + j];
    for (int i = 0; i < kMlpHid; ++i) s += w[kMlpW2 + j * kMlpHid + i] * h1[i];
    h2[j] = s > 0.0f ? s : 0.0f;
  }
  for (int j = 0; j < kMlpOut; ++j) {
    float s = w[kMlpB3 + j];
    for (int i = 0; i < kMlpHid; ++i) s += w[kMlpW3 + j * kMlpHid + i] * h2[i];
    out[j] = s;
  }
}

// Forward pass for one feature vector (FP32 accumulate over BF16 weights).
__host__ __device__ __forceinline__ void mlp_forward1(const uint16_t* __restrict__ w, const float* x, float* out) {
  float h1[kMlpHid], h2[kMlpHid];
  for (int j = 0; j < kMlpHid; ++j) {
    float s = bf16_to_f32(w[kMlpB1 + j]);
    for (int i = 0; i < kMlpIn; ++i) s += bf16_to_f32(w[kMlpW1 + j * kMlpIn + i]) * x[i];
    h1[j] = s > 0.0f ? s : 0.0f;
  }
  for (int j = 0; j < kMlpHid; ++j) {
    float s = bf16_to_f32(w[kMlpB2 + j]);
    for (int i = 0; i < kMlpHid; ++i) s += bf16_to_f32(w[kMlpW2 + j * kMlpHid + i]) * h1[i];
    h2[j] = s > 0.0f ? s : 0.0f;
  }
  for (int j = 0; j < kMlpOut; ++j) {
    float s = bf16_to_f32(w[kMlpB3 + j]);
    for (int i = 0; i < kMlpHid; ++i) s += bf16_to_f32(w[kMlpW3 + j * kMlpHid + i]) * h2[i];
    out[j] = s;
  }
}

}  // namespace csaw
