// Learned cost model: HIP forward, backward and SGD kernels for the
// 16 -> 64 -> 64 -> CSAW_NUM_VARIANTS MLP (BF16 weights, FP32 accumulate).
//
// The model is tiny, so the kernels are plain VALU code: forward and the
// activation gradients run one lane per sample; the weight gradients run one
// lane per parameter summing over the minibatch in sample order, which keeps
// every gradient deterministic (no atomics).  SGD keeps an FP32 master copy
// and re-rounds the BF16 working weights after each step.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "graph.hpp"
#include "mlp.hpp"
#include "plan.hpp"

namespace csaw {

constexpr int kMlpThreads = 256;

__global__ __launch_bounds__(kMlpThreads) void k_mlp_fwd(const uint16_t* __restrict__ w, const float* __restrict__ X,
                                                        const int32_t* __restrict__ idx, int64_t n,
                                                        float* __restrict__ H1, float* __restrict__ H2,
                                                        float* __restrict__ Y) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx ? idx[s] : s;
    float x[kMlpIn];
    for (int i = 0; i < kMlpIn; ++i) x[i] = X[r * kMlpIn + i];
    float h1[kMlpHid];
    for (int j = 0; j < kMlpHid; ++j) {
      float a = bf16_to_f32(w[kMlpB1 + j]);
      for (int i = 0; i < kMlpIn; ++i) a = fmaf(bf16_to_f32(w[kMlpW1 + j * kMlpIn + i]), x[i], a);
      h1[j] = a > 0.f ? a : 0.f;
      if (H1) H1[s * kMlpHid + j] = h1[j];
    }
    float h2[kMlpHid];
    for (int j = 0; j < kMlpHid; ++j) {
      float a = bf16_to_f32(w[kMlpB2 + j]);
      for (int i = 0; i < kMlpHid; ++i) a = fmaf(bf16_to_f32(w[kMlpW2 + j * kMlpHid + i]), h1[i], a);
      h2[j] = a > 0.f ? a : 0.f;
      if (H2) H2[s * kMlpHid + j] = h2[j];
    }
    for (int j = 0; j < kMlpOut; ++j) {
      float a = bf16_to_f32(w[kMlpB3 + j]);
      for (int i = 0; i < kMlpHid; ++i) a = fmaf(bf16_to_f32(w[kMlpW3 + j * kMlpHid + i]), h2[i], a);
      Y[s * kMlpOut + j] = a;
    }
  }
}

// dY = M (Y - T) / inv_cnt; dH2 = (W3^T dY) * [H2 > 0]; dH1 = (W2^T dH2) * [H1 > 0]
__global__ __launch_bounds__(kMlpThreads) void k_mlp_bwd_act(const uint16_t* __restrict__ w, const int32_t* __restrict__ idx,
                                                            int64_t n, const float* __restrict__ T,
                                                            const float* __restrict__ M, const float* __restrict__ Y,
                                                            const float* __restrict__ H1, const float* __restrict__ H2,
                                                            float inv_cnt, float* __restrict__ dY,
                                                            float* __restrict__ dH1, float* __restrict__ dH2,
                                                            float* __restrict__ loss_s) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = idx ? idx[s] : s;
    float dy[kMlpOut];
    float l = 0.f;
    for (int j = 0; j < kMlpOut; ++j) {
      const float m = M[r * kMlpOut + j];
      const float e = Y[s * kMlpOut + j] - T[r * kMlpOut + j];
      dy[j] = m != 0.f ? e * inv_cnt : 0.f;
      l += m != 0.f ? 0.5f * e * e : 0.f;
      dY[s * kMlpOut + j] = dy[j];
    }
    loss_s[s] = l * inv_cnt;
    float d2[kMlpHid];
    for (int i = 0; i < kMlpHid; ++i) {
      float a = 0.f;
      for (int j = 0; j < kMlpOut; ++j) a = fmaf(bf16_to_f32(w[kMlpW3 + j * kMlpHid + i]), dy[j], a);
      d2[i] = H2[s * kMlpHid + i] > 0.f ? a : 0.f;
      dH2[s * kMlpHid + i] = d2[i];
    }
    for (int i = 0; i < kMlpHid; ++i) {
      float a = 0.f;
      for (int j = 0; j < kMlpHid; ++j) a = fmaf(bf16_to_f32(w[kMlpW2 + j * kMlpHid + i]), d2[j], a);
      dH1[s * kMlpHid + i] = H1[s * kMlpHid + i] > 0.f ? a : 0.f;
    }
  }
}

// one lane per parameter; sums over the minibatch in order
__global__ __launch_bounds__(kMlpThreads) void k_mlp_bwd_w(const float* __restrict__ X, const int32_t* __restrict__ idx,
                                                          int64_t n, const float* __restrict__ H1,
                                                          const float* __restrict__ H2, const float* __restrict__ dY,
                                                          const float* __restrict__ dH1, const float* __restrict__ dH2,
                                                          float* __restrict__ grad) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < kMlpParams; p += gridDim.x * blockDim.x) {
    float acc = 0.f;
    if (p < kMlpB1) {
      const int j = p / kMlpIn, i = p % kMlpIn;
      for (int64_t s = 0; s < n; ++s) acc = fmaf(dH1[s * kMlpHid + j], X[(idx ? idx[s] : s) * kMlpIn + i], acc);
    } else if (p < kMlpW2) {
      const int j = p - kMlpB1;
      for (int64_t s = 0; s < n; ++s) acc += dH1[s * kMlpHid + j];
    } else if (p < kMlpB2) {
      const int q = p - kMlpW2, j = q / kMlpHid, i = q % kMlpHid;
      for (int64_t s = 0; s < n; ++s) acc = fmaf(dH2[s * kMlpHid + j], H1[s * kMlpHid + i], acc);
    } else if (p < kMlpW3) {
      const int j = p - kMlpB2;
      for (int64_t s = 0; s < n; ++s) acc += dH2[s * kMlpHid + j];
    } else if (p < kMlpB3) {
      const int q = p - kMlpW3, j = q / kMlpHid, i = q % kMlpHid;
      for (int64_t s = 0; s < n; ++s) acc = fmaf(dY[s * kMlpOut + j], H2[s * kMlpHid + i], acc);
    } else {
      const int j = p - kMlpB3;
      for (int64_t s = 0; s < n; ++s) acc += dY[s * kMlpOut + j];
    }
    grad[p] = acc;
  }
}

__global__ void k_mlp_sgd(float* __restrict__ master, uint16_t* __restrict__ wb, const float* __restrict__ grad, float lr) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < kMlpParams; p += gridDim.x * blockDim.x) {
    const float v = master[p] - lr * grad[p];
    master[p] = v;
    wb[p] = f32_to_bf16(v);
  }
}

__global__ void k_mlp_round(const float* __restrict__ master, uint16_t* __restrict__ wb) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < kMlpParams; p += gridDim.x * blockDim.x)
    wb[p] = f32_to_bf16(master[p]);
}

// single-workgroup ordered sum
__global__ __launch_bounds__(kMlpThreads) void k_sum(const float* __restrict__ x, int64_t n, float* __restrict__ out) {
  __shared__ float part[kMlpThreads];
  float a = 0.f;
  for (int64_t i = threadIdx.x; i < n; i += kMlpThreads) a += x[i];
  part[threadIdx.x] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < kMlpThreads; ++i) s += part[i];
    *out = s;
  }
}

}  // namespace csaw

using namespace csaw;

struct csaw_mlp {
  std::vector<float> master;  // host mirror of the FP32 master weights
  float* d_master = nullptr;
  uint16_t* d_w = nullptr;
  hipStream_t stream = nullptr;
  int num_cu = 0;
};

namespace csaw {
int mlp_upload_bf16(const csaw_mlp* m, uint16_t* dst) {
  CSAW_HIP(hipMemcpy(dst, m->d_w, sizeof(uint16_t) * kMlpParams, hipMemcpyDeviceToDevice));
  return CSAW_OK;
}
}  // namespace csaw

struct MlpWork {
  float *X = nullptr, *T = nullptr, *M = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr;
  float *dY = nullptr, *dH1 = nullptr, *dH2 = nullptr, *grad = nullptr, *loss_s = nullptr, *loss = nullptr;
  int32_t* idx = nullptr;
  ~MlpWork() {
    for (float* p : {X, T, M, H1, H2, Y, dY, dH1, dH2, grad, loss_s, loss}) (void)hipFree(p);
    (void)hipFree(idx);
  }
};

static int mlp_alloc_work(MlpWork& w, int64_t n_data, int64_t n_batch, bool targets) {
  const int64_t nd = std::max<int64_t>(n_data, 1), nb = std::max<int64_t>(n_batch, 1);
  CSAW_HIP(hipMalloc(&w.X, sizeof(float) * nd * kMlpIn));
  if (targets) {
    CSAW_HIP(hipMalloc(&w.T, sizeof(float) * nd * kMlpOut));
    CSAW_HIP(hipMalloc(&w.M, sizeof(float) * nd * kMlpOut));
    CSAW_HIP(hipMalloc(&w.dY, sizeof(float) * nb * kMlpOut));
    CSAW_HIP(hipMalloc(&w.dH1, sizeof(float) * nb * kMlpHid));
    CSAW_HIP(hipMalloc(&w.dH2, sizeof(float) * nb * kMlpHid));
    CSAW_HIP(hipMalloc(&w.grad, sizeof(float) * kMlpParams));
    CSAW_HIP(hipMalloc(&w.loss_s, sizeof(float) * nb));
    CSAW_HIP(hipMalloc(&w.loss, sizeof(float)));
    CSAW_HIP(hipMalloc(&w.idx, sizeof(int32_t) * nd));
  }
  CSAW_HIP(hipMalloc(&w.H1, sizeof(float) * nb * kMlpHid));
  CSAW_HIP(hipMalloc(&w.H2, sizeof(float) * nb * kMlpHid));
  CSAW_HIP(hipMalloc(&w.Y, sizeof(float) * nb * kMlpOut));
  return CSAW_OK;
}

static int grid_n(const csaw_mlp* m, int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + kMlpThreads - 1) / kMlpThreads, (int64_t)m->num_cu * 8));
}

// forward + backward over rows idx[0..n) (idx == nullptr: rows 0..n)
static int mlp_step_grad(csaw_mlp* m, MlpWork& w, const int32_t* idx, int64_t n, float inv_cnt) {
  k_mlp_fwd<<<grid_n(m, n), kMlpThreads, 0, m->stream>>>(m->d_w, w.X, idx, n, w.H1, w.H2, w.Y);
  CSAW_LAUNCH_CHECK();
  k_mlp_bwd_act<<<grid_n(m, n), kMlpThreads, 0, m->stream>>>(m->d_w, idx, n, w.T, w.M, w.Y, w.H1, w.H2, inv_cnt, w.dY,
                                                             w.dH1, w.dH2, w.loss_s);
  CSAW_LAUNCH_CHECK();
  k_mlp_bwd_w<<<(kMlpParams + kMlpThreads - 1) / kMlpThreads, kMlpThreads, 0, m->stream>>>(w.X, idx, n, w.H1, w.H2, w.dY,
                                                                                          w.dH1, w.dH2, w.grad);
  CSAW_LAUNCH_CHECK();
  k_sum<<<1, kMlpThreads, 0, m->stream>>>(w.loss_s, n, w.loss);
  CSAW_LAUNCH_CHECK();
  return CSAW_OK;
}

extern "C" int csaw_mlp_num_params(void) { return kMlpParams; }

extern "C" int csaw_mlp_create(uint64_t init_key, csaw_mlp** out) {
  if (!out) return CSAW_ERR_INVALID_ARG;
  *out = nullptr;
  csaw_mlp* m = new csaw_mlp();
  m->master.assign(kMlpParams, 0.f);
  // He-uniform initialisation from Philox (domain kDomExplore, word 2 = 1)
  auto fill = [&](int off, int rows, int cols) {
    const float lim = std::sqrt(6.0f / (float)cols);
    for (int i = 0; i < rows * cols; ++i) {
      const U4 o = philox4x32_10((uint32_t)(off + i), 0u, 1u, kDomExplore << 28, (uint32_t)init_key,
                                 (uint32_t)(init_key >> 32));
      const float u = (float)(o.x >> 8) * 5.9604644775390625e-08f;
      m->master[off + i] = (2.f * u - 1.f) * lim;
    }
  };
  fill(kMlpW1, kMlpHid, kMlpIn);
  fill(kMlpW2, kMlpHid, kMlpHid);
  fill(kMlpW3, kMlpOut, kMlpHid);
  auto body = [&]() -> int {
    CSAW_TRY(device_init(&m->num_cu));
    CSAW_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    CSAW_HIP(hipMalloc(&m->d_master, sizeof(float) * kMlpParams));
    CSAW_HIP(hipMalloc(&m->d_w, sizeof(uint16_t) * kMlpParams));
    return csaw_mlp_set_params(m, m->master.data());
  };
  int st = body();
  if (st != CSAW_OK) {
    csaw_mlp_destroy(m);
    return st;
  }
  *out = m;
  return CSAW_OK;
}

extern "C" int csaw_mlp_destroy(csaw_mlp* m) {
  if (!m) return CSAW_OK;
  (void)hipFree(m->d_master);
  (void)hipFree(m->d_w);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return CSAW_OK;
}

extern "C" int csaw_mlp_get_params(const csaw_mlp* m, float* params) {
  if (!m || !params) return CSAW_ERR_INVALID_ARG;
  CSAW_HIP(hipMemcpy(params, m->d_master, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  return CSAW_OK;
}

extern "C" int csaw_mlp_set_params(csaw_mlp* m, const float* params) {
  if (!m || !params) return CSAW_ERR_INVALID_ARG;
  if (params != m->master.data()) m->master.assign(params, params + kMlpParams);
  CSAW_HIP(hipMemcpy(m->d_master, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
  k_mlp_round<<<(kMlpParams + 255) / 256, 256, 0, m->stream>>>(m->d_master, m->d_w);
  CSAW_LAUNCH_CHECK();
  CSAW_HIP(hipStreamSynchronize(m->stream));
  return CSAW_OK;
}

extern "C" int csaw_mlp_forward(csaw_mlp* m, const float* X, int64_t n, float* Y) {
  if (!m || n < 0 || (n > 0 && (!X || !Y))) return CSAW_ERR_INVALID_ARG;
  if (n == 0) return CSAW_OK;
  MlpWork w;
  CSAW_TRY(mlp_alloc_work(w, n, n, false));
  CSAW_HIP(hipMemcpy(w.X, X, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice));
  k_mlp_fwd<<<grid_n(m, n), kMlpThreads, 0, m->stream>>>(m->d_w, w.X, nullptr, n, nullptr, nullptr, w.Y);
  CSAW_LAUNCH_CHECK();
  CSAW_HIP(hipStreamSynchronize(m->stream));
  CSAW_HIP(hipMemcpy(Y, w.Y, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToHost));
  return CSAW_OK;
}

static double mask_count(const float* M, int64_t n) {
  double c = 0;
  for (int64_t i = 0; i < n * kMlpOut; ++i) c += M[i] != 0.f ? 1.0 : 0.0;
  return c;
}

extern "C" int csaw_mlp_grad(csaw_mlp* m, const float* X, const float* T, const float* M, int64_t n, float* loss,
                             float* grad) {
  if (!m || n <= 0 || !X || !T || !M) return CSAW_ERR_INVALID_ARG;
  MlpWork w;
  CSAW_TRY(mlp_alloc_work(w, n, n, true));
  CSAW_HIP(hipMemcpy(w.X, X, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice));
  CSAW_HIP(hipMemcpy(w.T, T, sizeof(float) * n * kMlpOut, hipMemcpyHostToDevice));
  CSAW_HIP(hipMemcpy(w.M, M, sizeof(float) * n * kMlpOut, hipMemcpyHostToDevice));
  const double cnt = mask_count(M, n);
  CSAW_TRY(mlp_step_grad(m, w, nullptr, n, cnt > 0 ? (float)(1.0 / cnt) : 0.f));
  CSAW_HIP(hipStreamSynchronize(m->stream));
  if (loss) CSAW_HIP(hipMemcpy(loss, w.loss, sizeof(float), hipMemcpyDeviceToHost));
  if (grad) CSAW_HIP(hipMemcpy(grad, w.grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  return CSAW_OK;
}

extern "C" int csaw_mlp_train(csaw_mlp* m, const float* X, const float* T, const float* M, int64_t n, int32_t epochs,
                              int32_t batch, float lr, uint64_t shuffle_key, float* final_loss) {
  if (!m || n <= 0 || !X || !T || !M || epochs < 0 || batch <= 0) return CSAW_ERR_INVALID_ARG;
  MlpWork w;
  const int64_t nb = std::min<int64_t>(batch, n);
  CSAW_TRY(mlp_alloc_work(w, n, nb, true));
  CSAW_HIP(hipMemcpy(w.X, X, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice));
  CSAW_HIP(hipMemcpy(w.T, T, sizeof(float) * n * kMlpOut, hipMemcpyHostToDevice));
  CSAW_HIP(hipMemcpy(w.M, M, sizeof(float) * n * kMlpOut, hipMemcpyHostToDevice));
  std::vector<int32_t> perm(n);
  for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
  float last = 0.f;
  for (int32_t ep = 0; ep < epochs; ++ep) {
    // Fisher-Yates with Philox draws
    for (int64_t i = n - 1; i > 0; --i) {
      const U4 o = philox4x32_10((uint32_t)i, (uint32_t)ep, 2u, kDomExplore << 28, (uint32_t)shuffle_key,
                                 (uint32_t)(shuffle_key >> 32));
      const int64_t j = (int64_t)mulhi64(philox_t64(o), (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    CSAW_HIP(hipMemcpyAsync(w.idx, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, m->stream));
    for (int64_t b = 0; b < n; b += nb) {
      const int64_t cnt_rows = std::min<int64_t>(nb, n - b);
      double cnt = 0;
      for (int64_t s = 0; s < cnt_rows; ++s)
        for (int j = 0; j < kMlpOut; ++j) cnt += M[(int64_t)perm[b + s] * kMlpOut + j] != 0.f ? 1.0 : 0.0;
      if (cnt == 0) continue;
      CSAW_TRY(mlp_step_grad(m, w, w.idx + b, cnt_rows, (float)(1.0 / cnt)));
      k_mlp_sgd<<<(kMlpParams + 255) / 256, 256, 0, m->stream>>>(m->d_master, m->d_w, w.grad, lr);
      CSAW_LAUNCH_CHECK();
    }
  }
  // full-data loss
  if (final_loss) {
    const double cnt = mask_count(M, n);
    MlpWork w2;
    CSAW_TRY(mlp_alloc_work(w2, n, n, true));
    CSAW_HIP(hipMemcpyAsync(w2.X, w.X, sizeof(float) * n * kMlpIn, hipMemcpyDeviceToDevice, m->stream));
    CSAW_HIP(hipMemcpyAsync(w2.T, w.T, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToDevice, m->stream));
    CSAW_HIP(hipMemcpyAsync(w2.M, w.M, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToDevice, m->stream));
    CSAW_TRY(mlp_step_grad(m, w2, nullptr, n, cnt > 0 ? (float)(1.0 / cnt) : 0.f));
    CSAW_HIP(hipStreamSynchronize(m->stream));
    CSAW_HIP(hipMemcpy(&last, w2.loss, sizeof(float), hipMemcpyDeviceToHost));
    *final_loss = last;
  }
  CSAW_HIP(hipStreamSynchronize(m->stream));
  m->master.resize(kMlpParams);
  CSAW_HIP(hipMemcpy(m->master.data(), m->d_master, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  return CSAW_OK;
}
