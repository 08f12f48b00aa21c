// hvm_mlp.hip — training kernels of the scheduler MLP (32 -> 64 -> 64 -> 4).
//
// The network maps the features of a finished turn to a predicted reward
// (log2 interactions per µs) for each of the four turn lengths; the learned
// scheduler in hvm_between_kernel takes the argmax.  Parameters are stored in
// fp16 and every product is accumulated in fp32.  All reductions are
// fixed-order (one thread per parameter or one block), so a gradient is
// bit-reproducible run to run.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../../include/hvm_hip.h"
#include "hvm_dev.hpp"

namespace {

typedef uint32_t u32;

constexpr u32 F = HVM_FEAT, H = HVM_HID, K = HVM_ACTIONS;
constexpr u32 OW1 = 0, OB1 = OW1 + F * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H * K;
constexpr u32 NP = OB3 + K;
static_assert(NP == HVM_MLP_PARAMS, "parameter packing");
static_assert(NP == HVM_HIP_MLP_PARAMS, "parameter packing (C ABI)");

__device__ __forceinline__ float ph(const __half* p, u32 i) { return __half2float(p[i]); }

// One block (64 threads) per sample.
__global__ __launch_bounds__(64) void mlp_fwd_kernel(const __half* P, const float* X, u32 n, float* H1,
                                                      float* H2, float* Y) {
  __shared__ float x[F], h1[H], h2[H];
  const u32 b = blockIdx.x, t = threadIdx.x;
  if (b >= n) return;
  if (t < F) x[t] = X[(size_t)b * F + t];
  __syncthreads();
  float acc = ph(P, OB1 + t);
  for (u32 i = 0; i < F; ++i) acc += x[i] * ph(P, OW1 + i * H + t);
  h1[t] = fmaxf(acc, 0.0f);
  H1[(size_t)b * H + t] = h1[t];
  __syncthreads();
  acc = ph(P, OB2 + t);
  for (u32 i = 0; i < H; ++i) acc += h1[i] * ph(P, OW2 + i * H + t);
  h2[t] = fmaxf(acc, 0.0f);
  H2[(size_t)b * H + t] = h2[t];
  __syncthreads();
  if (t < K) {
    acc = ph(P, OB3 + t);
    for (u32 i = 0; i < H; ++i) acc += h2[i] * ph(P, OW3 + i * K + t);
    Y[(size_t)b * K + t] = acc;
  }
}

// Per-sample deltas of the loss 0.5 * (y[a] - r)^2.
__global__ __launch_bounds__(64) void mlp_bwd_kernel(const __half* P, const float* H1, const float* H2,
                                                      const float* Y, const u32* A, const float* R, u32 n,
                                                      float* dY, float* dH2, float* dH1) {
  __shared__ float dy[K], dh2[H];
  const u32 b = blockIdx.x, t = threadIdx.x;
  if (b >= n) return;
  if (t < K) {
    float d = (t == A[b]) ? (Y[(size_t)b * K + t] - R[b]) : 0.0f;
    dy[t] = d;
    dY[(size_t)b * K + t] = d;
  }
  __syncthreads();
  float acc = 0.0f;
  for (u32 k = 0; k < K; ++k) acc += ph(P, OW3 + t * K + k) * dy[k];
  float g2 = H2[(size_t)b * H + t] > 0.0f ? acc : 0.0f;
  dh2[t] = g2;
  dH2[(size_t)b * H + t] = g2;
  __syncthreads();
  acc = 0.0f;
  for (u32 j = 0; j < H; ++j) acc += ph(P, OW2 + t * H + j) * dh2[j];
  dH1[(size_t)b * H + t] = H1[(size_t)b * H + t] > 0.0f ? acc : 0.0f;
}

// One thread per parameter: mean over the batch, in sample order.
__global__ void mlp_wgrad_kernel(const float* X, const float* H1, const float* H2, const float* dY,
                                 const float* dH2, const float* dH1, u32 n, float* G) {
  const u32 p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NP) return;
  float acc = 0.0f;
  if (p < OB1) {
    u32 i = (p - OW1) / H, j = (p - OW1) % H;
    for (u32 b = 0; b < n; ++b) acc += X[(size_t)b * F + i] * dH1[(size_t)b * H + j];
  } else if (p < OW2) {
    u32 j = p - OB1;
    for (u32 b = 0; b < n; ++b) acc += dH1[(size_t)b * H + j];
  } else if (p < OB2) {
    u32 i = (p - OW2) / H, j = (p - OW2) % H;
    for (u32 b = 0; b < n; ++b) acc += H1[(size_t)b * H + i] * dH2[(size_t)b * H + j];
  } else if (p < OW3) {
    u32 j = p - OB2;
    for (u32 b = 0; b < n; ++b) acc += dH2[(size_t)b * H + j];
  } else if (p < OB3) {
    u32 i = (p - OW3) / K, k = (p - OW3) % K;
    for (u32 b = 0; b < n; ++b) acc += H2[(size_t)b * H + i] * dY[(size_t)b * K + k];
  } else {
    u32 k = p - OB3;
    for (u32 b = 0; b < n; ++b) acc += dY[(size_t)b * K + k];
  }
  G[p] = acc / (float)n;
}

// Mean loss, one block of 256 threads with a fixed reduction tree.
__global__ __launch_bounds__(256) void mlp_loss_kernel(const float* dY, u32 n, float* loss) {
  __shared__ float s[256];
  float acc = 0.0f;
  for (u32 b = threadIdx.x; b < n; b += 256) {
    for (u32 k = 0; k < K; ++k) {
      float d = dY[(size_t)b * K + k];
      acc += 0.5f * d * d;
    }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (u32 w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = s[0] / (float)n;
}

__global__ void mlp_adam_kernel(float* Pm, const float* G, float* M, float* V, float lr, float b1t, float b2t,
                                __half* P16) {
  const u32 p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NP) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float g = G[p];
  float m = b1 * M[p] + (1.0f - b1) * g;
  float v = b2 * V[p] + (1.0f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  float w = Pm[p] - lr * (m / (1.0f - b1t)) / (sqrtf(v / (1.0f - b2t)) + eps);
  Pm[p] = w;
  P16[p] = __float2half(w);
}

__global__ void mlp_to_half_kernel(const float* Pm, __half* P16) {
  const u32 p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NP) P16[p] = __float2half(Pm[p]);
}

// Device scratch for one call; freed on destruction.
struct Scratch {
  std::vector<void*> ptrs;
  bool ok = true;
  template <typename T>
  T* get(size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, count * sizeof(T) + 16) != hipSuccess) { ok = false; return nullptr; }
    ptrs.push_back(p);
    return (T*)p;
  }
  ~Scratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

struct Batch {
  float *X, *H1, *H2, *Y, *dY, *dH2, *dH1, *R, *G, *loss;
  u32* A;
};

bool alloc_batch(Scratch& s, u32 n, Batch& B) {
  B.X = s.get<float>((size_t)n * F);
  B.H1 = s.get<float>((size_t)n * H);
  B.H2 = s.get<float>((size_t)n * H);
  B.Y = s.get<float>((size_t)n * K);
  B.dY = s.get<float>((size_t)n * K);
  B.dH2 = s.get<float>((size_t)n * H);
  B.dH1 = s.get<float>((size_t)n * H);
  B.R = s.get<float>(n);
  B.A = s.get<u32>(n);
  B.G = s.get<float>(NP);
  B.loss = s.get<float>(1);
  return s.ok;
}

void grad_step(const __half* P, const Batch& B, u32 n) {
  hipLaunchKernelGGL(mlp_fwd_kernel, dim3(n), dim3(64), 0, 0, P, B.X, n, B.H1, B.H2, B.Y);
  hipLaunchKernelGGL(mlp_bwd_kernel, dim3(n), dim3(64), 0, 0, P, B.H1, B.H2, B.Y, B.A, B.R, n, B.dY, B.dH2, B.dH1);
  hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((NP + 255) / 256), dim3(256), 0, 0, B.X, B.H1, B.H2, B.dY, B.dH2,
                     B.dH1, n, B.G);
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, 0, B.dY, n, B.loss);
}

bool has_device() {
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

}  // namespace

extern "C" int hvm_mlp_forward(const uint16_t* params, const float* x, uint32_t n, float* y) {
  if (!params || !x || !y || n == 0) return HVM_HIP_E_ARG;
  if (!has_device()) return HVM_HIP_E_DEVICE;
  Scratch s;
  Batch B;
  __half* P = s.get<__half>(NP);
  if (!alloc_batch(s, n, B) || !P) return HVM_HIP_E_DEVICE;
  if (hipMemcpy(P, params, NP * 2, hipMemcpyHostToDevice) != hipSuccess) return HVM_HIP_E_DEVICE;
  if (hipMemcpy(B.X, x, (size_t)n * F * 4, hipMemcpyHostToDevice) != hipSuccess) return HVM_HIP_E_DEVICE;
  hipLaunchKernelGGL(mlp_fwd_kernel, dim3(n), dim3(64), 0, 0, P, B.X, n, B.H1, B.H2, B.Y);
  if (hipMemcpy(y, B.Y, (size_t)n * K * 4, hipMemcpyDeviceToHost) != hipSuccess) return HVM_HIP_E_DEVICE;
  return hipDeviceSynchronize() == hipSuccess ? HVM_HIP_OK : HVM_HIP_E_DEVICE;
}

extern "C" int hvm_mlp_grad(const uint16_t* params, const float* x, const uint32_t* actions, const float* rewards,
                            uint32_t n, float* grad, float* loss) {
  if (!params || !x || !actions || !rewards || !grad || n == 0) return HVM_HIP_E_ARG;
  for (u32 b = 0; b < n; ++b) if (actions[b] >= K) return HVM_HIP_E_ARG;
  if (!has_device()) return HVM_HIP_E_DEVICE;
  Scratch s;
  Batch B;
  __half* P = s.get<__half>(NP);
  if (!alloc_batch(s, n, B) || !P) return HVM_HIP_E_DEVICE;
  bool ok = hipMemcpy(P, params, NP * 2, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.X, x, (size_t)n * F * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.A, actions, (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.R, rewards, (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) return HVM_HIP_E_DEVICE;
  grad_step(P, B, n);
  ok = hipMemcpy(grad, B.G, NP * 4, hipMemcpyDeviceToHost) == hipSuccess;
  if (loss) ok = ok && hipMemcpy(loss, B.loss, 4, hipMemcpyDeviceToHost) == hipSuccess;
  return ok && hipDeviceSynchronize() == hipSuccess ? HVM_HIP_OK : HVM_HIP_E_DEVICE;
}

extern "C" int hvm_mlp_train(const float* params_in, const float* x, const uint32_t* actions, const float* rewards,
                             uint32_t n, uint32_t epochs, float lr, uint16_t* params_out, float* params_fp32_out,
                             float* loss) {
  if (!params_in || !x || !actions || !rewards || !params_out || n == 0) return HVM_HIP_E_ARG;
  for (u32 b = 0; b < n; ++b) if (actions[b] >= K) return HVM_HIP_E_ARG;
  if (!has_device()) return HVM_HIP_E_DEVICE;
  Scratch s;
  Batch B;
  float* Pm = s.get<float>(NP);
  float* M = s.get<float>(NP);
  float* V = s.get<float>(NP);
  __half* P = s.get<__half>(NP);
  if (!alloc_batch(s, n, B) || !Pm || !M || !V || !P) return HVM_HIP_E_DEVICE;
  bool ok = hipMemcpy(Pm, params_in, NP * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemset(M, 0, NP * 4) == hipSuccess && hipMemset(V, 0, NP * 4) == hipSuccess &&
            hipMemcpy(B.X, x, (size_t)n * F * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.A, actions, (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(B.R, rewards, (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) return HVM_HIP_E_DEVICE;
  hipLaunchKernelGGL(mlp_to_half_kernel, dim3((NP + 255) / 256), dim3(256), 0, 0, Pm, P);
  double b1t = 1.0, b2t = 1.0;
  for (u32 e = 0; e < epochs; ++e) {
    grad_step(P, B, n);
    b1t *= 0.9;
    b2t *= 0.999;
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((NP + 255) / 256), dim3(256), 0, 0, Pm, B.G, M, V, lr,
                       (float)b1t, (float)b2t, P);
  }
  grad_step(P, B, n);  // final loss at the trained parameters
  ok = hipMemcpy(params_out, P, NP * 2, hipMemcpyDeviceToHost) == hipSuccess;
  if (params_fp32_out) ok = ok && hipMemcpy(params_fp32_out, Pm, NP * 4, hipMemcpyDeviceToHost) == hipSuccess;
  if (loss) ok = ok && hipMemcpy(loss, B.loss, 4, hipMemcpyDeviceToHost) == hipSuccess;
  return ok && hipDeviceSynchronize() == hipSuccess ? HVM_HIP_OK : HVM_HIP_E_DEVICE;
}
