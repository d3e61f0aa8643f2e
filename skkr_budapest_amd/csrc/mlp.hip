// mlp.hip -- forward, backward and Adam kernels of the error-prediction MLP
// (nin -> nhid tanh -> nhid tanh -> 1, FP32) used by the adaptive BZ mesh.
//
// Forward and the per-sample backward pass run one thread per sample with the
// parameters staged in LDS; the weight gradients are sums of outer products
// over samples, one thread per parameter summing the samples in order (so a
// gradient is the same on every run).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/skkr_mlp.h"

namespace {
constexpr int MAXIN = 16;
constexpr int NH = 64;

struct Layout {
  int nin, nh;
  int w1, b1, w2, b2, w3, b3, total;
};

__host__ __device__ inline Layout layout(int nin, int nh) {
  Layout L;
  L.nin = nin;
  L.nh = nh;
  L.w1 = 0;
  L.b1 = L.w1 + nh * nin;
  L.w2 = L.b1 + nh;
  L.b2 = L.w2 + nh * nh;
  L.w3 = L.b2 + nh;
  L.b3 = L.w3 + nh;
  L.total = L.b3 + 1;
  return L;
}
}  // namespace

#define HCHK(x)                                                                   \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      fprintf(stderr, "skkr_mlp: %s: %s\n", #x, hipGetErrorString(e_));           \
      return 2;                                                                   \
    }                                                                             \
  } while (0)

// one thread per sample; h1, h2 (n x NH) kept for the backward pass
__global__ __launch_bounds__(256) void mlp_forward_kernel(int n, int nin, const float* __restrict__ p,
                                                          const float* __restrict__ x,
                                                          float* __restrict__ h1,
                                                          float* __restrict__ h2,
                                                          float* __restrict__ y) {
  __shared__ float sp[NH * MAXIN + NH + NH * NH + NH + NH + 1];
  const Layout L = layout(nin, NH);
  for (int i = threadIdx.x; i < L.total; i += blockDim.x) sp[i] = p[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float xin[MAXIN];
  for (int j = 0; j < nin; ++j) xin[j] = x[(long long)s * nin + j];
  float a1[NH];
#pragma unroll 8
  for (int i = 0; i < NH; ++i) {
    float acc = sp[L.b1 + i];
    for (int j = 0; j < nin; ++j) acc += sp[L.w1 + i * nin + j] * xin[j];
    a1[i] = tanhf(acc);
    if (h1) h1[(long long)s * NH + i] = a1[i];
  }
  float out = sp[L.b3];
  for (int i = 0; i < NH; ++i) {
    float acc = sp[L.b2 + i];
#pragma unroll 8
    for (int j = 0; j < NH; ++j) acc += sp[L.w2 + i * NH + j] * a1[j];
    const float a2 = tanhf(acc);
    if (h2) h2[(long long)s * NH + i] = a2;
    out += sp[L.w3 + i] * a2;
  }
  y[s] = out;
}

// per-sample deltas of the loss (1/n) sum (y - t)^2
__global__ __launch_bounds__(256) void mlp_delta_kernel(int n, int nin, const float* __restrict__ p,
                                                        const float* __restrict__ h1,
                                                        const float* __restrict__ h2,
                                                        const float* __restrict__ y,
                                                        const float* __restrict__ t,
                                                        float* __restrict__ d1,
                                                        float* __restrict__ d2,
                                                        float* __restrict__ d3) {
  __shared__ float sp[NH * MAXIN + NH + NH * NH + NH + NH + 1];
  const Layout L = layout(nin, NH);
  for (int i = threadIdx.x; i < L.total; i += blockDim.x) sp[i] = p[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float g3 = 2.0f * (y[s] - t[s]) / (float)n;
  d3[s] = g3;
  float g2[NH];
  for (int i = 0; i < NH; ++i) {
    const float a2 = h2[(long long)s * NH + i];
    g2[i] = g3 * sp[L.w3 + i] * (1.0f - a2 * a2);
    d2[(long long)s * NH + i] = g2[i];
  }
  for (int j = 0; j < NH; ++j) {
    float acc = 0.0f;
#pragma unroll 8
    for (int i = 0; i < NH; ++i) acc += g2[i] * sp[L.w2 + i * NH + j];
    const float a1 = h1[(long long)s * NH + j];
    d1[(long long)s * NH + j] = acc * (1.0f - a1 * a1);
  }
}

// Weight gradients in two deterministic stages: one thread per parameter sums
// the samples of one chunk in order (grid.y = chunk), then one thread per
// parameter sums the chunk partials in order.
constexpr int GRAD_CHUNK = 512;

__global__ __launch_bounds__(256) void mlp_grad_partial_kernel(
    int n, int nin, const float* __restrict__ x, const float* __restrict__ h1,
    const float* __restrict__ h2, const float* __restrict__ d1, const float* __restrict__ d2,
    const float* __restrict__ d3, float* __restrict__ part) {
  const Layout L = layout(nin, NH);
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L.total) return;
  const int s0 = blockIdx.y * GRAD_CHUNK;
  const int s1 = min(n, s0 + GRAD_CHUNK);
  float acc = 0.0f;
  if (q < L.b1) {
    const int i = (q - L.w1) / nin, j = (q - L.w1) % nin;
    for (int s = s0; s < s1; ++s) acc += d1[(long long)s * NH + i] * x[(long long)s * nin + j];
  } else if (q < L.w2) {
    const int i = q - L.b1;
    for (int s = s0; s < s1; ++s) acc += d1[(long long)s * NH + i];
  } else if (q < L.b2) {
    const int i = (q - L.w2) / NH, j = (q - L.w2) % NH;
    for (int s = s0; s < s1; ++s) acc += d2[(long long)s * NH + i] * h1[(long long)s * NH + j];
  } else if (q < L.w3) {
    const int i = q - L.b2;
    for (int s = s0; s < s1; ++s) acc += d2[(long long)s * NH + i];
  } else if (q < L.b3) {
    const int i = q - L.w3;
    for (int s = s0; s < s1; ++s) acc += d3[s] * h2[(long long)s * NH + i];
  } else {
    for (int s = s0; s < s1; ++s) acc += d3[s];
  }
  part[(long long)blockIdx.y * L.total + q] = acc;
}

__global__ __launch_bounds__(256) void mlp_grad_reduce_kernel(int np, int nchunk,
                                                              const float* __restrict__ part,
                                                              float* __restrict__ grad) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  float acc = 0.0f;
  for (int c = 0; c < nchunk; ++c) acc += part[(long long)c * np + q];
  grad[q] = acc;
}

// (1/n) sum (y - t)^2 in one work-group, fixed order
__global__ __launch_bounds__(256) void mlp_loss_kernel(int n, const float* __restrict__ y,
                                                       const float* __restrict__ t,
                                                       float* __restrict__ loss) {
  __shared__ float red[256];
  float acc = 0.0f;
  for (int s = threadIdx.x; s < n; s += 256) {
    const float d = y[s] - t[s];
    acc += d * d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] / (float)n;
}

__global__ __launch_bounds__(256) void mlp_adam_kernel(int np, float* __restrict__ p,
                                                       const float* __restrict__ g,
                                                       float* __restrict__ m, float* __restrict__ v,
                                                       float lr, float b1t, float b2t) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  m[q] = b1 * m[q] + (1.0f - b1) * g[q];
  v[q] = b2 * v[q] + (1.0f - b2) * g[q] * g[q];
  const float mh = m[q] / (1.0f - b1t), vh = v[q] / (1.0f - b2t);
  p[q] -= lr * mh / (sqrtf(vh) + eps);
}

struct skkr_mlp {
  int nin, nh, np, device;
  float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;
  float *x = nullptr, *t = nullptr, *y = nullptr, *h1 = nullptr, *h2 = nullptr, *d1 = nullptr,
        *d2 = nullptr, *d3 = nullptr, *loss = nullptr, *part = nullptr;
  int cap = 0;
  long long steps_done = 0;
};

static int mlp_reserve(skkr_mlp* M, int n) {
  if (n <= M->cap) return 0;
  M->cap = 0;  // until every buffer of the new size exists
  float** bufs[] = {&M->x, &M->t, &M->y, &M->h1, &M->h2, &M->d1, &M->d2, &M->d3, &M->part};
  for (float** b : bufs)
    if (*b) {
      hipFree(*b);
      *b = nullptr;
    }
  const size_t N = n;
  HCHK(hipMalloc(&M->x, sizeof(float) * N * M->nin));
  HCHK(hipMalloc(&M->t, sizeof(float) * N));
  HCHK(hipMalloc(&M->y, sizeof(float) * N));
  HCHK(hipMalloc(&M->h1, sizeof(float) * N * NH));
  HCHK(hipMalloc(&M->h2, sizeof(float) * N * NH));
  HCHK(hipMalloc(&M->d1, sizeof(float) * N * NH));
  HCHK(hipMalloc(&M->d2, sizeof(float) * N * NH));
  HCHK(hipMalloc(&M->d3, sizeof(float) * N));
  const size_t nchunk = (N + GRAD_CHUNK - 1) / GRAD_CHUNK;
  HCHK(hipMalloc(&M->part, sizeof(float) * nchunk * M->np));
  M->cap = n;
  return 0;
}

static int mlp_fwd_bwd(skkr_mlp* M, int n) {
  const int blocks = (n + 255) / 256;
  mlp_forward_kernel<<<blocks, 256>>>(n, M->nin, M->p, M->x, M->h1, M->h2, M->y);
  mlp_delta_kernel<<<blocks, 256>>>(n, M->nin, M->p, M->h1, M->h2, M->y, M->t, M->d1, M->d2, M->d3);
  const int nchunk = (n + GRAD_CHUNK - 1) / GRAD_CHUNK;
  mlp_grad_partial_kernel<<<dim3((M->np + 255) / 256, nchunk), 256>>>(
      n, M->nin, M->x, M->h1, M->h2, M->d1, M->d2, M->d3, M->part);
  mlp_grad_reduce_kernel<<<(M->np + 255) / 256, 256>>>(M->np, nchunk, M->part, M->g);
  mlp_loss_kernel<<<1, 256>>>(n, M->y, M->t, M->loss);
  HCHK(hipGetLastError());
  return 0;
}

extern "C" {

int skkr_mlp_create(skkr_mlp** out, int nin, int nhid, int device, unsigned long long seed) {
  if (!out || nin < 1 || nin > MAXIN || nhid != NH) return 1;
  HCHK(hipSetDevice(device));
  skkr_mlp* M = new skkr_mlp();
  M->nin = nin;
  M->nh = nhid;
  M->device = device;
  const Layout L = layout(nin, nhid);
  M->np = L.total;
  HCHK(hipMalloc(&M->p, sizeof(float) * M->np));
  HCHK(hipMalloc(&M->m, sizeof(float) * M->np));
  HCHK(hipMalloc(&M->v, sizeof(float) * M->np));
  HCHK(hipMalloc(&M->g, sizeof(float) * M->np));
  HCHK(hipMalloc(&M->loss, sizeof(float)));
  // Glorot-uniform initial weights, zero biases
  std::mt19937_64 rng(seed);
  std::vector<float> h(M->np, 0.0f);
  auto fill = [&](int off, int fan_in, int fan_out, int count) {
    const float a = std::sqrt(6.0f / (fan_in + fan_out));
    std::uniform_real_distribution<float> U(-a, a);
    for (int i = 0; i < count; ++i) h[off + i] = U(rng);
  };
  fill(L.w1, nin, NH, NH * nin);
  fill(L.w2, NH, NH, NH * NH);
  fill(L.w3, NH, 1, NH);
  HCHK(hipMemcpy(M->p, h.data(), sizeof(float) * M->np, hipMemcpyHostToDevice));
  HCHK(hipMemset(M->m, 0, sizeof(float) * M->np));
  HCHK(hipMemset(M->v, 0, sizeof(float) * M->np));
  *out = M;
  return 0;
}

void skkr_mlp_destroy(skkr_mlp* M) {
  if (!M) return;
  hipSetDevice(M->device);
  float* bufs[] = {M->p, M->m, M->v, M->g, M->loss, M->x, M->t, M->y,
                   M->h1, M->h2, M->d1, M->d2, M->d3, M->part};
  for (float* b : bufs)
    if (b) hipFree(b);
  delete M;
}

int skkr_mlp_nparams(const skkr_mlp* M) { return M ? M->np : 0; }

int skkr_mlp_get_params(skkr_mlp* M, float* p) {
  if (!M || !p) return 1;
  HCHK(hipMemcpy(p, M->p, sizeof(float) * M->np, hipMemcpyDeviceToHost));
  return 0;
}

int skkr_mlp_set_params(skkr_mlp* M, const float* p) {
  if (!M || !p) return 1;
  HCHK(hipMemcpy(M->p, p, sizeof(float) * M->np, hipMemcpyHostToDevice));
  HCHK(hipMemset(M->m, 0, sizeof(float) * M->np));
  HCHK(hipMemset(M->v, 0, sizeof(float) * M->np));
  M->steps_done = 0;
  return 0;
}

int skkr_mlp_predict(skkr_mlp* M, int n, const float* x, float* y) {
  if (!M || n < 0 || (n > 0 && (!x || !y))) return 1;
  if (n == 0) return 0;
  HCHK(hipSetDevice(M->device));
  if (mlp_reserve(M, n)) return 2;
  HCHK(hipMemcpy(M->x, x, sizeof(float) * n * M->nin, hipMemcpyHostToDevice));
  mlp_forward_kernel<<<(n + 255) / 256, 256>>>(n, M->nin, M->p, M->x, nullptr, nullptr, M->y);
  HCHK(hipGetLastError());
  HCHK(hipMemcpy(y, M->y, sizeof(float) * n, hipMemcpyDeviceToHost));
  return 0;
}

int skkr_mlp_loss_grad(skkr_mlp* M, int n, const float* x, const float* t, float* loss,
                       float* grad) {
  if (!M || n <= 0 || !x || !t) return 1;
  HCHK(hipSetDevice(M->device));
  if (mlp_reserve(M, n)) return 2;
  HCHK(hipMemcpy(M->x, x, sizeof(float) * n * M->nin, hipMemcpyHostToDevice));
  HCHK(hipMemcpy(M->t, t, sizeof(float) * n, hipMemcpyHostToDevice));
  if (mlp_fwd_bwd(M, n)) return 2;
  if (loss) HCHK(hipMemcpy(loss, M->loss, sizeof(float), hipMemcpyDeviceToHost));
  if (grad) HCHK(hipMemcpy(grad, M->g, sizeof(float) * M->np, hipMemcpyDeviceToHost));
  return 0;
}

int skkr_mlp_train(skkr_mlp* M, int n, const float* x, const float* t, int steps, float lr,
                   float* loss_history) {
  if (!M || n <= 0 || !x || !t || steps < 0) return 1;
  HCHK(hipSetDevice(M->device));
  if (mlp_reserve(M, n)) return 2;
  HCHK(hipMemcpy(M->x, x, sizeof(float) * n * M->nin, hipMemcpyHostToDevice));
  HCHK(hipMemcpy(M->t, t, sizeof(float) * n, hipMemcpyHostToDevice));
  for (int it = 0; it < steps; ++it) {
    if (mlp_fwd_bwd(M, n)) return 2;
    M->steps_done++;
    const float b1t = std::pow(0.9f, (float)M->steps_done);
    const float b2t = std::pow(0.999f, (float)M->steps_done);
    mlp_adam_kernel<<<(M->np + 255) / 256, 256>>>(M->np, M->p, M->g, M->m, M->v, lr, b1t, b2t);
    HCHK(hipGetLastError());
    if (loss_history) HCHK(hipMemcpy(&loss_history[it], M->loss, sizeof(float), hipMemcpyDeviceToHost));
  }
  HCHK(hipDeviceSynchronize());
  return 0;
}

}  // extern "C"
