// mlp.hip -- forward, backward and Adam kernels of the warm-start network
// (include/couplb_mlp.h) on the gfx950 matrix cores.
//
// One block of 256 lanes (four waves) takes 64 samples.  Every product is a
// set of 32x32 output tiles of v_mfma_f32_32x32x16_bf16, one tile per wave:
// lane l feeds A[l%32][8(l/32) .. 8(l/32)+7] and B[8(l/32) .. +7][l%32] of a
// 16-deep K block and receives D[8(i/4) + 4(l/32) + i%4][l%32] in element i
// (layout checked on the device before this code was written).  Operands sit
// in LDS with K contiguous, so each lane loads its eight BF16 values with one
// 16-byte read.  Gradients are per-block partials summed over blocks in a
// fixed order: training is deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/couplb_hip.h"
#include "../../include/couplb_mlp.h"
#include "mlp.hpp"

namespace couplb {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ inline __bf16 bf(float f) { return (__bf16)f; }  // round to nearest even
__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// one 32x32 tile: rows of A (lda) and rows of B^T (ldb), K = 16*KB
template <int KB>
__device__ inline f32x16 tile_mma(const __bf16* A, int lda, const __bf16* Bt, int ldb) {
  const int l = threadIdx.x & 63;
  const int r = l & 31, ko = (l >> 5) * 8;
  f32x16 acc;
#pragma unroll
  for (int i = 0; i < 16; i++) acc[i] = 0.0f;
#pragma unroll
  for (int kb = 0; kb < KB; kb++) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(A + r * lda + kb * 16 + ko);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(Bt + r * ldb + kb * 16 + ko);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  return acc;
}
__device__ inline int tile_row(int i) { return 8 * (i >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (i & 3); }
__device__ inline int tile_col() { return threadIdx.x & 31; }

// sum over the 32 rows of one tile column (the same column in lanes l and
// l^32): every lane ends with its column's sum, in a fixed order
__device__ inline float tile_colsum(const float v[16]) {
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; i++) s += v[i];
  const float o = __shfl_xor(s, 32);
  return ((threadIdx.x & 63) < 32) ? s + o : o + s;
}

// Forward pass of the 64 samples of this block; rows come from idx when it
// is given.  Saves pre-activations z1, z2 (per batch position) when asked.
__global__ void __launch_bounds__(256) k_forward(int n, const float* X, const int* idx,
                                                 const float* P, float* Y, float* Z1g, float* Z2g) {
  __shared__ __attribute__((aligned(16))) __bf16 Xs[TILE * INP], W1t[H * INP], W2t[H * H],
      W3t[OUTP * H], A1s[TILE * H], A2s[TILE * H];
  __shared__ float b1s[H], b2s[H], b3s[OUTP];
  const int tid = threadIdx.x, s0 = blockIdx.x * TILE;
  for (int e = tid; e < H * INP; e += 256) {
    const int c = e / INP, k = e % INP;
    W1t[e] = bf(k < IN ? P[OFF_W1 + k * H + c] : 0.0f);
  }
  for (int e = tid; e < H * H; e += 256) {
    const int c = e / H, k = e % H;
    W2t[e] = bf(P[OFF_W2 + k * H + c]);
  }
  for (int e = tid; e < OUTP * H; e += 256) {
    const int o = e / H, k = e % H;
    W3t[e] = bf(o < OUT ? P[OFF_W3 + k * OUT + o] : 0.0f);
  }
  if (tid < H) { b1s[tid] = P[OFF_B1 + tid]; b2s[tid] = P[OFF_B2 + tid]; }
  if (tid < OUTP) b3s[tid] = tid < OUT ? P[OFF_B3 + tid] : 0.0f;
  for (int e = tid; e < TILE * INP; e += 256) {
    const int j = e / INP, k = e % INP, s = s0 + j;
    float v = 0.0f;
    if (s < n && k < IN) v = X[(size_t)(idx ? idx[s] : s) * IN + k];
    Xs[e] = bf(v);
  }
  __syncthreads();
  const int w = tid >> 6, rt = w >> 1, ct = w & 1;
  {
    const f32x16 acc = tile_mma<INP / 16>(Xs + 32 * rt * INP, INP, W1t + 32 * ct * INP, INP);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int m = 32 * rt + tile_row(i), c = 32 * ct + tile_col();
      const float z = acc[i] + b1s[c];
      A1s[m * H + c] = bf(gelu(z));
      if (Z1g && s0 + m < n) Z1g[(size_t)(s0 + m) * H + c] = z;
    }
  }
  __syncthreads();
  {
    const f32x16 acc = tile_mma<H / 16>(A1s + 32 * rt * H, H, W2t + 32 * ct * H, H);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int m = 32 * rt + tile_row(i), c = 32 * ct + tile_col();
      const float z = acc[i] + b2s[c];
      A2s[m * H + c] = bf(gelu(z));
      if (Z2g && s0 + m < n) Z2g[(size_t)(s0 + m) * H + c] = z;
    }
  }
  __syncthreads();
  if (w < 2) {
    const f32x16 acc = tile_mma<H / 16>(A2s + 32 * w * H, H, W3t, H);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int m = 32 * w + tile_row(i), o = tile_col(), s = s0 + m;
      if (o < OUT && s < n) Y[(size_t)s * OUT + o] = acc[i] + b3s[o];
    }
  }
}

// Backward pass of the 64 samples of this block; writes the block's
// gradient partial (NPARAM values) and loss partial to part[blk][NPARAM+1].
__global__ void __launch_bounds__(256) k_backward(int n, float inv_n, const float* X, const float* T,
                                                  const int* idx, const float* P, const float* Y,
                                                  const float* Z1g, const float* Z2g, float* part) {
  __shared__ __attribute__((aligned(16))) __bf16 XT[96 * TILE], A1T[H * TILE], A2T[H * TILE],
      dY[TILE * OUTP], dYT[OUTP * TILE], W3n[H * OUTP], W2s[H * H], dZ2[TILE * H], dZ2T[H * TILE],
      dZ1T[H * TILE];
  __shared__ float red[4][H], lsum[256];
  const int tid = threadIdx.x, s0 = blockIdx.x * TILE;
  float* out = part + (size_t)blockIdx.x * (NPARAM + 1);
  for (int e = tid; e < 96 * TILE; e += 256) {
    const int k = e / TILE, j = e % TILE, s = s0 + j;
    float v = 0.0f;
    if (s < n && k < IN) v = X[(size_t)(idx ? idx[s] : s) * IN + k];
    XT[e] = bf(v);
  }
  for (int e = tid; e < H * TILE; e += 256) {
    const int h = e / TILE, j = e % TILE, s = s0 + j;
    A1T[e] = bf(s < n ? gelu(Z1g[(size_t)s * H + h]) : 0.0f);
    A2T[e] = bf(s < n ? gelu(Z2g[(size_t)s * H + h]) : 0.0f);
  }
  float lp = 0.0f;
  for (int e = tid; e < TILE * OUTP; e += 256) {
    const int j = e / OUTP, o = e % OUTP, s = s0 + j;
    float g = 0.0f;
    if (s < n && o < OUT) {
      const float d = Y[(size_t)s * OUT + o] - T[(size_t)(idx ? idx[s] : s) * OUT + o];
      g = 2.0f * d * inv_n;
      lp += d * d;
    }
    dY[j * OUTP + o] = bf(g);
    dYT[o * TILE + j] = bf(g);
  }
  for (int e = tid; e < H * OUTP; e += 256) {
    const int h = e / OUTP, o = e % OUTP;
    W3n[e] = bf(o < OUT ? P[OFF_W3 + h * OUT + o] : 0.0f);
  }
  for (int e = tid; e < H * H; e += 256) W2s[e] = bf(P[OFF_W2 + e]);
  lsum[tid] = lp;
  __syncthreads();
  // loss and db3 partials (thread order, fixed)
  if (tid == 0) {
    float a = 0.0f;
    for (int t = 0; t < 256; t++) a += lsum[t];
    out[NPARAM] = a * inv_n;
  }
  if (tid < OUT) {
    float a = 0.0f;
    for (int j = 0; j < TILE; j++) {
      const int s = s0 + j;
      if (s < n) a += 2.0f * (Y[(size_t)s * OUT + tid] - T[(size_t)(idx ? idx[s] : s) * OUT + tid]) * inv_n;
    }
    out[OFF_B3 + tid] = a;
  }
  const int w = tid >> 6, rt = w >> 1, ct = w & 1;
  if (w < 2) {  // dW3 = A2^T dY
    const f32x16 acc = tile_mma<TILE / 16>(A2T + 32 * w * TILE, TILE, dYT, TILE);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int h = 32 * w + tile_row(i), o = tile_col();
      if (o < OUT) out[OFF_W3 + h * OUT + o] = acc[i];
    }
  }
  {  // dZ2 = (dY W3^T) * gelu'(z2)
    const f32x16 acc = tile_mma<OUTP / 16>(dY + 32 * rt * OUTP, OUTP, W3n + 32 * ct * OUTP, OUTP);
    float g[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int m = 32 * rt + tile_row(i), c = 32 * ct + tile_col(), s = s0 + m;
      g[i] = s < n ? acc[i] * gelu_d(Z2g[(size_t)s * H + c]) : 0.0f;
      dZ2[m * H + c] = bf(g[i]);
      dZ2T[c * TILE + m] = bf(g[i]);
    }
    const float cs = tile_colsum(g);
    if ((tid & 63) < 32) red[rt][32 * ct + tile_col()] = cs;
  }
  __syncthreads();
  if (tid < H) out[OFF_B2 + tid] = red[0][tid] + red[1][tid];
  {  // dW2 = A1^T dZ2
    const f32x16 acc = tile_mma<TILE / 16>(A1T + 32 * rt * TILE, TILE, dZ2T + 32 * ct * TILE, TILE);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int h1 = 32 * rt + tile_row(i), h2 = 32 * ct + tile_col();
      out[OFF_W2 + h1 * H + h2] = acc[i];
    }
  }
  {  // dZ1 = (dZ2 W2^T) * gelu'(z1)
    const f32x16 acc = tile_mma<H / 16>(dZ2 + 32 * rt * H, H, W2s + 32 * ct * H, H);
    float g[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int m = 32 * rt + tile_row(i), c = 32 * ct + tile_col(), s = s0 + m;
      g[i] = s < n ? acc[i] * gelu_d(Z1g[(size_t)s * H + c]) : 0.0f;
      dZ1T[c * TILE + m] = bf(g[i]);
    }
    const float cs = tile_colsum(g);
    if ((tid & 63) < 32) red[2 + rt][32 * ct + tile_col()] = cs;
  }
  __syncthreads();
  if (tid < H) out[OFF_B1 + tid] = red[2][tid] + red[3][tid];
  // dW1 = X^T dZ1: 3 x 2 tiles of (input, hidden)
  for (int t = w; t < 6; t += 4) {
    const int kt = t >> 1, ht = t & 1;
    const f32x16 acc = tile_mma<TILE / 16>(XT + 32 * kt * TILE, TILE, dZ1T + 32 * ht * TILE, TILE);
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int k = 32 * kt + tile_row(i), h = 32 * ht + tile_col();
      if (k < IN) out[OFF_W1 + k * H + h] = acc[i];
    }
  }
}

// sum the block partials (fixed order), then one Adam step if asked
__global__ void __launch_bounds__(256) k_adam(int nblk, const float* part, float* P, float* m,
                                              float* v, float* grad, double* loss, int update,
                                              float lr, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > NPARAM) return;
  double g = 0.0;
  for (int b = 0; b < nblk; b++) g += (double)part[(size_t)b * (NPARAM + 1) + i];
  if (i == NPARAM) {
    *loss = g;
    return;
  }
  const float gf = (float)g;
  if (grad) grad[i] = gf;
  if (!update) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float mi = b1 * m[i] + (1.0f - b1) * gf;
  const float vi = b2 * v[i] + (1.0f - b2) * gf * gf;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

void forward(int n, const float* X, const float* params, float* Y, hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_forward, dim3((n + TILE - 1) / TILE), dim3(256), 0, stream, n, X,
                     (const int*)nullptr, params, Y, (float*)nullptr, (float*)nullptr);
}

}  // namespace mlp
}  // namespace couplb

using namespace couplb::mlp;

namespace {

int mfail(couplb_mlp* m, int code, const std::string& msg) {
  if (m) m->err = msg;
  return code;
}
#define MCHK(net, call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return mfail(net, COUPLB_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int ensure_scratch(couplb_mlp* m, int batch) {
  if (batch <= m->cap) return COUPLB_OK;
  MCHK(m, hipStreamSynchronize(m->stream));
  (void)hipFree(m->z1); (void)hipFree(m->z2); (void)hipFree(m->y); (void)hipFree(m->part);
  m->z1 = m->z2 = m->y = m->part = nullptr;
  m->cap = 0;
  const int nblk = (batch + TILE - 1) / TILE;
  MCHK(m, hipMalloc(&m->z1, sizeof(float) * (size_t)batch * H));
  MCHK(m, hipMalloc(&m->z2, sizeof(float) * (size_t)batch * H));
  MCHK(m, hipMalloc(&m->y, sizeof(float) * (size_t)batch * OUT));
  MCHK(m, hipMalloc(&m->part, sizeof(float) * (size_t)nblk * (NPARAM + 1)));
  m->cap = batch;
  return COUPLB_OK;
}

// one mini-batch: forward with saves, backward, reduction (+ Adam step)
int batch_step(couplb_mlp* m, int n, const float* X, const float* T, const int* idx, int update,
               float lr, float* grad) {
  int rc = ensure_scratch(m, n);
  if (rc) return rc;
  const int nblk = (n + TILE - 1) / TILE;
  hipLaunchKernelGGL(k_forward, dim3(nblk), dim3(256), 0, m->stream, n, X, idx, (const float*)m->params,
                     m->y, m->z1, m->z2);
  hipLaunchKernelGGL(k_backward, dim3(nblk), dim3(256), 0, m->stream, n, 1.0f / (float)n, X, T, idx,
                     (const float*)m->params, (const float*)m->y, (const float*)m->z1,
                     (const float*)m->z2, m->part);
  float c1 = 1.0f, c2 = 1.0f;
  if (update) {
    m->step += 1;
    c1 = 1.0f - std::pow(0.9f, (float)m->step);
    c2 = 1.0f - std::pow(0.999f, (float)m->step);
  }
  hipLaunchKernelGGL(k_adam, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, m->stream, nblk,
                     (const float*)m->part, m->params, m->adam_m, m->adam_v, grad, m->loss, update, lr, c1,
                     c2);
  MCHK(m, hipGetLastError());
  return COUPLB_OK;
}

}  // namespace

extern "C" {

int couplb_mlp_create(int device, unsigned long long seed, couplb_mlp** out) {
  if (!out) return COUPLB_EINVAL;
  auto* m = new couplb_mlp();
  *out = m;
  m->device = device;
  MCHK(m, hipSetDevice(device));
  MCHK(m, hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  MCHK(m, hipMalloc(&m->params, sizeof(float) * NPARAM));
  MCHK(m, hipMalloc(&m->adam_m, sizeof(float) * NPARAM));
  MCHK(m, hipMalloc(&m->adam_v, sizeof(float) * NPARAM));
  MCHK(m, hipMalloc(&m->grad, sizeof(float) * NPARAM));
  MCHK(m, hipMalloc(&m->loss, sizeof(double)));
  std::mt19937_64 rng(seed);
  std::vector<float> p(NPARAM, 0.0f);
  auto fill = [&](int off, int fan_in, int count) {
    std::normal_distribution<double> nd(0.0, 1.0 / std::sqrt((double)fan_in));
    for (int i = 0; i < count; i++) p[off + i] = (float)nd(rng);
  };
  fill(OFF_W1, IN, IN * H);
  fill(OFF_W2, H, H * H);
  fill(OFF_W3, H, H * OUT);
  return couplb_mlp_set_params(m, p.data());
}

int couplb_mlp_destroy(couplb_mlp* m) {
  if (!m) return COUPLB_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  (void)hipFree(m->params); (void)hipFree(m->adam_m); (void)hipFree(m->adam_v); (void)hipFree(m->grad);
  (void)hipFree(m->loss); (void)hipFree(m->z1); (void)hipFree(m->z2); (void)hipFree(m->y);
  (void)hipFree(m->part); (void)hipFree(m->idx);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return COUPLB_OK;
}

const char* couplb_mlp_last_error(const couplb_mlp* m) { return m ? m->err.c_str() : "null network"; }

int couplb_mlp_get_params(couplb_mlp* m, float* params) {
  if (!m || !params) return COUPLB_EINVAL;
  MCHK(m, hipMemcpyAsync(params, m->params, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return COUPLB_OK;
}

int couplb_mlp_set_params(couplb_mlp* m, const float* params) {
  if (!m || !params) return COUPLB_EINVAL;
  MCHK(m, hipMemcpyAsync(m->params, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemsetAsync(m->adam_m, 0, sizeof(float) * NPARAM, m->stream));
  MCHK(m, hipMemsetAsync(m->adam_v, 0, sizeof(float) * NPARAM, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  m->step = 0;
  return COUPLB_OK;
}

int couplb_mlp_forward_device(couplb_mlp* m, int n, const float* X, float* Y) {
  if (!m || n < 0 || (n > 0 && (!X || !Y))) return COUPLB_EINVAL;
  forward(n, X, m->params, Y, m->stream);
  MCHK(m, hipGetLastError());
  MCHK(m, hipStreamSynchronize(m->stream));
  return COUPLB_OK;
}

int couplb_mlp_gradient_device(couplb_mlp* m, int n, const float* X, const float* T, float* grad,
                               double* loss) {
  if (!m || n <= 0 || !X || !T || !grad || !loss) return COUPLB_EINVAL;
  int rc = batch_step(m, n, X, T, nullptr, 0, 0.0f, m->grad);
  if (rc) return rc;
  MCHK(m, hipMemcpyAsync(grad, m->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipMemcpyAsync(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return COUPLB_OK;
}

int couplb_mlp_train_device(couplb_mlp* m, int n, const float* X, const float* T, int batch,
                            int epochs, float lr, unsigned long long seed, double* loss) {
  if (!m || n <= 0 || !X || !T || batch <= 0 || epochs < 0 || !(lr > 0)) return COUPLB_EINVAL;
  batch = std::min(batch, n);
  if (m->idx_cap < n) {
    MCHK(m, hipStreamSynchronize(m->stream));
    (void)hipFree(m->idx);
    m->idx = nullptr;
    MCHK(m, hipMalloc(&m->idx, sizeof(int) * (size_t)n));
    m->idx_cap = n;
  }
  std::vector<int> perm(n);
  std::mt19937_64 rng(seed);
  double last = 0.0;
  for (int ep = 0; ep < epochs; ep++) {
    for (int i = 0; i < n; i++) perm[i] = i;
    std::shuffle(perm.begin(), perm.end(), rng);
    MCHK(m, hipStreamSynchronize(m->stream));
    MCHK(m, hipMemcpy(m->idx, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
    double sum = 0.0;
    int nb = 0;
    for (int b0 = 0; b0 < n; b0 += batch) {
      const int nbat = std::min(batch, n - b0);
      int rc = batch_step(m, nbat, X, T, m->idx + b0, 1, lr, nullptr);
      if (rc) return rc;
      double l = 0.0;
      MCHK(m, hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
      MCHK(m, hipStreamSynchronize(m->stream));
      sum += l;
      nb++;
    }
    last = nb ? sum / nb : 0.0;
  }
  if (loss) *loss = last;
  return COUPLB_OK;
}

}  // extern "C"
