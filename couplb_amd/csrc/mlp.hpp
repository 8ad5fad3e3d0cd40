// mlp.hpp -- the warm-start network of the multi-direct-forcing IBM:
// 67 inputs -> 64 -> 64 -> 3 outputs, GELU (erf form) after both hidden
// layers, FP32 master weights, BF16 operands on the gfx950 matrix cores
// (v_mfma_f32_32x32x16_bf16) with FP32 accumulation.
//
// Parameter vector (FP32), row-major with the input index first:
//   W1[67][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3]   (8707 values)
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace couplb {
namespace mlp {

constexpr int IN = 67, INP = 80, H = 64, OUT = 3, OUTP = 32;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + IN * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H * OUT, NPARAM = OFF_B3 + OUT;
constexpr int TILE = 64;  // samples per block

// Forward pass of n samples: X [n][67] FP32 (device), params FP32 (device),
// Y [n][3] FP32 (device); enqueued on stream.
void forward(int n, const float* X, const float* params, float* Y, hipStream_t stream);

}  // namespace mlp
}  // namespace couplb

struct couplb_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float* params = nullptr;            // FP32 master weights (device)
  float *adam_m = nullptr, *adam_v = nullptr;
  long step = 0;
  // training scratch, sized for `cap` samples per mini-batch
  int cap = 0;
  float *z1 = nullptr, *z2 = nullptr, *y = nullptr, *part = nullptr, *grad = nullptr;
  double* loss = nullptr;
  int* idx = nullptr;
  long idx_cap = 0;
  std::string err;
};
