// mlp.hpp -- the learned starting point of newton (kernel.cu:635-680).
//
// A 6 -> 32 -> 32 -> 32 -> 1 MLP with tanh hidden units maps the normalised
// primitives of a two-shock face to log(p* / p_ref).  Parameters live in one
// device block (ast_mlp_t):
//   [0, NP) floats         FP32 masters (trained): W1[32][6] b1 W2[32][32] b2
//                          W3[32][32] b3 w4[32] b4
//   [BF_OFF, BYTES) bytes  BF16 MFMA operands built from the masters
// Inference on a wave of 64 faces runs on MFMA (v_mfma_f32_16x16x32_bf16,
// FP32 accumulate) in the transposed form H_out^T = W * H_in^T, with the
// hidden units of each layer stored in a slot order chosen so that the
// accumulator layout of one layer is exactly the B-operand layout of the
// next: no data moves between layers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev.hpp"

namespace ast {
namespace mlp {

constexpr int IN = 6, H = 32;
constexpr int oW1 = 0, oB1 = oW1 + H * IN, oW2 = oB1 + H, oB2 = oW2 + H * H, oW3 = oB2 + H,
              oB3 = oW3 + H * H, oW4 = oB3 + H, oB4 = oW4 + H, NP = oB4 + 1;
static_assert(NP == AST_MLP_NPARAM, "parameter layout");
constexpr int BF_OFF = 9728;                    // bytes, 256-aligned after the masters
constexpr int OPS = 7;                          // A1[2], A2[2], A3[2], A4
constexpr int BYTES = BF_OFF + OPS * 64 * 8 * 2;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// k-slot s (0..31) of a B/A operand <-> hidden unit: lane group q = s/8
// holds the units its accumulators hold after the previous layer, rows
// 4q..4q+3 of output tile 0 and of output tile 1.
AST_DEV __host__ int slot_unit(int s) {
  const int q = s >> 3, v = s & 7;
  return v < 4 ? 4 * q + v : 16 + 4 * q + (v - 4);
}

// tanh(x) = 1 - 2 / (exp(2x) + 1) with the hardware exp2 and reciprocal; the
// same function serves training and both inference paths.
AST_DEV float tanh_f(float x) {
  const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * x);  // exp(2x)
  return 1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f);
}

// Normalised features of a face in the post-swap frame of the exact solver
// (pI <= pII, kernel.cu:156-184): log rho and log p relative to the
// geometric means, normal velocity minus the mean over the mean sound speed.
// Returns log(p_ref) (the target is log(p* / p_ref)).
template <typename R>
AST_DEV float features(R rI, R pI, R uI, R rII, R pII, R uII, R gamma, float* f) {
  const float lrI = logf((float)rI), lrII = logf((float)rII);
  const float lpI = logf((float)pI), lpII = logf((float)pII);
  const float lr = 0.5f * (lrI + lrII), lp = 0.5f * (lpI + lpII);
  const float cref = sqrtf((float)gamma * expf(lp - lr));
  const float um = 0.5f * ((float)uI + (float)uII);
  const float ic = 1.f / cref;
  f[0] = lrI - lr;
  f[1] = ((float)uI - um) * ic;
  f[2] = lpI - lp;
  f[3] = lrII - lr;
  f[4] = ((float)uII - um) * ic;
  f[5] = lpII - lp;
  return lp;
}

// FP32 forward of one sample on the vector ALU (the masters).
AST_DEV float forward_f32(const float* __restrict__ P, const float* x) {
  float h1[H], h2[H];
#pragma unroll
  for (int i = 0; i < H; i++) {
    float a = P[oB1 + i];
#pragma unroll
    for (int k = 0; k < IN; k++) a += P[oW1 + i * IN + k] * x[k];
    h1[i] = tanh_f(a);
  }
#pragma unroll
  for (int i = 0; i < H; i++) {
    float a = P[oB2 + i];
#pragma unroll
    for (int k = 0; k < H; k++) a += P[oW2 + i * H + k] * h1[k];
    h2[i] = tanh_f(a);
  }
  float y = P[oB4];
#pragma unroll
  for (int i = 0; i < H; i++) {
    float a = P[oB3 + i];
#pragma unroll
    for (int k = 0; k < H; k++) a += P[oW3 + i * H + k] * h2[k];
    y += P[oW4 + i] * tanh_f(a);
  }
  return y;
}

AST_DEV float lane_shfl(float v, int src) { return __shfl(v, src, 64); }

// Forward of the 64 samples of a wave on MFMA; every lane passes its own
// sample's features and receives its prediction.  Must be called by all 64
// lanes of the wave (EXEC full).
AST_DEV float forward_wave_bf16(const unsigned char* __restrict__ buf, const float* x) {
  const float* P = (const float*)buf;
  const bf16x8* A = (const bf16x8*)(buf + BF_OFF);
  const int lane = threadIdx.x & 63, q = lane >> 4, j = lane & 15;
  // layer-1 B operands: lane group 0 holds the 6 features of sample
  // nb*16 + j in slots 0..5; the other slots are zero
  bf16x8 B[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    bf16x8 b;
#pragma unroll
    for (int v = 0; v < 8; v++) {
      const float xv = lane_shfl(v < IN ? x[v < IN ? v : 0] : 0.f, nb * 16 + j);
      b[v] = (__bf16)((q == 0 && v < IN) ? xv : 0.f);
    }
    B[nb] = b;
  }
  const float* bias[3] = {P + oB1, P + oB2, P + oB3};
#pragma unroll
  for (int layer = 0; layer < 3; layer++) {
    const bf16x8 a0 = A[(2 * layer + 0) * 64 + lane];
    const bf16x8 a1 = A[(2 * layer + 1) * 64 + lane];
    float bv[8];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      bv[v] = bias[layer][4 * q + v];
      bv[4 + v] = bias[layer][16 + 4 * q + v];
    }
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const f32x4 z = {0.f, 0.f, 0.f, 0.f};
      const f32x4 c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, B[nb], z, 0, 0, 0);
      const f32x4 c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, B[nb], z, 0, 0, 0);
      bf16x8 nbv;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        nbv[v] = (__bf16)tanh_f(c0[v] + bv[v]);
        nbv[4 + v] = (__bf16)tanh_f(c1[v] + bv[4 + v]);
      }
      B[nb] = nbv;
    }
  }
  // output row 0 of the last tile: y of sample nb*16 + j sits in lane j
  const bf16x8 a4 = A[6 * 64 + lane];
  float yv[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    const f32x4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a4, B[nb], z, 0, 0, 0);
    yv[nb] = c[0];
  }
  float y = 0.f;
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    const float t = lane_shfl(yv[nb], lane & 15);
    y = ((lane >> 4) == nb) ? t : y;
  }
  return y + P[oB4];
}

}  // namespace mlp
}  // namespace ast
