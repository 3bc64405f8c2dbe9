// mlp.hip -- training and inference kernels of the learned p* guess.
//
//  * features: the solver's own face log.  For each face of a batch the FP64
//    exact solver's classification and newton (kernel.cu:200-231, 635-680)
//    run at the reference tolerance; two-shock faces get mask 1 and the
//    target log(p*/p_ref).
//  * forward: FP32 masters on the vector ALU (one sample per lane), or the
//    BF16 MFMA path (one 64-sample tile per wave, mlp.hpp).
//  * loss + gradient: mean squared error over the masked rows, forward and
//    backward per lane in FP32, per-workgroup weight gradients reduced
//    through LDS, per-workgroup partials summed in a fixed order by a second
//    kernel (deterministic).
//  * Adam on the FP32 masters, then the BF16 MFMA operands are rebuilt.
#include "common.hpp"
#include "flux64.hpp"
#include "mlp.hpp"

namespace ast {
namespace mlp {

// ---- features from the FP64 exact solve ---------------------------------
__global__ void features_kernel(size_t n, const double* __restrict__ in, double n1, double n2,
                                d64::GasD gas, float* __restrict__ feats,
                                float* __restrict__ target, unsigned char* __restrict__ mask) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double rL = in[i], pL = in[n + i], uL = in[2 * n + i] * n1 + in[3 * n + i] * n2;
    const double rR = in[4 * n + i], pR = in[5 * n + i],
                 uR = in[6 * n + i] * n1 + in[7 * n + i] * n2;
    // post-swap states, kernel.cu:156-184
    const bool sw = pL > pR;
    const double enI = sw ? -uR : uL, enII = sw ? -uL : uR;
    const double pI = sw ? pR : pL, pII = sw ? pL : pR, rI = sw ? rR : rL, rII = sw ? rL : rR;
    float f[IN];
    const float lp = features<double>(rI, pI, enI, rII, pII, enII, gas.g, f);
    for (int k = 0; k < IN; k++) feats[k * n + i] = f[k];
    // two shocks when the closing speed exceeds f_I(p_II) (kernel.cu:200-231)
    const double du = enI - enII;
    const bool two_shock = du > (pII - pI) / sqrt(0.5 * rI * (gas.g2 * pII + gas.g1 * pI));
    if (mask) mask[i] = two_shock ? 1 : 0;
    if (target) {
      float t = 0.f;
      if (two_shock) {
        int it;
        bool capped = false;
        const double zI = rI * sqrt(gas.g * pI / rI), zII = rII * sqrt(gas.g * pII / rII);
        const double p =
            d64::two_shock_root(gas, du, pI, rI, zI, pII, rII, zII, -1.0, false, it, capped);
        t = (float)(log(p) - (double)lp);
      }
      target[i] = t;
    }
  }
}

// ---- forward ----------------------------------------------------------------
__global__ void forward_f32_kernel(const float* __restrict__ P, size_t n,
                                   const float* __restrict__ feats, float* __restrict__ pred) {
  __shared__ float Ps[NP];
  for (int k = threadIdx.x; k < NP; k += blockDim.x) Ps[k] = P[k];
  __syncthreads();
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    float x[IN];
    for (int k = 0; k < IN; k++) x[k] = feats[k * n + i];
    pred[i] = forward_f32(Ps, x);
  }
}

// every wave walks whole 64-sample tiles (EXEC stays full for the MFMAs);
// lanes past n compute on clamped rows and do not store
__global__ void forward_bf16_kernel(const unsigned char* __restrict__ buf, size_t n,
                                    const float* __restrict__ feats, float* __restrict__ pred) {
  const size_t waves = (size_t)gridDim.x * (blockDim.x / 64);
  const size_t w0 = (size_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const size_t tiles = (n + 63) / 64;
  for (size_t t = w0; t < tiles; t += waves) {
    const size_t i = t * 64 + (threadIdx.x & 63);
    const size_t ic = i < n ? i : n - 1;
    float x[IN];
    for (int k = 0; k < IN; k++) x[k] = feats[k * n + ic];
    const float y = forward_wave_bf16(buf, x);
    if (i < n) pred[i] = y;
  }
}

// ---- loss and gradient -----------------------------------------------------
constexpr int GNT = 128;       // threads (samples per tile) of the gradient kernel
constexpr int GBLOCKS = 256;   // workgroups: partial gradients to reduce
constexpr int PW = NP + 2;     // partial: gradient, loss sum, row count

__global__ __launch_bounds__(GNT) void grad_kernel(const float* __restrict__ P, size_t n,
                                                   const float* __restrict__ feats,
                                                   const float* __restrict__ target,
                                                   const unsigned char* __restrict__ mask,
                                                   float* __restrict__ partial) {
  __shared__ float Ps[NP];
  __shared__ float G[GNT][H + 1];
  __shared__ float Hh[GNT][H + 1];
  const int tid = threadIdx.x;
  for (int k = tid; k < NP; k += GNT) Ps[k] = P[k];
  __syncthreads();
  // owned gradient entries: dW3/dW2 rows i = tid/4, columns 8*(tid%4)..+7
  const int oi = tid >> 2, oj = (tid & 3) * 8;
  float aW3[8], aW2[8], aW1[2] = {0.f, 0.f}, aB = 0.f, aW4 = 0.f, aB4 = 0.f;
  for (int c = 0; c < 8; c++) aW3[c] = aW2[c] = 0.f;
  float lsum = 0.f, cnt = 0.f;
  const size_t tiles = (n + GNT - 1) / GNT;
  for (size_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const size_t i = t * GNT + tid;
    const bool valid = i < n && (mask == nullptr || mask[i]);
    float x[IN], h1[H], h2[H], h3[H], g[H];
    for (int k = 0; k < IN; k++) x[k] = i < n ? feats[k * n + i] : 0.f;
    for (int a = 0; a < H; a++) {
      float s = Ps[oB1 + a];
      for (int k = 0; k < IN; k++) s += Ps[oW1 + a * IN + k] * x[k];
      h1[a] = tanh_f(s);
    }
    for (int a = 0; a < H; a++) {
      float s = Ps[oB2 + a];
      for (int k = 0; k < H; k++) s += Ps[oW2 + a * H + k] * h1[k];
      h2[a] = tanh_f(s);
    }
    float y = Ps[oB4];
    for (int a = 0; a < H; a++) {
      float s = Ps[oB3 + a];
      for (int k = 0; k < H; k++) s += Ps[oW3 + a * H + k] * h2[k];
      h3[a] = tanh_f(s);
      y += Ps[oW4 + a] * h3[a];
    }
    const float e = valid ? (y - target[i]) : 0.f;
    lsum += e * e;
    cnt += valid ? 1.f : 0.f;
    const float dy = 2.f * e;  // d(e^2)/dy; the mean is taken in the reduction
    // layer 4 and layer-3 deltas
    for (int a = 0; a < H; a++) g[a] = dy * Ps[oW4 + a] * (1.f - h3[a] * h3[a]);
    __syncthreads();
    for (int a = 0; a < H; a++) {
      G[tid][a] = g[a];
      Hh[tid][a] = h2[a];
    }
    G[tid][H] = dy;
    Hh[tid][H] = 0.f;
    __syncthreads();
    for (int s = 0; s < GNT; s++) {
      const float gi = G[s][oi];
      for (int c = 0; c < 8; c++) aW3[c] += gi * Hh[s][oj + c];
    }
    if (tid < H) {
      float sb = 0.f;
      for (int s = 0; s < GNT; s++) sb += G[s][tid];
      aB += sb;  // db3 (tid < 32)
    }
    // dW4 = sum dy h3 and db4: h3 is not in LDS; accumulate with a second pass
    __syncthreads();
    for (int a = 0; a < H; a++) Hh[tid][a] = h3[a];
    __syncthreads();
    if (tid >= 32 && tid < 64) {
      float s4 = 0.f;
      for (int s = 0; s < GNT; s++) s4 += G[s][H] * Hh[s][tid - 32];
      aW4 += s4;
    }
    if (tid == 64) {
      float s4 = 0.f;
      for (int s = 0; s < GNT; s++) s4 += G[s][H];
      aB4 += s4;
    }
    // layer-2 deltas: g2 = (W3^T g3) * (1 - h2^2)
    float g2[H];
    for (int k = 0; k < H; k++) {
      float s = 0.f;
      for (int a = 0; a < H; a++) s += Ps[oW3 + a * H + k] * g[a];
      g2[k] = s * (1.f - h2[k] * h2[k]);
    }
    __syncthreads();
    for (int a = 0; a < H; a++) {
      G[tid][a] = g2[a];
      Hh[tid][a] = h1[a];
    }
    __syncthreads();
    for (int s = 0; s < GNT; s++) {
      const float gi = G[s][oi];
      for (int c = 0; c < 8; c++) aW2[c] += gi * Hh[s][oj + c];
    }
    if (tid >= 96 && tid < 128) {
      float sb = 0.f;
      for (int s = 0; s < GNT; s++) sb += G[s][tid - 96];
      aB += sb;  // db2 (96 <= tid < 128)
    }
    // layer-1 deltas: g1 = (W2^T g2) * (1 - h1^2)
    for (int k = 0; k < H; k++) {
      float s = 0.f;
      for (int a = 0; a < H; a++) s += Ps[oW2 + a * H + k] * g2[a];
      g[k] = s * (1.f - h1[k] * h1[k]);
    }
    __syncthreads();
    for (int a = 0; a < H; a++) G[tid][a] = g[a];
    for (int k = 0; k < IN; k++) Hh[tid][k] = x[k];
    __syncthreads();
    for (int r = 0; r < 2; r++) {  // dW1: 192 entries, tid and tid + 128
      const int e1 = tid + r * GNT;
      if (e1 < H * IN) {
        const int a = e1 / IN, k = e1 % IN;
        float s1 = 0.f;
        for (int s = 0; s < GNT; s++) s1 += G[s][a] * Hh[s][k];
        aW1[r] += s1;
      }
    }
    if (tid >= 64 && tid < 96) {
      float sb = 0.f;
      for (int s = 0; s < GNT; s++) sb += G[s][tid - 64];
      aB += sb;  // db1 (64 <= tid < 96)
    }
  }
  // this workgroup's partial gradient
  float* out = partial + (size_t)blockIdx.x * PW;
  for (int c = 0; c < 8; c++) {
    out[oW3 + oi * H + oj + c] = aW3[c];
    out[oW2 + oi * H + oj + c] = aW2[c];
  }
  for (int r = 0; r < 2; r++)
    if (tid + r * GNT < H * IN) out[oW1 + tid + r * GNT] = aW1[r];
  if (tid < 32) out[oB3 + tid] = aB;
  if (tid >= 96) out[oB2 + tid - 96] = aB;
  if (tid >= 64 && tid < 96) out[oB1 + tid - 64] = aB;
  if (tid >= 32 && tid < 64) out[oW4 + tid - 32] = aW4;
  if (tid == 64) out[oB4] = aB4;
  // loss and count: wave sums, then one lane per wave adds in LDS order
  __syncthreads();
  G[tid][0] = lsum;
  G[tid][1] = cnt;
  __syncthreads();
  if (tid == 0) {
    float ls = 0.f, cs = 0.f;
    for (int s = 0; s < GNT; s++) {
      ls += G[s][0];
      cs += G[s][1];
    }
    out[NP] = ls;
    out[NP + 1] = cs;
  }
}

// sums the partials in block order; grad = sum / count, loss = mean
__global__ void grad_reduce_kernel(const float* __restrict__ partial, int nblocks,
                                   float* __restrict__ grad, float* __restrict__ loss) {
  float cnt = 0.f;
  for (int b = 0; b < nblocks; b++) cnt += partial[(size_t)b * PW + NP + 1];
  const float inv = cnt > 0.f ? 1.f / cnt : 0.f;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < NP + 1; k += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int b = 0; b < nblocks; b++) s += partial[(size_t)b * PW + k];
    if (k < NP) grad[k] = s * inv;
    else if (loss) *loss = s * inv;
  }
}

// rebuilds the BF16 MFMA operands from the FP32 masters
__global__ void refresh_bf16_kernel(unsigned char* buf) {
  const float* P = (const float*)buf;
  bf16x8* A = (bf16x8*)(buf + BF_OFF);
  const int lane = threadIdx.x & 63, op = threadIdx.x >> 6;  // 7 waves
  if (op >= OPS) return;
  const int i = lane & 15, q = lane >> 4;
  bf16x8 a;
  for (int v = 0; v < 8; v++) {
    const int s = 8 * q + v;
    float w = 0.f;
    if (op < 2) {  // layer 1: feature slots 0..5
      if (s < IN) w = P[oW1 + (op * 16 + i) * IN + s];
    } else if (op < 6) {  // layers 2, 3: permuted hidden slots
      const int layer = (op - 2) / 2, mb = (op - 2) % 2;
      const int off = layer == 0 ? oW2 : oW3;
      w = P[off + (mb * 16 + i) * H + slot_unit(s)];
    } else {  // output: row 0 = w4
      if (i == 0) w = P[oW4 + slot_unit(s)];
    }
    a[v] = (__bf16)w;
  }
  A[op * 64 + lane] = a;
}

__global__ void adam_kernel(float* __restrict__ P, const float* __restrict__ g,
                            float* __restrict__ st, int step, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < NP; k += gridDim.x * blockDim.x) {
    float m = st[k], v = st[NP + k];
    m = b1 * m + (1.f - b1) * g[k];
    v = b2 * v + (1.f - b2) * g[k] * g[k];
    st[k] = m;
    st[NP + k] = v;
    P[k] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
  }
}

}  // namespace mlp
}  // namespace ast

using namespace ast;
using namespace ast::mlp;

extern "C" size_t ast_mlp_bytes(void) { return (size_t)BYTES; }
extern "C" size_t ast_mlp_work_floats(void) { return (size_t)GBLOCKS * PW; }

extern "C" int ast_mlp_features(size_t n, const double* in, double n1, double n2, double gamma,
                                float* feats, float* target, unsigned char* mask, void* stream) {
  if (n == 0) return AST_OK;
  if (!in || !feats) return ast_fail(AST_ERR_ARG, "ast_mlp_features: null pointer");
  if (!(gamma > 1.0)) return ast_fail(AST_ERR_ARG, "ast_mlp_features: gamma must be > 1");
  const d64::GasD gas = d64::make_gas(gamma);
  size_t blocks = (n + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(features_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, n,
                     in, n1, n2, gas, feats, target, mask);
  AST_LAUNCH_CHECK("ast_mlp_features launch");
  return AST_OK;
}

extern "C" int ast_mlp_forward(const ast_mlp_t* mlp, size_t n, const float* feats, float* pred,
                               int precision, void* stream) {
  if (n == 0) return AST_OK;
  if (!mlp || !feats || !pred) return ast_fail(AST_ERR_ARG, "ast_mlp_forward: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (precision == 0) {
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(forward_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, st,
                       (const float*)mlp, n, feats, pred);
  } else if (precision == 1) {
    size_t tiles = (n + 63) / 64, blocks = (tiles + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(forward_bf16_kernel, dim3((unsigned)blocks), dim3(256), 0, st,
                       (const unsigned char*)mlp, n, feats, pred);
  } else {
    return ast_fail(AST_ERR_ARG, "ast_mlp_forward: precision %d (0 FP32, 1 BF16 MFMA)", precision);
  }
  AST_LAUNCH_CHECK("ast_mlp_forward launch");
  return AST_OK;
}

extern "C" int ast_mlp_loss_grad(const ast_mlp_t* mlp, size_t n, const float* feats,
                                 const float* target, const unsigned char* mask, float* grad,
                                 float* loss, float* work, void* stream) {
  if (!mlp || !feats || !target || !grad || !work)
    return ast_fail(AST_ERR_ARG, "ast_mlp_loss_grad: null pointer");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(grad_kernel, dim3(GBLOCKS), dim3(GNT), 0, st, (const float*)mlp, n, feats,
                     target, mask, work);
  AST_LAUNCH_CHECK("ast_mlp_loss_grad launch");
  hipLaunchKernelGGL(grad_reduce_kernel, dim3(10), dim3(256), 0, st, work, GBLOCKS, grad, loss);
  AST_LAUNCH_CHECK("ast_mlp_loss_grad reduce");
  return AST_OK;
}

extern "C" int ast_mlp_refresh(ast_mlp_t* mlp, void* stream) {
  if (!mlp) return ast_fail(AST_ERR_ARG, "ast_mlp_refresh: null pointer");
  hipLaunchKernelGGL(refresh_bf16_kernel, dim3(1), dim3(64 * OPS), 0, (hipStream_t)stream,
                     (unsigned char*)mlp);
  AST_LAUNCH_CHECK("ast_mlp_refresh launch");
  return AST_OK;
}

extern "C" int ast_mlp_adam(ast_mlp_t* mlp, const float* grad, float* state, int step, float lr,
                            void* stream) {
  if (!mlp || !grad || !state || step < 1) return ast_fail(AST_ERR_ARG, "ast_mlp_adam: bad argument");
  hipLaunchKernelGGL(adam_kernel, dim3(10), dim3(256), 0, (hipStream_t)stream, (float*)mlp, grad,
                     state, step, lr);
  AST_LAUNCH_CHECK("ast_mlp_adam launch");
  return ast_mlp_refresh(mlp, stream);
}

extern "C" int ast_mlp_set_params(ast_mlp_t* mlp, const float* host_params, void* stream) {
  if (!mlp || !host_params) return ast_fail(AST_ERR_ARG, "ast_mlp_set_params: null pointer");
  hipStream_t st = (hipStream_t)stream;
  AST_HIP_CHECK(hipMemcpyAsync(mlp, host_params, NP * sizeof(float), hipMemcpyHostToDevice, st));
  int rc = ast_mlp_refresh(mlp, stream);
  if (rc != AST_OK) return rc;
  AST_HIP_CHECK(hipStreamSynchronize(st));
  return AST_OK;
}

extern "C" int ast_mlp_get_params(const ast_mlp_t* mlp, float* host_params, void* stream) {
  if (!mlp || !host_params) return ast_fail(AST_ERR_ARG, "ast_mlp_get_params: null pointer");
  hipStream_t st = (hipStream_t)stream;
  AST_HIP_CHECK(hipMemcpyAsync(host_params, mlp, NP * sizeof(float), hipMemcpyDeviceToHost, st));
  AST_HIP_CHECK(hipStreamSynchronize(st));
  return AST_OK;
}
