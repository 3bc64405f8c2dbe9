// flux_f32.hip -- batched FP32 face-flux operators over SoA face arrays:
// HLL, HLLC (both wave-speed estimates) and the exact solver (FP64 islands)
// of the fast path applied to
// n independent faces with a common normal (n1, n2).  52 B/face of
// algorithmic traffic (8 words in, 5 out).
#include <type_traits>

#include "common.hpp"
#include "exact_f32.hpp"
#include "fast_flux.hpp"

namespace ast {
namespace f32 {

constexpr int FNT = 256;

// STATS: the variant that counts the exact solver's work into *stats;
// LEARNED: the exact solver with the learned newton start compiled in
template <int METHOD, bool STATS, bool LEARNED>
__global__ __launch_bounds__(FNT) void flux_kernel(size_t n, const float* __restrict__ in,
                                                   float n1, float n2, float krad,
                                                   ExactCtx ex, float* __restrict__ out,
                                                   ast_stats_t* stats, int* __restrict__ info) {
  __shared__ uint32_t redi[(FNT / 64) * 10];
  typename std::conditional<STATS, StatsF32, NoStats>::type st;
  st.clear();
  // whole 64-face tiles per wave: every lane runs every iteration (EXEC full,
  // as the wave-level learned guess needs); lanes past n compute on a clamped
  // face and do not store
  const size_t stride = (size_t)gridDim.x * FNT;
  const size_t nt = (n + 63) / 64 * 64;
  const int lane = threadIdx.x & 63;
  const size_t base0 = (size_t)blockIdx.x * FNT + 64 * __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (size_t base = base0; base < nt; base += stride) {  // wave-uniform
    const size_t ii = base + lane;
    const bool valid = ii < n;
    const size_t i = valid ? ii : n - 1;
    const float rl = in[i], pl = in[n + i], vxl = in[2 * n + i], vyl = in[3 * n + i];
    const float rr = in[4 * n + i], pr = in[5 * n + i], vxr = in[6 * n + i], vyr = in[7 * n + i];
    // face frame: normal (n1, n2), tangent (-n2, n1)
    const PrimF L = {rl, pl, vxl * n1 + vyl * n2, -vxl * n2 + vyl * n1};
    const PrimF R = {rr, pr, vxr * n1 + vyr * n2, -vxr * n2 + vyr * n1};
    float UU;
    FluxF f;
    if constexpr (METHOD == AST_HLL) f = hll_f32(L, R, ex.g, ex.ig1, UU);
    else if constexpr (METHOD == AST_HLLC) f = hllc_f32<false>(L, R, ex.g, ex.ig1, UU);
    else if constexpr (METHOD == AST_HLLC_A) f = hllc_f32<true>(L, R, ex.g, ex.ig1, UU);
    else {
      FaceInfo fi;
      f = exact_f32<LEARNED>(L, R, ex, UU, fi);
      if (!valid) fi = FaceInfo{kNoFace, 0, false, false, false, false};
      st.add(fi);
      if (info && valid) {
        info[2 * i] = fi.it;
        info[2 * i + 1] = (fi.pat == kNoPattern ? -1 : fi.pat + 1) | (fi.capped ? 8 : 0) |
                          (fi.island ? 16 : 0);
      }
    }
    if (valid) {
      out[i] = f.m;
      out[n + i] = f.e;
      out[2 * n + i] = f.un * n1 - f.ut * n2;
      out[3 * n + i] = f.un * n2 + f.ut * n1;
      out[4 * n + i] = krad * frcp(UU);
    }
  }
  if constexpr (STATS) block_stats_f32_to<FNT>(st, redi, stats);
}

template <int METHOD>
static int launch(size_t n, const float* in, float n1, float n2, float krad, const ExactCtx& ex,
                  float* out, ast_stats_t* stats, int* info, hipStream_t stream) {
  size_t blocks = (n + FNT - 1) / FNT;
  if (blocks > 256 * 64) blocks = 256 * 64;
  const bool st = METHOD == AST_EXACT && stats != nullptr;
  if (METHOD == AST_EXACT && ex.mlp != nullptr) {
    if constexpr (METHOD == AST_EXACT) {
      if (st)
        hipLaunchKernelGGL((flux_kernel<METHOD, true, true>), dim3((unsigned)blocks), dim3(FNT), 0,
                           stream, n, in, n1, n2, krad, ex, out, stats, info);
      else
        hipLaunchKernelGGL((flux_kernel<METHOD, false, true>), dim3((unsigned)blocks), dim3(FNT), 0,
                           stream, n, in, n1, n2, krad, ex, out, stats, info);
    }
  } else if (st) {
    hipLaunchKernelGGL((flux_kernel<METHOD, true, false>), dim3((unsigned)blocks), dim3(FNT), 0,
                       stream, n, in, n1, n2, krad, ex, out, stats, info);
  } else {
    hipLaunchKernelGGL((flux_kernel<METHOD, false, false>), dim3((unsigned)blocks), dim3(FNT), 0,
                       stream, n, in, n1, n2, krad, ex, out, stats, info);
  }
  AST_LAUNCH_CHECK("ast_flux_f32 launch");
  return AST_OK;
}

}  // namespace f32
}  // namespace ast

extern "C" int ast_flux_f32_info(int method, size_t n, const float* in, float n1, float n2,
                                 float rad, float gamma, float krit, const ast_mlp_t* mlp,
                                 float* out, ast_stats_t* stats, int* info, void* stream) {
  using namespace ast::f32;
  if (n == 0) return AST_OK;
  if (!in || !out) return ast_fail(AST_ERR_ARG, "ast_flux_f32: null in/out");
  if (!(gamma > 1.0f)) return ast_fail(AST_ERR_ARG, "ast_flux_f32: gamma must be > 1");
  const ast::ExactCtx ex = ast::make_exact_ctx((double)gamma, mlp);
  const float krad = krit * rad;
  hipStream_t st = (hipStream_t)stream;
  switch (method) {
    case AST_HLL: return launch<AST_HLL>(n, in, n1, n2, krad, ex, out, stats, nullptr, st);
    case AST_HLLC: return launch<AST_HLLC>(n, in, n1, n2, krad, ex, out, stats, nullptr, st);
    case AST_EXACT: return launch<AST_EXACT>(n, in, n1, n2, krad, ex, out, stats, info, st);
    case AST_HLLC_A: return launch<AST_HLLC_A>(n, in, n1, n2, krad, ex, out, stats, nullptr, st);
    default: return ast_fail(AST_ERR_ARG, "ast_flux_f32: unknown method %d", method);
  }
}

extern "C" int ast_flux_f32(int method, size_t n, const float* in, float n1, float n2, float rad,
                            float gamma, float krit, float* out, ast_stats_t* stats,
                            void* stream) {
  return ast_flux_f32_info(method, n, in, n1, n2, rad, gamma, krit, nullptr, out, stats, nullptr,
                           stream);
}
