// step_f32_kernel.hpp -- the fast step kernel (described in step_f32.hip)
// and its launch over geometry, window and counter variants.  Each method's
// kernels are instantiated in a source of their own (step_f32_<method>.hip),
// so that the build compiles them in parallel; step_f32.hip holds the host
// entry point.
#pragma once

#include <type_traits>

#include "strip_f32.hpp"

namespace ast {
namespace f32 {

// the arguments of one step launch, checked by the host entry point
struct StepArgs {
  ast_grid_t g;
  Consts c;
  ExactCtx ex;
  bool window;
  const float2* s;
  const float2* u;
  float2* s2;
  float2* u2;
  double* T;
  const double* T_do;
  KTime kt;
  ast_stats_t* stats;
  int krows;
  hipStream_t stream;
  const SrcF* src;  // neutral-H planes, or null
  const double* TT;  // simulation time for pulsed sides without kt (else unused)
};

// launches one method's kernel for a.g's geometry and a.window; defined
// below, instantiated once per method in its own source
template <int METHOD>
int step_launch(const StepArgs& a);

extern template int step_launch<AST_HLL>(const StepArgs&);
extern template int step_launch<AST_HLLC>(const StepArgs&);
extern template int step_launch<AST_HLLC_A>(const StepArgs&);
extern template int step_launch<AST_EXACT>(const StepArgs&);
extern template int step_launch<kExactLearned>(const StepArgs&);
extern template int step_launch<kExactDetect>(const StepArgs&);

#ifdef AST_STEP_F32_KERNELS

// STATS: the variant that counts the exact solver's work into *stats; SRC:
// the variants with add_MK's neutral-hydrogen source planes (kSrcF64, the
// reference's planes; kSrcF32, FP32 copies of them); PULSE: the variant for
// grids with a pulsed Dirichlet side (HLL family), whose fast edge strips
// apply the pulse -- the kernels without it keep their code unchanged
template <int METHOD, int GEOM, bool WINDOW, bool STATS, int SRC, bool PULSE>
__global__ __launch_bounds__(NT) void step_kernel(ast_grid_t g, Consts c, ExactCtx ex,
                                                  const float2* __restrict__ s,
                                                  const float2* __restrict__ u,
                                                  float2* __restrict__ s2,
                                                  float2* __restrict__ u2, double* T,
                                                  const double* __restrict__ T_do, KTime kt,
                                                  ast_stats_t* stats, int nbands, int krows,
                                                  int nwg, SrcF src, const double* TT) {
  __shared__ double red[WPB];
  __shared__ uint32_t redi[WPB * 10];
  const double tdo = step_dt_load(T_do, kt);
  // pulsed Dirichlet sides: the density factor at the time the reference's
  // step reads (TT after the previous step's funk_time).  With the runtime's
  // slots (kt), workgroup 0 advances T[2] during this step, so the step
  // takes the time the previous step of its run used (word 2 of that step's
  // slot set) plus its own dt -- the sum T[2] receives -- and records its
  // own for the next step; the first step of a run reads T[2].
  double pulse = 1.0;
  if (c.pulsed) {  // wave-uniform
    double tt;
    if (kt.cur != nullptr) {
      tt = kt.prev != nullptr ? kt.prev[2] + tdo : T[2];
      if (blockIdx.x == 0 && threadIdx.x == 0) kt.cur[2] = tt;
    } else {
      tt = *TT;
    }
    pulse = pulse_factor(g, tt);
  }
  step_dt_commit(T, kt, tdo);
  constexpr bool kDetects = METHOD == kExactLearned || METHOD == kExactDetect;
  typename std::conditional<STATS, StatsF32,
                            typename std::conditional<kDetects, StrongStats, NoStats>::type>::type st;
  st.clear();
  double tmin = INFINITY;
  // wave index made explicitly uniform: rows and the loop stay scalar
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int band, m0, m1;
  strip_rows(g.M, nbands, krows, xcd_tile(blockIdx.x, nwg), wave, band, m0, m1);
  tmin = strip_of_wave<METHOD, GEOM, WINDOW, SRC, PULSE>(g, c, ex, s, u, s2, u2, tdo, band, m0,
                                                         m1, st, src, pulse);
  if (ex.strong != nullptr && st.any_strong() && (threadIdx.x & 63) == 0) atomicOr(ex.strong, 1u);
  if constexpr (STATS) block_stats_f32_to<NT>(st, redi, stats);
  block_min_uniform<NT>(tmin, red, kt.cur ? kt.cur : T, kt.cur ? kt.nslots : 0);
}

template <int METHOD, int GEOM, bool WINDOW>
int launch_f32(const ast_grid_t& g, const Consts& c, const ExactCtx& ex, const float2* s,
                      const float2* u, float2* s2, float2* u2, double* T, const double* T_do,
                      const KTime& kt, ast_stats_t* stats, int krows, hipStream_t stream,
                      const SrcF* src, const double* TT) {
  // krows > 0: strips of krows rows; krows = -S: S strips per band (strip_rows)
  const int nbands = (g.N + WOUT - 1) / WOUT;
  long long nwg;
  if (krows > 0) {
    const int nstrips = (g.M + krows - 1) / krows;
    nwg = (long long)nbands * ((nstrips + WPB - 1) / WPB);
  } else {
    nwg = ((long long)nbands * -(long long)krows + WPB - 1) / WPB;
  }
  if (nwg > 0x7fffffffLL / NT) return ast_fail(AST_ERR_SIZE, "ast_step_f32: grid too large");
  const unsigned grid = (unsigned)nwg;
  const SrcF none{nullptr, nullptr, nullptr, nullptr, 0.f, 0.f};
  constexpr bool kSrc = METHOD != kExactLearned && METHOD != kExactDetect;
  // pulsed sides: the PULSE kernels of the HLL family, and of the plain
  // exact solver without counters or source planes (the other exact-solver
  // kernels serve pulsed sides on the boundary strip path)
  constexpr bool kPulseKernels = !is_exact(METHOD) || METHOD == AST_EXACT;
  constexpr bool kPulseSrcKernels = !is_exact(METHOD);
  const bool pulsed = kPulseKernels && c.pulsed;
#define AST_LAUNCH_K(STATS_, SRC_, PULSE_, SRCV)                                                    \
  hipLaunchKernelGGL((step_kernel<METHOD, GEOM, WINDOW, STATS_, SRC_, PULSE_>), dim3(grid),        \
                     dim3(NT), 0, stream, g, c, ex, s, u, s2, u2, T, T_do, kt, stats, nbands, krows, \
                     (int)nwg, SRCV, TT)
#define AST_LAUNCH_P(HAS_, STATS_, SRC_, SRCV)  \
  do {                                          \
    if constexpr (HAS_) {                       \
      if (pulsed) {                             \
        AST_LAUNCH_K(STATS_, SRC_, true, SRCV); \
        break;                                  \
      }                                         \
    }                                           \
    AST_LAUNCH_K(STATS_, SRC_, false, SRCV);    \
  } while (0)
  if (src != nullptr) {
    if constexpr (kSrc) {
      const bool f32 = src->xy != nullptr || src->e != nullptr;
      const bool counting = is_exact(METHOD) && stats != nullptr;
      if (counting) {
        if constexpr (is_exact(METHOD)) {
          if (f32) AST_LAUNCH_K(true, kSrcF32, false, *src);
          else AST_LAUNCH_K(true, kSrcF64, false, *src);
        }
      } else if (f32) {
        AST_LAUNCH_P(kPulseSrcKernels, false, kSrcF32, *src);
      } else {
        AST_LAUNCH_P(kPulseSrcKernels, false, kSrcF64, *src);
      }
    } else {
      return ast_fail(AST_ERR_ARG, "ast_step_f32: source planes with the learned start");
    }
  } else if (is_exact(METHOD) && stats != nullptr) {
    if constexpr (is_exact(METHOD)) AST_LAUNCH_K(true, kSrcNone, false, none);
  } else {
    AST_LAUNCH_P(kPulseKernels, false, kSrcNone, none);
  }
#undef AST_LAUNCH_P
#undef AST_LAUNCH_K
  AST_LAUNCH_CHECK("ast_step_f32 launch");
  return AST_OK;
}

template <int METHOD>
int step_launch(const StepArgs& a) {
  const bool axi = a.g.geometry == AST_AXISYMMETRIC;
#define AST_LAUNCH(G, W)                                                                     \
  launch_f32<METHOD, G, W>(a.g, a.c, a.ex, a.s, a.u, a.s2, a.u2, a.T, a.T_do, a.kt, a.stats, \
                           a.krows, a.stream, a.src, a.TT)
  if constexpr (METHOD == kExactLearned || METHOD == kExactDetect) {  // without a window only
    return axi ? AST_LAUNCH(AST_AXISYMMETRIC, false) : AST_LAUNCH(AST_PLANAR, false);
  } else {
    if (axi) return a.window ? AST_LAUNCH(AST_AXISYMMETRIC, true) : AST_LAUNCH(AST_AXISYMMETRIC, false);
    return a.window ? AST_LAUNCH(AST_PLANAR, true) : AST_LAUNCH(AST_PLANAR, false);
  }
#undef AST_LAUNCH
}

#endif  // AST_STEP_F32_KERNELS

}  // namespace f32
}  // namespace ast
