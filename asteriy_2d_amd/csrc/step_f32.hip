// step_f32.hip -- the fast step: FP32 state, every face evaluated once.
//
// Same scheme as the reference step (Kernel_TVD, kernel.cu:2513-3073): minmod
// piecewise-linear reconstruction with the positivity fallback, a Riemann
// flux per face, the planar or axisymmetric (KorolkovSolver.cu:13-32) update
// and the CFL dt -- reorganised for gfx950:
//
//  * "wave strip": a wave owns a band of 64 columns (60 output columns plus
//    2 halo columns on each side) and walks down KR rows.  The x-neighbours
//    of a cell live in the adjacent lanes and are read with DPP wave shifts
//    (wave_shr:1 / wave_shl:1, a register move, no LDS); the y-neighbours
//    are a 3-row sliding window in registers.  Each lane evaluates one
//    x-face (its west face) and one y-face (its north face) per row, and
//    the east face flux arrives from the next lane by another DPP shift, so
//    every face of the grid is computed once (plus the strip/band seams).
//  * state traffic: each cell is read once per strip (+ 4 halo rows) and
//    written once: 16 B in + 16 B out per cell-update in FP32.
//  * slopes use the minmod of neighbour differences (the reference divides
//    by coordinate differences and multiplies by the half spacing: the
//    same number in exact arithmetic, kernel.cu:35-50).
//  * dt: each lane keeps the largest |wave speed| of its x- and y-faces;
//    the wave reduces them (shuffles), dt = krit*min(dx/max_x, dy/max_y) in
//    FP64, one 64-bit atomicMin per workgroup (deterministic).  The runtime
//    spreads these atomics over 64 slots (one address would serialise
//    them), and the next step folds the slots into its own dt and advances
//    the time itself (time_slots.h): a step is one kernel.
//  * the exact solver runs in FP32 with its p* solve promoted to FP64 near
//    vacuum and in strong rarefactions, optionally seeded by the learned
//    guess (exact_f32.hpp).
#include <type_traits>

#include "strip_f32.hpp"

namespace ast {
namespace f32 {

// STATS: the variant that counts the exact solver's work into *stats
template <int METHOD, int GEOM, bool WINDOW, bool STATS>
__global__ __launch_bounds__(NT) void step_kernel(ast_grid_t g, Consts c, ExactCtx ex,
                                                  const float2* __restrict__ s,
                                                  const float2* __restrict__ u,
                                                  float2* __restrict__ s2,
                                                  float2* __restrict__ u2, double* T,
                                                  const double* __restrict__ T_do, KTime kt,
                                                  ast_stats_t* stats, int nbands, int krows,
                                                  int nwg) {
  __shared__ double red[WPB];
  __shared__ uint32_t redi[WPB * 10];
  // gated launch: this kernel's mode is not the selected one (grid-uniform)
  if (ex.gate != nullptr && *ex.gate != ex.gate_mode) return;
  const double tdo = step_dt(T, T_do, kt);
  constexpr bool kDetects = METHOD == kExactLearned || METHOD == kExactDetect;
  typename std::conditional<STATS, StatsF32,
                            typename std::conditional<kDetects, StrongStats, NoStats>::type>::type st;
  st.clear();
  double tmin = INFINITY;
  if constexpr (METHOD == kExactLearned) {
    // several strip groups per workgroup when the grid is smaller than the
    // work: the learned kernel runs on its co-resident capacity, so that it
    // costs little in the steps in which it only reads the gate (the loop
    // costs the other kernels registers, so they take one group each)
    for (int item = blockIdx.x; item < nwg; item += gridDim.x) {  // workgroup-uniform
      const int wg = xcd_tile(item, nwg);
      tmin = fmin(tmin, strip_of_wave<METHOD, GEOM, WINDOW>(g, c, ex, s, u, s2, u2, tdo,
                                                            wg % nbands, wg / nbands, krows, st));
    }
  } else {
    const int wg = xcd_tile(blockIdx.x, nwg);
    tmin = strip_of_wave<METHOD, GEOM, WINDOW>(g, c, ex, s, u, s2, u2, tdo, wg % nbands,
                                               wg / nbands, krows, st);
  }
  if (ex.strong != nullptr && st.any_strong() && (threadIdx.x & 63) == 0) atomicOr(ex.strong, 1u);
  if constexpr (STATS) block_stats_f32_to<NT>(st, redi, stats);
  block_min_uniform<NT>(tmin, red, kt.cur ? kt.cur : T, kt.cur ? kt.nslots : 0);
}

}  // namespace f32
}  // namespace ast

using namespace ast;
using namespace ast::f32;

static int g_krows = 8;

extern "C" int ast_set_tile_rows(int krows) {
  if (krows < 1 || krows > 1024) return ast_fail(AST_ERR_ARG, "ast_set_tile_rows: %d", krows);
  g_krows = krows;
  return AST_OK;
}
extern "C" int ast_get_tile_rows(void) { return g_krows; }

template <int METHOD, int GEOM, bool WINDOW>
static int launch_f32(const ast_grid_t& g, const Consts& c, const ExactCtx& ex, const float2* s,
                      const float2* u, float2* s2, float2* u2, double* T, const double* T_do,
                      const KTime& kt, ast_stats_t* stats, int krows, hipStream_t stream) {
  const int nbands = (g.N + WOUT - 1) / WOUT;
  const int nstrips = (g.M + krows - 1) / krows;
  const int nsg = (nstrips + WPB - 1) / WPB;
  const long long nwg = (long long)nbands * nsg;
  if (nwg > 0x7fffffffLL / NT) return ast_fail(AST_ERR_SIZE, "ast_step_f32: grid too large");
  // the gated learned kernel: 512 workgroups (two waves per SIMD, what its
  // registers allow to be resident at once) that loop over the strip groups
  const unsigned grid =
      (unsigned)(METHOD == kExactLearned && ex.gate != nullptr && nwg > 512 ? 512 : nwg);
  if (is_exact(METHOD) && stats != nullptr) {
    if constexpr (is_exact(METHOD))
      hipLaunchKernelGGL((step_kernel<METHOD, GEOM, WINDOW, true>), dim3(grid), dim3(NT), 0,
                         stream, g, c, ex, s, u, s2, u2, T, T_do, kt, stats, nbands, krows,
                         (int)nwg);
  } else {
    hipLaunchKernelGGL((step_kernel<METHOD, GEOM, WINDOW, false>), dim3(grid), dim3(NT), 0,
                       stream, g, c, ex, s, u, s2, u2, T, T_do, kt, stats, nbands, krows,
                       (int)nwg);
  }
  AST_LAUNCH_CHECK("ast_step_f32 launch");
  return AST_OK;
}

template <int METHOD>
static int dispatch_geom(const ast_grid_t& g, const Consts& c, const ExactCtx& ex, bool window,
                         const float2* s, const float2* u, float2* s2, float2* u2, double* T,
                         const double* T_do, const KTime& kt, ast_stats_t* stats, int krows,
                         hipStream_t st) {
  const bool axi = g.geometry == AST_AXISYMMETRIC;
  if constexpr (METHOD == kExactLearned || METHOD == kExactDetect) {  // without a window only
    (void)window;
    return axi ? launch_f32<METHOD, AST_AXISYMMETRIC, false>(g, c, ex, s, u, s2, u2, T, T_do, kt, stats, krows, st)
               : launch_f32<METHOD, AST_PLANAR, false>(g, c, ex, s, u, s2, u2, T, T_do, kt, stats, krows, st);
  }
  if (axi) {
    return window ? launch_f32<METHOD, AST_AXISYMMETRIC, true>(g, c, ex, s, u, s2, u2, T, T_do, kt, stats, krows, st)
                  : launch_f32<METHOD, AST_AXISYMMETRIC, false>(g, c, ex, s, u, s2, u2, T, T_do, kt, stats, krows, st);
  }
  return window ? launch_f32<METHOD, AST_PLANAR, true>(g, c, ex, s, u, s2, u2, T, T_do, kt, stats, krows, st)
                : launch_f32<METHOD, AST_PLANAR, false>(g, c, ex, s, u, s2, u2, T, T_do, kt, stats, krows, st);
}

// tm (may be null): the runtime's time protocol (time_slots.h).  Null: the
// reference protocol (the step lowers *T, dt is *T_do, a learned guess is
// always on).  Otherwise the step lowers the slots of tm->cur and, with
// tm->prev, takes its dt from the previous step's slots and advances the
// time itself.  With the exact method, a learned guess, no window and
// tm->gate, the step is two gated launches: the plain solver (mode 0), which
// flags strong two-shock faces in the strong word of tm->cur, and the solver
// with the learned start (mode 1); the mode is the strong flag of tm->prev,
// or tm->gate[0] without one.
int ast_step_f32_rows(const ast_grid_t* g, const float* s, const float* u, float* s2, float* u2,
                      double* T, const double* T_do, int method, const ast_mlp_t* mlp,
                      ast_stats_t* stats, int krows, const StepTime* tm, void* stream) {
  int rc = ast_grid_check(g);
  if (rc != AST_OK) return rc;
  if (!s || !u || !s2 || !u2 || !T || !T_do)
    return ast_fail(AST_ERR_ARG, "ast_step_f32: null state or time pointer");
  if (s == s2 || u == u2 || s == u2 || u == s2)
    return ast_fail(AST_ERR_ARG, "ast_step_f32: output planes must not alias the input planes");
  if (krows < 1) return ast_fail(AST_ERR_ARG, "ast_step_f32: krows %d", krows);
  KTime kt = {nullptr, nullptr, nullptr, 0};
  if (tm != nullptr) {
    if (tm->nslots < 1 || tm->nslots > kMaxSlots || tm->cur == nullptr)
      return ast_fail(AST_ERR_ARG, "ast_step_f32: %d dt slots", tm->nslots);
    kt = KTime{tm->prev, tm->cur, tm->rst, tm->nslots};
  }
  if (!strip_fits(*g, krows))
    return ast_fail(AST_ERR_SIZE, "ast_step_f32: %d columns x %d rows per strip exceed one buffer range",
                    g->N, krows + 4);
  rc = check_positive_bc(*g);
  if (rc != AST_OK) return rc;
  const Consts c = make_consts(*g);
  const ExactCtx ex = make_exact_ctx(g->gamma, mlp);
  const bool window = !(g->ho_lo == -INFINITY && g->ho_hi == INFINITY);
  hipStream_t st = (hipStream_t)stream;
  const float2* S = (const float2*)s;
  const float2* U = (const float2*)u;
  float2* S2 = (float2*)s2;
  float2* U2 = (float2*)u2;
  switch (method) {
    case AST_HLL: return dispatch_geom<AST_HLL>(*g, c, ex, window, S, U, S2, U2, T, T_do, kt, stats, krows, st);
    case AST_HLLC: return dispatch_geom<AST_HLLC>(*g, c, ex, window, S, U, S2, U2, T, T_do, kt, stats, krows, st);
    case AST_HLLC_A: return dispatch_geom<AST_HLLC_A>(*g, c, ex, window, S, U, S2, U2, T, T_do, kt, stats, krows, st);
    case AST_EXACT:  // the learned start only where it can run (no window: full waves)
      if (mlp != nullptr && !window && tm != nullptr && tm->gate != nullptr) {
        ExactCtx e0 = ex, e1 = ex;
        e0.gate = e1.gate = tm->prev ? reinterpret_cast<const uint32_t*>(tm->prev + 1) : tm->gate;
        e0.strong = e1.strong = reinterpret_cast<uint32_t*>(tm->cur + 1);
        e0.gate_mode = 0;
        e1.gate_mode = 1;
        rc = dispatch_geom<kExactDetect>(*g, c, e0, window, S, U, S2, U2, T, T_do, kt, stats, krows, st);
        if (rc != AST_OK) return rc;
        return dispatch_geom<kExactLearned>(*g, c, e1, window, S, U, S2, U2, T, T_do, kt, stats, krows, st);
      }
      if (mlp != nullptr && !window)
        return dispatch_geom<kExactLearned>(*g, c, ex, window, S, U, S2, U2, T, T_do, kt, stats, krows, st);
      return dispatch_geom<AST_EXACT>(*g, c, ex, window, S, U, S2, U2, T, T_do, kt, stats, krows, st);
    default: return ast_fail(AST_ERR_ARG, "ast_step_f32: unknown method %d", method);
  }
}

extern "C" int ast_step_f32(const ast_grid_t* g, const float* s, const float* u, float* s2,
                            float* u2, double* T, const double* T_do, double* TT, int method,
                            const ast_mlp_t* mlp, ast_stats_t* stats, void* stream) {
  (void)TT;
  return ast_step_f32_rows(g, s, u, s2, u2, T, T_do, method, mlp, stats, g_krows, nullptr, stream);
}
