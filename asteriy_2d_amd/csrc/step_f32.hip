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
//
// This source holds the host entry point; the kernels are instantiated per
// method in step_f32_hll.hip, step_f32_exact.hip, step_f32_learned.hip and
// step_f32_detect.hip (step_f32_kernel.hpp).
#include "step_f32_kernel.hpp"
#include "step_sources.h"

using namespace ast;
using namespace ast::f32;

static int g_krows = 8;

// krows > 0: rows per strip; krows = -S: S strips of (nearly) equal height
// per band (strip_rows)
extern "C" int ast_set_tile_rows(int krows) {
  if (krows == 0 || krows > 1024 || krows < -(1 << 20))
    return ast_fail(AST_ERR_ARG, "ast_set_tile_rows: %d", krows);
  g_krows = krows;
  return AST_OK;
}
extern "C" int ast_get_tile_rows(void) { return g_krows; }

// tm (may be null): the runtime's time protocol (time_slots.h).  Null: the
// reference protocol (the step lowers *T, dt is *T_do, a learned guess is
// always on).  Otherwise the step lowers the slots of tm->cur and, with
// tm->prev, takes its dt from the previous step's slots and advances the
// time itself.  With the exact method, a learned guess, no window and
// tm->strong, the runtime has chosen the mode of the step: the solver with
// the learned start (tm->learned) or the plain solver; both flag strong
// two-shock faces in *tm->strong.
int ast_step_f32_rows(const ast_grid_t* g, const float* s, const float* u, float* s2, float* u2,
                      double* T, const double* T_do, const double* TT, int method,
                      const ast_mlp_t* mlp, ast_stats_t* stats, int krows, const StepTime* tm,
                      const StepSources* src, void* stream) {
  int rc = ast_grid_check(g);
  if (rc != AST_OK) return rc;
  if (!s || !u || !s2 || !u2 || !T || !T_do)
    return ast_fail(AST_ERR_ARG, "ast_step_f32: null state or time pointer");
  if (s == s2 || u == u2 || s == u2 || u == s2)
    return ast_fail(AST_ERR_ARG, "ast_step_f32: output planes must not alias the input planes");
  if (krows == 0 || krows > 1024 || (krows < 0 && -(long long)krows > g->M))
    return ast_fail(AST_ERR_ARG, "ast_step_f32: strip tiling %d for %d rows", krows, g->M);
  const int height = tallest_strip(g->M, krows);
  KTime kt = {nullptr, nullptr, nullptr, 0};
  if (tm != nullptr) {
    if (tm->nslots < 1 || tm->nslots > kMaxSlots || tm->cur == nullptr)
      return ast_fail(AST_ERR_ARG, "ast_step_f32: %d dt slots", tm->nslots);
    kt = KTime{tm->prev, tm->cur, tm->rst, tm->nslots};
  }
  if (!strip_fits(*g, height))
    return ast_fail(AST_ERR_SIZE, "ast_step_f32: %d columns x %d rows per strip exceed one buffer range",
                    g->N, height + 4);
  rc = check_positive_bc(*g);
  if (rc != AST_OK) return rc;
  const Consts c = make_consts(*g);
  if (c.pulsed && tm == nullptr && TT == nullptr)
    return ast_fail(AST_ERR_ARG, "ast_step_f32: a pulsed Dirichlet side needs the time TT");
  const ExactCtx ex = make_exact_ctx(g->gamma, mlp);
  const bool window = !(g->ho_lo == -INFINITY && g->ho_hi == INFINITY);
  const bool has_src = src != nullptr && (src->nn2 != nullptr || src->nn3 != nullptr ||
                                          src->xy != nullptr || src->e != nullptr);
  const bool f32_src = has_src && (src->xy != nullptr || src->e != nullptr);
  const SrcF sf = has_src ? SrcF{f32_src ? nullptr : src->nn2, f32_src ? nullptr : src->nn3,
                                 (const float2*)src->xy, src->e, (float)src->kmom,
                                 (float)src->ken}
                          : SrcF{nullptr, nullptr, nullptr, nullptr, 0.f, 0.f};
  StepArgs a{*g, c, ex, window, (const float2*)s, (const float2*)u, (float2*)s2, (float2*)u2,
             T, T_do, kt, stats, krows, (hipStream_t)stream, has_src ? &sf : nullptr,
             c.pulsed && tm == nullptr ? TT : nullptr};
  if (has_src) mlp = nullptr;  // the learned start is not served with source planes
  switch (method) {
    case AST_HLL: return step_launch<AST_HLL>(a);
    case AST_HLLC: return step_launch<AST_HLLC>(a);
    case AST_HLLC_A: return step_launch<AST_HLLC_A>(a);
    case AST_EXACT:  // the learned start only where it can run (no window: full waves)
      if (mlp != nullptr && !window && tm != nullptr && tm->strong != nullptr) {
        a.ex.strong = tm->strong;
        return tm->learned ? step_launch<kExactLearned>(a) : step_launch<kExactDetect>(a);
      }
      if (mlp != nullptr && !window) return step_launch<kExactLearned>(a);
      return step_launch<AST_EXACT>(a);
    default: return ast_fail(AST_ERR_ARG, "ast_step_f32: unknown method %d", method);
  }
}

extern "C" int ast_step_f32(const ast_grid_t* g, const float* s, const float* u, float* s2,
                            float* u2, double* T, const double* T_do, double* TT, int method,
                            const ast_mlp_t* mlp, ast_stats_t* stats, void* stream) {
  return ast_step_f32_rows(g, s, u, s2, u2, T, T_do, TT, method, mlp, stats, g_krows, nullptr,
                           nullptr, stream);
}

extern "C" int ast_step_f32_src(const ast_grid_t* g, const float* s, const float* u, float* s2,
                                float* u2, double* T, const double* T_do, double* TT,
                                int method, const ast_sources_t* src, ast_stats_t* stats,
                                void* stream) {
  const StepSources ss = src ? StepSources{src->nn2, src->nn3, nullptr, nullptr, src->kmom, src->ken}
                             : StepSources{nullptr, nullptr, nullptr, nullptr, 0.0, 0.0};
  return ast_step_f32_rows(g, s, u, s2, u2, T, T_do, TT, method, nullptr, stats, g_krows,
                           nullptr, src ? &ss : nullptr, stream);
}
