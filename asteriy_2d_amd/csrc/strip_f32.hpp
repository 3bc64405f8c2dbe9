// strip_f32.hpp -- the "wave strip" of the FP32 step, shared by the
// single-grid step (step_f32.hip) and the batched persistent launch
// (batch.hip).  See step_f32.hip for the scheme.
#pragma once
#include "common.hpp"
#include "exact_f32.hpp"
#include "fast_flux.hpp"

namespace ast {
namespace f32 {

constexpr int WOUT = 60;  // output columns per wave
constexpr int WPB = 4;    // waves per workgroup (stacked strips)
constexpr int NT = 64 * WPB;

struct Q {
  float r, p, vx, vy;
};

// lane i <- lane i-1 (lane 0 keeps its own value)
AST_DEV float shr1(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138,
                                                    0xf, 0xf, false));
}
// lane i <- lane i+1 (lane 63 keeps its own value)
AST_DEV float shl1(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130,
                                                    0xf, 0xf, false));
}
AST_DEV Q shr1(const Q& q) { return Q{shr1(q.r), shr1(q.p), shr1(q.vx), shr1(q.vy)}; }
AST_DEV Q shl1(const Q& q) { return Q{shl1(q.r), shl1(q.p), shl1(q.vx), shl1(q.vy)}; }
AST_DEV FluxF shl1(const FluxF& f) { return FluxF{shl1(f.m), shl1(f.e), shl1(f.un), shl1(f.ut)}; }

// half the minmod slope, 0.5*minmod(a, b) (kernel.cu:22-33): minmod(a, b) is
// the median of {a, b, 0} -- one v_med3_f32.
AST_DEV float hmm(float a, float b) { return 0.5f * __builtin_amdgcn_fmed3f(a, b, 0.f); }

// Face values at +-1/2 of a cell from its half slopes, with the positivity
// fallback of kernel.cu:2706-2788 (rho, p <= 0 -> the cell value).
AST_DEV void face_values(const Q& c, const Q& h, Q& plus, Q& minus) {
  plus.r = c.r + h.r;
  minus.r = c.r - h.r;
  plus.p = c.p + h.p;
  minus.p = c.p - h.p;
  plus.vx = c.vx + h.vx;
  minus.vx = c.vx - h.vx;
  plus.vy = c.vy + h.vy;
  minus.vy = c.vy - h.vy;
  if (plus.r <= 0.f) plus.r = c.r;
  if (minus.r <= 0.f) minus.r = c.r;
  if (plus.p <= 0.f) plus.p = c.p;
  if (minus.p <= 0.f) minus.p = c.p;
}

AST_DEV Q half_slopes(const Q& a, const Q& c, const Q& b) {
  return Q{hmm(c.r - a.r, b.r - c.r), hmm(c.p - a.p, b.p - c.p), hmm(c.vx - a.vx, b.vx - c.vx),
           hmm(c.vy - a.vy, b.vy - c.vy)};
}

// Cell (n, m) in FP32 with the ghost rules; corner/out-of-halo positions are
// clamped into the grid (never read by an output cell, kept finite).
// `interior` (wave-uniform: every lane's column and the row lie inside the
// grid) takes the plain load without any ghost logic.
AST_DEV Q load(const ast_grid_t& g, const float2* __restrict__ s, const float2* __restrict__ u,
               int n, int m, bool cols_in) {
  float2 a, b;
  if (cols_in && m >= 0 && m < g.M) {
    const size_t k = (size_t)m * g.N + n;
    a = s[k];
    b = u[k];
    return Q{a.x, a.y, b.x, b.y};
  }
  const bool xin = (n >= 0) & (n < g.N), yin = (m >= 0) & (m < g.M);
  if ((!xin & !yin) | (n < -2) | (n > g.N + 1) | (m < -2) | (m > g.M + 1)) {
    const int cn = n < 0 ? 0 : (n >= g.N ? g.N - 1 : n);
    const int cm = m < 0 ? 0 : (m >= g.M ? g.M - 1 : m);
    const size_t k = (size_t)cm * g.N + cn;
    a = s[k];
    b = u[k];
  } else {
    fetch_cell(g, s, u, n, m, a, b);
  }
  return Q{a.x, a.y, b.x, b.y};
}

struct Consts {
  float g, ig1, g1, gig1;  // gamma, 1/(gamma-1), gamma-1, gamma/(gamma-1)
  float fx0, iax2, iay2;   // frozen ellipse
  int frozen;
};

static inline Consts make_consts(const ast_grid_t& g) {
  Consts c;
  c.g = (float)g.gamma;
  c.g1 = (float)(g.gamma - 1.0);
  c.ig1 = (float)(1.0 / (g.gamma - 1.0));
  c.gig1 = (float)(g.gamma / (g.gamma - 1.0));
  c.frozen = g.frozen_ax > 0.0;
  c.fx0 = (float)g.frozen_x0;
  c.iax2 = c.frozen ? (float)(1.0 / (g.frozen_ax * g.frozen_ax)) : 0.f;
  c.iay2 = c.frozen ? (float)(1.0 / (g.frozen_ay * g.frozen_ay)) : 0.f;
  return c;
}

// Strip heights / band counts of one grid.
struct StripShape {
  int nbands, krows, nsg;
};
static inline StripShape strip_shape(const ast_grid_t& g, int krows) {
  StripShape sh;
  sh.krows = krows;
  sh.nbands = (g.N + WOUT - 1) / WOUT;
  const int nstrips = (g.M + krows - 1) / krows;
  sh.nsg = (nstrips + WPB - 1) / WPB;
  return sh;
}

// The strip of one wave given its workgroup's (band, strip group): fills
// StripArgs and runs the fast or the boundary variant.  Returns the lane's
// dt candidate.
template <int METHOD, int GEOM, bool WINDOW>
AST_DEV double strip_of_wave(const ast_grid_t& g, const Consts& c, const ExactCtx& ex,
                             const float2* __restrict__ s, const float2* __restrict__ u,
                             float2* __restrict__ s2, float2* __restrict__ u2, double tdo,
                             int band, int sgrp, int krows, FaceStats& st);

template <int METHOD>
AST_DEV FluxF face_flux(const PrimF& L, const PrimF& R, const Consts& c, float& UU,
                        const ExactCtx& ex, FaceStats& st) {
  if constexpr (METHOD == AST_HLL) return hll_f32(L, R, c.g, c.ig1, UU);
  else if constexpr (METHOD == AST_HLLC) return hllc_f32(L, R, c.g, c.ig1, UU);
  else return exact_f32(L, R, ex, UU, st);
}

// One face in the x (NORMAL_X) or y direction: rotate to the face frame,
// flux, and back.  With WINDOW, faces outside the high-order window use
// HLL on the cell states (kernel.cu:2812, 2838-2839, 2856).
template <int METHOD, bool WINDOW, bool NORMAL_X>
AST_DEV FluxF face(const Q& l, const Q& r, const Q& lc, const Q& rc, bool ho, const Consts& c,
                   float& UU, const ExactCtx& ex, FaceStats& st) {
  PrimF L, R;
  if constexpr (WINDOW) {  // faces diverge here: no wave-level learned guess
    if (!ho) {
      L = NORMAL_X ? PrimF{lc.r, lc.p, lc.vx, lc.vy} : PrimF{lc.r, lc.p, lc.vy, lc.vx};
      R = NORMAL_X ? PrimF{rc.r, rc.p, rc.vx, rc.vy} : PrimF{rc.r, rc.p, rc.vy, rc.vx};
      return hll_f32(L, R, c.g, c.ig1, UU);
    }
    ExactCtx e2 = ex;
    e2.mlp = nullptr;
    L = NORMAL_X ? PrimF{l.r, l.p, l.vx, l.vy} : PrimF{l.r, l.p, l.vy, l.vx};
    R = NORMAL_X ? PrimF{r.r, r.p, r.vx, r.vy} : PrimF{r.r, r.p, r.vy, r.vx};
    return face_flux<METHOD>(L, R, c, UU, e2, st);
  } else {
    L = NORMAL_X ? PrimF{l.r, l.p, l.vx, l.vy} : PrimF{l.r, l.p, l.vy, l.vx};
    R = NORMAL_X ? PrimF{r.r, r.p, r.vx, r.vy} : PrimF{r.r, r.p, r.vy, r.vx};
    return face_flux<METHOD>(L, R, c, UU, ex, st);
  }
}

AST_DEV bool frozen_at(const Consts& c, float x, float y) {
  if (!c.frozen) return false;
  const float d = x - c.fx0;
  return d * d * c.iax2 + y * y * c.iay2 < 1.f;
}

// Can any cell of columns [n0, n1] x rows [m0, m1] lie in the frozen
// ellipse?  (wave-uniform; false lets the wave skip the per-cell tests)
AST_DEV bool frozen_possible(const ast_grid_t& g, int n0, int n1, int m0, int m1) {
  if (!(g.frozen_ax > 0.0)) return false;
  const double x0 = g.x_min + n0 * g.dx, x1 = g.x_min + n1 * g.dx;
  const double y0 = g.y_min + m0 * g.dy, y1 = g.y_min + m1 * g.dy;
  const double ax = g.frozen_ax * 1.01, ay = g.frozen_ay * 1.01;
  return x1 >= g.frozen_x0 - ax && x0 <= g.frozen_x0 + ax && y1 >= -ay && y0 <= ay;
}

struct StripArgs {
  int n, m0, m1;
  bool cols_in, out_col, xvote, ho_x, ho_y;
  float tdV, tdof, dxf, dyf, x, ymin;
};

// One wave's strip: prologue + the row loop.  FAST (wave-uniform): every
// load of the strip lies inside the grid, so it is a plain load with no
// ghost logic; otherwise load() applies the boundary rules.
template <int METHOD, int GEOM, bool WINDOW, bool FAST>
AST_DEV void run_strip(const ast_grid_t& g, const Consts& c, const Consts& cw, const ExactCtx& ex,
                       const float2* __restrict__ s, const float2* __restrict__ u,
                       float2* __restrict__ s2, float2* __restrict__ u2, const StripArgs& a,
                       float& UUx, float& UUy, FaceStats& st) {
  const int n = a.n, m0 = a.m0, m1 = a.m1, N = g.N;
  const float dxf = a.dxf, dyf = a.dyf, x = a.x;
  auto LD = [&](int m) -> Q {
    if constexpr (FAST) {
      const size_t k = (size_t)m * N + n;
      const float2 p = s[k], q = u[k];
      return Q{p.x, p.y, q.x, q.y};
    } else {
      return load(g, s, u, n, m, a.cols_in);
    }
  };
  // prologue: rows m0-2 .. m0+1, the south face of row m0
  const Q qm2 = LD(m0 - 2);
  const Q qm1 = LD(m0 - 1);
  Q q0 = LD(m0);
  Q q1 = LD(m0 + 1);
  Q nvm1, svm1, nv0, sv0;
  face_values(qm1, half_slopes(qm2, qm1, q0), nvm1, svm1);
  face_values(q0, half_slopes(qm1, q0, q1), nv0, sv0);
  float uu;
  FluxF Gs = face<METHOD, WINDOW, false>(nvm1, sv0, qm1, q0, a.ho_y, c, uu, ex, st);
  bool fr0 = frozen_at(cw, x, a.ymin + m0 * dyf);
  const bool frm1 = frozen_at(cw, x, a.ymin + (m0 - 1) * dyf);
  if (a.out_col && !(fr0 && frm1)) UUy = fmaxf(UUy, uu);

  // rows are fetched one iteration ahead of their use (latency hiding)
  Q qnext = LD(m0 + 2);
  size_t k = (size_t)m0 * N + n;
  for (int m = m0; m < m1; m++, k += N) {
    const Q q2 = qnext;
    if (m + 1 < m1) qnext = LD(m + 3);
    Q nv1, sv1;
    face_values(q1, half_slopes(q0, q1, q2), nv1, sv1);
    const float y = a.ymin + m * dyf;
    const bool fr1 = frozen_at(cw, x, y + dyf);
    // north face of row m
    const FluxF Gn = face<METHOD, WINDOW, false>(nv0, sv1, q0, q1, a.ho_y, c, uu, ex, st);
    if (a.out_col && !(fr0 && fr1)) UUy = fmaxf(UUy, uu);
    // west face of row m: left state = east value of lane-1
    const Q ql = shr1(q0), qr = shl1(q0);
    Q ev, wv;
    face_values(q0, half_slopes(ql, q0, qr), ev, wv);
    const Q evl = shr1(ev);
    float ux;
    const FluxF Fw = face<METHOD, WINDOW, true>(evl, wv, ql, q0, a.ho_x, c, ux, ex, st);
    const FluxF Fe = shl1(Fw);
    const bool frl = frozen_at(cw, x - dxf, y);
    if (a.xvote && !(fr0 && frl)) UUx = fmaxf(UUx, ux);

    // outward flux sums, kernel.cu:3027-3032 (x faces: Fe - Fw, y faces: Gn - Gs)
    const float PSx = (Fe.m - Fw.m) * dyf + (Gn.m - Gs.m) * dxf;
    const float PSy = (Fe.e - Fw.e) * dyf + (Gn.e - Gs.e) * dxf;
    const float PUx = (Fe.un - Fw.un) * dyf + (Gn.ut - Gs.ut) * dxf;
    const float PUy = (Fe.ut - Fw.ut) * dyf + (Gn.un - Gs.un) * dxf;
    const float ke = 0.5f * q0.r * (q0.vx * q0.vx + q0.vy * q0.vy);
    float dr = -a.tdV * PSx, dmx = -a.tdV * PUx, dmy = -a.tdV * PUy, dE = -a.tdV * PSy;
    if constexpr (GEOM == AST_AXISYMMETRIC) {
      // KorolkovSolver.cu:13-32: -(T_do/y) v {rho, rho u, rho v, E + p}
      const float ty = a.tdof * frcp(y);
      const float rv = q0.r * q0.vy;
      dr -= ty * rv;
      dmx -= ty * rv * q0.vx;
      dmy -= ty * rv * q0.vy;
      dE -= ty * q0.vy * (c.gig1 * q0.p + ke);
    }
    float r2 = q0.r + dr;
    r2 = (r2 <= 0.f) ? q0.r : r2;  // kernel.cu:3060-3064
    const float ir = frcp(r2);
    float vx2 = (q0.r * q0.vx + dmx) * ir;
    float vy2 = (q0.r * q0.vy + dmy) * ir;
    // p' = (E' - rho'|u'|^2/2)(gamma-1), written so that the O(E) terms
    // cancel exactly: E' - KE' = p/(g-1) + dE - (KE' - KE) with
    // KE' - KE = (m.dm + |dm|^2/2 - KE*dr) / rho'.
    const float mx = q0.r * q0.vx, my = q0.r * q0.vy;
    const float dke = (mx * dmx + my * dmy + 0.5f * (dmx * dmx + dmy * dmy) - ke * dr) * ir;
    float p2 = q0.p + c.g1 * (dE - dke);
    p2 = (p2 <= 0.f) ? 0.000001f : p2;  // kernel.cu:3070-3073
    if (fr0) {  // kernel.cu:2317-2324
      r2 = q0.r; p2 = q0.p; vx2 = q0.vx; vy2 = q0.vy;
    }
    if (a.out_col) {
      s2[k] = make_float2(r2, p2);
      u2[k] = make_float2(vx2, vy2);
    }
    // slide the window
    q0 = q1;
    q1 = q2;
    nv0 = nv1;
    Gs = Gn;
    fr0 = fr1;
  }
}

template <int METHOD, int GEOM, bool WINDOW>
AST_DEV double strip_of_wave(const ast_grid_t& g, const Consts& c, const ExactCtx& ex,
                             const float2* __restrict__ s, const float2* __restrict__ u,
                             float2* __restrict__ s2, float2* __restrict__ u2, double tdo,
                             int band, int sgrp, int krows, FaceStats& st) {
  // wave index made explicitly uniform: rows and the loop stay scalar
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  StripArgs a;
  a.m0 = (sgrp * WPB + wave) * krows;
  a.m1 = min(a.m0 + krows, g.M);
  a.n = band * WOUT + lane - 2;
  a.cols_in = (band * WOUT - 2 >= 0) && (band * WOUT + 61 < g.N);  // wave-uniform
  a.out_col = (lane >= 2) & (lane < 2 + WOUT) & (a.n < g.N);
  a.xvote = (lane >= 2) & (lane <= 2 + WOUT) & (a.n <= g.N);
  a.tdV = (float)(tdo / (g.dx * g.dy));
  a.tdof = (float)tdo;
  a.dxf = (float)g.dx;
  a.dyf = (float)g.dy;
  a.x = (float)(g.x_min + a.n * g.dx);
  a.ymin = (float)g.y_min;
  a.ho_x = high_order(g, g.x_min + a.n * g.dx - 0.5 * g.dx);  // west face window
  a.ho_y = high_order(g, g.x_min + a.n * g.dx);               // y faces: cell x
  float UUx = 0.f, UUy = 0.f;
  Consts cw = c;  // frozen tests only where the wave can meet the region
  cw.frozen = c.frozen &&
              frozen_possible(g, band * WOUT - 3, band * WOUT + 62, a.m0 - 2, a.m0 + krows + 1);
  if (a.m0 < g.M) {
    const bool fast = a.cols_in && a.m0 - 2 >= 0 && a.m1 + 1 < g.M;  // wave-uniform
    if (fast)
      run_strip<METHOD, GEOM, WINDOW, true>(g, c, cw, ex, s, u, s2, u2, a, UUx, UUy, st);
    else
      run_strip<METHOD, GEOM, WINDOW, false>(g, c, cw, ex, s, u, s2, u2, a, UUx, UUy, st);
  }
  // per-face dt = krit*rad/UU (monotone in UU), capped at 1000 (kernel.cu:2530)
  double tmin = (a.out_col && a.m0 < g.M) ? 1000.0 : INFINITY;
  if (UUx > 0.f) tmin = fmin(tmin, g.krit * g.dx / (double)UUx);
  if (UUy > 0.f) tmin = fmin(tmin, g.krit * g.dy / (double)UUy);
  return tmin;
}

}  // namespace f32
}  // namespace ast
