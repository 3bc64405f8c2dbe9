// ref_flux.hpp -- device face-flux operators in the reference's own form.
//
// HLL (kernel.cu:1399-1523), HLLC_2d_Korolkov (Gas_Dinamic.cu:1604-1826),
// HLLC_Aleksashov (kernel.cu:767-1079) and the exact Godunov_Solver_Alexashov with newton/devtwo/lev
// (kernel.cu:128-765), written for the GPU with the reference's formulas,
// constants, branch structure and operation order, so that a translation
// unit compiled with -ffp-contract=off reproduces the reference arithmetic
// bit for bit (HLL, HLLC) or to the last ulp of pow (exact solver).
// Templated on the real type R: R=double is the parity path, R=float the
// straight FP32 variant of the batched flux operators.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/asteriy.h"

#define AST_DEV __device__ __forceinline__

namespace ast {

template <typename R> struct V2 { R x, y; };

template <typename R> struct Mth;
template <> struct Mth<double> {
  static AST_DEV double sqrt_(double a) { return ::sqrt(a); }
  static AST_DEV double pow_(double a, double b) { return ::pow(a, b); }
  static AST_DEV double fabs_(double a) { return ::fabs(a); }
  static AST_DEV double fmin_(double a, double b) { return ::fmin(a, b); }
  static AST_DEV double fmax_(double a, double b) { return ::fmax(a, b); }
  static AST_DEV double nan_() { return __longlong_as_double(0x7ff8000000000000LL); }
};
template <> struct Mth<float> {
  static AST_DEV float sqrt_(float a) { return ::sqrtf(a); }
  static AST_DEV float pow_(float a, float b) { return ::powf(a, b); }
  static AST_DEV float fabs_(float a) { return ::fabsf(a); }
  static AST_DEV float fmin_(float a, float b) { return ::fminf(a, b); }
  static AST_DEV float fmax_(float a, float b) { return ::fmaxf(a, b); }
  static AST_DEV float nan_() { return __int_as_float(0x7fc00000); }
};

// gamma-derived constants, Header.h:17-26.
template <typename R> struct Phys {
  R ga, g1, g2, gp, gm, krit;
};
template <typename R>
AST_DEV Phys<R> make_phys(R gamma, R krit) {
  Phys<R> p;
  p.ga = gamma;
  p.g1 = gamma - R(1.0);
  p.g2 = gamma + R(1.0);
  p.gp = (p.g2 / gamma) / R(2.0);
  p.gm = (p.g1 / gamma) / R(2.0);
  p.krit = krit;
  return p;
}

// Per-face solver statistics kept in registers, summed per block.
struct FaceStats {
  uint32_t exact, newton, newton_it, bisect, bisect_it, raref, vacuum, capped, fp64, mlp;
  AST_DEV void clear() {
    exact = newton = newton_it = bisect = bisect_it = raref = vacuum = capped = fp64 = mlp = 0;
  }
};

template <typename R> AST_DEV R my_min(R a, R b) { return (a <= b) ? a : b; }  // kernel.cu:53-63
template <typename R> AST_DEV R my_max(R a, R b) { return (a >= b) ? a : b; }  // kernel.cu:65-75

// Gas_Dinamic.cu:858-872
template <typename R> AST_DEV int sign_of(R x) { return (x > 0) ? 1 : ((x < 0) ? -1 : 0); }

// kernel.cu:22-33 (CUDA min(double,double) is fmin)
template <typename R> AST_DEV R minmod(R x, R y) {
  int sx = sign_of(x), sy = sign_of(y);
  if (sx + sy == 0) return R(0.0);
  return (R(sx + sy) / R(2.0)) * Mth<R>::fmin_(Mth<R>::fabs_(x), Mth<R>::fabs_(y));
}

// kernel.cu:35-39
template <typename R>
AST_DEV R linear(R x1, R t1, R x2, R t2, R x3, R t3, R y) {
  R d = minmod((t1 - t2) / (x1 - x2), (t2 - t3) / (x2 - x3));
  return (d * (y - x2) + t2);
}

// kernel.cu:41-50
template <typename R>
AST_DEV void linear2(R x1, R t1, R x2, R t2, R x3, R t3, R y1, R y2, R& A, R& B) {
  R d = minmod((t1 - t2) / (x1 - x2), (t2 - t3) / (x2 - x3));
  A = (d * (y1 - x2) + t2);
  B = (d * (y2 - x2) + t2);
}

// HLL, kernel.cu:1399-1523
template <typename R>
AST_DEV R HLL(const Phys<R>& ph, const V2<R>& Ls, const V2<R>& Lu, const V2<R>& Rs,
              const V2<R>& Ru, R n1, R n2, V2<R>& Ps, V2<R>& Pu, R rad) {
  const R ggg = ph.ga;
  const R Omega = R(0.0);
  R ro_L = Ls.x, p_L = Ls.y, v1_L = Lu.x, v2_L = Lu.y;
  R ro_R = Rs.x, p_R = Rs.y, v1_R = Ru.x, v2_R = Ru.y;
  R Vkv_L = v1_L * v1_L + v2_L * v2_L;
  R Vkv_R = v1_R * v1_R + v2_R * v2_R;
  R c_L = (ro_L <= 0) ? R(0.0) : Mth<R>::sqrt_(ggg * p_L / ro_L);
  R c_R = (ro_R <= 0) ? R(0.0) : Mth<R>::sqrt_(ggg * p_R / ro_R);
  R e_L = p_L / (ggg - R(1.0)) + ro_L * Vkv_L / R(2.0);
  R e_R = p_R / (ggg - R(1.0)) + ro_R * Vkv_R / R(2.0);
  R Vn_L = v1_L * n1 + v2_L * n2;
  R Vn_R = v1_R * n1 + v2_R * n2;
  R D_L = my_min(Vn_L, Vn_R) - my_max(c_L, c_R);
  R D_R = my_max(Vn_L, Vn_R) + my_max(c_L, c_R);
  R t = R(10000000);
  t = my_min(t, ph.krit * rad / my_max(Mth<R>::fabs_(D_L), Mth<R>::fabs_(D_R)));

  R fx1 = ro_L * v1_L;
  R fx2 = ro_L * v1_L * v1_L + p_L;
  R fx3 = ro_L * v1_L * v2_L;
  R fx5 = (e_L + p_L) * v1_L;
  R fy1 = ro_L * v2_L;
  R fy2 = ro_L * v1_L * v2_L;
  R fy3 = ro_L * v2_L * v2_L + p_L;
  R fy5 = (e_L + p_L) * v2_L;
  R fl_1 = fx1 * n1 + fy1 * n2;
  R fl_2 = fx2 * n1 + fy2 * n2;
  R fl_3 = fx3 * n1 + fy3 * n2;
  R fl_5 = fx5 * n1 + fy5 * n2;
  if (D_L > Omega) {
    Ps.x = fl_1; Pu.x = fl_2; Pu.y = fl_3; Ps.y = fl_5;
    return t;
  }
  R hx1 = ro_R * v1_R;
  R hx2 = ro_R * v1_R * v1_R + p_R;
  R hx3 = ro_R * v1_R * v2_R;
  R hx5 = (e_R + p_R) * v1_R;
  R hy1 = ro_R * v2_R;
  R hy2 = ro_R * v1_R * v2_R;
  R hy3 = ro_R * v2_R * v2_R + p_R;
  R hy5 = (e_R + p_R) * v2_R;
  R fr_1 = hx1 * n1 + hy1 * n2;
  R fr_2 = hx2 * n1 + hy2 * n2;
  R fr_3 = hx3 * n1 + hy3 * n2;
  R fr_5 = hx5 * n1 + hy5 * n2;
  if (D_R < Omega) {
    Ps.x = fr_1; Pu.x = fr_2; Pu.y = fr_3; Ps.y = fr_5;
    return t;
  }
  R dq1 = ro_R - ro_L;
  R dq2 = ro_R * v1_R - ro_L * v1_L;
  R dq3 = ro_R * v2_R - ro_L * v2_L;
  R dq5 = e_R - e_L;
  Ps.x = (D_R * fl_1 - D_L * fr_1 + D_L * D_R * dq1) / (D_R - D_L);
  Pu.x = (D_R * fl_2 - D_L * fr_2 + D_L * D_R * dq2) / (D_R - D_L);
  Pu.y = (D_R * fl_3 - D_L * fr_3 + D_L * D_R * dq3) / (D_R - D_L);
  Ps.y = (D_R * fl_5 - D_L * fr_5 + D_L * D_R * dq5) / (D_R - D_L);
  return t;
}

// HLLC_2d_Korolkov (Gas_Dinamic.cu:1814-1826) -> HLLC_2d_Korolkov_b_s
// (Gas_Dinamic.cu:1604-1811) with Q=1, W=0, metod=2, nul_potok=false.  The
// passive scalar and the HLL-average state UZ* do not reach the outputs under
// these arguments and are not computed.
template <typename R>
AST_DEV R HLLC(const Phys<R>& ph, const V2<R>& Ls, const V2<R>& Lu, const V2<R>& Rs,
               const V2<R>& Ru, R n1, R n2, V2<R>& Ps, V2<R>& Pu, R rad) {
  const R ggg = ph.ga, g1 = ph.g1;
  const R W = R(0.0);
  R ro_L = Ls.x, p_L = Ls.y, v1_L = Lu.x, v2_L = Lu.y;
  R ro_R = Rs.x, p_R = Rs.y, v1_R = Ru.x, v2_R = Ru.y;
  R t1 = n2, t2 = -n1;
  R u1 = v1_L * n1 + v2_L * n2;
  R v1 = v1_L * t1 + v2_L * t2;
  R u2 = v1_R * n1 + v2_R * n2;
  R v2 = v1_R * t1 + v2_R * t2;
  R cL = Mth<R>::sqrt_(ggg * p_L / ro_L);
  R cR = Mth<R>::sqrt_(ggg * p_R / ro_R);
  R uu_L = ((v1_L * v1_L) + (v2_L * v2_L)) / R(2.0);
  R uu_R = ((v1_R * v1_R) + (v2_R * v2_R)) / R(2.0);
  R SL = Mth<R>::fmin_(u1, u2) - Mth<R>::fmax_(cL, cR);
  R SR = Mth<R>::fmax_(u1, u2) + Mth<R>::fmax_(cL, cR);
  R suR = (SR - u2);
  R suL = (SL - u1);
  R SM = (suR * ro_R * u2 - suL * ro_L * u1 - p_R + p_L) / (suR * ro_R - suL * ro_L);
  R UU = Mth<R>::fmax_(Mth<R>::fabs_(SL), Mth<R>::fabs_(SR));
  R time = ph.krit * rad / UU;

  R e1 = p_L / g1 + ro_L * uu_L;
  R e2 = p_R / g1 + ro_R * uu_R;
  R FL0 = ro_L * u1, FL1 = ro_L * u1 * u1 + p_L, FL2 = ro_L * u1 * v1, FL3 = (e1 + p_L) * u1;
  R FR0 = ro_R * u2, FR1 = ro_R * u2 * u2 + p_R, FR2 = ro_R * u2 * v2, FR3 = (e2 + p_R) * u2;
  R UL0 = ro_L, UL1 = ro_L * u1, UL2 = ro_L * v1, UL3 = e1;
  R UR0 = ro_R, UR1 = ro_R * u2, UR2 = ro_R * v2, UR3 = e2;
  R P0, P1, P2, P3;
  if (SL >= W) {
    P1 = n1 * (FL1 - W * UL1) + t1 * (FL2 - W * UL2);
    P2 = n2 * (FL1 - W * UL1) + t2 * (FL2 - W * UL2);
    P0 = FL0 - W * UL0;
    P3 = FL3 - W * UL3;
  } else if (SR <= W) {
    P1 = n1 * (FR1 - W * UR1) + t1 * (FR2 - W * UR2);
    P2 = n2 * (FR1 - W * UR1) + t2 * (FR2 - W * UR2);
    P0 = FR0 - W * UR0;
    P3 = FR3 - W * UR3;
  } else {
    R ro_LL = ro_L * (SL - u1) / (SL - SM);
    R ro_RR = ro_R * (SR - u2) / (SR - SM);
    R suRm = suR / (SR - SM);
    R suLm = suL / (SL - SM);
    R ptzR = p_R + ro_R * suR * (SM - u2);
    R ptzL = p_L + ro_L * suL * (SM - u1);
    R ptz = (ptzR + ptzL) / R(2.0);
    R vLL = v1, vRR = v2;
    R ee2 = e2 * suRm + (ptz * SM - p_R * u2) / (SR - SM);
    R ee1 = e1 * suLm + (ptz * SM - p_L * u1) / (SL - SM);
    R PO0, PO1, PO2, PO3;
    if (SL < W && SM >= W) {
      R U0 = ro_LL, U1 = ro_LL * SM, U2 = ro_LL * vLL, U3 = ee1;
      PO0 = FL0 + SL * U0 - SL * UL0 - W * U0;
      PO1 = FL1 + SL * U1 - SL * UL1 - W * U1;
      PO2 = FL2 + SL * U2 - SL * UL2 - W * U2;
      PO3 = FL3 + SL * U3 - SL * UL3 - W * U3;
    } else if (SR > W && SM < W) {
      R U0 = ro_RR, U1 = ro_RR * SM, U2 = ro_RR * vRR, U3 = ee2;
      PO0 = FR0 + SR * U0 - SR * UR0 - W * U0;
      PO1 = FR1 + SR * U1 - SR * UR1 - W * U1;
      PO2 = FR2 + SR * U2 - SR * UR2 - W * U2;
      PO3 = FR3 + SR * U3 - SR * UR3 - W * U3;
    } else {
      PO0 = PO1 = PO2 = PO3 = Mth<R>::nan_();
    }
    P1 = n1 * PO1 + t1 * PO2;
    P2 = n2 * PO1 + t2 * PO2;
    P0 = PO0;
    P3 = PO3;
  }
  Ps.x = P0;  // output map, Gas_Dinamic.cu:1821-1824
  Ps.y = P3;
  Pu.x = P1;
  Pu.y = P2;
  return time;
}

// HLLC_Aleksashov (kernel.cu:767-1079) as add_MK calls it: zero magnetic
// field and out-of-plane velocity, so every term carrying them is an exact
// zero and is left out -- the frame's else branch (kernel.cu:819-861) gives
// the in-plane tangent (-n2, n1), the fast speeds reduce to
// sqrt(c*(c + 2*0)), pt = p, the HLL state feeds only B terms and the star
// tangential velocities are the sides' own (kernel.cu:1009-1019).  Wave
// speeds min(vL - cL, vR - cR), max(vL + cL, vR + cR) (kernel.cu:895-896),
// averaged ptz in the star energy (kernel.cu:981-1007), wv = 0.
template <typename R>
AST_DEV R HLLC_A(const Phys<R>& ph, const V2<R>& Ls, const V2<R>& Lu, const V2<R>& Rs,
                 const V2<R>& Ru, R n1, R n2, V2<R>& Ps, V2<R>& Pu, R rad) {
  typedef Mth<R> M;
  const R ga = ph.ga, g1 = ph.g1;
  const R r1 = Ls.x, p1 = Ls.y, u1 = Lu.x, v1 = Lu.y;
  const R r2 = Rs.x, p2 = Rs.y, u2 = Ru.x, v2 = Ru.y;
  const R vL0 = n1 * u1 + n2 * v1, vL2 = -n2 * u1 + n1 * v1;  // kernel.cu:861-867
  const R vR0 = n1 * u2 + n2 * v2, vR2 = -n2 * u2 + n1 * v2;
  const R cL = M::sqrt_(ga * p1 / r1);  // kernel.cu:869-891
  R qp = M::sqrt_(cL * (cL + R(2.0) * R(0.0))), qm = M::sqrt_(cL * (cL - R(2.0) * R(0.0)));
  const R cfL = (qp + qm) / R(2.0);
  const R cR = M::sqrt_(ga * p2 / r2);
  qp = M::sqrt_(cR * (cR + R(2.0) * R(0.0)));
  qm = M::sqrt_(cR * (cR - R(2.0) * R(0.0)));
  const R cfR = (qp + qm) / R(2.0);
  const R ptL = p1, ptR = p2;
  const R SL = M::fmin_(vL0 - cfL, vR0 - cfR);
  const R SR = M::fmax_(vL0 + cfL, vR0 + cfR);
  const R suR = SR - vR0;
  const R suL = SL - vL0;
  const R SM = (suR * r2 * vR0 - ptR + ptL - suL * r1 * vL0) / (suR * r2 - suL * r1);
  const R UU = M::fmax_(M::fabs_(SL), M::fabs_(SR));
  const R time = ph.krit * rad / UU;
  const R upt1 = (u1 * u1 + v1 * v1) / R(2.0);  // kernel.cu:928-975
  const R upt2 = (u2 * u2 + v2 * v2) / R(2.0);
  const R e1 = p1 / g1 + r1 * upt1;
  const R e2 = p2 / g1 + r2 * upt2;
  // mass, normal and tangential momentum, energy
  const R FL[4] = {r1 * vL0, r1 * vL0 * vL0 + ptL, r1 * vL0 * vL2, (e1 + ptL) * vL0};
  const R FR[4] = {r2 * vR0, r2 * vR0 * vR0 + ptR, r2 * vR0 * vR2, (e2 + ptR) * vR0};
  const R UL[4] = {r1, r1 * vL0, r1 * vL2, e1};
  const R UR[4] = {r2, r2 * vR0, r2 * vR2, e2};
  const R suRm = suR / (SR - SM);  // kernel.cu:977-1031
  const R suLm = suL / (SL - SM);
  const R rzR = r2 * suRm;
  const R rzL = r1 * suLm;
  const R ptzR = ptR + r2 * suR * (SM - vR0);
  const R ptzL = ptL + r1 * suL * (SM - vL0);
  const R ptz = (ptzR + ptzL) / R(2.0);
  const R ezR = e2 * suRm + (ptz * SM - ptR * vR0) / (SR - SM);
  const R ezL = e1 * suLm + (ptz * SM - ptL * vL0) / (SL - SM);
  const R UZL[4] = {rzL, SM * rzL, vL2 * rzL, ezL};
  const R UZR[4] = {rzR, SM * rzR, vR2 * rzR, ezR};
  R P[4];
  if (SL > R(0.0)) {  // kernel.cu:1033-1068
    for (int k = 0; k < 4; k++) P[k] = FL[k];
  } else if (SL <= R(0.0) && SM >= R(0.0)) {
    for (int k = 0; k < 4; k++) P[k] = FL[k] + SL * (UZL[k] - UL[k]);
  } else if (SM <= R(0.0) && SR >= R(0.0)) {
    for (int k = 0; k < 4; k++) P[k] = FR[k] + SR * (UZR[k] - UR[k]);
  } else if (SR < R(0.0)) {
    for (int k = 0; k < 4; k++) P[k] = FR[k];
  } else {  // NaN speeds
    for (int k = 0; k < 4; k++) P[k] = M::nan_();
  }
  Ps.x = P[0];
  Ps.y = P[3];
  Pu.x = n1 * P[1] - n2 * P[2];  // kernel.cu:1075-1076
  Pu.y = n2 * P[1] + n1 * P[2];
  return time;
}

// lev, kernel.cu:751-765
template <typename R>
AST_DEV R lev(const Phys<R>& ph, R enI, R pI, R rI, R enII, R pII, R rII, R uuu) {
  R cI = Mth<R>::sqrt_(ph.ga * pI / rI);
  R cII = Mth<R>::sqrt_(ph.ga * pII / rII);
  R fI = (uuu - pI) / (rI * cI * Mth<R>::sqrt_(ph.gp * uuu / pI + ph.gm));
  R fII = R(2.0) / ph.g1 * cII * (Mth<R>::pow_((uuu / pII), ph.gm) - R(1.0));
  R f1 = fI + fII;
  R f2 = enI - enII;
  return f1 - f2;
}

// newton, kernel.cu:635-680.  `pn0` replaces the starting point when
// use_guess is set; the stopping scale pee stays the PVRS-type estimate of
// kernel.cu:641-644.  Capped at AST_ITER_CAP iterations.
template <typename R>
AST_DEV R newton(const Phys<R>& ph, R enI, R pI, R rI, R enII, R pII, R rII, bool use_guess,
                 R pn0, R tol, int& iters, bool& capped) {
  const R ga = ph.ga, gp = ph.gp, gm = ph.gm;
  R cI = Mth<R>::sqrt_(ga * pI / rI);
  R cII = Mth<R>::sqrt_(ga * pII / rII);
  R pn = pI * rII * cII + pII * rI * cI + (enI - enII) * rI * cI * rII * cII;
  pn = pn / (rI * cI + rII * cII);
  R pee = pn;
  if (use_guess) pn = pn0;
  int kiter = 0;
  R p;
  for (;;) {
    p = pn;
    kiter = kiter + 1;
    R fI = (p - pI) / (rI * cI * Mth<R>::sqrt_(gp * p / pI + gm));
    R fIs = (ga + R(1.0)) * p / pI + (R(3.0) * ga - R(1.0));
    fIs = fIs / (R(4.0) * ga * rI * cI * Mth<R>::pow_((gp * p / pI + gm), (R(3.0) / R(2.0))));
    R fII = (p - pII) / (rII * cII * Mth<R>::sqrt_(gp * p / pII + gm));
    R fIIs = (ga + R(1.0)) * p / pII + (R(3.0) * ga - R(1.0));
    fIIs = fIIs / (R(4.0) * ga * rII * cII * Mth<R>::pow_((gp * p / pII + gm), (R(3.0) / R(2.0))));
    pn = p - (fI + fII - (enI - enII)) / (fIs + fIIs);
    if (!(Mth<R>::fabs_(pn / pee - p / pee) >= tol)) break;
    if (kiter >= AST_ITER_CAP) { capped = true; break; }
  }
  iters = kiter;
  return pn;
}

// devtwo, kernel.cu:682-749: bisection on [pI, pII] until |lev| <= epsil.
template <typename R>
AST_DEV R devtwo(const Phys<R>& ph, R enI, R pI, R rI, R enII, R pII, R rII, R epsil,
                 int& iters, bool& capped) {
  R kl = pI, kp = pII, kc;
  R ksi = lev(ph, enI, pI, rI, enII, pII, rII, kl);
  R ksir = lev(ph, enI, pI, rI, enII, pII, rII, kp);
  iters = 0;
  if (Mth<R>::fabs_(ksi) <= epsil) return kl;
  if (Mth<R>::fabs_(ksir) <= epsil) return kp;
  int kpizd = 0;
  for (;;) {
    kpizd = kpizd + 1;
    kc = (kl + kp) / R(2.0);
    R ksit = lev(ph, enI, pI, rI, enII, pII, rII, kc);
    if (Mth<R>::fabs_(ksit) <= epsil) break;
    if ((ksi * ksit) <= R(0.0)) {
      kp = kc;
    } else {
      kl = kc;
      ksi = ksit;
    }
    if (kpizd >= AST_ITER_CAP) { capped = true; break; }
  }
  iters = kpizd;
  return kc;
}

// Godunov_Solver_Alexashov, kernel.cu:128-570: the exact solution sampled at
// the face (w = 0), then the face flux.  `guess`/`use_guess`: optional
// starting point of newton (the learned guess).
template <typename R>
AST_DEV R GODUNOV(const Phys<R>& ph, const V2<R>& Ls, const V2<R>& Lu, const V2<R>& Rs,
                  const V2<R>& Ru, R n1, R n2, V2<R>& Ps, V2<R>& Pu, R rad, FaceStats& st,
                  bool use_guess = false, R guess = R(0)) {
  typedef Mth<R> F;
  const R ga = ph.ga, g1 = ph.g1, g2 = ph.g2, gm = ph.gm;
  const R gg1 = g1, gg2 = g2, gga = ga;
  const R krit = ph.krit;
  const R eps = R(10e-10), eps8 = R(10e-8);  // Header.h:62-63
  R w = R(0.0);
  R al = n1, be = n2, ge = R(0.0);
  R time = R(0.0);
  R al2 = -n2, be2 = n1, ge2 = R(0.0);
  R al3 = R(0.0), be3 = R(0.0), ge3 = R(1.0);
  R enI = al * Lu.x + be * Lu.y;
  R teI2 = al2 * Lu.x + be2 * Lu.y;
  R teI3 = al3 * Lu.x + be3 * Lu.y;
  R enII = al * Ru.x + be * Ru.y;
  R teII2 = al2 * Ru.x + be2 * Ru.y;
  R teII3 = al3 * Ru.x + be3 * Ru.y;
  R pI = Ls.y, pII = Rs.y, rI = Ls.x, rII = Rs.x;
  int ipiz = 0;
  if (pI > pII) {
    R eno2 = enII, teo22 = teII2, teo23 = teII3, p2 = pII, r2 = rII;
    R eno1 = enI, teo12 = teI2, teo13 = teI3, p1 = pI, r1 = rI;
    enI = -eno2; teI2 = teo22; teI3 = teo23; pI = p2; rI = r2;
    enII = -eno1; teII2 = teo12; teII3 = teo13; pII = p1; rII = r1;
    w = -w;
    ipiz = 1;
  }
  R cI = R(0.0), cII = R(0.0);
  if (rI != R(0.0)) cI = F::sqrt_(ga * pI / rI);
  if (rII != R(0.0)) cII = F::sqrt_(ga * pII / rII);
  R a = F::sqrt_(rI * (g2 * pII + g1 * pI) / R(2.0));
  R Uud = (pII - pI) / a;
  R Urz = R(-2.0) * cII / g1 * (R(1.0) - F::pow_((pI / pII), gm));
  R Uvk = R(-2.0) * (cII + cI) / g1;
  R Udf = enI - enII;

  int pat = 9;  // -1 vacuum, 0 two rarefactions, 1 shock+rarefaction, 2 two shocks
  R p = F::nan_(), r = F::nan_(), te2 = F::nan_(), te3 = F::nan_(), en = F::nan_();
  int iters = 0;
  bool capped = false;
  st.exact += 1;
  if (Udf < Uvk) {
    pat = -1;
    st.vacuum += 1;
  } else if ((Udf >= Uvk) && (Udf <= Urz)) {
    p = pI * F::pow_(((Udf - Uvk) / (Urz - Uvk)), (R(1.0) / gm));
    pat = 0;
    st.raref += 1;
  } else if ((Udf > Urz) && (Udf <= Uud)) {
    p = devtwo(ph, enI, pI, rI, enII, pII, rII, R(10e-10), iters, capped);
    pat = 1;
    st.bisect += 1;
    st.bisect_it += iters;
  } else if (Udf > Uud) {
    p = newton(ph, enI, pI, rI, enII, pII, rII, use_guess, guess, eps, iters, capped);
    pat = 2;
    st.newton += 1;
    st.newton_it += iters;
  }
  st.capped += capped ? 1 : 0;

  if (pat == 2) {  // TWO SHOCKS, kernel.cu:234-288
    R aI = F::sqrt_(rI * (g2 / R(2.0) * p + g1 / R(2.0) * pI));
    R aII = F::sqrt_(rII * (g2 / R(2.0) * p + g1 / R(2.0) * pII));
    R u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
    R dI = enI - aI / rI;
    R dII = enII + aII / rII;
    R UU = F::fmax_(F::fabs_(dI), F::fabs_(dII));
    time = (UU > eps8) ? krit * rad / UU : krit * rad / eps8;
    if (w <= dI) {
      en = enI; p = pI; r = rI; te2 = teI2; te3 = teI3;
    } else if ((w > dI) && (w <= u)) {
      en = u; r = rI * aI / (aI - rI * (enI - u)); te2 = teI2; te3 = teI3;
    } else if ((w > u) && (w < dII)) {
      en = u; r = rII * aII / (aII + rII * (enII - u)); te2 = teII2; te3 = teII3;
    } else if (w >= dII) {
      en = enII; p = pII; r = rII; te2 = teII2; te3 = teII3;
    }
  } else if (pat == 1) {  // LEFT SHOCK, RIGHT FAN, kernel.cu:292-364
    R aI = F::sqrt_(rI * (g2 / R(2.0) * p + g1 / R(2.0) * pI));
    R aII;
    if (F::fabs_(p - pII) < eps) aII = rII * cII;
    else aII = gm * rII * cII * (R(1.0) - p / pII) / (R(1.0) - F::pow_((p / pII), gm));
    R u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
    R dI = enI - aI / rI;
    R dII = enII + cII;
    R ddII = u + cII - g1 * (enII - u) / R(2.0);
    R UU = F::fmax_(F::fabs_(dI), F::fabs_(dII));
    UU = F::fmax_(UU, F::fabs_(ddII));
    time = (UU > eps8) ? krit * rad / UU : krit * rad / eps8;
    if (w <= dI) { en = enI; p = pI; r = rI; te2 = teI2; te3 = teI3; }
    if ((w > dI) && (w <= u)) {
      en = u; r = rI * aI / (aI - rI * (enI - u)); te2 = teI2; te3 = teI3;
    }
    if ((w > u) && (w <= ddII)) {
      R ce = cII - g1 / R(2.0) * (enII - u);
      en = u; r = ga * p / ce / ce; te2 = teII2; te3 = teII3;
    }
    if ((w > ddII) && (w < dII)) {
      R ce = -g1 / g2 * (enII - w) + R(2.0) / g2 * cII;
      en = w - ce;
      p = pII * F::pow_((ce / cII), (R(1.0) / gm));
      r = ga * p / ce / ce; te2 = teII2; te3 = teII3;
    }
    if (w >= dII) { en = enII; p = pII; r = rII; te2 = teII2; te3 = teII3; }
  } else if (pat == 0) {  // TWO FANS, kernel.cu:366-477
    R aI, aII;
    if (F::fabs_(p - pI) < eps) aI = rI * cI;
    else aI = gm * rI * cI * (R(1.0) - p / pI) / (R(1.0) - F::pow_((p / pI), gm));
    if (F::fabs_(p - pII) < eps) aII = rII * cII;
    else aII = gm * rII * cII * (R(1.0) - p / pII) / (R(1.0) - F::pow_((p / pII), gm));
    R u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
    R dI = enI - cI;
    R ddI = u - cI - g1 * (enI - u) / R(2.0);
    R dII = enII + cII;
    R ddII = u + cII - g1 * (enII - u) / R(2.0);
    R UU = F::fmax_(F::fabs_(dI), F::fabs_(dII));
    UU = F::fmax_(UU, F::fabs_(ddII));
    UU = F::fmax_(UU, F::fabs_(ddI));
    time = (UU > eps8) ? krit * rad / UU : krit * rad / eps8;
    if (w <= dI) { en = enI; p = pI; r = rI; te2 = teI2; te3 = teI3; }
    if ((w > dI) && (w < ddI)) {
      R ce = g1 / g2 * (enI - w) + R(2.0) / g2 * cI;
      en = w + ce;
      p = pI * F::pow_((ce / cI), (R(1.0) / gm));
      r = ga * p / ce / ce; te2 = teI2; te3 = teI3;
    }
    if ((w >= ddI) && (w <= u)) {
      R ce = cI + g1 / R(2.0) * (enI - u);
      en = u; r = ga * p / ce / ce; te2 = teI2; te3 = teI3;
    }
    if ((w > u) && (w <= ddII)) {
      R ce = cII - g1 / R(2.0) * (enII - u);
      en = u; r = ga * p / ce / ce; te2 = teII2; te3 = teII3;
    }
    if ((w > ddII) && (w < dII)) {
      R ce = -g1 / g2 * (enII - w) + R(2.0) / g2 * cII;
      en = w - ce;
      p = pII * F::pow_((ce / cII), (R(1.0) / gm));
      r = ga * p / ce / ce; te2 = teII2; te3 = teII3;
    }
    if (w >= dII) { en = enII; p = pII; r = rII; te2 = teII2; te3 = teII3; }
  } else if (pat == -1) {  // VACUUM, kernel.cu:480-544
    R dI = enI - cI;
    R ddI = enI + R(2.0) / gg1 * cI;
    R dII = enII + cII;
    R ddII = enII - R(2.0) / gg1 * cII;
    R UU = F::fmax_(F::fabs_(dI), F::fabs_(dII));
    UU = F::fmax_(UU, F::fabs_(ddII));
    UU = F::fmax_(UU, F::fabs_(ddI));
    time = (UU > eps8) ? krit * rad / UU : krit * rad / eps8;
    if (w <= dI) { en = enI; p = pI; r = rI; te2 = teI2; te3 = teI3; }
    if ((w > dI) && (w < ddI)) {
      R ce = gg1 / gg2 * (enI - w) + R(2.0) / gg2 * cI;
      en = w + ce;
      p = pI * F::pow_((ce / cI), (R(1.0) / gm));
      r = gga * p / ce / ce; te2 = teI2; te3 = teI3;
    }
    if ((w >= ddI) && (w <= ddII)) {
      en = w; p = R(0.0); r = R(0.0); te2 = R(0.0); te3 = R(0.0);
    }
    if ((w > ddII) && (w < dII)) {
      R ce = -gg1 / gg2 * (enII - w) + R(2.0) / gg2 * cII;
      en = w - ce;
      p = pII * F::pow_((ce / cII), (R(1.0) / gm));
      r = gga * p / ce / ce; te2 = teII2; te3 = teII3;
    }
    if (w >= dII) { en = enII; p = pII; r = rII; te2 = teII2; te3 = teII3; }
  }
  if (ipiz == 1) {
    en = -en;
    w = -w;
  }
  R uo = al * en + al2 * te2 + al3 * te3;
  R vo = be * en + be2 * te2 + be3 * te3;
  R wo = ge * en + ge2 * te2 + ge3 * te3;
  R eo = p / g1 + R(0.5) * r * (uo * uo + vo * vo + wo * wo);
  en = al * uo + be * vo + ge * wo;
  Ps.x = (r * (en - w));
  Pu.x = (r * (en - w) * uo + al * p);
  Pu.y = (r * (en - w) * vo + be * p);
  Ps.y = ((en - w) * eo + en * p);
  return time;
}

}  // namespace ast
