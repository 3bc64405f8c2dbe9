// flux64.hpp -- FP64 face fluxes of the drop-in step and of the batched
// flux operators, in the face frame.
//
// A face sees two primitive states (density, pressure, normal velocity u,
// tangential velocity v) with the normal pointing from the left state to the
// right one, and returns the mass, energy, normal- and tangential-momentum
// fluxes along that normal plus the largest |wave speed| of the face (the
// caller forms dt = krit * rad / that, capped at 1000).  The reference's
// operators (HLL kernel.cu:1399-1523, HLLC_2d_Korolkov Gas_Dinamic.cu:1604-1826,
// HLLC_Aleksashov kernel.cu:767-1079, Godunov_Solver_Alexashov
// kernel.cu:128-765) compute the same quantities from lab-frame states and a
// normal; these are the same Riemann solvers written for a face-centric step:
//
//  * HLL: Davis wave speeds S_L = min(u_L, u_R) - max(c_L, c_R),
//    S_R = max(u_L, u_R) + max(c_L, c_R), upwinding when one of them does
//    not straddle the face, else the HLL average.
//  * HLLC: the same speeds (or the per-side speeds min(u_L - c_L, u_R - c_R),
//    max(u_L + c_L, u_R + c_R) of HLLC_Aleksashov), the contact speed S_M,
//    and the star flux in contact form (rho*_K S_M, rho*_K S_M^2 + p*,
//    rho*_K S_M v_K, S_M (E*_K + p*)), which equals F_K + S_K (U*_K - U_K).
//  * exact: the gamma-law Riemann problem sampled at x/t = 0 (below).
//
// Everything is IEEE double (correctly rounded division and square root).
// The functions compile for the host too, so host code can evaluate the
// same arithmetic.
#pragma once
#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define AST_HD __host__ __device__ __forceinline__
#else
#define AST_HD static inline
#endif

#include "../../include/asteriy.h"

namespace ast {
namespace d64 {

struct PrimD {
  double r, p, u, v;  // density, pressure, normal and tangential velocity
};
struct FluxD {
  double m, e, un, ut;  // mass, energy, normal and tangential momentum
};

// gamma and the combinations the solvers use.  n = 1/gm = 2 gamma/(gamma-1)
// when that is an integer in [1, 16] (5 for gamma = 5/3, 7 for 7/5), else 0.
struct GasD {
  double g, g1, g2, ig1, gp, gm, igm, c2g1, g1g2, c2g2;
  int n;
};

AST_HD GasD make_gas(double gamma) {
  GasD a;
  a.g = gamma;
  a.g1 = gamma - 1.0;
  a.g2 = gamma + 1.0;
  a.ig1 = 1.0 / a.g1;
  a.gp = a.g2 / (2.0 * gamma);
  a.gm = a.g1 / (2.0 * gamma);
  a.igm = 1.0 / a.gm;
  a.c2g1 = 2.0 / a.g1;
  a.g1g2 = a.g1 / a.g2;
  a.c2g2 = 2.0 / a.g2;
  const double k = floor(a.igm + 0.5);
  a.n = (fabs(a.igm - k) <= 1e-12 * a.igm && k >= 1.0 && k <= 16.0) ? (int)k : 0;
  return a;
}

AST_HD double max2(double a, double b) { return a >= b ? a : b; }
AST_HD double min2(double a, double b) { return a <= b ? a : b; }

// Physical flux of one side in the face frame; E = p/(g-1) + rho |u|^2/2.
AST_HD FluxD side_flux(const PrimD& q, double E) {
  FluxD f;
  f.m = q.r * q.u;
  f.un = f.m * q.u + q.p;
  f.ut = f.m * q.v;
  f.e = (E + q.p) * q.u;
  return f;
}
AST_HD double total_energy(const PrimD& q, double ig1) {
  return q.p * ig1 + 0.5 * q.r * (q.u * q.u + q.v * q.v);
}

// Largest sound speed of the two sides (a side with rho <= 0 counts as
// c = 0, as in kernel.cu:1418-1434); sqrt is monotone, so the maximum is
// taken before it.
AST_HD double max_sound_speed(const PrimD& L, const PrimD& R, double g) {
  const double aL = L.r > 0.0 ? g * L.p / L.r : 0.0;
  const double aR = R.r > 0.0 ? g * R.p / R.r : 0.0;
  return sqrt(max2(aL, aR));
}

// ---- HLL ----------------------------------------------------------------
AST_HD FluxD hll(const PrimD& L, const PrimD& R, const GasD& a, double& UU) {
  const double cm = max_sound_speed(L, R, a.g);
  const double SL = min2(L.u, R.u) - cm;
  const double SR = max2(L.u, R.u) + cm;
  UU = max2(fabs(SL), fabs(SR));
  const double EL = total_energy(L, a.ig1), ER = total_energy(R, a.ig1);
  const FluxD fl = side_flux(L, EL);
  if (SL > 0.0) return fl;
  const FluxD fr = side_flux(R, ER);
  if (SR < 0.0) return fr;
  const double w = 1.0 / (SR - SL), sls = SL * SR;
  FluxD h;
  h.m = (SR * fl.m - SL * fr.m + sls * (R.r - L.r)) * w;
  h.un = (SR * fl.un - SL * fr.un + sls * (R.r * R.u - L.r * L.u)) * w;
  h.ut = (SR * fl.ut - SL * fr.ut + sls * (R.r * R.v - L.r * L.v)) * w;
  h.e = (SR * fl.e - SL * fr.e + sls * (ER - EL)) * w;
  return h;
}

// ---- HLLC ---------------------------------------------------------------
// PER_SIDE selects the wave speeds of HLLC_Aleksashov (kernel.cu:895-896);
// S_K - u_K is then formed without the cancellation of S_K next to u_K on a
// cold side: S_L - u_L = min(-c_L, (u_R - u_L) - c_R), S_R - u_R =
// max(c_L - (u_R - u_L), c_R).
//
// SCALAR adds a passive scalar Q (a conserved density, carried at Q/rho):
// its flux *PQ is Q_K u_K outside the waves and Q*_K S_M in the star region,
// Q*_K = Q_K (S_K - u_K)/(S_K - S_M) (Q_LL, Q_RR of Gas_Dinamic.cu:1700-1701;
// the star values of HLLCQ_Korolkov_2D, Gas_Dinamic.cu:305-331).
template <bool PER_SIDE, bool SCALAR = false>
AST_HD FluxD hllc(const PrimD& L, const PrimD& R, const GasD& a, double& UU, double QL = 0.0,
                  double QR = 0.0, double* PQ = nullptr) {
  double SL, SR, wL, wR;  // wave speeds and S_K - u_K
  if (PER_SIDE) {
    const double cL = sqrt(a.g * L.p / L.r), cR = sqrt(a.g * R.p / R.r);
    const double d = R.u - L.u;
    wL = min2(-cL, d - cR);
    wR = max2(cL - d, cR);
    SL = L.u + wL;
    SR = R.u + wR;
  } else {
    const double cm = sqrt(max2(a.g * L.p / L.r, a.g * R.p / R.r));
    SL = min2(L.u, R.u) - cm;
    SR = max2(L.u, R.u) + cm;
    wL = SL - L.u;
    wR = SR - R.u;
  }
  UU = max2(fabs(SL), fabs(SR));
  // contact speed from the two mass fluxes through the outer waves
  const double mL = L.r * wL, mR = R.r * wR;
  const double SM = (mR * R.u - mL * L.u + L.p - R.p) / (mR - mL);
  // the side whose (star) flux crosses the face
  const bool supL = SL >= 0.0, supR = SR <= 0.0;
  const bool left = supL || (!supR && SM >= 0.0);
  const PrimD& K = left ? L : R;
  const double EK = total_energy(K, a.ig1);
  if (supL || supR) {
    if constexpr (SCALAR) *PQ = (left ? QL : QR) * K.u;
    return side_flux(K, EK);
  }
  const double SK = left ? SL : SR, mK = left ? mL : mR, wK = left ? wL : wR;
  const double dK = 1.0 / (SK - SM);
  if constexpr (SCALAR) *PQ = (left ? QL : QR) * wK * dK * SM;
  const double smK = SM - K.u;
  const double q = mK * smK;  // p* - p_K
  const double ps = K.p + q;
  // E*_K (S_K - S_M) = E_K (S_K - u_K) + p_K (S_M - u_K) + (p* - p_K) S_M
  const double Es = (EK * wK + K.p * smK + q * SM) * dK;
  const double rs = mK * dK;  // rho*_K
  FluxD f;
  f.m = rs * SM;
  f.un = f.m * SM + ps;
  f.ut = f.m * K.v;
  f.e = SM * (Es + ps);
  return f;
}

// ---- the exact Riemann solver ------------------------------------------
//
// Sides are ordered so that the low-pressure side comes first (I), the
// high-pressure side second (II), mirroring the normal when they swap.  With
// zI = rho_I c_I, zII = rho_II c_II and the closing speed du = u_I - u_II,
// the pressure functions of the two waves decide the pattern (Toro ch. 4):
//   du < -2 (c_I + c_II)/(g-1)                   vacuum between two fans
//   du <= f_II(p_I)  (two-rarefaction bound)     two rarefactions, closed form
//   du <= f_I(p_II)  (shock-rarefaction bound)   shock into I, fan into II
//   otherwise                                    two shocks
// The shock + rarefaction root is bracketed by [p_I, p_II] and found by
// bisection to |f| <= 1e-9; the two-shock root by Newton from the acoustic
// (PVRS-type) estimate to |dp| < 1e-9 times that estimate -- the reference's
// two iterations and stopping rules (kernel.cu:635-749), kept because
// bisection stops 1e-9 short of the root and everything sampled from p*
// inherits that: another root finder would move the flux by ~1e-9.  For the
// same reason the star velocity and densities are formed from the waves'
// mass fluxes a_K as the reference forms them (u* = (a_I u_I + a_II u_II +
// p_I - p_II)/(a_I + a_II), shock density from a_K, fan density from the
// Riemann invariant).  Both loops stop at AST_ITER_CAP iterations (the
// reference has no bound) and report it.

// What the exact solver did on one face.
struct ExactInfo {
  int pat;  // -1 vacuum, 0 two rarefactions, 1 shock + rarefaction, 2 two shocks, -9 none (NaN)
  int it;   // newton / bisection iterations
  bool capped, guessed;
};

constexpr double kRootTol = 1e-9;  // both stopping rules (Header.h:62, kernel.cu:685, 644)
constexpr double kMinSpeed = 1e-7;  // floor of the dt speed (Header.h:63, kernel.cu:246-253)

// Shock pressure function f_K(p) = (p - pK) sqrt(A_K/(p + B_K)) of Toro
// eq. 4.6, A_K = 2/((g+1) rK), B_K = (g-1)/(g+1) pK, and its derivative.
AST_HD double shock_fn(const GasD& a, double p, double pK, double rK, double& df) {
  const double AK = a.c2g2 / rK, BK = a.g1g2 * pK;
  const double s = sqrt(AK / (p + BK));
  df = s * (1.0 - 0.5 * (p - pK) / (p + BK));
  return (p - pK) * s;
}
// Rarefaction pressure function 2 cK/(g-1) ((p/pK)^gm - 1).
AST_HD double raref_fn(const GasD& a, double p, double pK, double cK) {
  return a.c2g1 * cK * (pow(p / pK, a.gm) - 1.0);
}

// x^n for a small integer n, and 1 + x + ... + x^(n-1).
AST_HD double ipow(double x, int n) {
  double r = x;
  for (int k = 1; k < n; k++) r *= x;
  return r;
}
AST_HD double geom_sum(double x, int n) {
  double s = 1.0;
  for (int k = 1; k < n; k++) s = fma(s, x, 1.0);
  return s;
}
// x^(1/gm): products for an integer 1/gm, else pow.
AST_HD double pow_igm(const GasD& a, double x) { return a.n ? ipow(x, a.n) : pow(x, a.igm); }

// Two shocks: Newton on f_I(p) + f_II(p) - du from the acoustic estimate,
// or from `guess` (finite, > 0); a guess and its iterates are kept at or
// above p_II, where the root lies and from where Newton on this concave,
// increasing function climbs to it monotonically.  The stopping scale stays
// the acoustic estimate, so a guess changes the iteration count only.
AST_HD double two_shock_root(const GasD& a, double du, double pI, double rI, double zI,
                             double pII, double rII, double zII, double guess, bool use_guess,
                             int& it, bool& capped) {
  const double pee = (pI * zII + pII * zI + du * zI * zII) / (zI + zII);
  const double plo = use_guess ? pII : -INFINITY;
  double pn = use_guess ? max2(guess, pII) : pee;
  double p;
  int k = 0;
  for (;;) {
    p = pn;
    k++;
    double dI, dII;
    const double fI = shock_fn(a, p, pI, rI, dI);
    const double fII = shock_fn(a, p, pII, rII, dII);
    pn = max2(p - (fI + fII - du) / (dI + dII), plo);
    if (!(fabs(pn - p) >= kRootTol * pee)) break;  // also ends on NaN
    if (k >= AST_ITER_CAP) {
      capped = true;
      break;
    }
  }
  it = k;
  return pn;
}

// Shock into I, rarefaction into II: bisection of
// f(p) = f_I(p) + f_II(p) - du on [p_I, p_II] until |f| <= 1e-9.
AST_HD double shock_raref_root(const GasD& a, double du, double pI, double rI, double pII,
                               double cII, int& it, bool& capped) {
  auto f = [&](double p) {
    double d;
    return shock_fn(a, p, pI, rI, d) + raref_fn(a, p, pII, cII) - du;
  };
  double lo = pI, hi = pII;
  double flo = f(lo);
  it = 0;
  if (fabs(flo) <= kRootTol) return lo;
  if (fabs(f(hi)) <= kRootTol) return hi;
  double mid;
  int k = 0;
  for (;;) {
    k++;
    mid = 0.5 * (lo + hi);
    const double fm = f(mid);
    if (fabs(fm) <= kRootTol) break;
    if (flo * fm <= 0.0) {
      hi = mid;
    } else {
      lo = mid;
      flo = fm;
    }
    if (k >= AST_ITER_CAP) {
      capped = true;
      break;
    }
  }
  it = k;
  return mid;
}

// Mass flux of a rarefaction from pK to p: gm rK cK (1 - x)/(1 - x^gm) with
// x = p/pK, given x^gm = t.  For an integer 1/gm = n this is
// gm rK cK (1 + t + ... + t^(n-1)) (no cancellation as x -> 1); else the
// quotient, whose x -> 1 limit rK cK is taken within 1e-12 of it.
AST_HD double raref_mass_flux(const GasD& a, double rK, double cK, double x, double t) {
  if (a.n) return a.gm * rK * cK * geom_sum(t, a.n);
  if (fabs(1.0 - x) < 1e-12) return rK * cK;
  return a.gm * rK * cK * (1.0 - x) / (1.0 - t);
}

// Density, normal velocity and pressure inside a rarefaction fan at x/t = 0
// of a fan facing side K (sign = +1 left-facing side I, -1 side II): the
// sound speed there is (g-1)/(g+1) sign*uK + 2/(g+1) cK.
AST_HD void fan_state(const GasD& a, double uK, double pK, double cK, double sign, double& r,
                      double& u, double& p) {
  const double ce = sign * a.g1g2 * uK + a.c2g2 * cK;
  u = sign * ce;
  p = pK * pow_igm(a, ce / cK);
  r = a.g * p / (ce * ce);
}

// The exact flux; `guess` (> 0, finite) seeds the two-shock Newton.
AST_HD FluxD exact(const PrimD& L, const PrimD& R, const GasD& a, double& UU, ExactInfo& info,
                   double guess = -1.0) {
  const bool sw = L.p > R.p;
  // side I: the low-pressure side; the frame is mirrored when the sides swap
  const double rI = sw ? R.r : L.r, pI = sw ? R.p : L.p, uI = sw ? -R.u : L.u;
  const double rII = sw ? L.r : R.r, pII = sw ? L.p : R.p, uII = sw ? -L.u : R.u;
  const double vI = sw ? R.v : L.v, vII = sw ? L.v : R.v;
  const double cI = sqrt(a.g * pI / rI), cII = sqrt(a.g * pII / rII);
  const double zI = rI * cI, zII = rII * cII;
  const double du = uI - uII;
  const double q = pow(pI / pII, a.gm);            // (pI/pII)^gm
  const double du_vac = -a.c2g1 * (cI + cII);      // vacuum bound
  const double du_rr = a.c2g1 * cII * (q - 1.0);   // f_II(p_I)
  const double du_sr = (pII - pI) / sqrt(0.5 * rI * (a.g2 * pII + a.g1 * pI));  // f_I(p_II)
  info.it = 0;
  info.capped = false;
  info.guessed = false;
  double p = NAN, t = 0.0;
  // each pattern's whole interval is tested, so NaN input matches none
  if (du < du_vac) {
    info.pat = -1;
  } else if (du >= du_vac && du <= du_rr) {
    // two rarefactions: p* = pI t^(1/gm), t = (du - du_vac)/(2/(g-1)(cII q + cI))
    info.pat = 0;
    t = (du - du_vac) / (a.c2g1 * (cII * q + cI));
    p = pI * pow_igm(a, t);
  } else if (du > du_rr && du <= du_sr) {
    info.pat = 1;
    p = shock_raref_root(a, du, pI, rI, pII, cII, info.it, info.capped);
  } else if (du > du_sr) {
    info.pat = 2;
    info.guessed = guess > 0.0 && guess < INFINITY;
    p = two_shock_root(a, du, pI, rI, zI, pII, rII, zII, guess, info.guessed, info.it,
                       info.capped);
  } else {
    info.pat = -9;  // NaN input: no pattern
  }

  // sample at x/t = 0 (the reference's region edges: kernel.cu:256-544)
  double rs = NAN, us = NAN, ps = NAN, vs = NAN;
  if (info.pat == 2 || info.pat == 1) {
    const double aI = sqrt(rI * (0.5 * a.g2 * p + 0.5 * a.g1 * pI));  // shock mass flux
    double aII;
    if (info.pat == 2) {
      aII = sqrt(rII * (0.5 * a.g2 * p + 0.5 * a.g1 * pII));
    } else {
      const double x = p / pII;
      aII = raref_mass_flux(a, rII, cII, x, pow(x, a.gm));
    }
    const double u = (aI * uI + aII * uII + pI - pII) / (aI + aII);
    const double sI = uI - aI / rI;  // shock into I
    if (info.pat == 2) {
      const double sII = uII + aII / rII;
      UU = max2(fabs(sI), fabs(sII));
      if (0.0 <= sI) { rs = rI; us = uI; ps = pI; vs = vI; }
      else if (0.0 <= u) { rs = rI * aI / (aI - rI * (uI - u)); us = u; ps = p; vs = vI; }
      else if (0.0 < sII) { rs = rII * aII / (aII + rII * (uII - u)); us = u; ps = p; vs = vII; }
      else { rs = rII; us = uII; ps = pII; vs = vII; }
    } else {
      const double hII = uII + cII;                    // head of the fan into II
      const double cs = cII - 0.5 * a.g1 * (uII - u);  // sound speed behind it
      const double tII = u + cs;                       // its tail
      UU = max2(max2(fabs(sI), fabs(hII)), fabs(tII));
      if (0.0 <= sI) { rs = rI; us = uI; ps = pI; vs = vI; }
      else if (0.0 <= u) { rs = rI * aI / (aI - rI * (uI - u)); us = u; ps = p; vs = vI; }
      else if (0.0 <= tII) { rs = a.g * p / (cs * cs); us = u; ps = p; vs = vII; }
      else if (0.0 < hII) { fan_state(a, uII, pII, cII, -1.0, rs, us, ps); vs = vII; }
      else { rs = rII; us = uII; ps = pII; vs = vII; }
    }
  } else if (info.pat == 0) {
    const double aI = raref_mass_flux(a, rI, cI, p / pI, t);
    const double aII = raref_mass_flux(a, rII, cII, p / pII, t * q);
    const double u = (aI * uI + aII * uII + pI - pII) / (aI + aII);
    const double hI = uI - cI, hII = uII + cII;        // fan heads
    const double csI = cI + 0.5 * a.g1 * (uI - u);     // star sound speeds
    const double csII = cII - 0.5 * a.g1 * (uII - u);
    const double tI = u - csI, tII = u + csII;         // fan tails
    UU = max2(max2(fabs(hI), fabs(hII)), max2(fabs(tII), fabs(tI)));
    if (0.0 <= hI) { rs = rI; us = uI; ps = pI; vs = vI; }
    else if (0.0 < tI) { fan_state(a, uI, pI, cI, 1.0, rs, us, ps); vs = vI; }
    else if (0.0 <= u) { rs = a.g * p / (csI * csI); us = u; ps = p; vs = vI; }
    else if (0.0 <= tII) { rs = a.g * p / (csII * csII); us = u; ps = p; vs = vII; }
    else if (0.0 < hII) { fan_state(a, uII, pII, cII, -1.0, rs, us, ps); vs = vII; }
    else { rs = rII; us = uII; ps = pII; vs = vII; }
  } else if (info.pat == -1) {
    const double hI = uI - cI, fI = uI + a.c2g1 * cI;       // head, vacuum front of fan I
    const double hII = uII + cII, fII = uII - a.c2g1 * cII;  // head, vacuum front of fan II
    UU = max2(max2(fabs(hI), fabs(hII)), max2(fabs(fII), fabs(fI)));
    if (0.0 <= hI) { rs = rI; us = uI; ps = pI; vs = vI; }
    else if (0.0 < fI) { fan_state(a, uI, pI, cI, 1.0, rs, us, ps); vs = vI; }
    else if (0.0 <= fII) { rs = 0.0; us = 0.0; ps = 0.0; vs = 0.0; }
    else if (0.0 < hII) { fan_state(a, uII, pII, cII, -1.0, rs, us, ps); vs = vII; }
    else { rs = rII; us = uII; ps = pII; vs = vII; }
  } else {
    UU = INFINITY;  // no pattern: dt 0 (the reference's solver leaves its dt at 0)
  }
  if (!(UU > kMinSpeed)) UU = kMinSpeed;
  if (sw) us = -us;
  FluxD f;
  f.m = rs * us;
  f.un = f.m * us + ps;
  f.ut = f.m * vs;
  f.e = us * (ps * a.ig1 + 0.5 * rs * (us * us + vs * vs) + ps);
  return f;
}

// The wave pattern `exact` will take for a face, as a schedule key for
// sorting faces by pattern: 0 vacuum, 1 two rarefactions, 2 shock +
// rarefaction, 3 two shocks, 4 no pattern (NaN) -- the same tests.
constexpr int kPatternKeys = 5;
AST_HD int pattern_key(const PrimD& L, const PrimD& R, const GasD& a) {
  const bool sw = L.p > R.p;
  const double rI = sw ? R.r : L.r, pI = sw ? R.p : L.p, uI = sw ? -R.u : L.u;
  const double rII = sw ? L.r : R.r, pII = sw ? L.p : R.p, uII = sw ? -L.u : R.u;
  const double cI = sqrt(a.g * pI / rI), cII = sqrt(a.g * pII / rII);
  const double du = uI - uII;
  const double du_vac = -a.c2g1 * (cI + cII);
  const double du_rr = a.c2g1 * cII * (pow(pI / pII, a.gm) - 1.0);
  const double du_sr = (pII - pI) / sqrt(0.5 * rI * (a.g2 * pII + a.g1 * pI));
  if (du < du_vac) return 0;
  if (du >= du_vac && du <= du_rr) return 1;
  if (du > du_rr && du <= du_sr) return 2;
  if (du > du_sr) return 3;
  return 4;
}

// The method's flux (AST_HLL, AST_HLLC, AST_HLLC_A, AST_EXACT).
template <int METHOD>
AST_HD FluxD face_flux(const PrimD& L, const PrimD& R, const GasD& a, double& UU,
                       ExactInfo& info) {
  if constexpr (METHOD == AST_HLL) return hll(L, R, a, UU);
  else if constexpr (METHOD == AST_HLLC) return hllc<false>(L, R, a, UU);
  else if constexpr (METHOD == AST_HLLC_A) return hllc<true>(L, R, a, UU);
  else return exact(L, R, a, UU, info);
}

}  // namespace d64
}  // namespace ast
