// fast_flux.hpp -- FP32 face fluxes of the fast step path.
//
// The same HLL (kernel.cu:1399-1523) and HLLC_2d_Korolkov
// (Gas_Dinamic.cu:1604-1826) algebra as the reference -- Davis wave-speed
// estimates, supersonic upwinding with the reference's >= / <= edges, the
// HLLC contact speed SM, the averaged star pressure ptz, tangential velocity
// kept per side -- rewritten for throughput: states are given in the face
// frame (u normal, v tangential), divisions become hardware reciprocals,
// max(cL, cR) is taken before the square root (monotone), and the branchy
// four-way choice of HLLC becomes selects so a wave never diverges.  The
// return value is the largest |wave speed| of the face (the reference
// returns krit*rad/that; the step turns the maximum into dt once).
#pragma once
#include <hip/hip_runtime.h>

#include "dev.hpp"

namespace ast {

struct PrimF {
  float r, p, u, v;  // density, pressure, normal and tangential velocity
};
struct FluxF {
  float m, e, un, ut;  // mass, energy, normal and tangential momentum
};

AST_DEV float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
AST_DEV float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// min / max of finite values as one v_med3_f32 against -FLT_MAX / FLT_MAX
// (a plain min/max of a value the compiler cannot prove canonical costs an
// extra canonicalising instruction)
AST_DEV float fmin3(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, -3.40282347e38f); }
AST_DEV float fmax3(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 3.40282347e38f); }

// A face flux written in the frame that moves with a velocity (wn, wt) of
// the face (face frame components): the mass flux m, the momentum fluxes gn
// (normal) and gt (tangential) and the energy flux ge seen from that frame,
// so that the lab-frame fluxes are m, gn + m wn, gt + m wt and
// ge + wn gn + wt gt + m |w|^2/2 (with_frame_velocity below).  The step
// moves each face's flux into the frame of each of its two cells (the
// cell-frame update, strip_f32.hpp), where the energy flux is of the size of
// the internal-energy flux: in the lab frame it carries the kinetic energy,
// O(M^2) times larger at Mach M, and an FP32 rounding of it would be
// amplified by M^2 in the pressure the update recovers.
//   A flux that is the physical flux of one state (r, un, ut, p, internal
// energy e) -- HLLC's upwind or star state, the exact solver's sampled state
// -- is written in the frame of that state: m = r un, w = (un, ut), gn = p,
// gt = 0, ge = m e.
struct FaceF {
  float m, wn, wt, gn, gt, ge;
};

// The lab-frame (face-frame axes) conservative flux of a FaceF.
AST_DEV FluxF lab_flux(const FaceF& f) {
  FluxF r;
  r.m = f.m;
  r.un = fmaf(f.m, f.wn, f.gn);
  r.ut = fmaf(f.m, f.wt, f.gt);
  r.e = f.ge + f.wn * f.gn + f.wt * f.gt + 0.5f * f.m * (f.wn * f.wn + f.wt * f.wt);
  return r;
}

// Physical flux of one side in the face frame.
AST_DEV FluxF phys_flux(const PrimF& q, float E) {
  FluxF f;
  f.m = q.r * q.u;
  f.un = f.m * q.u + q.p;
  f.ut = f.m * q.v;
  f.e = (E + q.p) * q.u;
  return f;
}

// HLL with the reference's Davis speeds, kernel.cu:1436-1443, 1505-1519.
AST_DEV FluxF hll_f32(const PrimF& L, const PrimF& R, float g, float ig1, float& UU) {
  const float c2 = g * fmax3(L.p * frcp(L.r), R.p * frcp(R.r));
  const float cm = fsqrt(c2);
  const float DL = fmin3(L.u, R.u) - cm;
  const float DR = fmax3(L.u, R.u) + cm;
  UU = fmaxf(fabsf(DL), fabsf(DR));
  const float EL = L.p * ig1 + 0.5f * L.r * (L.u * L.u + L.v * L.v);
  const float ER = R.p * ig1 + 0.5f * R.r * (R.u * R.u + R.v * R.v);
  const FluxF fl = phys_flux(L, EL), fr = phys_flux(R, ER);
  const float id = frcp(DR - DL);
  const float dd = DL * DR;
  FluxF h;
  h.m = (DR * fl.m - DL * fr.m + dd * (R.r - L.r)) * id;
  h.un = (DR * fl.un - DL * fr.un + dd * (R.r * R.u - L.r * L.u)) * id;
  h.ut = (DR * fl.ut - DL * fr.ut + dd * (R.r * R.v - L.r * L.v)) * id;
  h.e = (DR * fl.e - DL * fr.e + dd * (ER - EL)) * id;
  const bool upL = DL > 0.f, upR = DR < 0.f;
  FluxF f;
  f.m = upL ? fl.m : (upR ? fr.m : h.m);
  f.un = upL ? fl.un : (upR ? fr.un : h.un);
  f.ut = upL ? fl.ut : (upR ? fr.ut : h.ut);
  f.e = upL ? fl.e : (upR ? fr.e : h.e);
  return f;
}

// HLL in the frame w of the two sides' mean velocity: the same Davis speeds
// and upwinding as hll_f32, each side's flux and state written in that
// frame (the HLL average is linear in them, so the frame change commutes
// with it).
AST_DEV FaceF hll_frame_f32(const PrimF& L, const PrimF& R, float g, float ig1, float& UU) {
  const float c2 = g * fmax3(L.p * frcp(L.r), R.p * frcp(R.r));
  const float cm = fsqrt(c2);
  const float DL = fmin3(L.u, R.u) - cm;
  const float DR = fmax3(L.u, R.u) + cm;
  UU = fmaxf(fabsf(DL), fabsf(DR));
  FaceF f;
  f.wn = 0.5f * (L.u + R.u);
  f.wt = 0.5f * (L.v + R.v);
  // side K in frame w: velocity (dn, dt) = u_K - w, mass flux mK = rK uK
  // (frame-free), momentum fluxes mK d + p n, energy flux
  // mK (e + |d|^2/2) + p dn; state rK, rK d, rK e + rK |d|^2/2
  const float dnL = 0.5f * (L.u - R.u), dnR = -dnL;
  const float dtL = 0.5f * (L.v - R.v), dtR = -dtL;
  const float kL = 0.5f * (dnL * dnL + dtL * dtL), kR = kL;
  const float mL = L.r * L.u, mR = R.r * R.u;
  const float iL = L.p * ig1, iR = R.p * ig1;  // rho e
  const float fnL = fmaf(mL, dnL, L.p), fnR = fmaf(mR, dnR, R.p);
  const float ftL = mL * dtL, ftR = mR * dtR;
  const float feL = fmaf(L.u, iL, fmaf(mL, kL, L.p * dnL));
  const float feR = fmaf(R.u, iR, fmaf(mR, kR, R.p * dnR));
  const float id = frcp(DR - DL);
  const float dd = DL * DR;
  const float hm = (DR * mL - DL * mR + dd * (R.r - L.r)) * id;
  const float hn = (DR * fnL - DL * fnR + dd * (R.r * dnR - L.r * dnL)) * id;
  const float ht = (DR * ftL - DL * ftR + dd * (R.r * dtR - L.r * dtL)) * id;
  const float he = (DR * feL - DL * feR + dd * (fmaf(R.r, kR, iR) - fmaf(L.r, kL, iL))) * id;
  const bool upL = DL > 0.f, upR = DR < 0.f;
  f.m = upL ? mL : (upR ? mR : hm);
  f.gn = upL ? fnL : (upR ? fnR : hn);
  f.gt = upL ? ftL : (upR ? ftR : ht);
  f.ge = upL ? feL : (upR ? feR : he);
  return f;
}

// HLLC_2d_Korolkov_b_s (metod 2, W = 0, nul_potok = false).
//
// The subsonic star flux F*_K = F_K + S_K (U*_K - U_K) of the reference
// (Gas_Dinamic.cu:1739-1802) is evaluated in its contact form, which is the
// same expression rearranged with rho*_K (S_K - S_M) = rho_K (S_K - u_K):
//   mass      rho*_K S_M
//   normal    rho*_K S_M^2 + p*      (p* = p_K + rho_K (S_K - u_K)(S_M - u_K))
//   tangent   rho*_K S_M v_K          (= v_K x mass flux, also when supersonic)
//   energy    S_M (E*_K + p*)
// with E*_K = (E_K (S_K - u_K) + p* S_M - p_K u_K) / (S_K - S_M).  The
// reference's energy uses the average of the two sides' p*, equal to p* in
// exact arithmetic (that equality is what defines S_M).
//
// PER_SIDE: HLLC_Aleksashov (kernel.cu:767-1079 with zero magnetic field):
// the same star states and energy, with the wave speeds
// SL = min(uL - cL, uR - cR), SR = max(uL + cL, uR + cR) (kernel.cu:895-896)
// instead of Davis' min/max of the velocities -/+ max(cL, cR); its side
// selection (SL > 0 / SM >= 0 / SR < 0) picks the same flux on the edges
// SL = 0 and SR = 0, where star and physical flux coincide.  Its speeds
// put S_K next to u_K when side K is cold (cK << |uK|), so S_K - u_K and the
// star energy are formed without that cancellation: S_L - u_L =
// min(-cL, (uR - uL) - cR), S_R - u_R = max(cL - (uR - uL), cR), and
// E*_K (S_K - S_M) = E_K (S_K - u_K) + p_K (S_M - u_K) + (p* - p_K) S_M,
// the reference's E_K (S_K - u_K) + p* S_M - p_K u_K rearranged (the FP32
// flux error against FP64 on the extreme microbenchmark pairs: p99.9 4e-3
// without this, 6e-6 with it).
//
// hllc_frame_f32 returns the flux as the physical flux of its state (the
// upwind side's on supersonic faces, the star state of side K otherwise) in
// the frame of that state: m, w = (S_M or u_K, v_K), gn = p* (p_K), gt = 0,
// ge = m e*.  The star internal energy follows from the energy relation
// above without cancellation: rho*_K e*_K = (p_K (S_K - u_K)/(g-1)
// + p_K (S_M - u_K)) / (S_K - S_M) + rho*_K (S_M - u_K)^2 / 2.
template <bool PER_SIDE>
AST_DEV FaceF hllc_frame_f32(const PrimF& L, const PrimF& R, float g, float ig1, float& UU) {
  float SL, SR, suL, suR;
  if constexpr (PER_SIDE) {
    const float cL = fsqrt(g * L.p * frcp(L.r)), cR = fsqrt(g * R.p * frcp(R.r));
    const float d = R.u - L.u;
    suL = fmin3(-cL, d - cR);
    suR = fmax3(cL - d, cR);
    SL = L.u + suL;
    SR = R.u + suR;
  } else {
    const float c2 = g * fmax3(L.p * frcp(L.r), R.p * frcp(R.r));
    const float cm = fsqrt(c2);
    SL = fmin3(L.u, R.u) - cm;                 // Gas_Dinamic.cu:1636
    SR = fmax3(L.u, R.u) + cm;                 // Gas_Dinamic.cu:1637
    suL = SL - L.u;
    suR = SR - R.u;
  }
  UU = fmaxf(fabsf(SL), fabsf(SR));
  const float mL = L.r * suL, mR = R.r * suR;
  const float SM = (mR * R.u - mL * L.u - R.p + L.p) * frcp(mR - mL);  // :1645-1646
  // side whose flux (or star flux) is taken, Gas_Dinamic.cu:1686-1785
  const bool supL = SL >= 0.f, supR = SR <= 0.f;
  const bool left = supL | (!supR & (SM >= 0.f));
  const float rK = left ? L.r : R.r, pK = left ? L.p : R.p;
  const float uK = left ? L.u : R.u, vK = left ? L.v : R.v;
  const float SK = left ? SL : SR;
  const float mK = left ? mL : mR, wK = left ? suL : suR;
  // star state of side K (subsonic faces; non-finite star terms of a
  // supersonic face are discarded by the selects, no branch)
  const float inv = frcp(SK - SM);
  const float dS = SM - uK;
  const float ps = fmaf(mK, dS, pK);  // p* = p_K + rho_K (S_K - u_K)(S_M - u_K)
  const float rs = mK * inv;          // rho*_K
  const float ies = fmaf(0.5f * rs * dS, dS, pK * fmaf(wK, ig1, dS) * inv);  // rho* e*
  const bool sub = !(supL | supR);
  FaceF f;
  f.wn = sub ? SM : uK;
  f.wt = vK;
  f.m = sub ? rs * SM : rK * uK;
  f.gn = sub ? ps : pK;
  f.gt = 0.f;
  f.ge = f.wn * (sub ? ies : pK * ig1);
  return f;
}

template <bool PER_SIDE>
AST_DEV FluxF hllc_f32(const PrimF& L, const PrimF& R, float g, float ig1, float& UU) {
  return lab_flux(hllc_frame_f32<PER_SIDE>(L, R, g, ig1, UU));
}

}  // namespace ast
