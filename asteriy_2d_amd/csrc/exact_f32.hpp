// exact_f32.hpp -- the exact Riemann face flux of the fast path.
//
// Godunov_Solver_Alexashov (kernel.cu:128-570) in FP32: the same side swap
// (pI <= pII), the same wave-pattern test on Uud/Urz/Uvk, the same sampling
// at the face (w = 0) with the reference's region edges, and the same flux
// and dt -- with the p* solve adapted to FP32:
//  * two shocks: the reference's newton (kernel.cu:635-680) in FP32, its
//    stopping scale pee the PVRS-type estimate as in the reference, the
//    tolerance raised from 1e-9 to 1e-6 (FP32 resolution) and the starting
//    point optionally a learned guess;
//  * shock + rarefaction: the root of lev (kernel.cu:751-765) on the same
//    bracket [pI, pII] as devtwo's bisection (kernel.cu:682-749), found by a
//    bracket-safeguarded Newton iteration to 1e-6 relative;
//  * FP64 islands: near vacuum and in strong rarefactions the p* equation
//    subtracts two nearly equal numbers, and FP32 would lose the digits of
//    p* there.  That difference -- and only it: every other term of these
//    equations is a sum of like-signed terms -- is formed in FP64 from FP64
//    sound speeds; the solve then runs in FP32 to FP32 accuracy.
//      two rarefactions, (Udf-Uvk)/(Urz-Uvk) < 0.05: the numerator
//        Udf - Uvk = (enI - enII) + 2/(g-1) (cI + cII);
//      shock + rarefaction, pI < 1e-4 pII: the constant D = c2 + du of the
//        pressure equation fI(p) + c2 q(p) - D = 0 (c2 = 2 cII/(g-1),
//        du = enI - enII, q = (p/pII)^((g-1)/2g)).
//    Counted in ast_stats_t.fp64_faces.  (No FP64 pow: the islands keep the
//    kernel's register budget, and so its occupancy, at the FP32 level.)
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "fast_flux.hpp"
#include "mlp.hpp"
#include "dev.hpp"

namespace ast {

struct ExactCtx {
  float g, g1, g2, gp, gm, igm, ig1, c2g1, g1g2, c2g2;
  double gd;
  int n;  // 1/gm = 2 gamma/(gamma - 1) when that is an integer in [1, 16] (5 for
          // gamma = 5/3, 7 for 7/5), else 0
  const unsigned char* mlp;  // learned-guess block (ast_mlp_t) in device memory, or null
  // When set, the kernel ORs 1 into *strong when it meets a strong two-shock
  // face (the runtime's learned mode, time_slots.h).
  uint32_t* strong;
};

static inline ExactCtx make_exact_ctx(double gamma, const void* mlp) {
  ExactCtx e;
  const double g1 = gamma - 1.0, g2 = gamma + 1.0;
  const double gp = (g2 / gamma) / 2.0, gm = (g1 / gamma) / 2.0;
  e.g = (float)gamma;
  e.g1 = (float)g1;
  e.g2 = (float)g2;
  e.gp = (float)gp;
  e.gm = (float)gm;
  e.igm = (float)(1.0 / gm);
  e.ig1 = (float)(1.0 / g1);
  e.c2g1 = (float)(2.0 / g1);
  e.g1g2 = (float)(g1 / g2);
  e.c2g2 = (float)(2.0 / g2);
  e.gd = gamma;
  const double ig = 1.0 / gm, ni = floor(ig + 0.5);
  // (an FP32 gamma of 5/3 or 7/5 still counts: 1e-6 is its resolution)
  e.n = (fabs(ig - ni) <= 1e-6 * ig && ni >= 1.0 && ni <= 16.0) ? (int)ni : 0;
  e.mlp = (const unsigned char*)mlp;
  e.strong = nullptr;
  return e;
}

// x^y for x > 0 via the hardware log2/exp2 (about 1 ulp of FP32 each)
AST_DEV float fpow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }

// For an integer n = 1/gm (e.n, wave-uniform): x^n by products and the
// geometric sum 1 + x + ... + x^(n-1) by Horner's rule, straight-line for
// n = 5 (gamma 5/3) and n = 7 (gamma 7/5), a loop otherwise.
AST_DEV float ipow(float x, int n) {
  if (n == 5) {
    const float x2 = x * x;
    return x2 * x2 * x;
  }
  if (n == 7) {
    const float x3 = x * x * x;
    return x3 * x3 * x;
  }
  float p = x;
  for (int k = 1; k < n; k++) p *= x;
  return p;
}
AST_DEV float geom_sum(float x, int n) {
  if (n == 5) return fmaf(fmaf(fmaf(fmaf(x, 1.f, 1.f), x, 1.f), x, 1.f), x, 1.f);
  if (n == 7)
    return fmaf(fmaf(fmaf(fmaf(fmaf(fmaf(x, 1.f, 1.f), x, 1.f), x, 1.f), x, 1.f), x, 1.f), x, 1.f);
  float s = 1.f;
  for (int k = 1; k < n; k++) s = fmaf(s, x, 1.f);
  return s;
}
// x^(1/gm): exact products for an integer 1/gm, else exp2/log2
AST_DEV float pow_igm(const ExactCtx& e, float x) { return e.n ? ipow(x, e.n) : fpow(x, e.igm); }

// What the exact solver did on one face (pat: -1 vacuum, 0 two
// rarefactions, 1 shock + rarefaction, 2 two shocks; kNoFace: not an exact
// face, e.g. a low-order face of the window).
constexpr int kNoFace = -3;
constexpr int kNoPattern = -9;  // an exact face whose input matches no wave pattern (NaN)
struct FaceInfo {
  int pat, it;
  bool capped, island, learned;
  bool strong;  // a strong two-shock face (the learned start's gate), when detected
};

AST_DEV uint32_t wcount(bool c) {
  return (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(c));
}

// The counter type of kernels that count nothing (counting costs the exact
// kernels a third of their time, so it is compiled only into the variants
// that are asked for counters).
struct NoStats {
  AST_DEV void clear() {}
  AST_DEV void add(const struct FaceInfo&) {}
  AST_DEV bool any_strong() const { return false; }
};
// Only whether the wave met a strong two-shock face (gated launches).
struct StrongStats {
  uint32_t strong;
  AST_DEV void clear() { strong = 0; }
  AST_DEV void add(const FaceInfo& f) { strong += wcount(f.strong); }
  AST_DEV bool any_strong() const { return strong != 0; }
};

// Solver counters of the FP32 exact solver.  Face counts are per wave
// (ballot + popcount: wave-uniform, kept in scalar registers), the iteration
// sums per lane.  Per-lane face counters would hold ten vector registers for
// the whole strip and cost the kernel a wave of occupancy.  add() runs in
// converged code (every lane of the wave active), after the face calls.
struct StatsF32 {
  uint32_t exact, newton, bisect, raref, vacuum, capped, fp64, mlp, strong;  // per wave
  uint32_t newton_it, bisect_it;                                             // per lane
  AST_DEV void clear() {
    exact = newton = bisect = raref = vacuum = capped = fp64 = mlp = strong = 0;
    newton_it = bisect_it = 0;
  }
  AST_DEV bool any_strong() const { return strong != 0; }
  AST_DEV void add(const FaceInfo& f) {
    strong += wcount(f.strong);
    exact += wcount(f.pat != kNoFace);
    raref += wcount(f.pat == 0);
    vacuum += wcount(f.pat == -1);
    bisect += wcount(f.pat == 1);
    newton += wcount(f.pat == 2);
    bisect_it += f.pat == 1 ? f.it : 0;
    newton_it += f.pat == 2 ? f.it : 0;
    capped += wcount(f.capped);
    fp64 += wcount(f.island);
    mlp += wcount(f.learned);
  }
};

// Adds a block's StatsF32 to *stats (ast_stats_t order); full workgroup.
template <int NT>
AST_DEV void block_stats_f32_to(const StatsF32& st, uint32_t* red, ast_stats_t* stats) {
  constexpr int NW = NT / 64;
  const uint32_t v[10] = {st.exact, st.newton, wave_sum(st.newton_it), st.bisect,
                          wave_sum(st.bisect_it), st.raref, st.vacuum, st.capped, st.fp64,
                          st.mlp};
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int i = 0; i < 10; i++) red[wave * 10 + i] = v[i];
  __syncthreads();
  if (threadIdx.x < 10) {
    unsigned long long t = 0;
    for (int w = 0; w < NW; w++) t += red[w * 10 + threadIdx.x];
    if (t) atomicAdd(reinterpret_cast<unsigned long long*>(stats) + threadIdx.x, t);
  }
}

constexpr float kNewtonTol32 = 1e-6f;
// a two-shock face is served by the learned guess when the PVRS start is
// poor: PVRS estimate above kStrongShock * pII (strong shocks)
constexpr float kStrongShock = 3.f;
constexpr float kVacuumIsland = 0.05f;
constexpr float kRarefIsland = 1e-4f;

// FP32 newton of kernel.cu:635-680 (two shocks), x^1.5 as x*sqrt(x).
AST_DEV float newton_f32(const ExactCtx& e, float enI, float pI, float rI, float cI, float enII,
                         float pII, float rII, float cII, float guess, bool use_guess,
                         int& iters, bool& capped) {
  const float g = e.g, gp = e.gp, gm = e.gm;
  const float zI = rI * cI, zII = rII * cII;
  float pn = (pI * zII + pII * zI + (enI - enII) * zI * zII) * frcp(zI + zII);
  const float pee = pn;
  // a learned start is clamped into the two-shock range p* > pII: the
  // function is concave and increasing in p, so Newton from at or below the
  // root climbs to it monotonically, while a start far above the root would
  // jump below zero (and sqrt of a negative argument is NaN); every iterate
  // is clamped as well
  const float plo = use_guess ? pII : -INFINITY;
  if (use_guess) pn = fmaxf(guess, pII);
  const float ipI = frcp(pI), ipII = frcp(pII);
  const float k1 = g + 1.f, k3 = 3.f * g - 1.f, k4 = 4.f * g;
  int k = 0;
  float p, dprev = INFINITY;
  for (;;) {
    p = pn;
    k++;
    const float aI = gp * p * ipI + gm, aII = gp * p * ipII + gm;
    const float sI = fsqrt(aI), sII = fsqrt(aII);
    const float fI = (p - pI) * frcp(zI * sI);
    const float fII = (p - pII) * frcp(zII * sII);
    const float dI = (k1 * p * ipI + k3) * frcp(k4 * zI * aI * sI);
    const float dII = (k1 * p * ipII + k3) * frcp(k4 * zII * aII * sII);
    pn = fmaxf(p - (fI + fII - (enI - enII)) * frcp(dI + dII), plo);
    // the reference's rule |pn - p| < tol*pee, floored at the FP32
    // resolution of the residual (pee can lie far below p* in strong
    // shocks, where the residual's rounding moves p by ~2 ulp); close to
    // the root (|dp| < 1e-4 p) it also stops once the steps stop shrinking,
    // i.e. at the rounding floor
    const float d = fabsf(pn - p);
    if (!(d >= fmaxf(kNewtonTol32 * pee, 2e-6f * fabsf(pn)))) break;
    if (k > 2 && d < 1e-4f * fabsf(pn) && !(d < 0.5f * dprev)) break;
    dprev = d;
    if (k >= AST_ITER_CAP) { capped = true; break; }
  }
  iters = k;
  return pn;
}

template <typename R> struct FastM;
template <> struct FastM<float> {
  static AST_DEV float sqrt_(float a) { return fsqrt(a); }
  static AST_DEV float pow_(float a, float b) { return fpow(a, b); }
  static AST_DEV float div_(float a, float b) { return a * frcp(b); }
};
template <> struct FastM<double> {
  static AST_DEV double sqrt_(double a) { return ::sqrt(a); }
  static AST_DEV double pow_(double a, double b) { return ::pow(a, b); }
  static AST_DEV double div_(double a, double b) { return a / b; }
};

// Root of lev(p) (left shock + right rarefaction) on [pI, pII]:
// safeguarded Newton on fI(p) + c2 q(p) - D with D = c2 + du supplied by the
// caller (FP64-formed in strong rarefactions).  R is float or double.
template <typename R>
AST_DEV R shock_raref(R g, R g1, R gp, R gm, R enI, R pI, R rI, R enII, R pII, R rII, R cI,
                      R cII, R D, R tol, int& iters, bool& capped) {
  typedef FastM<R> F;
  const R zI = rI * cI, zII = rII * cII, du = enI - enII, c2 = F::div_(R(2), g1) * cII;
  const R ipI = F::div_(R(1), pI), ipII = F::div_(R(1), pII);
  R lo = pI, hi = pII;
  // start: the linearised (PVRS-type) estimate, clipped into the bracket
  R p = F::div_(pI * zII + pII * zI + du * zI * zII, zI + zII);
  if (!(p > lo && p < hi)) p = R(0.5) * (lo + hi);
  int k = 0;
  for (;;) {
    k++;
    const R a = gp * p * ipI + gm;
    const R sa = F::sqrt_(a);
    const R q = F::pow_(p * ipII, gm);
    const R f = F::div_(p - pI, zI * sa) + c2 * q - D;
    const R df = F::div_((g + R(1)) * p * ipI + (R(3) * g - R(1)), R(4) * g * zI * a * sa) +
                 F::div_(cII * q, g * p);
    if (f < R(0)) lo = p;
    else hi = p;
    R pn = p - F::div_(f, df);
    if (!(pn > lo && pn < hi)) pn = R(0.5) * (lo + hi);
    const R dp = pn - p;
    p = pn;
    if (!((dp < R(0) ? -dp : dp) > tol * p) || !(hi - lo > tol * p)) break;
    if (k >= AST_ITER_CAP) { capped = true; break; }
  }
  iters = k;
  return p;
}

// The wave pattern exact_f32 will find for a face, as a schedule key for
// sorting faces by pattern: 0 vacuum, 1 two rarefactions, 2 shock +
// rarefaction, 3 two shocks, 4 two shocks that are strong (acoustic estimate
// above kStrongShock * pII: the learned start's faces), 5 no pattern (NaN).
// The same tests as exact_f32's, so the key is the pattern the solver takes.
constexpr int kPatternKeys = 6;
AST_DEV int pattern_key(const PrimF& L, const PrimF& R, const ExactCtx& e) {
  const bool sw = L.p > R.p;
  const float enI = sw ? -R.u : L.u, enII = sw ? -L.u : R.u;
  const float pI = sw ? R.p : L.p, pII = sw ? L.p : R.p;
  const float rI = sw ? R.r : L.r, rII = sw ? L.r : R.r;
  const float cI = fsqrt(e.g * pI * frcp(rI)), cII = fsqrt(e.g * pII * frcp(rII));
  const float q = fpow(pI * frcp(pII), e.gm);
  const float a = fsqrt(rI * (e.g2 * pII + e.g1 * pI) * 0.5f);
  const float Uud = (pII - pI) * frcp(a);
  const float Urz = -e.c2g1 * cII * (1.f - q);
  const float Uvk = -e.c2g1 * (cII + cI);
  const float Udf = enI - enII;
  if (Udf < Uvk) return 0;
  if (Udf >= Uvk && Udf <= Urz) return 1;
  if (Udf > Urz && Udf <= Uud) return 2;
  if (Udf > Uud) {
    const float zI = rI * cI, zII = rII * cII;
    return (pI * zII + pII * zI + Udf * zI * zII) * frcp(zI + zII) > kStrongShock * pII ? 4 : 3;
  }
  return 5;
}

// FP64 sound speed of an FP32 state (for the FP64-formed differences)
AST_DEV double sound_speed_d(double g, float p, float r) { return sqrt(g * (double)p / (double)r); }

// The exact face flux in the face frame (u normal, v tangential), kernel.cu:
// 128-570 in FP32 (see the header), as the physical flux of the state it
// samples at the face, written in that state's frame (FaceF: m = r un,
// w = (un, ut), gn = p, gt = 0, ge = un p/(g-1)); UU is the largest |wave
// speed| used for dt; fi reports the wave pattern, iterations and flags.
// `guess` (finite, > 0) seeds newton.
// LEARNED (compiled only into the kernels that serve a learned guess; its
// network needs registers the plain solver does not) with e.mlp set: the
// caller guarantees a full wave (EXEC all ones); if any lane of the wave
// holds a strong two-shock face, the wave evaluates the learned guess for
// all its lanes on MFMA (a wave-uniform branch) and those lanes start
// newton from it.
// DETECT: report strong two-shock faces (fi.strong) without the network.
template <bool LEARNED, bool DETECT = false>
AST_DEV FaceF exact_frame_f32(const PrimF& L, const PrimF& R, const ExactCtx& e, float& UU,
                              FaceInfo& fi, float guess = -1.f) {
  const float g = e.g, g1 = e.g1, g2 = e.g2, gm = e.gm;
  // swap sides so that pI <= pII, kernel.cu:156-184
  const bool sw = L.p > R.p;
  const float enI = sw ? -R.u : L.u, enII = sw ? -L.u : R.u;
  const float teI = sw ? R.v : L.v, teII = sw ? L.v : R.v;
  const float pI = sw ? R.p : L.p, pII = sw ? L.p : R.p;
  const float rI = sw ? R.r : L.r, rII = sw ? L.r : R.r;
  const float cI = fsqrt(g * pI * frcp(rI)), cII = fsqrt(g * pII * frcp(rII));
  const float q = fpow(pI * frcp(pII), gm);  // (pI/pII)^gm
  const float a = fsqrt(rI * (g2 * pII + g1 * pI) * 0.5f);
  const float Uud = (pII - pI) * frcp(a);
  const float Urz = -e.c2g1 * cII * (1.f - q);
  const float Uvk = -e.c2g1 * (cII + cI);
  const float Udf = enI - enII;
  // strong two-shock face: PVRS estimate above kStrongShock * pII (the
  // learned kernel needs it before the branches for its wave ballot, the
  // detecting one only in the two-shock branch)
  bool strong = false;
  auto strong_pvrs = [&]() {
    const float zI = rI * cI, zII = rII * cII;
    return (pI * zII + pII * zI + Udf * zI * zII) * frcp(zI + zII) > kStrongShock * pII;
  };
  if constexpr (LEARNED) strong = (Udf > Uud) && strong_pvrs();
  if (LEARNED && e.mlp != nullptr) {
    if (__builtin_amdgcn_ballot_w64(strong) != 0) {  // wave-uniform
      float fe[mlp::IN];
      const float lp = mlp::features<float>(rI, pI, enI, rII, pII, enII, g, fe);
      const float y = mlp::forward_wave_bf16(e.mlp, fe);
      if (strong) guess = __expf(y + lp);
    }
  }
  int pat;
  float p = 0.f, t = 1.f;
  int it = 0;
  bool capped = false, island = false, learned = false;
  // each pattern's whole interval is tested, so NaN input matches none
  if (Udf < Uvk) {
    pat = -1;
  } else if (Udf >= Uvk && Udf <= Urz) {
    // two rarefactions, Urz - Uvk = 2/(g-1) (cII q + cI) (no cancellation)
    pat = 0;
    const float den = e.c2g1 * (cII * q + cI);
    t = (Udf - Uvk) * frcp(den);
    if (t < kVacuumIsland) {  // FP64 island: near vacuum
      const double gd = e.gd;
      const double num = ((double)enI - (double)enII) +
                         2.0 / (gd - 1.0) * (sound_speed_d(gd, pI, rI) + sound_speed_d(gd, pII, rII));
      t = fmaxf((float)num, 0.f) * frcp(den);
      island = true;
    }
    p = pI * pow_igm(e, t);
  } else if (Udf > Urz && Udf <= Uud) {
    pat = 1;
    float D;
    if (pI < kRarefIsland * pII) {  // FP64 island: strong rarefaction
      const double gd = e.gd;
      D = (float)(2.0 / (gd - 1.0) * sound_speed_d(gd, pII, rII) + ((double)enI - (double)enII));
      island = true;
    } else {
      D = e.c2g1 * cII + Udf;
    }
    p = shock_raref<float>(g, g1, e.gp, gm, enI, pI, rI, enII, pII, rII, cI, cII, D, kNewtonTol32,
                           it, capped);
  } else if (Udf > Uud) {
    pat = 2;
    if constexpr (DETECT && !LEARNED) strong = strong_pvrs();
    learned = guess > 0.f && guess < 3.0e38f;
    p = newton_f32(e, enI, pI, rI, cI, enII, pII, rII, cII, guess, learned, it, capped);
  } else {
    pat = kNoPattern;  // NaN input
  }
  fi.pat = pat;
  fi.it = it;
  fi.capped = capped;
  fi.island = island;
  fi.learned = learned;
  fi.strong = strong;

  // sample the solution at w = 0, kernel.cu:234-544
  float en, pp, r, te;
  if (pat == 2) {
    const float aI = fsqrt(rI * (g2 * 0.5f * p + g1 * 0.5f * pI));
    const float aII = fsqrt(rII * (g2 * 0.5f * p + g1 * 0.5f * pII));
    const float u = (aI * enI + aII * enII + pI - pII) * frcp(aI + aII);
    const float dI = enI - aI * frcp(rI), dII = enII + aII * frcp(rII);
    UU = fmaxf(fabsf(dI), fabsf(dII));
    if (0.f <= dI) { en = enI; pp = pI; r = rI; te = teI; }
    else if (0.f <= u) { en = u; pp = p; r = rI * aI * frcp(aI - rI * (enI - u)); te = teI; }
    else if (0.f < dII) { en = u; pp = p; r = rII * aII * frcp(aII + rII * (enII - u)); te = teII; }
    else { en = enII; pp = pII; r = rII; te = teII; }
  } else if (pat == 1) {
    const float aI = fsqrt(rI * (g2 * 0.5f * p + g1 * 0.5f * pI));
    float aII;
    if (fabsf(p - pII) < 1e-6f * pII) aII = rII * cII;
    else aII = gm * rII * cII * (1.f - p * frcp(pII)) * frcp(1.f - fpow(p * frcp(pII), gm));
    const float u = (aI * enI + aII * enII + pI - pII) * frcp(aI + aII);
    const float dI = enI - aI * frcp(rI), dII = enII + cII;
    const float ddII = u + cII - g1 * (enII - u) * 0.5f;
    UU = fmaxf(fmaxf(fabsf(dI), fabsf(dII)), fabsf(ddII));
    if (0.f <= dI) { en = enI; pp = pI; r = rI; te = teI; }
    else if (0.f <= u) { en = u; pp = p; r = rI * aI * frcp(aI - rI * (enI - u)); te = teI; }
    else if (0.f <= ddII) {
      const float ce = cII - g1 * 0.5f * (enII - u);
      en = u; pp = p; r = g * p * frcp(ce * ce); te = teII;
    } else if (0.f < dII) {
      const float ce = -e.g1g2 * enII + e.c2g2 * cII;
      en = -ce; pp = pII * pow_igm(e, ce * frcp(cII)); r = g * pp * frcp(ce * ce); te = teII;
    } else { en = enII; pp = pII; r = rII; te = teII; }
  } else if (pat == 0) {
    // rarefaction impedances aK = gm rK cK (1 - x)/(1 - x^gm), x = p/pK
    // (kernel.cu:369-388), with x^gm = t (side I) and t q (side II)
    float aI, aII;
    if (e.n) {
      aI = gm * rI * cI * geom_sum(t, e.n);
      aII = gm * rII * cII * geom_sum(t * q, e.n);
    } else {
      if (fabsf(p - pI) < 1e-6f * pI) aI = rI * cI;
      else aI = gm * rI * cI * (1.f - p * frcp(pI)) * frcp(1.f - t);
      if (fabsf(p - pII) < 1e-6f * pII) aII = rII * cII;
      else aII = gm * rII * cII * (1.f - p * frcp(pII)) * frcp(1.f - t * q);
    }
    const float u = (aI * enI + aII * enII + pI - pII) * frcp(aI + aII);
    const float dI = enI - cI, ddI = u - cI - g1 * (enI - u) * 0.5f;
    const float dII = enII + cII, ddII = u + cII - g1 * (enII - u) * 0.5f;
    UU = fmaxf(fmaxf(fabsf(dI), fabsf(dII)), fmaxf(fabsf(ddII), fabsf(ddI)));
    if (0.f <= dI) { en = enI; pp = pI; r = rI; te = teI; }
    else if (0.f < ddI) {
      const float ce = e.g1g2 * enI + e.c2g2 * cI;
      en = ce; pp = pI * pow_igm(e, ce * frcp(cI)); r = g * pp * frcp(ce * ce); te = teI;
    } else if (0.f <= u) {
      const float ce = cI + g1 * 0.5f * (enI - u);
      en = u; pp = p; r = g * p * frcp(ce * ce); te = teI;
    } else if (0.f <= ddII) {
      const float ce = cII - g1 * 0.5f * (enII - u);
      en = u; pp = p; r = g * p * frcp(ce * ce); te = teII;
    } else if (0.f < dII) {
      const float ce = -e.g1g2 * enII + e.c2g2 * cII;
      en = -ce; pp = pII * pow_igm(e, ce * frcp(cII)); r = g * pp * frcp(ce * ce); te = teII;
    } else { en = enII; pp = pII; r = rII; te = teII; }
  } else if (pat == -1) {  // vacuum, kernel.cu:480-544
    const float dI = enI - cI, ddI = enI + e.c2g1 * cI;
    const float dII = enII + cII, ddII = enII - e.c2g1 * cII;
    UU = fmaxf(fmaxf(fabsf(dI), fabsf(dII)), fmaxf(fabsf(ddII), fabsf(ddI)));
    if (0.f <= dI) { en = enI; pp = pI; r = rI; te = teI; }
    else if (0.f < ddI) {
      const float ce = e.g1g2 * enI + e.c2g2 * cI;
      en = ce; pp = pI * pow_igm(e, ce * frcp(cI)); r = g * pp * frcp(ce * ce); te = teI;
    } else if (0.f <= ddII) { en = 0.f; pp = 0.f; r = 0.f; te = 0.f; }
    else if (0.f < dII) {
      const float ce = -e.g1g2 * enII + e.c2g2 * cII;
      en = -ce; pp = pII * pow_igm(e, ce * frcp(cII)); r = g * pp * frcp(ce * ce); te = teII;
    } else { en = enII; pp = pII; r = rII; te = teII; }
  } else {  // no pattern (NaN input): NaN fluxes, dt 0
    UU = INFINITY;
    en = pp = r = te = __int_as_float(0x7fc00000);
  }
  if (!(UU > 1e-7f)) UU = 1e-7f;  // eps8 floor of the reference's dt, kernel.cu:246-253
  if (sw) en = -en;
  FaceF f;
  f.m = r * en;
  f.wn = en;
  f.wt = te;
  f.gn = pp;
  f.gt = 0.f;
  f.ge = en * pp * e.ig1;
  return f;
}

// The same flux in the lab frame (face-frame axes).
template <bool LEARNED, bool DETECT = false>
AST_DEV FluxF exact_f32(const PrimF& L, const PrimF& R, const ExactCtx& e, float& UU,
                        FaceInfo& fi, float guess = -1.f) {
  return lab_flux(exact_frame_f32<LEARNED, DETECT>(L, R, e, UU, fi, guess));
}

}  // namespace ast
