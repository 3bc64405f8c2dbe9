// flux_mhd_f64.hip -- batched FP64 MHD face fluxes with a passive scalar:
// HLLDQ_Korolkov (Gas_Dinamic.cu:874-1415) with metod 2 (HLLC for MHD) and
// metod 3 (HLLD), applied to n independent faces with a common normal.
//
// Memory: 18 input planes and 10 output planes (SoA), one face per lane per
// iteration of a grid-stride loop; every load and store is a coalesced
// 8-byte access of consecutive faces: 224 B per face of algorithmic traffic.
//
// Each face is solved in the frame (n, t, m) of the normal.  The wave fan
// (S_L <= S_L* <= S_M <= S_R* <= S_R for HLLD; S_L* = S_R* = S_M for HLLC)
// is built once; the region of x/t = 0 is chosen by the reference's chain of
// tests, in its order (Gas_Dinamic.cu:1160-1187, 1346-1393), so that faces
// whose estimated speeds leave the fan unordered land in the same region;
// the flux is then assembled on that region's side K as
//   F = F_K + a (U*_K - U_K) + b (U**_K - U*_K),
// a = S_K in the star regions, b = S_K* in the Alfven regions (the
// Rankine-Hugoniot chain from the side's outer state), and rotated back.
#include "common.hpp"

namespace ast {
namespace mhd {

constexpr int MNT = 256;
constexpr double kSqrt4Pi = 3.544907701811032;  // spi4, Header.h:69, as written there

// The face frame, chosen on the host by the reference's rule.
struct Frame {
  double n[3], t[3], m[3];
};

struct Gas {
  double g, g1, krad;
};

// One state in the frame of the face: b = B / sqrt(4 pi).
struct Side {
  double r, q, p;       // density, scalar, pressure
  double un, ut, um;    // velocity
  double bn, bt, bm;    // field
  double pT, e, cf;     // total pressure, total energy, fast speed along n
};

AST_DEV Side load_side(const double* __restrict__ in, size_t n, size_t i, int k0,
                       const Frame& f, const Gas& gas) {
  Side s;
  s.r = in[(k0 + 0) * n + i];
  s.q = in[(k0 + 1) * n + i];
  s.p = in[(k0 + 2) * n + i];
  const double vx = in[(k0 + 3) * n + i], vy = in[(k0 + 4) * n + i], vz = in[(k0 + 5) * n + i];
  // true divisions: the fast speed's inner root cancels for a field along
  // the normal with gamma p ~ b_n^2, where a one-ulp change of b moves c_f
  // by up to ~1e-8; dividing as the reference does keeps c_f and dt exact
  const double bx = in[(k0 + 6) * n + i] / kSqrt4Pi;
  const double by = in[(k0 + 7) * n + i] / kSqrt4Pi;
  const double bz = in[(k0 + 8) * n + i] / kSqrt4Pi;
  s.un = vx * f.n[0] + vy * f.n[1] + vz * f.n[2];
  s.ut = vx * f.t[0] + vy * f.t[1] + vz * f.t[2];
  s.um = vx * f.m[0] + vy * f.m[1] + vz * f.m[2];
  s.bn = bx * f.n[0] + by * f.n[1] + bz * f.n[2];
  s.bt = bx * f.t[0] + by * f.t[1] + bz * f.t[2];
  s.bm = bx * f.m[0] + by * f.m[1] + bz * f.m[2];
  const double bb = bx * bx + by * by + bz * bz;
  const double kin = (vx * vx + vy * vy + vz * vz) / 2.0;
  s.pT = s.p + bb / 2.0;
  s.e = s.p / gas.g1 + s.r * kin + bb / 2.0;
  // c_f^2 = (gamma p + b^2 + sqrt((gamma p + b^2)^2 - 4 gamma p b_n^2)) / (2 rho)
  const double gpb = gas.g * s.p + bb;
  s.cf = sqrt((gpb + sqrt(gpb * gpb - 4.0 * gas.g * s.p * (s.bn * s.bn))) / (2.0 * s.r));
  return s;
}

// Physical flux components of a side that the fan needs beyond F_K: the
// tangential-field fluxes enter the HLL average of b.
AST_DEV double flux_bt(const Side& s) { return s.un * s.bt - s.ut * s.bn; }
AST_DEV double flux_bm(const Side& s) { return s.un * s.bm - s.um * s.bn; }
AST_DEV double vdotb(const Side& s) { return s.un * s.bn + s.ut * s.bt + s.um * s.bm; }

// The nine normal-frame components (rho, rho u_n, rho u_t, rho u_m, e, b_n,
// b_t, b_m, Q) of a flux.
struct Nine {
  double c[9];
};

AST_DEV Nine phys_flux(const Side& s) {
  Nine f;
  f.c[0] = s.r * s.un;
  f.c[1] = s.r * s.un * s.un + s.pT - s.bn * s.bn;
  f.c[2] = s.r * s.un * s.ut - s.bn * s.bt;
  f.c[3] = s.r * s.un * s.um - s.bn * s.bm;
  f.c[4] = (s.e + s.pT) * s.un - s.bn * vdotb(s);
  f.c[5] = 0.0;
  f.c[6] = flux_bt(s);
  f.c[7] = flux_bm(s);
  f.c[8] = s.q * s.un;
  return f;
}

// Outer star state of one side (rho*, tangential v*, b*, e*) and the
// density weights of HLLD's inner states.
struct Star {
  double r, q, v, w, bt, bm, e;
};

template <int METHOD>
__global__ __launch_bounds__(MNT) void flux_mhd_kernel(size_t n, const double* __restrict__ in,
                                                       Frame fr, Gas gas,
                                                       double* __restrict__ out) {
  const size_t stride = (size_t)gridDim.x * MNT;
  for (size_t i = (size_t)blockIdx.x * MNT + threadIdx.x; i < n; i += stride) {
    const Side L = load_side(in, n, i, 0, fr, gas);
    const Side R = load_side(in, n, i, 9, fr, gas);
    // fan bounds and the contact speed (Gas_Dinamic.cu:970-990)
    const double SL = fmin(L.un, R.un) - fmax(L.cf, R.cf);
    const double SR = fmax(L.un, R.un) + fmax(L.cf, R.cf);
    const double sL = SL - L.un, sR = SR - R.un;
    const double den = sR * R.r - sL * L.r;
    const double SM = (sR * R.r * R.un - sL * L.r * L.un - R.pT + L.pT) / den;
    const double dLR = SR - SL;
    const double dL = SL - SM, dR = SR - SM;
    // HLL averages of the field (b_n, b_t, b_m); b_n by a true division, as
    // it decides the Alfven switch |b_n| > 0.001 exactly as the reference
    const double bn = (SR * R.bn - SL * L.bn) / dLR;
    const double idLR = 1.0 / dLR;
    const double bt = (SR * R.bt - SL * L.bt + flux_bt(L) - flux_bt(R)) * idLR;
    const double bm = (SR * R.bm - SL * L.bm + flux_bm(L) - flux_bm(R)) * idLR;
    Star sl, sr;
    const double cL = sL / dL, cR = sR / dR;  // rho*_K / rho_K
    sl.r = L.r * cL;
    sr.r = R.r * cR;
    sl.q = L.q * cL;
    sr.q = R.q * cR;
    double SLs = SM, SRs = SM;  // Alfven speeds (HLLD)
    double v2 = 0.0, w2 = 0.0, bt2 = 0.0, bm2 = 0.0, de2L = 0.0, de2R = 0.0;
    int region;
    if constexpr (METHOD == AST_HLLD) {
      const double PT = (sR * R.r * L.pT - sL * L.r * R.pT + L.r * R.r * sR * sL * (R.un - L.un)) / den;
      const double bbn = bn * bn;
      // outer star states (Gas_Dinamic.cu:1214-1259)
      auto outer = [&](const Side& K, double sK, double dK, Star& s) {
        const double tt = K.r * sK * dK - bbn;
        if (fabs(tt) >= 0.000001) {
          const double itt = 1.0 / tt;
          const double f = (SM - K.un) * itt;
          const double g = (K.r * sK * sK - bbn) * itt;
          s.v = K.ut - bn * K.bt * f;
          s.w = K.um - bn * K.bm * f;
          s.bt = K.bt * g;
          s.bm = K.bm * g;
        } else {
          s.v = K.ut;
          s.w = K.um;
          s.bt = 0.0;
          s.bm = 0.0;
        }
        s.e = (K.e * sK + PT * SM - K.pT * K.un + bn * (vdotb(K) - (SM * bn + s.v * s.bt + s.w * s.bm))) / dK;
      };
      outer(L, sL, dL, sl);
      outer(R, sR, dR, sr);
      // inner (Alfven) states, shared velocity and field (Gas_Dinamic.cu:1261-1296)
      const double qL = sqrt(sl.r), qR = sqrt(sr.r);
      double sig = 0.0;
      if (fabs(bn) > 0.001) {
        sig = bn > 0.0 ? 1.0 : (bn < 0.0 ? -1.0 : 0.0);
        SLs = SM - fabs(bn) / qL;
        SRs = SM + fabs(bn) / qR;
      }
      const double iqs = 1.0 / (qL + qR);
      v2 = (qL * sl.v + qR * sr.v + sig * (sr.bt - sl.bt)) * iqs;
      w2 = (qL * sl.w + qR * sr.w + sig * (sr.bm - sl.bm)) * iqs;
      bt2 = (qL * sr.bt + qR * sl.bt + sig * qL * qR * (sr.v - sl.v)) * iqs;
      bm2 = (qL * sr.bm + qR * sl.bm + sig * qL * qR * (sr.w - sl.w)) * iqs;
      const double vb2 = v2 * bt2 + w2 * bm2;
      de2L = -sig * qL * ((sl.v * sl.bt + sl.w * sl.bm) - vb2);
      de2R = sig * qR * ((sr.v * sr.bt + sr.w * sr.bm) - vb2);
      // Gas_Dinamic.cu:1346-1393, in order
      if (SL >= 0.0) region = 0;
      else if (SL < 0.0 && SLs >= 0.0) region = 1;
      else if (SLs <= 0.0 && SM >= 0.0) region = 2;
      else if (SM < 0.0 && SRs > 0.0) region = 3;
      else if (SR > 0.0 && SRs <= 0.0) region = 4;
      else if (SR <= 0.0) region = 5;
      else region = -1;
    } else {
      // HLLC for MHD: tangential velocity and field of the HLL average on
      // both sides, total pressure the mean of the two sides' star values
      // (Gas_Dinamic.cu:1051-1088)
      const double z0 = (SR * R.r - SL * L.r + L.r * L.un - R.r * R.un) / dLR;
      const double z1 = (SR * R.r * R.un - SL * L.r * L.un + (L.r * L.un * L.un + L.pT - L.bn * L.bn) -
                         (R.r * R.un * R.un + R.pT - R.bn * R.bn)) / dLR;
      const double z2 = (SR * R.r * R.ut - SL * L.r * L.ut + (L.r * L.un * L.ut - L.bn * L.bt) -
                         (R.r * R.un * R.ut - R.bn * R.bt)) / dLR;
      const double z3 = (SR * R.r * R.um - SL * L.r * L.um + (L.r * L.un * L.um - L.bn * L.bm) -
                         (R.r * R.un * R.um - R.bn * R.bm)) / dLR;
      const double ptz = ((R.pT + R.r * sR * (SM - R.un)) + (L.pT + L.r * sL * (SM - L.un))) / 2.0;
      const double vz = z2 / z0, wz = z3 / z0;
      const double vbz = (bn * z1 + bt * z2 + bm * z3) / z0;
      sl.v = sr.v = vz;
      sl.w = sr.w = wz;
      sl.bt = sr.bt = bt;
      sl.bm = sr.bm = bm;
      sl.e = L.e * (sL / dL) + (ptz * SM - L.pT * L.un + bn * (vdotb(L) - vbz)) / dL;
      sr.e = R.e * (sR / dR) + (ptz * SM - R.pT * R.un + bn * (vdotb(R) - vbz)) / dR;
      // Gas_Dinamic.cu:1160-1187, in order
      if (SL >= 0.0) region = 0;
      else if (SL < 0.0 && SM >= 0.0) region = 1;
      else if (SR > 0.0 && SM < 0.0) region = 4;
      else if (SR <= 0.0) region = 5;
      else region = -1;
    }
    // assemble on the region's side; the star terms only where the region
    // is a star region (an outer region takes F_K whatever the star states)
    const bool left = region <= 2;
    const Side& K = left ? L : R;
    const Star& s = left ? sl : sr;
    Nine F = phys_flux(K);
    if (region >= 1 && region <= 4) {
      const double a = left ? SL : SR;
      F.c[0] += a * (s.r - K.r);
      F.c[1] += a * (s.r * SM - K.r * K.un);
      F.c[2] += a * (s.r * s.v - K.r * K.ut);
      F.c[3] += a * (s.r * s.w - K.r * K.um);
      F.c[4] += a * (s.e - K.e);
      F.c[6] += a * (s.bt - K.bt);
      F.c[7] += a * (s.bm - K.bm);
      F.c[8] += a * (s.q - K.q);
      if (region == 2 || region == 3) {
        const double b = left ? SLs : SRs;
        F.c[2] += b * (s.r * (v2 - s.v));
        F.c[3] += b * (s.r * (w2 - s.w));
        F.c[4] += b * (left ? de2L : de2R);
        F.c[6] += b * (bt2 - s.bt);
        F.c[7] += b * (bm2 - s.bm);
      }
    }
    if (region < 0)
      for (int k = 0; k < 9; k++) F.c[k] = __longlong_as_double(0x7ff8000000000000LL);
    // the normal-field term (Gas_Dinamic.cu:1189-1191) and the lab frame
    const double SN = fmax(fabs(SL), fabs(SR));
    F.c[5] = -SN * (R.bn - L.bn);
    out[i] = F.c[0];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      out[(1 + k) * n + i] = fr.n[k] * F.c[1] + fr.t[k] * F.c[2] + fr.m[k] * F.c[3];
      out[(4 + k) * n + i] = kSqrt4Pi * (fr.n[k] * F.c[5] + fr.t[k] * F.c[6] + fr.m[k] * F.c[7]);
    }
    out[7 * n + i] = F.c[4];
    out[8 * n + i] = F.c[8];
    out[9 * n + i] = gas.krad / SN;
  }
}

// The reference's frame rule (Gas_Dinamic.cu:894-931): the first of n1, n2,
// n3 above 0.1, then the first below -0.1, picks the tangent pair.
static bool pick_frame(double n1, double n2, double n3, Frame& f) {
  f = Frame{{n1, n2, n3}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  if (n1 > 0.1) { f.t[1] = 1.0; f.m[2] = 1.0; }
  else if (n2 > 0.1) { f.t[2] = 1.0; f.m[0] = 1.0; }
  else if (n3 > 0.1) { f.t[0] = 1.0; f.m[1] = 1.0; }
  else if (n1 < -0.1) { f.t[2] = -1.0; f.m[1] = -1.0; }
  else if (n2 < -0.1) { f.t[0] = -1.0; f.m[2] = -1.0; }
  else if (n3 < -0.1) { f.t[0] = -1.0; f.m[1] = -1.0; }
  else return false;
  return true;
}

// One face per lane up to 2^28 faces (measured: a grid of 16384 workgroups
// iterating four times left the loads of one wave's next faces behind its
// arithmetic; one face per lane lets the dispatcher keep every SIMD's
// waves at different phases), a grid-stride loop beyond.
static unsigned mhd_blocks(size_t n) {
  size_t blocks = (n + MNT - 1) / MNT;
  if (blocks > (size_t(1) << 20)) blocks = size_t(1) << 20;
  return (unsigned)blocks;
}

}  // namespace mhd
}  // namespace ast

extern "C" int ast_flux_mhd_f64(int method, size_t n, const double* in, double n1, double n2,
                                double n3, double rad, double gamma, double krit, double* out,
                                void* stream) {
  using namespace ast::mhd;
  if (method != AST_HLLC_MHD && method != AST_HLLD)
    return ast_fail(AST_ERR_ARG, "ast_flux_mhd_f64: method %d (AST_HLLC_MHD, AST_HLLD)", method);
  if (!(gamma > 1.0)) return ast_fail(AST_ERR_ARG, "ast_flux_mhd_f64: gamma must be > 1");
  Frame fr;
  if (!pick_frame(n1, n2, n3, fr))
    return ast_fail(AST_ERR_ARG,
                    "ast_flux_mhd_f64: normal (%g, %g, %g) has no component beyond +-0.1", n1,
                    n2, n3);
  if (n == 0) return AST_OK;
  if (!in || !out) return ast_fail(AST_ERR_ARG, "ast_flux_mhd_f64: null in/out");
  const char* a0 = (const char*)in;
  const char* b0 = (const char*)out;
  if (a0 < b0 + 10 * n * sizeof(double) && b0 < a0 + 18 * n * sizeof(double))
    return ast_fail(AST_ERR_ARG, "ast_flux_mhd_f64: in and out overlap");
  const Gas gas{gamma, gamma - 1.0, krit * rad};
  hipStream_t st = (hipStream_t)stream;
  if (method == AST_HLLD)
    hipLaunchKernelGGL(flux_mhd_kernel<AST_HLLD>, dim3(mhd_blocks(n)), dim3(MNT), 0, st, n, in,
                       fr, gas, out);
  else
    hipLaunchKernelGGL(flux_mhd_kernel<AST_HLLC_MHD>, dim3(mhd_blocks(n)), dim3(MNT), 0, st, n,
                       in, fr, gas, out);
  AST_LAUNCH_CHECK("ast_flux_mhd_f64 launch");
  return AST_OK;
}
