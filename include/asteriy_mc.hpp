/* asteriy_mc.hpp -- the host helpers of the reference's Monte-Carlo neutral-
 * hydrogen code as __host__ __device__ functions (header-only), so that a
 * particle tracer can run them on the GPU per particle.  Batched forms over
 * device arrays are ast_mc_* in asteriy.h.
 *
 * Replaces (Gas_Dinamic.cu, declared Header.h:135-146):
 *   double minplus(double x, double y)                            :1468
 *   bool Peresechenie(x0, y0, x, y, z, Vx, Vy, Vz, int& mode, double& t)
 *                                                                 :1492
 *   void peresich(y, z, Vy, Vz, R, double& t1, double& t2)        :1543
 *   void Belong_point(x, y, int& n, int& m)                       :1559
 *   double Velosity_1 / Velosity_2 / Velosity_3(u, cp)            :1566-1602
 * The grid's cell sizes and origin (the Header.h macros dx, dy, x_min) come
 * from an McGrid built from an ast_grid_t.  Results are the reference's:
 * its thresholds (1e-11 in minplus, 1e-9 and 1e-7 in Peresechenie, the
 * u < 0.001 series switch of the rates) and its nine-digit sqrt(pi),
 * 1.77245385 (Header.h:65), are kept.
 *
 * What the functions compute, for a particle at (x, y, z) with velocity
 * (Vx, Vy, Vz) in a cell of centre (x0, r0 = y0) of the axisymmetric grid:
 *   peresich        the times at which the ray (y, z) + t (Vy, Vz) meets the
 *                   circle of radius R (t1 >= t2), or (-1, -1);
 *   minplus         the smaller of two times above 1e-11, -1 if neither;
 *   Peresechenie    the first time t > 1e-9 at which the ray leaves the cell
 *                   and the side (mode 1: x = x0 + dx/2, 2: x0 - dx/2,
 *                   3: r = y0 + dy/2, 4: r = y0 - dy/2); false when none;
 *   Belong_point    the cell (n, m) of a point (x, r);
 *   Velosity_1      <|w|>, w = u - v over the Maxwellian exp(-|v|^2/cp^2)
 *                   (normalised), u the particle's speed in the gas frame;
 *   Velosity_2      Int |w| w_par exp(-|v|^2/cp^2) d^3v (not normalised):
 *                   the momentum exchange;
 *   Velosity_3      <|w|^3> / <|w|^2>: the energy exchange. */
#ifndef ASTERIY_MC_HPP
#define ASTERIY_MC_HPP

#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define AST_MC_HD __host__ __device__ __forceinline__
#else
#define AST_MC_HD static inline
#endif

#include "asteriy.h"

namespace asteriy {
namespace mc {

constexpr double kPi = 3.14159265358979323846;  // Header.h:64
constexpr double kSqrtPi = 1.77245385;          // Header.h:65, as written there

struct McGrid {
  double x_min, dx, y_min, dy;
};

AST_MC_HD McGrid make_grid(const ast_grid_t& g) { return McGrid{g.x_min, g.dx, g.y_min, g.dy}; }

AST_MC_HD double minplus(double x, double y) {
  constexpr double eps = 0.00000000001;
  if (x < eps && y < eps) return -1.0;
  if (x < eps) return y;
  if (y < eps) return x;
  return x < y ? x : y;
}

AST_MC_HD void peresich(double y, double z, double Vy, double Vz, double R, double& t1,
                        double& t2) {
  const double b = (2.0 * y * Vy + 2.0 * z * Vz);
  const double a = (Vy * Vy + Vz * Vz);
  const double D = b * b - 4.0 * (y * y + z * z - R * R) * a;
  if (D < 0) {
    t1 = -1.0;
    t2 = -1.0;
    return;
  }
  const double s = sqrt(D);
  t1 = (-b + s) / (2.0 * a);
  t2 = (-b - s) / (2.0 * a);
}

AST_MC_HD bool peresechenie(const McGrid& g, double x0, double y0, double x, double y, double z,
                            double Vx, double Vy, double Vz, int& mode, double& t) {
  constexpr double tmin = 0.000000001, vmin = 0.0000001;
  double tx_hi = -1.0, tx_lo = -1.0, tr_hi = -1.0, tr_lo = -1.0;
  if (fabs(Vx) > vmin) {
    tx_hi = (x0 + g.dx / 2.0 - x) / Vx;
    tx_lo = (x0 - g.dx / 2.0 - x) / Vx;
  }
  if (Vy * Vy + Vz * Vz > vmin) {
    double a, b;
    peresich(y, z, Vy, Vz, y0 + g.dy / 2.0, a, b);
    tr_hi = minplus(a, b);
    peresich(y, z, Vy, Vz, y0 - g.dy / 2.0, a, b);
    tr_lo = minplus(a, b);
  }
  // the first positive crossing; ties keep the earlier side, in this order
  mode = 0;
  t = 100000000.0;
  const double cand[4] = {tx_hi, tx_lo, tr_hi, tr_lo};
  for (int k = 0; k < 4; k++)
    if (t > cand[k] && cand[k] > tmin) {
      t = cand[k];
      mode = k + 1;
    }
  if (fabs(Vx) > vmin && tx_hi < tmin && tx_lo < tmin) return false;
  return mode != 0;
}

AST_MC_HD void belong_point(const McGrid& g, double x, double y, int& n, int& m) {
  // the reference's (int)(y/dy) for its cell centres (m + 1/2) dy, from the
  // grid's own first centre
  m = (int)((y - (g.y_min - g.dy / 2.0)) / g.dy);
  n = (int)((x - g.x_min + g.dx / 2.0) / g.dx);
}

AST_MC_HD double velosity_1(double u, double cp) {
  if (u < 0.001)
    return 2.0 * cp / kSqrtPi + 2.0 * u * u / (3.0 * cp * kSqrtPi) -
           u * u * u * u / (15.0 * cp * cp * cp * kSqrtPi);
  return exp(-u * u / (cp * cp)) * cp / kSqrtPi + (u + (cp * cp) / (2.0 * u)) * erf(u / cp);
}

AST_MC_HD double velosity_2(double u, double cp) {
  const double c2 = cp * cp, u2 = u * u;
  if (u < 0.001)
    return (8.0 / 3.0) * c2 * c2 * kPi * u + (8.0 / 15.0) * c2 * kPi * u * u * u -
           (4.0 / 105.0) * kPi * u2 * u2 * u;
  return cp * cp * cp * kPi *
         (exp(-u * u / c2) * cp * u * 2.0 * (c2 + 2.0 * u2) +
          kSqrtPi * (4.0 * u2 * u2 + 4.0 * cp * cp * u2 - c2 * c2) * erf(u / cp)) /
         (4.0 * u * u);
}

AST_MC_HD double velosity_3(double u, double cp) {
  const double c2 = cp * cp, u2 = u * u;
  if (u < 0.001)
    return 8.0 * cp / (3.0 * kSqrtPi) + 8.0 * u * u / (9.0 * cp * kSqrtPi) -
           44.0 * u * u * u * u / (135.0 * cp * cp * cp * kSqrtPi);
  return exp(-u * u / c2) * cp * (5.0 * c2 + 2.0 * u2) / (kSqrtPi * (3.0 * c2 + 2.0 * u2)) +
         (4.0 * u2 * u2 + 12.0 * cp * cp * u2 + 3.0 * c2 * c2) * erf(u / cp) /
             (2.0 * u * (3.0 * c2 + 2.0 * u2));
}

// The three rates of one pair with exp(-u^2/cp^2) and erf(u/cp) evaluated
// once: the same operations as velosity_1/2/3, so the same bits.
AST_MC_HD void velosities(double u, double cp, double& v1, double& v2, double& v3) {
  if (u < 0.001) {
    v1 = velosity_1(u, cp);
    v2 = velosity_2(u, cp);
    v3 = velosity_3(u, cp);
    return;
  }
  const double c2 = cp * cp, u2 = u * u;
  const double e = exp(-u * u / c2), f = erf(u / cp);
  v1 = e * cp / kSqrtPi + (u + c2 / (2.0 * u)) * f;
  v2 = cp * cp * cp * kPi *
       (e * cp * u * 2.0 * (c2 + 2.0 * u2) +
        kSqrtPi * (4.0 * u2 * u2 + 4.0 * cp * cp * u2 - c2 * c2) * f) /
       (4.0 * u * u);
  v3 = e * cp * (5.0 * c2 + 2.0 * u2) / (kSqrtPi * (3.0 * c2 + 2.0 * u2)) +
       (4.0 * u2 * u2 + 12.0 * cp * cp * u2 + 3.0 * c2 * c2) * f / (2.0 * u * (3.0 * c2 + 2.0 * u2));
}

}  // namespace mc
}  // namespace asteriy

#endif
