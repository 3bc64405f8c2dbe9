/* asteriy_device.hpp -- the per-face flux operators as device functions, for
 * a caller's own HIP kernels (header-only; host-callable too).
 *
 * The reference's operators are __device__ functions that its kernels call
 * face by face:
 *   double HLL(double2& Ls, double2& Lu, double2& Rs, double2& Ru,
 *              double n1, double n2, double2& Ps, double2& Pu, double rad)
 *                                                    kernel.cu:1399-1400
 *   double HLLC_2d_Korolkov(..., double rad, bool nul_potok)
 *                                                    Gas_Dinamic.cu:1814-1815
 *   double HLLC_Aleksashov(...)                      kernel.cu:767
 *   double Godunov_Solver_Alexashov(...)             kernel.cu:128-129
 * A kernel of the caller that includes this header calls the same
 * arguments through asteriy::HLL, asteriy::HLLC_2d_Korolkov,
 * asteriy::HLLC_Aleksashov and asteriy::Godunov_Solver_Alexashov.  The one
 * addition is a leading `const asteriy::Gas&` (gamma and krit, which the
 * reference takes from the Header.h macros ggg/ga and krit), built once with
 * asteriy::make_gas(gamma, krit).  The solvers are those of the library's
 * FP64 kernels (ast_flux_f64, ast_step_f64), inlined into the caller's code:
 * Ps = {mass, energy}, Pu = {momentum x, y} along the normal (n1, n2), and
 * the return value krit*rad/max|wave speed| (HLL: at most 1e7,
 * kernel.cu:1442-1443).  HLLC_2d_Korolkov's nul_potok = true (the tangential
 * velocity of the HLL average, which no live call of the reference selects)
 * is not offered: the wrapper has no such argument.
 *
 * Compile the caller with hipcc for gfx950; -ffp-contract=off gives the
 * library's FP64 kernels' arithmetic. */
#ifndef ASTERIY_DEVICE_HPP
#define ASTERIY_DEVICE_HPP

#include <hip/hip_vector_types.h>

#include "../asteriy_2d_amd/csrc/flux64.hpp"

namespace asteriy {

struct Gas {
  ast::d64::GasD g;
  double krit;
};

AST_HD Gas make_gas(double gamma, double krit) { return Gas{ast::d64::make_gas(gamma), krit}; }

// One face of METHOD (AST_HLL, AST_HLLC, AST_HLLC_A, AST_EXACT) in the
// frame of the normal (n1, n2), tangent (-n2, n1).
template <int METHOD>
AST_HD double face(const Gas& gas, const double2& Ls, const double2& Lu, const double2& Rs,
                   const double2& Ru, double n1, double n2, double2& Ps, double2& Pu,
                   double rad) {
  const ast::d64::PrimD L{Ls.x, Ls.y, Lu.x * n1 + Lu.y * n2, -Lu.x * n2 + Lu.y * n1};
  const ast::d64::PrimD R{Rs.x, Rs.y, Ru.x * n1 + Ru.y * n2, -Ru.x * n2 + Ru.y * n1};
  double UU;
  ast::d64::ExactInfo info;
  const ast::d64::FluxD f = ast::d64::face_flux<METHOD>(L, R, gas.g, UU, info);
  Ps.x = f.m;
  Ps.y = f.e;
  Pu.x = f.un * n1 - f.ut * n2;
  Pu.y = f.un * n2 + f.ut * n1;
  const double t = (gas.krit * rad) / UU;
  return METHOD == AST_HLL ? ast::d64::min2(1e7, t) : t;
}

AST_HD double HLL(const Gas& gas, const double2& Ls, const double2& Lu, const double2& Rs,
                  const double2& Ru, double n1, double n2, double2& Ps, double2& Pu,
                  double rad) {
  return face<AST_HLL>(gas, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
}
AST_HD double HLLC_2d_Korolkov(const Gas& gas, const double2& Ls, const double2& Lu,
                               const double2& Rs, const double2& Ru, double n1, double n2,
                               double2& Ps, double2& Pu, double rad) {
  return face<AST_HLLC>(gas, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
}
AST_HD double HLLC_Aleksashov(const Gas& gas, const double2& Ls, const double2& Lu,
                              const double2& Rs, const double2& Ru, double n1, double n2,
                              double2& Ps, double2& Pu, double rad) {
  return face<AST_HLLC_A>(gas, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
}
AST_HD double Godunov_Solver_Alexashov(const Gas& gas, const double2& Ls, const double2& Lu,
                                       const double2& Rs, const double2& Ru, double n1,
                                       double n2, double2& Ps, double2& Pu, double rad) {
  return face<AST_EXACT>(gas, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
}

}  // namespace asteriy

#endif
