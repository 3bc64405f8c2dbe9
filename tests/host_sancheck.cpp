// host_sancheck.cpp -- host-side C++ of the library under AddressSanitizer
// and UndefinedBehaviorSanitizer (built and run by tests/test_sanitize.py,
// no GPU): the grid validation and error reporting of api.cpp over valid and
// invalid grids (messages longer than the error buffer included), and the
// FP64 face solvers of flux64.hpp -- the code the FP64 HIP kernels run,
// compiled for the host -- over random, NaN, vacuum and extreme faces in
// every pattern, with and without a learned newton start and a passive
// scalar; and the Monte-Carlo helpers of asteriy_mc.hpp over random
// particles, rays and rate pairs (both branches of the rates, rays along
// the axes and through the axis).  Exit status 0 and "host sancheck ok" mean
// every call returned.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../asteriy_2d_amd/csrc/flux64.hpp"
#include "../include/asteriy.h"
#include "../include/asteriy_mc.hpp"

int ast_fail(int code, const char* fmt, ...);

using namespace ast::d64;

static unsigned long long st = 0x2545F4914F6CDD1Dull;
static double urand() {
  st ^= st >> 12;
  st ^= st << 25;
  st ^= st >> 27;
  return (double)((st * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
}
static double uexp(double lo, double hi) { return exp(lo + (hi - lo) * urand()); }

static void check_api() {
  ast_grid_t g;
  memset(&g, 0, sizeof g);
  g.N = 16;
  g.M = 8;
  g.dx = 0.1;
  g.dy = 0.1;
  g.y_min = 0.05;
  g.gamma = 5.0 / 3.0;
  g.krit = 0.3;
  g.ho_lo = -INFINITY;
  g.ho_hi = INFINITY;
  int bad = 0;
  for (int k = 0; k < 64; k++) {
    ast_grid_t h = g;
    h.bc[k & 3] = (k >> 2) % 7 - 1;  // -1 .. 5: unknown kinds included
    h.bc_state[k & 3][0] = (k & 1) ? 1.0 : -1.0;
    h.bc_state[k & 3][1] = (k & 2) ? 1.0 : NAN;
    h.bc_pulse[0] = (k & 4) ? 0.5 : 2.0;
    h.bc_pulse[1] = (k & 8) ? 6.28 : INFINITY;
    h.geometry = (k & 16) ? AST_AXISYMMETRIC : AST_PLANAR;
    h.y_min = (k & 32) ? 0.0 : 0.05;
    if (ast_grid_check(&h) != AST_OK) {
      bad++;
      if (strlen(ast_last_error()) == 0) return (void)printf("empty error message\n");
    }
  }
  if (bad == 0 || ast_grid_check(&g) != AST_OK) {
    printf("grid check: %d rejected, plain grid %s\n", bad, ast_last_error());
    return;
  }
  ast_grid_t h = g;
  h.N = 1 << 20;
  h.M = 1 << 20;  // N*M beyond 2^31
  if (ast_grid_check(&h) == AST_OK) printf("oversized grid accepted\n");
  std::vector<char> big(4000, 'x');
  big.back() = 0;
  ast_fail(AST_ERR_ARG, "%s %s", big.data(), big.data());  // truncated, not overrun
  if (strlen(ast_last_error()) >= 512) printf("error message not truncated\n");
}

static void check_faces() {
  const double gammas[3] = {5.0 / 3.0, 1.4, 1.3};
  for (double gm : gammas) {
    const GasD a = make_gas(gm);
    for (int i = 0; i < 20000; i++) {
      PrimD L{uexp(-8, 4), uexp(-10, 6), 20 * urand() - 10, 4 * urand() - 2};
      PrimD R{uexp(-8, 4), uexp(-10, 6), 20 * urand() - 10, 4 * urand() - 2};
      switch (i % 97) {
        case 0: L.r = NAN; break;
        case 1: R.p = NAN; break;
        case 2: L.u = -1e3; R.u = 1e3; break;
        case 3: L.p = 1e-300; R.p = 1e300; break;
        case 4: L.r = INFINITY; break;
        case 5: L.u = 1e300; R.u = -1e300; break;
        default: break;
      }
      double UU, PQ;
      ExactInfo e;
      volatile double sink = 0;
      FluxD f = exact(L, R, a, UU, e);
      sink = sink + f.m + UU + e.it;
      f = exact(L, R, a, UU, e, uexp(-10, 6));  // learned start
      sink = sink + f.e;
      f = hll(L, R, a, UU);
      sink = sink + f.un;
      f = hllc<false>(L, R, a, UU);
      sink = sink + f.ut;
      f = hllc<true, true>(L, R, a, UU, 0.5 * L.r, -0.3 * R.r, &PQ);
      sink = sink + PQ;
      sink = sink + pattern_key(L, R, a);
      (void)sink;
    }
  }
}

static void check_mc() {
  namespace mc = asteriy::mc;
  ast_grid_t g;
  memset(&g, 0, sizeof g);
  g.N = 200;
  g.M = 40;
  g.x_min = -10.0;
  g.dx = 0.1;
  g.dy = 0.025;
  g.y_min = 0.0125;
  const mc::McGrid G = mc::make_grid(g);
  int bad = 0;
  for (int k = 0; k < 20000; k++) {
    // rates: both sides of the u < 0.001 switch, fused equal to separate
    const double cp = 0.05 + 3.0 * urand();
    const double u = (k & 3) == 0 ? 0.0 : ((k & 3) == 1 ? 2e-3 * urand() : 10.0 * urand());
    double v1, v2, v3;
    mc::velosities(u, cp, v1, v2, v3);
    if (!(v1 == mc::velosity_1(u, cp)) || !(v2 == mc::velosity_2(u, cp)) ||
        !(v3 == mc::velosity_3(u, cp)))
      bad++;
    if (!isfinite(v1) || !isfinite(v2) || !isfinite(v3)) bad++;
    // a particle inside cell (n, m), a ray in any direction (some along an
    // axis, some through the axis r = 0), its exit and the cell of a point
    const int n = (int)(urand() * g.N), m = (int)(urand() * g.M);
    const double x0 = g.x_min + n * g.dx, y0 = g.y_min + m * g.dy;
    const double x = x0 + (urand() - 0.5) * 0.9 * g.dx;
    const double r = y0 + (urand() - 0.5) * 0.9 * g.dy;
    const double ph = 6.283185307179586 * urand();
    const double y = r * cos(ph), z = r * sin(ph);
    double Vx = urand() - 0.5, Vy = urand() - 0.5, Vz = urand() - 0.5;
    if (k % 7 == 0) Vy = Vz = 0.0;
    if (k % 11 == 0) Vx = 0.0;
    if (k % 13 == 0) { Vy = -y; Vz = -z; }
    int mode = -1;
    double t = -1.0;
    (void)mc::peresechenie(G, x0, y0, x, y, z, Vx, Vy, Vz, mode, t);
    if (mode < 0 || mode > 4) bad++;
    int bn = -1, bm = -1;
    mc::belong_point(G, x, r, bn, bm);
    if (bn < n - 1 || bn > n + 1 || bm < m - 1 || bm > m + 1) bad++;
    if (mc::minplus(-1.0, -2.0) != -1.0) bad++;
  }
  if (bad) {
    printf("mc helpers: %d bad results\n", bad);
    exit(1);
  }
}

int main() {
  check_api();
  check_faces();
  check_mc();
  printf("host sancheck ok\n");
  return 0;
}
