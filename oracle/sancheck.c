/* sancheck.c -- CPU ORACLE test driver (test infrastructure only), built
 * with AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile,
 * target sancheck): runs every entry point of the oracle over edge cases --
 * NaN and near-vacuum faces, the unbounded newton/devtwo loops the
 * reference would spin in, every boundary rule (pulsed inflow included), the
 * frozen region, the floors, the neutral-H source planes and the
 * passive-scalar fluxes -- so that an out-of-bounds access, a signed
 * overflow or a misaligned load in the restatement aborts the run.  Exit
 * status 0 and the line "sancheck ok" mean every call returned. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "asteriy_oracle.h"

static unsigned long long rng_state = 0x9e3779b97f4a7c15ull;
static double urand(void) { /* xorshift64*, uniform in [0, 1) */
  rng_state ^= rng_state >> 12;
  rng_state ^= rng_state << 25;
  rng_state ^= rng_state >> 27;
  return (double)((rng_state * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
}
static double uexp(double lo, double hi) { return exp(lo + (hi - lo) * urand()); }

static void faces(long n, double* in, int special) {
  for (long i = 0; i < n; i++) {
    in[0 * n + i] = uexp(-8, 4);
    in[1 * n + i] = uexp(-10, 6);
    in[2 * n + i] = 20 * urand() - 10;
    in[3 * n + i] = 4 * urand() - 2;
    in[4 * n + i] = uexp(-8, 4);
    in[5 * n + i] = uexp(-10, 6);
    in[6 * n + i] = 20 * urand() - 10;
    in[7 * n + i] = 4 * urand() - 2;
  }
  if (special && n >= 8) {
    in[0 * n + 0] = NAN;                                 /* NaN density */
    in[5 * n + 1] = NAN;                                 /* NaN pressure */
    in[2 * n + 2] = -1e3; in[6 * n + 2] = 1e3;           /* vacuum generation */
    in[1 * n + 3] = 1e-300; in[5 * n + 3] = 1e300;       /* extreme pressure ratio */
    in[0 * n + 4] = INFINITY;                            /* infinite density */
    in[2 * n + 5] = 1e300; in[6 * n + 5] = -1e300;       /* colliding at 1e300 */
  }
}

static void check_fluxes(void) {
  const long n = 4096;
  ora_phys ph0;
  ora_phys_init(&ph0, 5.0 / 3.0, 0.05);
  double* in = malloc(sizeof(double) * 10 * n);
  double* out = malloc(sizeof(double) * 6 * n);
  int* it = malloc(sizeof(int) * 2 * n);
  double* guess = malloc(sizeof(double) * n);
  double* ps = malloc(sizeof(double) * n);
  faces(n, in, 1);
  for (long i = 0; i < n; i++) {
    in[8 * n + i] = (2 * urand() - 1) * in[i];
    in[9 * n + i] = (2 * urand() - 1) * in[4 * n + i];
    guess[i] = (i % 3 == 0) ? NAN : uexp(-10, 6);
  }
  in[8 * n + 6] = NAN;
  static const double normals[4][2] = {{1, 0}, {0, -1}, {-0.6, 0.8}, {0, 1}};
  for (int k = 0; k < 4; k++) {
    for (int m = AST_HLL; m <= AST_HLLC_A; m++)
      ora_flux_batch(m, 5.0 / 3.0, 0.05, n, in, normals[k][0], normals[k][1], 0.5, out,
                     m == AST_EXACT ? it : NULL, 1);
    ora_flux_batch(AST_EXACT, 1.4, 0.3, n, in, normals[k][0], normals[k][1], 0.5, out, it, 1);
    const int qm[3] = {AST_HLLC, AST_HLLCQ, AST_HLLCQ_A};
    for (int j = 0; j < 3; j++)
      ora_flux_q_batch(qm[j], 5.0 / 3.0, 0.05, n, in, normals[k][0], normals[k][1], 0.5, out, 1);
  }
  ora_godunov_guess_batch(5.0 / 3.0, 0.05, n, in, 1, 0, 0.5, guess, out, it, ps);
  /* MHD faces: the gas faces as rho/p/velocity, random fields, the normal
   * field switched off, at the 0.001 switch and along the normal; every
   * normal the frame rule knows and one it does not */
  {
    double* mh = malloc(sizeof(double) * 18 * n);
    double* mo = malloc(sizeof(double) * 10 * n);
    for (long i = 0; i < n; i++)
      for (int s = 0; s < 2; s++) {
        double* w = mh + 9 * s * n;
        w[0 * n + i] = in[(4 * s + 0) * n + i];
        w[1 * n + i] = 2 * urand() - 1;
        w[2 * n + i] = in[(4 * s + 1) * n + i];
        w[3 * n + i] = in[(4 * s + 2) * n + i];
        w[4 * n + i] = in[(4 * s + 3) * n + i];
        w[5 * n + i] = 2 * urand() - 1;
        for (int c = 6; c < 9; c++) w[c * n + i] = (i % 5 == 0) ? 0.0 : 6 * urand() - 3;
        if (i % 7 == 1) w[6 * n + i] = 0.0009 * 3.544907701811032;
        if (i % 11 == 2) { w[7 * n + i] = 0.0; w[8 * n + i] = 0.0; }
      }
    static const double mn[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0},
                                    {0, -1, 0}, {0, 0, -1}, {0.05, 0.05, 0.05}};
    for (int k = 0; k < 7; k++) {
      ora_flux_mhd_batch(AST_HLLC_MHD, 5.0 / 3.0, 0.05, n, mh, mn[k][0], mn[k][1], mn[k][2],
                         0.5, mo, 1);
      ora_flux_mhd_batch(AST_HLLD, 5.0 / 3.0, 0.05, n, mh, mn[k][0], mn[k][1], mn[k][2], 0.5,
                         mo, 1);
    }
    ora_mhd_fan fan;
    double Lm[9] = {1, 0.5, 1, 0.1, 0.2, 0.3, 2, 1, -1}, Rm[9] = {0.2, 0.1, 0.1, 0, 0, 0, 2, -1, 1};
    double P[8], PQ;
    ora_hlldq(&ph0, Lm, Rm, 1, 0, 0, 1.0, AST_HLLD, P, &PQ, &fan);
    ora_hlldq(&ph0, Lm, Rm, 0, 0, -1, 1.0, AST_HLLC_MHD, P, &PQ, &fan);
    free(mh);
    free(mo);
  }
  /* the Monte-Carlo helpers: rates across the series switch and at u = 0,
   * cell exits from random, axial, radial and resting particles, points */
  {
    ast_grid_t gm;
    memset(&gm, 0, sizeof gm);
    gm.N = 16;
    gm.M = 8;
    gm.x_min = -1.0;
    gm.dx = 0.125;
    gm.dy = 0.125;
    gm.y_min = 0.0625;
    double acc = 0.0;
    for (long i = 0; i < n; i++) {
      const double cp = uexp(-5, 5), u = (i % 3 == 0) ? 0.001 * urand() : cp * uexp(-8, 3);
      acc += ora_velosity_1(u, cp) + ora_velosity_2(u, cp) + ora_velosity_3(u, cp);
      int mode, cn, cm;
      double t;
      const double x0 = gm.x_min + (i % 16) * gm.dx, y0 = gm.y_min + (i % 8) * gm.dy;
      const double vx = (i % 5 == 0) ? 0.0 : 2 * urand() - 1, vy = (i % 7 == 0) ? 0.0 : 2 * urand() - 1;
      acc += ora_peresechenie(&gm, x0, y0, x0 + 0.01 * (urand() - 0.5), y0, 0.0, vx, vy,
                              (i % 11 == 0) ? 0.0 : 2 * urand() - 1, &mode, &t);
      ora_belong_point(&gm, x0 + 0.05 * urand(), y0, &cn, &cm);
      acc += t + mode + cn + cm + ora_minplus(2 * urand() - 1, 2 * urand() - 1);
    }
    if (acc != acc) printf("(nan in the helper sums)\n");
  }
  /* the reference's test kernel inputs and one face per exact pattern */
  ora_phys ph;
  ora_phys_init(&ph, 5.0 / 3.0, 0.05);
  const double L[4][4] = {{1, 1, 0, 0}, {1, 1, -3, 0}, {1, 1000, 0, 0}, {1e-6, 1e-9, 5, 1}};
  const double R[4][4] = {{0.125, 0.1, 0, 0}, {1, 1, 3, 0}, {1, 0.01, 0, 0}, {1, 1, -5, 0}};
  for (int k = 0; k < 4; k++) {
    double Ps[2], Pu[2], PQ;
    ora_info inf;
    ora_godunov_ex(&ph, L[k], L[k] + 2, R[k], R[k] + 2, 1, 0, Ps, Pu, 0.5, NULL, &inf);
    ora_hll(&ph, L[k], L[k] + 2, R[k], R[k] + 2, 0, 1, Ps, Pu, 0.5);
    ora_hllc(&ph, L[k], L[k] + 2, R[k], R[k] + 2, 0, 1, Ps, Pu, 0.5);
    ora_hllc_a(&ph, L[k], L[k] + 2, R[k], R[k] + 2, 0, 1, Ps, Pu, 0.5);
    ora_hllcq(&ph, L[k], L[k] + 2, R[k], R[k] + 2, 0.3, -0.2, 0, 1, Ps, Pu, &PQ, 0.5);
    ora_hllcq_a(&ph, L[k], L[k] + 2, R[k], R[k] + 2, 0.3, -0.2, 0, 1, Ps, Pu, &PQ, 0.5);
    double pm, um, d, u, p, sp[4];
    int its;
    if (toro_star(5.0 / 3.0, L[k][0], L[k][2], L[k][1], R[k][0], R[k][2], R[k][1], &pm, &um,
                  &its) == 0)
      toro_sample(5.0 / 3.0, L[k][0], L[k][2], L[k][1], R[k][0], R[k][2], R[k][1], pm, um, 0.1,
                  &d, &u, &p);
    toro_sample_vacuum(5.0 / 3.0, L[k][0], L[k][2], L[k][1], R[k][0], R[k][2], R[k][1], 0.0, &d,
                       &u, &p, sp);
  }
  double x[64], dd[64], uu[64], pp[64];
  for (int i = 0; i < 64; i++) x[i] = i / 63.0;
  toro_sod_profile(1.4, 1, 0, 1, 0.125, 0, 0.1, 0.5, 0.2, 64, x, dd, uu, pp);
  free(in);
  free(out);
  free(it);
  free(guess);
  free(ps);
}

static void check_steps(void) {
  const int N = 23, M = 11;
  const long K = (long)N * M;
  double* s = malloc(sizeof(double) * 2 * K);
  double* u = malloc(sizeof(double) * 2 * K);
  double* s2 = malloc(sizeof(double) * 2 * K);
  double* u2 = malloc(sizeof(double) * 2 * K);
  double* nn2 = malloc(sizeof(double) * 3 * K);
  double* nn3 = malloc(sizeof(double) * K);
  for (long k = 0; k < K; k++) {
    s[2 * k] = uexp(-3, 2);
    s[2 * k + 1] = uexp(-3, 3);
    u[2 * k] = 4 * urand() - 2;
    u[2 * k + 1] = 4 * urand() - 2;
    nn2[3 * k] = urand() - 0.5;
    nn2[3 * k + 1] = urand() - 0.5;
    nn2[3 * k + 2] = urand();
    nn3[k] = urand() - 0.5;
  }
  s[2 * 5] = 1e-300;  /* drives the density floor */
  s[2 * 7 + 1] = 1e-300;  /* and the pressure floor */
  ora_src src = {nn2, nn3, 3.0 / 0.419231, 0.02};
  for (int geom = AST_PLANAR; geom <= AST_AXISYMMETRIC; geom++) {
    for (int bcv = AST_BC_OUTFLOW; bcv <= AST_BC_DIRICHLET_PULSED; bcv++) {
      ast_grid_t g;
      memset(&g, 0, sizeof g);
      g.N = N;
      g.M = M;
      g.geometry = geom;
      for (int k = 0; k < 4; k++) {
        g.bc[k] = (bcv + k) % 5;
        g.bc_state[k][0] = 1.0;
        g.bc_state[k][1] = 0.7;
        g.bc_state[k][2] = 0.3;
        g.bc_state[k][3] = -0.1;
        g.bc_vclamp[k] = 0.2;
      }
      g.bc_pulse[0] = 0.01;
      g.bc_pulse[1] = 2 * M_PI;
      g.bc_pulse[2] = 0.5;
      g.bc_pulse[3] = 3.0;
      g.x_min = -1.0;
      g.dx = 2.0 / (N - 1);
      g.dy = 1.0 / M;
      g.y_min = g.dy / 2;
      g.gamma = 5.0 / 3.0;
      g.krit = 0.3;
      g.ho_lo = (bcv & 1) ? -0.3 : -INFINITY;
      g.ho_hi = (bcv & 1) ? 0.4 : INFINITY;
      g.frozen_x0 = 0.1;
      g.frozen_ax = (bcv & 2) ? 0.25 : 0.0;
      g.frozen_ay = 0.3;
      for (int method = AST_HLL; method <= AST_HLLC_A; method++) {
        double T = 1e7, T_do = 0.01;
        ora_step(&g, s, u, s2, u2, &T, &T_do, method, 1);
        T = 1e7;
        ora_step_src(&g, s, u, s2, u2, &T, &T_do, 0.2, method, &src, 1);
        ora_step_rows(&g, s, u, s2, u2, &T_do, method, 2, 7, 1);
      }
    }
  }
  double T = 1e7, T_do = 1e-9, TT = 0;
  int i = 0;
  ora_funk_time(&T, &T_do, &TT, &i);
  free(s);
  free(u);
  free(s2);
  free(u2);
  free(nn2);
  free(nn3);
}

int main(void) {
  check_fluxes();
  check_steps();
  printf("sancheck ok\n");
  return 0;
}
