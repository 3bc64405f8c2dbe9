/* asteriy_oracle.h -- CPU ORACLE (test infrastructure only; see
 * asteriy_oracle.c).  Not part of the product: only tests/, the smoke check
 * and the bench's CPU-baseline leg load it. */
#ifndef ASTERIY_ORACLE_H
#define ASTERIY_ORACLE_H
#include "../include/asteriy.h"

typedef struct ora_phys {
  double ga, g1, g2, gp, gm, krit;
} ora_phys;

typedef struct ora_info {
  int pattern;   /* -1 vacuum, 0 two rarefactions, 1 shock+rarefaction,
                    2 two shocks, -9 none matched (NaN input)            */
  int iters;     /* newton or devtwo iterations                           */
  int capped;    /* AST_ITER_CAP reached                                  */
  int swapped;   /* sides swapped so that pI <= pII (ipiz)                */
  double p_star, u_star;
  double enI, pI, rI, enII, pII, rII;   /* post-swap normal states        */
} ora_info;

void ora_phys_init(ora_phys* ph, double gamma, double krit);
double ora_minmod(double x, double y);
double ora_linear(double x1, double t1, double x2, double t2, double x3,
                  double t3, double y);
void ora_linear2(double x1, double t1, double x2, double t2, double x3,
                 double t3, double y1, double y2, double* A, double* B);
double ora_hll(const ora_phys*, const double* Ls, const double* Lu,
               const double* Rs, const double* Ru, double n1, double n2,
               double* Ps, double* Pu, double rad);
double ora_hllc(const ora_phys*, const double* Ls, const double* Lu,
                const double* Rs, const double* Ru, double n1, double n2,
                double* Ps, double* Pu, double rad);
double ora_hllc_a(const ora_phys*, const double* Ls, const double* Lu,
                const double* Rs, const double* Ru, double n1, double n2,
                double* Ps, double* Pu, double rad);
/* passive-scalar fluxes: (Ps, Pu, dt) as the operators above plus PQ */
double ora_hllc_q(const ora_phys*, const double* Ls, const double* Lu,
                  const double* Rs, const double* Ru, double QL, double QR,
                  double n1, double n2, double* Ps, double* Pu, double* PQ,
                  double rad);
double ora_hllcq(const ora_phys*, const double* Ls, const double* Lu,
                 const double* Rs, const double* Ru, double QL, double QR,
                 double n1, double n2, double* Ps, double* Pu, double* PQ,
                 double rad);
double ora_hllcq_a(const ora_phys*, const double* Ls, const double* Lu,
                   const double* Rs, const double* Ru, double QL, double QR,
                   double n1, double n2, double* Ps, double* Pu, double* PQ,
                   double rad);
void ora_flux_q_batch(int method, double gamma, double krit, long n,
                      const double* in, double n1, double n2, double rad,
                      double* out, int nthreads);
/* MHD fluxes with a passive scalar (HLLDQ_Korolkov, Gas_Dinamic.cu:874):
 * L, R = rho, Q, p, v1, v2, v3, Bx, By, Bz; P = mass, momentum x/y/z,
 * field x/y/z, energy; the fan in the frame of the normal: S = SL, SLL, SM,
 * SRR, SR; U = UL, U*L, U**L, U**R, U*R, UR; F = FL, FR (9 components: rho,
 * rho u_n, rho u_t, rho u_m, e, b_n, b_t, b_m, Q with b = B/sqrt(4 pi));
 * branch 0..5 = the fan region of the face, -1 none */
typedef struct {
  double S[5];
  double U[6][9];
  double F[2][9];
  int branch;
} ora_mhd_fan;
double ora_hlldq(const ora_phys*, const double* L, const double* R, double n1,
                 double n2, double n3, double rad, int method, double* P,
                 double* PQ, ora_mhd_fan* fan);
void ora_flux_mhd_batch(int method, double gamma, double krit, long n,
                        const double* in, double n1, double n2, double n3,
                        double rad, double* out, int nthreads);
/* host helpers of the Monte-Carlo code (Gas_Dinamic.cu:1468-1602) */
double ora_minplus(double x, double y);
void ora_peresich(double y, double z, double Vy, double Vz, double R, double* t1, double* t2);
int ora_peresechenie(const ast_grid_t* g, double x0, double y0, double x, double y, double z,
                     double Vx, double Vy, double Vz, int* mode, double* t);
void ora_belong_point(const ast_grid_t* g, double x, double y, int* n, int* m);
double ora_velosity_1(double u, double cp);
double ora_velosity_2(double u, double cp);
double ora_velosity_3(double u, double cp);
double ora_godunov(const ora_phys*, const double* Ls, const double* Lu,
                   const double* Rs, const double* Ru, double n1, double n2,
                   double* Ps, double* Pu, double rad);
double ora_godunov_ex(const ora_phys*, const double* Ls, const double* Lu,
                      const double* Rs, const double* Ru, double n1, double n2,
                      double* Ps, double* Pu, double rad, const double* guess,
                      ora_info* info);
double ora_flux(int method, const ora_phys*, const double* Ls,
                const double* Lu, const double* Rs, const double* Ru,
                double n1, double n2, double* Ps, double* Pu, double rad);
void ora_flux_batch(int method, double gamma, double krit, long n,
                    const double* in, double n1, double n2, double rad,
                    double* out, int* iters, int nthreads);
void ora_godunov_guess_batch(double gamma, double krit, long n, const double* in,
                             double n1, double n2, double rad,
                             const double* guess, double* out, int* iters,
                             double* pstar);
void ora_step(const ast_grid_t* g, const double* s, const double* u,
              double* s2, double* u2, double* T, const double* T_do,
              int method, int nthreads);
/* neutral-H source planes of add_MK (kernel.cu:2495-2498): nn2 holds 3
 * doubles per cell (x, y used), nn3 one; either may be NULL */
typedef struct {
  const double* nn2;
  const double* nn3;
  double kmom, ken;
} ora_src;
void ora_step_src(const ast_grid_t* g, const double* s, const double* u,
                  double* s2, double* u2, double* T, const double* T_do,
                  double TT, int method, const ora_src* src, int nthreads);
double ora_step_rows(const ast_grid_t* g, const double* s, const double* u,
                     double* s2, double* u2, const double* T_do, int method,
                     int m0, int m1, int nthreads);
void ora_funk_time(double* T, double* T_do, double* TT, int* i);
int ora_max_threads(void);

/* toro.c: independent exact Riemann solver (Toro, "Riemann Solvers and
 * Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4). */
int toro_star(double gamma, double dl, double ul, double pl, double dr,
              double ur, double pr, double* pm, double* um, int* iters);
void toro_sample(double gamma, double dl, double ul, double pl, double dr,
                 double ur, double pr, double pm, double um, double S,
                 double* d, double* u, double* p);
int toro_sample_vacuum(double gamma, double dl, double ul, double pl,
                       double dr, double ur, double pr, double S, double* d,
                       double* u, double* p, double* speeds);
void toro_sod_profile(double gamma, double dl, double ul, double pl, double dr,
                      double ur, double pr, double x0, double t, long n,
                      const double* x, double* d, double* u, double* p);
#endif
