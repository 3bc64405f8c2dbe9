/*
 * toro.c -- CPU ORACLE (test infrastructure only): an exact Riemann solver
 * for the gamma-law Euler equations written from the textbook method of
 * E. F. Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics",
 * 3rd ed., ch. 4 (pressure functions f_K of eq. 4.6-4.7, Newton-Raphson on
 * f(p) = f_L(p) + f_R(p) + u_R - u_L with the two-rarefaction / PVRS starting
 * value, sampling of sect. 4.5).  It shares no code with the restatement of
 * the reference's solver (asteriy_oracle.c) and is used as an independent
 * second opinion on p*, u* and the sampled state.
 */
#include <math.h>
#include "asteriy_oracle.h"

static void t_prefun(double g, double p, double dk, double pk, double ck,
                     double* f, double* fd) {
  if (p <= pk) { /* rarefaction */
    double prat = p / pk;
    *f = 2.0 / (g - 1.0) * ck * (pow(prat, (g - 1.0) / (2.0 * g)) - 1.0);
    *fd = (1.0 / (dk * ck)) * pow(prat, -(g + 1.0) / (2.0 * g));
  } else { /* shock */
    double ak = 2.0 / ((g + 1.0) * dk);
    double bk = (g - 1.0) / (g + 1.0) * pk;
    double qrt = sqrt(ak / (bk + p));
    *f = (p - pk) * qrt;
    *fd = (1.0 - 0.5 * (p - pk) / (bk + p)) * qrt;
  }
}

/* Star-region pressure and velocity.  Returns 0, or -1 when the data
 * generate vacuum (pressure positivity condition violated). */
int toro_star(double g, double dl, double ul, double pl, double dr, double ur,
              double pr, double* pm, double* um, int* iters) {
  double cl = sqrt(g * pl / dl), cr = sqrt(g * pr / dr);
  double du = ur - ul;
  if (2.0 / (g - 1.0) * (cl + cr) <= du) return -1;
  /* starting value: PVRS, switched to the two-rarefaction or two-shock
   * approximations as in Toro sect. 9.5.1 */
  double ppv = 0.5 * (pl + pr) - 0.125 * du * (dl + dr) * (cl + cr);
  if (ppv < 0.0) ppv = 0.0;
  double pmin = fmin(pl, pr), pmax = fmax(pl, pr);
  double p0;
  if (pmax / pmin <= 2.0 && pmin <= ppv && ppv <= pmax) {
    p0 = ppv;
  } else if (ppv < pmin) {
    double z = (g - 1.0) / (2.0 * g);
    double pq = pow(pl / pr, z);
    double umq = (pq * ul / cl + ur / cr + 2.0 / (g - 1.0) * (pq - 1.0)) /
                 (pq / cl + 1.0 / cr);
    double ptl = 1.0 + (g - 1.0) / 2.0 * (ul - umq) / cl;
    double ptr = 1.0 + (g - 1.0) / 2.0 * (umq - ur) / cr;
    p0 = 0.5 * (pl * pow(ptl, 1.0 / z) + pr * pow(ptr, 1.0 / z));
  } else {
    double gel = sqrt((2.0 / ((g + 1.0) * dl)) / ((g - 1.0) / (g + 1.0) * pl + ppv));
    double ger = sqrt((2.0 / ((g + 1.0) * dr)) / ((g - 1.0) / (g + 1.0) * pr + ppv));
    p0 = (gel * pl + ger * pr - du) / (gel + ger);
  }
  if (p0 < 1e-300) p0 = 1e-300;
  double p = p0, fl, fld, fr, frd;
  int k;
  for (k = 1; k <= 200; k++) {
    t_prefun(g, p, dl, pl, cl, &fl, &fld);
    t_prefun(g, p, dr, pr, cr, &fr, &frd);
    double pn = p - (fl + fr + du) / (fld + frd);
    if (pn < 0.0) pn = 1e-300;
    double change = 2.0 * fabs((pn - p) / (pn + p));
    p = pn;
    if (change <= 1e-15) break;
  }
  t_prefun(g, p, dl, pl, cl, &fl, &fld);
  t_prefun(g, p, dr, pr, cr, &fr, &frd);
  *pm = p;
  *um = 0.5 * (ul + ur + fr - fl);
  if (iters) *iters = k;
  return 0;
}

/* Solution sampled at speed S = x/t (Toro sect. 4.5). */
void toro_sample(double g, double dl, double ul, double pl, double dr,
                 double ur, double pr, double pm, double um, double S,
                 double* d, double* u, double* p) {
  double cl = sqrt(g * pl / dl), cr = sqrt(g * pr / dr);
  double g1 = (g - 1.0) / (2.0 * g), g2 = (g + 1.0) / (2.0 * g);
  double g3 = 2.0 * g / (g - 1.0), g4 = 2.0 / (g - 1.0), g5 = 2.0 / (g + 1.0);
  double g6 = (g - 1.0) / (g + 1.0), g7 = (g - 1.0) / 2.0;
  (void)g3;
  if (S <= um) {
    if (pm <= pl) { /* left rarefaction */
      double shl = ul - cl;
      if (S <= shl) { *d = dl; *u = ul; *p = pl; return; }
      double cml = cl * pow(pm / pl, g1);
      double stl = um - cml;
      if (S > stl) { *d = dl * pow(pm / pl, 1.0 / g); *u = um; *p = pm; return; }
      *u = g5 * (cl + g7 * ul + S);
      double c = g5 * (cl + g7 * (ul - S));
      *d = dl * pow(c / cl, g4);
      *p = pl * pow(c / cl, 1.0 / g1);
      return;
    }
    { /* left shock */
      double pml = pm / pl;
      double sl = ul - cl * sqrt(g2 * pml + g1);
      if (S <= sl) { *d = dl; *u = ul; *p = pl; return; }
      *d = dl * (pml + g6) / (pml * g6 + 1.0); *u = um; *p = pm; return;
    }
  }
  if (pm > pr) { /* right shock */
    double pmr = pm / pr;
    double sr = ur + cr * sqrt(g2 * pmr + g1);
    if (S >= sr) { *d = dr; *u = ur; *p = pr; return; }
    *d = dr * (pmr + g6) / (pmr * g6 + 1.0); *u = um; *p = pm; return;
  }
  { /* right rarefaction */
    double shr = ur + cr;
    if (S >= shr) { *d = dr; *u = ur; *p = pr; return; }
    double cmr = cr * pow(pm / pr, g1);
    double str = um + cmr;
    if (S <= str) { *d = dr * pow(pm / pr, 1.0 / g); *u = um; *p = pm; return; }
    *u = g5 * (-cr + g7 * ur + S);
    double c = g5 * (cr - g7 * (ur - S));
    *d = dr * pow(c / cr, g4);
    *p = pr * pow(c / cr, 1.0 / g1);
  }
}

/* Solution sampled at speed S = x/t when the data generate vacuum
 * (Toro sect. 4.6, 2(cL + cR)/(g - 1) <= uR - uL): a left rarefaction from
 * its head uL - cL to the vacuum front uL + 2 cL/(g - 1), vacuum (d = p = 0,
 * u = S), a right rarefaction from its front uR - 2 cR/(g - 1) to its head
 * uR + cR.  The fans follow the rarefaction relations of sect. 4.5
 * (eqs. 4.56 and 4.63).  Returns 0, or -1 when the data do not generate
 * vacuum. */
int toro_sample_vacuum(double g, double dl, double ul, double pl, double dr,
                       double ur, double pr, double S, double* d, double* u,
                       double* p, double* speeds) {
  double cl = sqrt(g * pl / dl), cr = sqrt(g * pr / dr);
  double g4 = 2.0 / (g - 1.0), g5 = 2.0 / (g + 1.0), g6 = (g - 1.0) / (g + 1.0);
  double g7 = (g - 1.0) / 2.0;
  if (g4 * (cl + cr) > ur - ul) return -1;
  double shl = ul - cl, sfl = ul + g4 * cl; /* left head, left vacuum front  */
  double sfr = ur - g4 * cr, shr = ur + cr; /* right vacuum front, right head */
  if (speeds) { speeds[0] = shl; speeds[1] = sfl; speeds[2] = sfr; speeds[3] = shr; }
  if (S <= shl) { *d = dl; *u = ul; *p = pl; return 0; }
  if (S < sfl) {
    double q = g5 + g6 / cl * (ul - S);
    *d = dl * pow(q, g4);
    *u = g5 * (cl + g7 * ul + S);
    *p = pl * pow(q, 2.0 * g / (g - 1.0));
    return 0;
  }
  if (S <= sfr) { *d = 0.0; *u = S; *p = 0.0; return 0; }
  if (S < shr) {
    double q = g5 - g6 / cr * (ur - S);
    *d = dr * pow(q, g4);
    *u = g5 * (-cr + g7 * ur + S);
    *p = pr * pow(q, 2.0 * g / (g - 1.0));
    return 0;
  }
  *d = dr; *u = ur; *p = pr;
  return 0;
}

/* Exact solution of a shock-tube problem at time t on points x
 * (discontinuity at x0). */
void toro_sod_profile(double g, double dl, double ul, double pl, double dr,
                      double ur, double pr, double x0, double t, long n,
                      const double* x, double* d, double* u, double* p) {
  double pm, um;
  toro_star(g, dl, ul, pl, dr, ur, pr, &pm, &um, 0);
  for (long i = 0; i < n; i++) {
    double S = (x[i] - x0) / t;
    toro_sample(g, dl, ul, pl, dr, ur, pr, pm, um, S, &d[i], &u[i], &p[i]);
  }
}
