/*
 * asteriy.h -- C-ABI of the MI355X (gfx950) implementation of Asteriy_2d's
 * explicit (z,r) finite-volume step.
 *
 * Every entry point takes plain pointers and sizes.  Device pointers are
 * caller-owned (hipMalloc'd or from any allocator); nothing here allocates
 * on the device except the objects that say so (ast_sim_*, ast_batch_*).
 * All calls are asynchronous on the given stream (a hipStream_t passed as
 * void*, NULL = default stream) and return an AST_* status code; no entry
 * point calls exit().  ast_last_error() gives the message of the last
 * failure on the calling thread.
 *
 * State layout at the boundary is the reference's (kernel.cu:3255-3265):
 * two planes of N*M pairs, s[k] = {rho, p} and u[k] = {vx, vy}, with the
 * flat index k = m*N + n (n along z/x, fastest; m along r/y),
 * kernel.cu:2515-2517.  The FP64 entry points take double pairs (the
 * reference's double2), the FP32 ones float pairs.
 */
#ifndef ASTERIY_H
#define ASTERIY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  AST_OK = 0,
  AST_ERR_ARG = 1,      /* invalid argument (null pointer, bad method, ...) */
  AST_ERR_HIP = 2,      /* a HIP runtime call failed                        */
  AST_ERR_SIZE = 3,     /* size out of the supported range                  */
  AST_ERR_STATE = 4,    /* object used in the wrong state                   */
  AST_ERR_NOGPU = 5     /* no HIP device visible                            */
};

/* ---- flux methods: the numbering of add_MK's `method` (kernel.cu:2385-2463),
 *      with method 1 bound to the live HLLC_2d_Korolkov (Gas_Dinamic.cu:1814);
 *      AST_HLLC_A is HLLC_Aleksashov (kernel.cu:767-1079, add_MK's method 1)
 *      with the zero magnetic field it is called with. */
enum { AST_HLL = 0, AST_HLLC = 1, AST_EXACT = 2, AST_HLLC_A = 3 };
/* passive-scalar fluxes (ast_flux_q_*): HLLCQ_Korolkov_2D
 * (Gas_Dinamic.cu:249) and HLLCQ_Aleksashov (Gas_Dinamic.cu:355, zero
 * magnetic field) */
enum { AST_HLLCQ = 4, AST_HLLCQ_A = 5 };
/* MHD fluxes with a passive scalar (ast_flux_mhd_f64): HLLDQ_Korolkov
 * (Gas_Dinamic.cu:874) with metod 2 (HLLC for MHD) and metod 3 (HLLD) */
enum { AST_HLLC_MHD = 6, AST_HLLD = 7 };

/* ---- geometry of the conservative update ------------------------------- */
enum {
  AST_PLANAR = 0,        /* kernel.cu:3059-3073                              */
  AST_AXISYMMETRIC = 1   /* r-weighted source, KorolkovSolver.cu:13-32       */
};

/* ---- boundary conditions (ghost-cell rules of Kernel_TVD/add_MK) -------- */
enum {
  AST_BC_OUTFLOW = 0,          /* zero gradient, kernel.cu:2549-2552        */
  AST_BC_REFLECT = 1,          /* mirror, normal velocity negated,
                                  kernel.cu:2564-2569, 2609-2614            */
  AST_BC_DIRICHLET = 2,        /* fixed state, kernel.cu:2577-2580          */
  AST_BC_OUTFLOW_CLAMPED = 3,  /* zero gradient with the normal velocity
                                  limited to bc_vclamp, kernel.cu:2340-2354 */
  AST_BC_DIRICHLET_PULSED = 4  /* fixed state whose density is multiplied by
                                  1 + a*sin(w*TT) in the rows with y < y_lim
                                  while TT < t_lim (bc_pulse = {a, w, y_lim,
                                  t_lim}; TT the simulation time before the
                                  step): the inflow perturbation of add2,
                                  kernel.cu:2022-2025 (a = 0.01, w = 2 pi,
                                  y_lim = 0.5, t_lim = 3).  FP64 step only */
};
enum { AST_XLO = 0, AST_XHI = 1, AST_YLO = 2, AST_YHI = 3 };

/* Grid and physics: replaces the compile-time macros of Header.h:6-63.
 * Cell centres are x = x_min + n*dx and y = y_min + m*dy (Header.h:11-16:
 * the reference uses dx = (x_max-x_min)/(N-1), dy = y_max/M, y_min = dy/2).
 * A face is "high order" (reconstructed states + `method`) when its
 * coordinate lies strictly inside (ho_lo, ho_hi); other faces use HLL on
 * the unreconstructed cell states, as Kernel_TVD does outside x in (10,40)
 * (kernel.cu:2799-2857).  ho_lo=-inf, ho_hi=+inf reconstructs everywhere.
 * Cells with ((x-frozen_x0)/frozen_ax)^2 + (y/frozen_ay)^2 < 1 are copied
 * unchanged (add_MK, kernel.cu:2317-2324); frozen_ax <= 0 disables it. */
typedef struct ast_grid {
  int32_t N, M;               /* cells along x (z) and y (r); N, M >= 2     */
  int32_t geometry;           /* AST_PLANAR / AST_AXISYMMETRIC              */
  int32_t bc[4];              /* per side AST_XLO..AST_YHI: AST_BC_*        */
  double x_min, dx;
  double y_min, dy;
  double gamma;               /* ratio of specific heats (Header.h:17)      */
  double krit;                /* CFL number per face (Header.h:47)          */
  double bc_state[4][4];      /* Dirichlet state per side: rho, p, vx, vy   */
  double bc_vclamp[4];        /* AST_BC_OUTFLOW_CLAMPED normal-velocity cap */
  double ho_lo, ho_hi;        /* high-order face window on x                */
  double frozen_x0, frozen_ax, frozen_ay;
  double bc_pulse[4];         /* AST_BC_DIRICHLET_PULSED: a, w, y_lim, t_lim */
} ast_grid_t;

/* Exact-solver and precision counters, accumulated on the device when a
 * device pointer to one is passed (NULL = not counted). */
typedef struct ast_stats {
  unsigned long long exact_faces;     /* faces solved by the exact solver  */
  unsigned long long newton_faces;    /* two-shock faces (newton)          */
  unsigned long long newton_iters;
  unsigned long long bisect_faces;    /* shock+rarefaction faces (devtwo)  */
  unsigned long long bisect_iters;
  unsigned long long rarefaction_faces;
  unsigned long long vacuum_faces;
  unsigned long long capped;          /* iteration cap reached             */
  unsigned long long fp64_faces;      /* FP32 path: faces solved in FP64   */
  unsigned long long mlp_faces;       /* faces seeded by the learned guess */
} ast_stats_t;

/* Largest iteration count of newton/devtwo (the reference loops without a
 * bound and prints an alarm at 1100, kernel.cu:665-668, 713-717). */
#define AST_ITER_CAP 1100

/* ---- library --------------------------------------------------------- */
const char* ast_last_error(void);
const char* ast_version(void);
/* Number of visible HIP devices (0 without a GPU); never fails. */
int ast_device_count(void);
/* Validates a grid description; AST_OK or AST_ERR_ARG with a message. */
int ast_grid_check(const ast_grid_t* g);

/* ---- face-flux operators, batched ---------------------------------------
 * Replaces the per-face device operators
 *   double HLL(double2& Ls, double2& Lu, double2& Rs, double2& Ru,
 *              double n1, double n2, double2& Ps, double2& Pu, double rad)
 *     kernel.cu:1399-1400,
 *   HLLC_2d_Korolkov(..., double rad, bool nul_potok=false)
 *     Gas_Dinamic.cu:1814-1815,
 *   Godunov_Solver_Alexashov(...) kernel.cu:128-129,
 * applied to n independent faces with a common normal (n1, n2) and `rad`.
 * in  : 8 planes of n values (SoA): rhoL, pL, vxL, vyL, rhoR, pR, vxR, vyR
 *       (= Ls.x, Ls.y, Lu.x, Lu.y, Rs.x, Rs.y, Ru.x, Ru.y).
 * out : 5 planes of n values: Ps.x (mass), Ps.y (energy), Pu.x, Pu.y
 *       (momentum, lab axes), and the operator's return value (local dt).
 * The FP32 variant computes in float; its exact solver keeps the p* solve
 * in FP64 near vacuum and in strong rarefactions. */
int ast_flux_f64(int method, size_t n, const double* in, double n1, double n2,
                 double rad, double gamma, double krit, double* out,
                 ast_stats_t* stats, void* stream);
int ast_flux_f32(int method, size_t n, const float* in, float n1, float n2,
                 float rad, float gamma, float krit, float* out,
                 ast_stats_t* stats, void* stream);

/* Passive-scalar fluxes, replacing
 *   double HLLCQ_Korolkov_2D(const double2& Ls, const double2& Lu,
 *       const double2& Rs, const double2& Ru, const double& LQ,
 *       const double& RQ, double n1, double n2, double2& Ps, double2& Pu,
 *       double& PQ, double rad)          Gas_Dinamic.cu:249-250
 *   HLLCQ_Aleksashov(same arguments)    Gas_Dinamic.cu:355-356 (zero field)
 * and the scalar of the live solver HLLC_2d_Korolkov_b_s (Gas_Dinamic.cu:
 * 1604, whose PQ HLLC_2d_Korolkov computes with Q = 1 and drops), over n
 * faces: in = 10 planes (the 8 of ast_flux_f64, then Q_L, Q_R), out = 6
 * planes (the 5 of ast_flux_f64, then PQ).  method: AST_HLLC (the live
 * solver's wave speeds), AST_HLLCQ or AST_HLLCQ_A (per-side wave speeds;
 * with a zero field the two reference operators are the same HLLC, and the
 * same as HLLC_Aleksashov, AST_HLLC_A). */
int ast_flux_q_f64(int method, size_t n, const double* in, double n1, double n2,
                   double rad, double gamma, double krit, double* out, void* stream);

/* MHD fluxes with a passive scalar, replacing
 *   double HLLDQ_Korolkov(ro_L, Q_L, p_L, v1_L, v2_L, v3_L, Bx_L, By_L, Bz_L,
 *       ro_R, Q_R, p_R, v1_R, v2_R, v3_R, Bx_R, By_R, Bz_R, double* P,
 *       double& PQ, n1, n2, n3, double& rad, int metod)
 *   Gas_Dinamic.cu:874-877 (metod 2: AST_HLLC_MHD, metod 3: AST_HLLD)
 * over n faces with a common normal (n1, n2, n3), gamma for both the sound
 * speed and the energy (the reference's ggg and ga), fields in the
 * reference's units (b = B / sqrt(4 pi) inside the solver).
 * in  : 18 planes of n values: rho, Q, p, v1, v2, v3, Bx, By, Bz of the
 *       left state, then the same 9 of the right state.
 * out : 10 planes: P[0..7] (mass; momentum x, y, z; field x, y, z; energy),
 *       PQ, and the operator's return value krit*rad/max(|SL|,|SR|).
 * The frame follows the reference: a normal component above 0.1 (then
 * below -0.1), tested in the order n1, n2, n3, picks the tangent pair; a
 * normal that picks none, which the reference reports as an error and then
 * solves with zero tangents, is refused here (AST_ERR_ARG). */
int ast_flux_mhd_f64(int method, size_t n, const double* in, double n1, double n2,
                     double n3, double rad, double gamma, double krit, double* out,
                     void* stream);

/* ---- host helpers of the Monte-Carlo neutral-hydrogen code --------------
 * The particle-tracing and rate functions of Gas_Dinamic.cu:1468-1602
 * (declared Header.h:135-146) over device arrays, one item per lane; the
 * per-particle __host__ __device__ forms are in asteriy_mc.hpp.  Results are
 * the reference's, its thresholds and its nine-digit sqrt(pi) (Header.h:65)
 * included.
 * ast_mc_rates_f64: Velosity_1, Velosity_2, Velosity_3 (Gas_Dinamic.cu:1566,
 *   1578, 1591) of n pairs (u[i], cp[i]); out = 3 planes of n.
 * ast_mc_cell_exit_f64: Peresechenie (Gas_Dinamic.cu:1492) of n particles
 *   with the grid's dx, dy; in = 8 planes (x0, y0, x, y, z, Vx, Vy, Vz: the
 *   cell centre, the position, the velocity); out: t[i], mode[i] (1..4: the
 *   side, 0 none) and hit[i] (the function's value, 0/1).
 * ast_mc_belong_point_f64: Belong_point (Gas_Dinamic.cu:1559) of n points;
 *   xy = 2 planes (x, r); nm = 2 int planes (n, m), the grid's cell indices
 *   (the reference's (int)(y/dy) with the grid's own first cell centre). */
int ast_mc_rates_f64(size_t n, const double* u, const double* cp, double* out, void* stream);
int ast_mc_cell_exit_f64(const ast_grid_t* g, size_t n, const double* in, double* t, int* mode,
                         int* hit, void* stream);
int ast_mc_belong_point_f64(const ast_grid_t* g, size_t n, const double* xy, int* nm,
                            void* stream);

/* ---- the step ------------------------------------------------------------
 * Replaces
 *   __global__ void Kernel_TVD(double2* s, double2* u, double2* s2,
 *       double2* u2, double* T, double* T_do, double* TT, int method)
 *   kernel.cu:2513
 * with the geometry of Header.h taken from *g and `method` honoured
 * (AST_HLL / AST_HLLC / AST_EXACT on high-order faces).  Reads s,u (N*M
 * pairs each), writes s2,u2, reads *T_do, and lowers *T to the minimum
 * over all faces of krit*rad/max|wave speed| (capped at 1000,
 * kernel.cu:2530): a deterministic min, not the reference's racy store
 * (kernel.cu:3013-3017).  TT, the simulation time, is read only when a
 * side is AST_BC_DIRICHLET_PULSED (add2's inflow perturbation); the
 * reference's live body leaves it unused.
 * ast_step_f64 is the step in IEEE double: each face solved once and shared
 * by its two cells, slopes from neighbour differences; it agrees with the
 * reference's step (which solves each face from both cells with
 * coordinate-difference slopes) to round-off, within 1e-12 relative per
 * step.  ast_step_f32 is the fast path: FP32 state, each face evaluated
 * once and shared by its two cells, FP64 kept for the dt min and (exact
 * solver) the p* solve near vacuum; `mlp` (may be NULL) seeds the exact
 * solver's Newton iteration with the learned guess.  The fast path needs
 * rho > 0 and p > 0 in every cell (a state the step itself keeps) and in
 * every Dirichlet side state (else AST_ERR_ARG): for such states the
 * reference's positivity fallback of the reconstructed face values can
 * never fire, and the fast path does not test for it. */
int ast_step_f64(const ast_grid_t* g, const double* s, const double* u,
                 double* s2, double* u2, double* T, const double* T_do,
                 double* TT, int method, ast_stats_t* stats, void* stream);

typedef struct ast_mlp ast_mlp_t;   /* the learned guess, see below */

int ast_step_f32(const ast_grid_t* g, const float* s, const float* u,
                 float* s2, float* u2, double* T, const double* T_do,
                 double* TT, int method, const ast_mlp_t* mlp,
                 ast_stats_t* stats, void* stream);

/* Replaces __global__ void funk_time(double* T, double* T_do, double* TT,
 * int* i), kernel.cu:1525-1536: T_do <- T; TT += T_do; T <- 1e7; i += 1. */
int ast_funk_time(double* T, double* T_do, double* TT, int* i, void* stream);

/* FP64 compensated (Neumaier) sums over the grid, for conservation
 * diagnostics: out[0] = sum rho*w*dV, out[1] = sum E*w*dV with
 * E = p/(gamma-1) + rho*|u|^2/2 and w = 1 (planar) or y (axisymmetric).
 * `fp32` selects float (1) or double (0) state planes.  out is a device
 * buffer of ast_conservation_work_doubles() doubles (results first). */
int ast_conservation_sums(const ast_grid_t* g, const void* s, const void* u,
                          int fp32, double* out, void* stream);

/* ast_conservation_sums without the workspace: the partial sums live in a
 * fixed 8 KiB device array of the library (one per GPU, held until the
 * process ends), so the call allocates nothing and needs only `out` (2
 * doubles, device).  No counterpart in the reference: it is for callers that
 * cannot allocate -- code under HIP graph capture, tight stepping or serving
 * loops, harnesses that forbid hipMalloc.  Its scratch use does not grow
 * with the grid (256 workgroups, 4 partials each), so every valid grid fits.
 * Calls on one GPU share that scratch: calls that overlap in time on one
 * GPU (other streams, other threads) race and must be ordered by the
 * caller. */
int ast_conservation_sums_noalloc(const ast_grid_t* g, const void* s, const void* u,
                                  int fp32, double* out, void* stream);

/* Converts between the FP64 and FP32 state planes (n pairs). */
int ast_convert_f64_to_f32(size_t n, const double* src, float* dst, void* stream);
int ast_convert_f32_to_f64(size_t n, const float* src, double* dst, void* stream);

/* Same as ast_flux_f32, with the learned newton start (mlp: device
 * ast_mlp_t, NULL = off; used by strong two-shock faces) and per-face
 * solver diagnostics for the exact solver: info (may be NULL) receives 2
 * ints per face: the newton/root iterations, and flags = (pattern + 1) |
 * 8*capped | 16*fp64_island with pattern -1 vacuum, 0 two rarefactions,
 * 1 shock+rarefaction, 2 two shocks. */
typedef struct ast_mlp ast_mlp_t;
int ast_flux_f32_info(int method, size_t n, const float* in, float n1, float n2,
                      float rad, float gamma, float krit, const ast_mlp_t* mlp,
                      float* out, ast_stats_t* stats, int* info, void* stream);

/* FP64 exact solver (Godunov_Solver_Alexashov, kernel.cu:128) over a batch
 * with per-face newton starting points: guess (device, n floats, may be
 * NULL) holds p* guesses (<= 0 or NaN = the reference's PVRS start).  The
 * stopping rule is the reference's (|dp/pee| < 1e-9, pee = PVRS estimate),
 * so the guess changes the iteration count only.  info as above (FP64
 * faces never set the island flag). */
/* The step with the neutral-H source terms of
 *   __global__ void add_MK(double2* s, double2* u, double2* s2, double2* u2,
 *       double* nn1, double3* nn2, double* nn3, double* T, double* T_do,
 *       int method, int step)   kernel.cu:2293
 * behind it (kernel.cu:2495-2498): the momentum of each cell gains
 * T_do*(rho*kmom)*nn2.{x,y} and its energy T_do*ken*nn3, before the density
 * and pressure floors; frozen cells are copied.  nn2 holds 3 doubles per cell
 * (the reference's double3 plane, x and y read), nn3 one double per cell
 * (device memory, N*M cells, row-major like s); either may be NULL (no such
 * term); nn1 is not read by the reference's step and has no slot.  The
 * reference's coefficients are kmom = n_H/Kn and ken = 0.0*n_H/Kn
 * (Header.h:84-85).  src == NULL is ast_step_f64. */
typedef struct ast_sources {
  const double* nn2;
  const double* nn3;
  double kmom;
  double ken;
} ast_sources_t;
int ast_step_f64_src(const ast_grid_t* g, const double* s, const double* u,
                     double* s2, double* u2, double* T, const double* T_do,
                     double* TT, int method, const ast_sources_t* src,
                     ast_stats_t* stats, void* stream);
/* The same source terms (kernel.cu:2495-2498) behind the FP32 fast step
 * (ast_step_f32 on FP32 state; the planes stay the reference's FP64 ones,
 * each value rounded to FP32 before use; the exact solver runs without the
 * learned start).  src == NULL is ast_step_f32 without a learned guess. */
int ast_step_f32_src(const ast_grid_t* g, const float* s, const float* u,
                     float* s2, float* u2, double* T, const double* T_do,
                     double* TT, int method, const ast_sources_t* src,
                     ast_stats_t* stats, void* stream);
int ast_flux_f64_guess(size_t n, const double* in, double n1, double n2,
                       double rad, double gamma, double krit, const float* guess,
                       double* out, ast_stats_t* stats, int* info, void* stream);

/* ---- learned p* initial guess ------------------------------------------
 * A 6 -> 32 -> 32 -> 32 -> 1 tanh MLP maps the normalised primitives of a
 * two-shock face to log(p_star / p_ref); it only replaces the starting point
 * of newton (kernel.cu:641-642), whose stopping rule keeps the PVRS guess as
 * its scale.  An ast_mlp_t is a block of ast_mlp_bytes() bytes of device
 * memory (caller-allocated): AST_MLP_NPARAM FP32 master parameters
 * (W1[32][6] b1[32] W2[32][32] b2 W3[32][32] b3 w4[32] b4) followed by the
 * BF16 MFMA operands derived from them. */
#define AST_MLP_IN 6
#define AST_MLP_HID 32
#define AST_MLP_NPARAM (AST_MLP_IN * AST_MLP_HID + AST_MLP_HID + \
                        2 * (AST_MLP_HID * AST_MLP_HID + AST_MLP_HID) + \
                        AST_MLP_HID + 1)
size_t ast_mlp_bytes(void);
/* Floats of workspace ast_mlp_loss_grad needs (per-workgroup partials). */
size_t ast_mlp_work_floats(void);

/* The solver's face log: normalised features of n faces given as the
 * 8-plane SoA of ast_flux_f64 (normal (n1,n2)): feats is 6 planes of n
 * floats; target (may be NULL) receives log(p_star / p_ref) from the FP64
 * exact solve at the reference tolerance; mask (may be NULL) is 1 for
 * two-shock faces (the only ones the guess serves), else 0. */
int ast_mlp_features(size_t n, const double* in, double n1, double n2,
                     double gamma, float* feats, float* target,
                     unsigned char* mask, void* stream);
/* Forward pass of n feature rows into pred (n floats).  precision 0: FP32
 * masters on the vector ALU; 1: BF16 operands on MFMA, FP32 accumulate. */
int ast_mlp_forward(const ast_mlp_t* mlp, size_t n, const float* feats,
                    float* pred, int precision, void* stream);
/* Mean-squared-error loss over the rows with mask != 0 (mask may be NULL)
 * and its gradient w.r.t. the FP32 masters: grad receives AST_MLP_NPARAM
 * floats, loss one float (device pointers); work holds
 * ast_mlp_work_floats() floats.  Deterministic. */
int ast_mlp_loss_grad(const ast_mlp_t* mlp, size_t n, const float* feats,
                      const float* target, const unsigned char* mask,
                      float* grad, float* loss, float* work, void* stream);
/* One Adam step (beta 0.9/0.999, eps 1e-8; state: 2*NPARAM floats,
 * zero-initialised by the caller, step >= 1) on the masters, then the BF16
 * operands are rebuilt. */
int ast_mlp_adam(ast_mlp_t* mlp, const float* grad, float* state, int step,
                 float lr, void* stream);
/* Rebuilds the BF16 operands from the masters. */
int ast_mlp_refresh(ast_mlp_t* mlp, void* stream);
/* Copies NPARAM host floats into / out of the masters (synchronous). */
int ast_mlp_set_params(ast_mlp_t* mlp, const float* host_params, void* stream);
int ast_mlp_get_params(const ast_mlp_t* mlp, float* host_params, void* stream);

/* ---- the host runtime: the reference's time loop ------------------------
 * Replaces main()'s protocol (kernel.cu:3868-3906: Kernel_TVD, funk_time,
 * ping-pong by swapping arguments, a device sync after every launch) with a
 * native object: it owns the two pairs of state planes and the device time
 * scalars (T, T_do, TT, i; initialised to 1e7, 1e-9, 0, 0 as in
 * kernel.cu:3321-3324) and replays runs of 20 steps from a captured HIP
 * graph on its own stream.  precision 0 = FP64 (ast_step_f64), 1 = FP32
 * (ast_step_f32).  Host state crosses in the reference layout (N*M double
 * pairs s = {rho,p}, u = {vx,vy}). */
typedef struct ast_sim ast_sim_t;
int ast_sim_create(const ast_grid_t* g, int precision, int method, ast_sim_t** out);
int ast_sim_destroy(ast_sim_t* sim);
int ast_sim_set_state(ast_sim_t* sim, const double* s, const double* u);
int ast_sim_get_state(ast_sim_t* sim, double* s, double* u);
/* Device planes holding the current state (FP64 or FP32 pairs). */
int ast_sim_device_state(ast_sim_t* sim, void** s, void** u);
int ast_sim_set_time(ast_sim_t* sim, double T, double T_do, double TT, int i);
int ast_sim_get_time(ast_sim_t* sim, double* T, double* T_do, double* TT, int* i);
/* Exact-solver counters accumulated by the simulation (reset on request),
 * while counting is on (ast_sim_collect_stats; default off: the counting
 * variants of the exact-solver kernels run markedly slower). */
int ast_sim_get_stats(ast_sim_t* sim, ast_stats_t* out, int reset);
int ast_sim_collect_stats(ast_sim_t* sim, int on);
/* Learned guess (device ast_mlp_t, NULL = off) for the FP32 exact step.
 * The network costs the solver kernel registers (occupancy) whether or not a
 * face uses it, so the simulation runs it only where it can serve: each run
 * of 20 steps launches one solver kernel, the plain one or the one with the
 * learned start, chosen from the strong-two-shock-face flag of the run two
 * before (read after that run's completion, never waiting for the run in
 * flight); learned runs and every fourth plain run set that flag (the test
 * costs the plain kernel ~4%); the first two runs after a create or load
 * take the plain solver (or the modes the checkpoint recorded). */
int ast_sim_set_mlp(ast_sim_t* sim, const ast_mlp_t* mlp);
/* Neutral-H source planes for every later step (ast_step_f64_src,
 * ast_step_f32_src), copied from host memory: nn2 N*M triples, nn3 N*M
 * doubles, either may be NULL, both NULL removes them.  An FP32 simulation
 * with planes runs the exact solver without the learned start. */
int ast_sim_set_sources(ast_sim_t* sim, const double* nn2, const double* nn3,
                        double kmom, double ken);
/* The learned modes (0 plain solver, 1 learned start) the next two runs
 * will take; waits for every launched run. */
int ast_sim_learned_modes(ast_sim_t* sim, int* modes);
/* -1 (default): the automatic choice above; 0 / 1: every run takes the plain
 * solver / the learned start (for strong faces that last too few steps for
 * the automatic choice to meet them). */
int ast_sim_set_learned_mode(ast_sim_t* sim, int mode);
/* Strip tiling of the step.  FP32: rows > 0 rows per wave strip, rows = -S
 * S strips of (nearly) equal height per band of 60 columns; strips of 4 to
 * 64 rows inside the grid take the fast strip path.  FP64: rows > 0 rows
 * per wave strip (default 6). */
int ast_sim_set_tile_rows(ast_sim_t* sim, int rows);
/* Times the step for each candidate tiling (NULL = a default list: FP32 4,
 * 8, 12, 16 rows and the tilings -S that give every SIMD 2, 3, 4 or 5 equal
 * waves; FP64 4, 6, 8, 12, 16, 32 rows), keeps the fastest; reports it (may
 * be NULL).  State and time are left unchanged. */
int ast_sim_autotune(ast_sim_t* sim, const int* candidates, int ncand, int reps,
                     int* best_rows, float* best_us);
/* The simulation's stream (hipStream_t) for event timing. */
void* ast_sim_stream(ast_sim_t* sim);
/* Advances `steps` steps asynchronously (use_graph: replay the captured
 * graph of 20 steps; a remainder is launched directly). */
int ast_sim_run(ast_sim_t* sim, int steps, int use_graph);
int ast_sim_sync(ast_sim_t* sim);
/* The reference's binary state file (kernel.cu:3414-3436 read,
 * 4111-4148 write): per cell in k order, 6 doubles x, y, rho, p, vx, vy. */
int ast_sim_write_b31(ast_sim_t* sim, const char* path);
int ast_sim_read_b31(ast_sim_t* sim, const char* path);
/* The reference's text outputs of a run (kernel.cu:4127-4215): a Tecplot
 * FEPOINT quadrilateral file of every stride-th node (the reference uses 2)
 * with X, Y, Ro, P, Vx, Vy, Mach, T = p/rho, Zav (vorticity), and the row
 * next to the axis (x y rho p vx vy 1/c per line). */
int ast_sim_write_tecplot(ast_sim_t* sim, const char* path, int stride);
int ast_sim_write_axis(ast_sim_t* sim, const char* path);
/* Checkpoint with the time state (T, T_do, TT, i) the reference does not
 * save: a resumed run continues bit for bit. */
int ast_sim_save(ast_sim_t* sim, const char* path);
int ast_sim_load(ast_sim_t* sim, const char* path);

/* ---- batched parameter sweep ------------------------------------------
 * Many independent FP32 instances (uneven grids, own physics, own dt) in
 * one persistent launch per step: a fixed grid of workgroups walks the
 * strip table of all instances; a second kernel advances every instance's
 * time (funk_time per instance); both are captured in one HIP graph.
 * Instances share the method.  Tile heights are chosen per grid shape by
 * ast_batch_autotune. */
typedef struct ast_batch ast_batch_t;
int ast_batch_create(int count, const ast_grid_t* grids, int method, ast_batch_t** out);
int ast_batch_destroy(ast_batch_t* b);
int ast_batch_set_state(ast_batch_t* b, int inst, const double* s, const double* u);
int ast_batch_get_state(ast_batch_t* b, int inst, double* s, double* u);
int ast_batch_get_time(ast_batch_t* b, int inst, double* T, double* T_do, double* TT, int* i);
/* Sets one instance's time scalars (as ast_sim_set_time). */
int ast_batch_set_time(ast_batch_t* b, int inst, double T, double T_do, double TT, int i);
/* Per-shape strip heights: times candidate heights on the batch itself. */
int ast_batch_autotune(ast_batch_t* b, int reps);
int ast_batch_run(ast_batch_t* b, int steps, int use_graph);
int ast_batch_sync(ast_batch_t* b);
void* ast_batch_stream(ast_batch_t* b);
/* Total cells of all instances (cell-updates per step). */
long long ast_batch_cells(const ast_batch_t* b);

/* ---- tile configuration of the stand-alone FP32 step entry point --------
 * Strip tiling as for ast_sim_set_tile_rows (default 8 rows per strip). */
int ast_set_tile_rows(int rows);
int ast_get_tile_rows(void);
/* Doubles of device memory ast_conservation_sums needs at `out` (the two
 * results followed by its partial sums). */
size_t ast_conservation_work_doubles(void);

#ifdef __cplusplus
}
#endif
#endif /* ASTERIY_H */
