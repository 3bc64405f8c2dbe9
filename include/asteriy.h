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
 *      with method 1 bound to the live HLLC_2d_Korolkov (Gas_Dinamic.cu:1814). */
enum { AST_HLL = 0, AST_HLLC = 1, AST_EXACT = 2 };

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
  AST_BC_OUTFLOW_CLAMPED = 3   /* zero gradient with the normal velocity
                                  limited to bc_vclamp, kernel.cu:2340-2354 */
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
 * (kernel.cu:3013-3017).  TT is accepted for signature parity and unused,
 * as in the reference's live body.
 * ast_step_f64 is the FP64, cell-by-cell restatement of the reference step
 * (each cell evaluates its own four faces, in the reference's operation
 * order).  ast_step_f32 is the fast path: FP32 state, each face evaluated
 * once and shared by its two cells, FP64 kept for the dt min and (exact
 * solver) the p* solve near vacuum; `mlp` (may be NULL) seeds the exact
 * solver's Newton iteration with the learned guess. */
int ast_step_f64(const ast_grid_t* g, const double* s, const double* u,
                 double* s2, double* u2, double* T, const double* T_do,
                 double* TT, int method, ast_stats_t* stats, void* stream);

typedef struct ast_mlp ast_mlp_t;   /* see the learned-guess section below */

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
 * pointer to 2 doubles. */
int ast_conservation_sums(const ast_grid_t* g, const void* s, const void* u,
                          int fp32, double* out, void* stream);

/* Converts between the FP64 and FP32 state planes (n pairs). */
int ast_convert_f64_to_f32(size_t n, const double* src, float* dst, void* stream);
int ast_convert_f32_to_f64(size_t n, const float* src, double* dst, void* stream);

/* ---- learned p* initial guess ------------------------------------------
 * A 6 -> 32 -> 32 -> 32 -> 1 tanh MLP maps the normalised primitives of a
 * two-shock face to log(p_star / p_ref); it only replaces the starting point of
 * newton (kernel.cu:641-642), whose stopping rule keeps the PVRS guess as
 * its scale.  Parameters are FP32 masters (trained) and BF16 copies
 * (inference on MFMA).  ast_mlp_t lives in device memory and is laid out
 * as AST_MLP_NPARAM floats followed by the BF16 copy; create it with
 * ast_mlp_alloc-free helpers below on caller-owned memory. */
#define AST_MLP_IN 6
#define AST_MLP_HID 32
#define AST_MLP_NPARAM (AST_MLP_IN * AST_MLP_HID + AST_MLP_HID + \
                        2 * (AST_MLP_HID * AST_MLP_HID + AST_MLP_HID) + \
                        AST_MLP_HID + 1)
/* Bytes of device memory an ast_mlp_t occupies. */
size_t ast_mlp_bytes(void);

/* Normalised features of n faces given as the 8-plane SoA of
 * ast_flux_f64 (normal (n1,n2)): feats is 6 planes of n floats, target
 * (may be NULL) receives log(p_star / p_ref) from the FP64 exact solve, and
 * mask (may be NULL) is 1 for two-shock faces (the only ones the guess
 * serves), else 0. */
int ast_mlp_features(size_t n, const double* in, double n1, double n2,
                     double gamma, float* feats, float* target,
                     unsigned char* mask, void* stream);
/* Forward pass of n feature rows; pred receives n floats.  precision 0:
 * FP32 master weights on the vector ALU; 1: BF16 weights on MFMA with
 * FP32 accumulation. */
int ast_mlp_forward(const ast_mlp_t* mlp, size_t n, const float* feats,
                    float* pred, int precision, void* stream);
/* Mean-squared-error loss and its gradient w.r.t. the FP32 parameters
 * over n rows (mask may be NULL): grad receives AST_MLP_NPARAM floats,
 * loss one float. */
int ast_mlp_loss_grad(const ast_mlp_t* mlp, size_t n, const float* feats,
                      const float* target, const unsigned char* mask,
                      float* grad, float* loss, void* stream);
/* One Adam step on the FP32 masters from grad (state: 2*NPARAM floats,
 * zero-initialised by the caller), then refreshes the BF16 copy. */
int ast_mlp_adam(ast_mlp_t* mlp, const float* grad, float* state, int step,
                 float lr, void* stream);
/* Copies NPARAM host floats into the masters and refreshes the BF16 copy
 * (synchronous). */
int ast_mlp_set_params(ast_mlp_t* mlp, const float* host_params, void* stream);
int ast_mlp_get_params(const ast_mlp_t* mlp, float* host_params, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* ASTERIY_H */
