// api.cpp -- host side of the C-ABI: status/error reporting, library
// information and argument validation shared by every entry point.  No
// entry point terminates the process (the reference exits on CUDA errors,
// kernel.cu:3875-3884); failures come back as AST_* codes with a message.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/asteriy.h"

static thread_local char g_err[512] = "";

int ast_fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

int ast_hip_fail(hipError_t e, const char* where) {
  return ast_fail(AST_ERR_HIP, "%s: %s (%d)", where, hipGetErrorString(e), (int)e);
}

extern "C" const char* ast_last_error(void) { return g_err; }

extern "C" const char* ast_version(void) { return "asteriy-mi355x 0.1 (gfx950)"; }

extern "C" int ast_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

extern "C" int ast_grid_check(const ast_grid_t* g) {
  if (!g) return ast_fail(AST_ERR_ARG, "grid: null pointer");
  if (g->N < 2 || g->M < 2) return ast_fail(AST_ERR_SIZE, "grid: N=%d M=%d, need N,M >= 2", g->N, g->M);
  if ((long long)g->N * g->M > (1LL << 31))
    return ast_fail(AST_ERR_SIZE, "grid: N*M = %lld cells exceeds 2^31", (long long)g->N * g->M);
  if (!(g->dx > 0) || !(g->dy > 0) || !isfinite(g->dx) || !isfinite(g->dy))
    return ast_fail(AST_ERR_ARG, "grid: dx, dy must be positive and finite");
  if (!(g->gamma > 1.0) || !isfinite(g->gamma)) return ast_fail(AST_ERR_ARG, "grid: gamma must be > 1");
  if (!(g->krit > 0) || !isfinite(g->krit)) return ast_fail(AST_ERR_ARG, "grid: krit must be > 0");
  if (g->geometry != AST_PLANAR && g->geometry != AST_AXISYMMETRIC)
    return ast_fail(AST_ERR_ARG, "grid: geometry %d unknown", g->geometry);
  if (g->geometry == AST_AXISYMMETRIC && !(g->y_min > 0))
    return ast_fail(AST_ERR_ARG, "grid: axisymmetric update needs cell radii y > 0 (y_min > 0)");
  for (int i = 0; i < 4; i++) {
    if (g->bc[i] < AST_BC_OUTFLOW || g->bc[i] > AST_BC_DIRICHLET_PULSED)
      return ast_fail(AST_ERR_ARG, "grid: bc[%d] = %d unknown", i, g->bc[i]);
    if ((g->bc[i] == AST_BC_DIRICHLET || g->bc[i] == AST_BC_DIRICHLET_PULSED) &&
        !(g->bc_state[i][0] > 0 && g->bc_state[i][1] > 0))
      return ast_fail(AST_ERR_ARG, "grid: Dirichlet state of side %d needs rho, p > 0", i);
    if (g->bc[i] == AST_BC_DIRICHLET_PULSED &&
        !(fabs(g->bc_pulse[0]) < 1.0 && isfinite(g->bc_pulse[1]) && !isnan(g->bc_pulse[2]) &&
          !isnan(g->bc_pulse[3])))
      return ast_fail(AST_ERR_ARG,
                      "grid: pulsed side %d needs |a| < 1 (a positive density), a finite w and "
                      "limits that are not NaN", i);
  }
  if (g->frozen_ax > 0 && !(g->frozen_ay > 0))
    return ast_fail(AST_ERR_ARG, "grid: frozen region needs frozen_ay > 0");
  return AST_OK;
}
