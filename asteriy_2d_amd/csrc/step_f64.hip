// step_f64.hip -- the FP64 step: a cell-by-cell restatement of the
// reference's Kernel_TVD (kernel.cu:2513-3073) for gfx950.
//
// Built with -ffp-contract=off: every cell evaluates its own four faces with
// the reference's coordinate-difference reconstruction (linear2/linear,
// kernel.cu:35-50) and operation order, so the result equals the reference
// arithmetic bit for bit for HLL/HLLC (exact solver: to the ulp of pow).
//
// Mapping: one workgroup = one TX x TY tile of cells, one thread per cell.
// The tile plus a 2-cell halo is staged once through LDS (SoA planes, ghost
// cells substituted while staging), so the 9-point cross gather of every
// cell reads LDS and each state word is fetched from HBM/L2 once per tile.
// The dt candidate is reduced by wave shuffles + one 64-bit atomicMin per
// tile (deterministic).
#include "common.hpp"

namespace ast {
namespace f64 {

constexpr int TX = 64, TY = 4, NT = TX * TY;
constexpr int LX = TX + 4, LY = TY + 4;

typedef V2<double> D2;

template <int METHOD>
AST_DEV double face_flux(const Phys<double>& ph, const D2& Ls, const D2& Lu, const D2& Rs,
                         const D2& Ru, double n1, double n2, D2& Ps, D2& Pu, double rad,
                         FaceStats& st) {
  if constexpr (METHOD == AST_HLL) {
    return HLL(ph, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
  } else if constexpr (METHOD == AST_HLLC) {
    return HLLC(ph, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
  } else if constexpr (METHOD == AST_HLLC_A) {
    return HLLC_A(ph, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
  } else {
    return GODUNOV(ph, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad, st);
  }
}

struct Tile {
  double r[LY][LX], p[LY][LX], vx[LY][LX], vy[LY][LX];
};

AST_DEV void ld(const Tile& L, int j, int i, D2& s, D2& u) {
  s.x = L.r[j][i];
  s.y = L.p[j][i];
  u.x = L.vx[j][i];
  u.y = L.vy[j][i];
}

// One cell of Kernel_TVD.  (j, i) is the cell's LDS position.
template <int METHOD, int GEOM>
AST_DEV double cell(const ast_grid_t& g, const Phys<double>& ph, const Tile& L, int j, int i,
                    int n, int m, double T_do, D2& so, D2& uo, FaceStats& st) {
  const double dx = g.dx, dy = g.dy, ggg = ph.ga;
  const double y = g.y_min + m * dy;
  const double x = g.x_min + n * dx;
  D2 s_1, s_2, s_3, s_4, s_5, u_1, u_2, u_3, u_4, u_5;
  D2 s21, s31, s41, s51, u21, u31, u41, u51;
  D2 s12, s13, s14, s15, u12, u13, u14, u15;
  D2 Ps12 = {0, 0}, Pu12 = {0, 0}, Ps13 = {0, 0}, Pu13 = {0, 0};
  D2 Ps14 = {0, 0}, Pu14 = {0, 0}, Ps15 = {0, 0}, Pu15 = {0, 0};
  double tmin = 1000;
  double A = 0, B = 0;
  ld(L, j, i, s_1, u_1);
  if (is_frozen(g, x, y)) {  // kernel.cu:2317-2324: copied, no dt vote
    so = s_1;
    uo = u_1;
    return INFINITY;
  }
  ld(L, j, i + 1, s_2, u_2);
  ld(L, j + 1, i, s_5, u_5);
  ld(L, j, i - 1, s_4, u_4);
  ld(L, j - 1, i, s_3, u_3);
  ld(L, j, i + 2, s21, u21);
  ld(L, j, i - 2, s41, u41);
  ld(L, j + 2, i, s51, u51);
  ld(L, j - 2, i, s31, u31);

  // own-side face states, kernel.cu:2701-2797 (grad_p = true)
  linear2(x - dx, s_4.x, x, s_1.x, x + dx, s_2.x, x - dx / 2.0, x + dx / 2.0, A, B);
  s12.x = (B <= 0) ? s_1.x : B;
  s14.x = (A <= 0) ? s_1.x : A;
  linear2(x - dx, s_4.y, x, s_1.y, x + dx, s_2.y, x - dx / 2.0, x + dx / 2.0, A, B);
  s12.y = (B <= 0) ? s_1.y : B;
  s14.y = (A <= 0) ? s_1.y : A;
  linear2(x - dx, u_4.x, x, u_1.x, x + dx, u_2.x, x - dx / 2.0, x + dx / 2.0, A, B);
  u12.x = B;
  u14.x = A;
  linear2(x - dx, u_4.y, x, u_1.y, x + dx, u_2.y, x - dx / 2.0, x + dx / 2.0, A, B);
  u12.y = B;
  u14.y = A;
  linear2(y - dy, s_3.x, y, s_1.x, y + dy, s_5.x, y - dy / 2.0, y + dy / 2.0, A, B);
  s15.x = (B <= 0) ? s_1.x : B;
  s13.x = (A <= 0) ? s_1.x : A;
  linear2(y - dy, s_3.y, y, s_1.y, y + dy, s_5.y, y - dy / 2.0, y + dy / 2.0, A, B);
  s15.y = (B <= 0) ? s_1.y : B;
  s13.y = (A <= 0) ? s_1.y : A;
  linear2(y - dy, u_3.x, y, u_1.x, y + dy, u_5.x, y - dy / 2.0, y + dy / 2.0, A, B);
  u15.x = B;
  u13.x = A;
  linear2(y - dy, u_3.y, y, u_1.y, y + dy, u_5.y, y - dy / 2.0, y + dy / 2.0, A, B);
  u15.y = B;
  u13.y = A;

  // east face, kernel.cu:2799-2813
  if (high_order(g, x + dx / 2.0)) {
    s21.x = linear(x, s_1.x, x + dx, s_2.x, x + 2.0 * dx, s21.x, x + dx / 2.0);
    if (s21.x <= 0) s21.x = s_2.x;
    s21.y = linear(x, s_1.y, x + dx, s_2.y, x + 2.0 * dx, s21.y, x + dx / 2.0);
    if (s21.y <= 0) s21.y = s_2.y;
    u21.x = linear(x, u_1.x, x + dx, u_2.x, x + 2.0 * dx, u21.x, x + dx / 2.0);
    u21.y = linear(x, u_1.y, x + dx, u_2.y, x + 2.0 * dx, u21.y, x + dx / 2.0);
    tmin = fmin(tmin, face_flux<METHOD>(ph, s12, u12, s21, u21, 1.0, 0.0, Ps12, Pu12, dx, st));
  } else {
    tmin = my_min(tmin, HLL(ph, s_1, u_1, s_2, u_2, 1.0, 0.0, Ps12, Pu12, dx));
  }
  // south and north faces, kernel.cu:2816-2840 (window tested at the cell x)
  if (high_order(g, x)) {
    s31.x = linear(y, s_1.x, y - dy, s_3.x, y - 2.0 * dy, s31.x, y - dy / 2.0);
    if (s31.x <= 0) s31.x = s_3.x;
    s31.y = linear(y, s_1.y, y - dy, s_3.y, y - 2.0 * dy, s31.y, y - dy / 2.0);
    if (s31.y <= 0) s31.y = s_3.y;
    u31.x = linear(y, u_1.x, y - dy, u_3.x, y - 2.0 * dy, u31.x, y - dy / 2.0);
    u31.y = linear(y, u_1.y, y - dy, u_3.y, y - 2.0 * dy, u31.y, y - dy / 2.0);
    s51.x = linear(y, s_1.x, y + dy, s_5.x, y + 2.0 * dy, s51.x, y + dy / 2.0);
    if (s51.x <= 0) s51.x = s_5.x;
    s51.y = linear(y, s_1.y, y + dy, s_5.y, y + 2.0 * dy, s51.y, y + dy / 2.0);
    if (s51.y <= 0) s51.y = s_5.y;
    u51.x = linear(y, u_1.x, y + dy, u_5.x, y + 2.0 * dy, u51.x, y + dy / 2.0);
    u51.y = linear(y, u_1.y, y + dy, u_5.y, y + 2.0 * dy, u51.y, y + dy / 2.0);
    tmin = fmin(tmin, face_flux<METHOD>(ph, s13, u13, s31, u31, 0.0, -1.0, Ps13, Pu13, dy, st));
    tmin = fmin(tmin, face_flux<METHOD>(ph, s15, u15, s51, u51, 0.0, 1.0, Ps15, Pu15, dy, st));
  } else {
    tmin = my_min(tmin, HLL(ph, s_1, u_1, s_3, u_3, 0.0, -1.0, Ps13, Pu13, dy));
    tmin = my_min(tmin, HLL(ph, s_1, u_1, s_5, u_5, 0.0, 1.0, Ps15, Pu15, dy));
  }
  // west face, kernel.cu:2843-2857
  if (high_order(g, x - dx / 2.0)) {
    s41.x = linear(x, s_1.x, x - dx, s_4.x, x - 2.0 * dx, s41.x, x - dx / 2.0);
    if (s41.x <= 0) s41.x = s_4.x;
    s41.y = linear(x, s_1.y, x - dx, s_4.y, x - 2.0 * dx, s41.y, x - dx / 2.0);
    if (s41.y <= 0) s41.y = s_4.y;
    u41.x = linear(x, u_1.x, x - dx, u_4.x, x - 2.0 * dx, u41.x, x - dx / 2.0);
    u41.y = linear(x, u_1.y, x - dx, u_4.y, x - 2.0 * dx, u41.y, x - dx / 2.0);
    tmin = fmin(tmin, face_flux<METHOD>(ph, s14, u14, s41, u41, -1.0, 0.0, Ps14, Pu14, dx, st));
  } else {
    tmin = my_min(tmin, HLL(ph, s_1, u_1, s_4, u_4, -1.0, 0.0, Ps14, Pu14, dx));
  }

  // flux assembly, kernel.cu:3027-3032
  double PSx = Ps12.x * dy + Ps13.x * dx + Ps14.x * dy + Ps15.x * dx;
  double PSy = Ps12.y * dy + Ps13.y * dx + Ps14.y * dy + Ps15.y * dx;
  double PUx = Pu12.x * dy + Pu13.x * dx + Pu14.x * dy + Pu15.x * dx;
  double PUy = Pu12.y * dy + Pu13.y * dx + Pu14.y * dy + Pu15.y * dx;
  double dV = dx * dy;
  D2 r, v;
  if constexpr (GEOM == AST_AXISYMMETRIC) {
    // EnergyConservationLaws, KorolkovSolver.cu:13-32
    r.x = s_1.x - (T_do / dV) * PSx - T_do * s_1.x * u_1.y / y;
    if (r.x <= 0) r.x = s_1.x;
    v.x = (s_1.x * u_1.x - (T_do / dV) * PUx - T_do * s_1.x * u_1.y * u_1.x / y) / r.x;
    v.y = (s_1.x * u_1.y - (T_do / dV) * PUy - T_do * s_1.x * u_1.y * u_1.y / y) / r.x;
    r.y = (((s_1.y / (ggg - 1) + s_1.x * (u_1.x * u_1.x + u_1.y * u_1.y) * 0.5) -
            (T_do / dV) * PSy -
            T_do * u_1.y * (ggg * s_1.y / (ggg - 1) + s_1.x * (u_1.x * u_1.x + u_1.y * u_1.y) * 0.5) / y) -
           0.5 * r.x * (v.x * v.x + v.y * v.y)) * (ggg - 1);
    if (r.y <= 0) r.y = 0.000001;
  } else {
    // kernel.cu:3059-3073 (body force Q2 = 0)
    r.x = s_1.x - (T_do / dV) * PSx;
    if (r.x <= 0.0) r.x = s_1.x;
    v.x = (s_1.x * u_1.x - (T_do / dV) * PUx) / r.x;
    v.y = (s_1.x * u_1.y - (T_do / dV) * PUy) / r.x;
    r.y = (((s_1.y / (ggg - 1) + s_1.x * (u_1.x * u_1.x + u_1.y * u_1.y) * 0.5) -
            (T_do / dV) * PSy) -
           0.5 * r.x * (v.x * v.x + v.y * v.y)) * (ggg - 1);
    if (r.y <= 0.0) r.y = 0.000001;
  }
  so = r;
  uo = v;
  return tmin;
}

// STATS: the variant that counts the exact solver's work into *stats (the
// per-face counters are dead code, and removed, in the other)
template <int METHOD, int GEOM, bool STATS>
__global__ __launch_bounds__(NT) void step_kernel(ast_grid_t g, const double2* __restrict__ s,
                                                  const double2* __restrict__ u,
                                                  double2* __restrict__ s2,
                                                  double2* __restrict__ u2, double* T,
                                                  const double* __restrict__ T_do,
                                                  ast_stats_t* stats, int tiles_x, int ntiles) {
  __shared__ Tile L;
  __shared__ double red[NT / 64];
  __shared__ uint32_t redi[(NT / 64) * 10];
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int n0 = (tile % tiles_x) * TX, m0 = (tile / tiles_x) * TY;

  for (int k = threadIdx.x; k < LX * LY; k += NT) {
    const int jy = k / LX, ix = k - jy * LX;
    double2 a, b;
    fetch_cell(g, s, u, n0 + ix - 2, m0 + jy - 2, a, b);
    L.r[jy][ix] = a.x;
    L.p[jy][ix] = a.y;
    L.vx[jy][ix] = b.x;
    L.vy[jy][ix] = b.y;
  }
  __syncthreads();

  const Phys<double> ph = make_phys<double>(g.gamma, g.krit);
  const int i = threadIdx.x % TX, j = threadIdx.x / TX;
  const int n = n0 + i, m = m0 + j;
  double tmin = INFINITY;
  FaceStats st;
  st.clear();
  if (n < g.N && m < g.M) {
    D2 so, uo;
    tmin = cell<METHOD, GEOM>(g, ph, L, j + 2, i + 2, n, m, *T_do, so, uo, st);
    const size_t k = (size_t)m * g.N + n;
    s2[k] = make_double2(so.x, so.y);
    u2[k] = make_double2(uo.x, uo.y);
  }
  block_min_to<NT>(tmin, red, T);
  if constexpr (STATS) block_stats_to<NT>(st, redi, stats);
}

template <int METHOD, int GEOM>
static int launch(const ast_grid_t& g, const double2* s, const double2* u, double2* s2,
                  double2* u2, double* T, const double* T_do, ast_stats_t* stats,
                  hipStream_t stream) {
  const int tiles_x = (g.N + TX - 1) / TX, tiles_y = (g.M + TY - 1) / TY;
  const int ntiles = tiles_x * tiles_y;
  if (METHOD == AST_EXACT && stats != nullptr)
    hipLaunchKernelGGL((step_kernel<METHOD, GEOM, true>), dim3(ntiles), dim3(NT), 0, stream, g, s,
                       u, s2, u2, T, T_do, stats, tiles_x, ntiles);
  else
    hipLaunchKernelGGL((step_kernel<METHOD, GEOM, false>), dim3(ntiles), dim3(NT), 0, stream, g, s,
                       u, s2, u2, T, T_do, stats, tiles_x, ntiles);
  AST_LAUNCH_CHECK("ast_step_f64 launch");
  return AST_OK;
}

}  // namespace f64
}  // namespace ast

extern "C" int ast_step_f64(const ast_grid_t* g, const double* s, const double* u, double* s2,
                            double* u2, double* T, const double* T_do, double* TT, int method,
                            ast_stats_t* stats, void* stream) {
  using namespace ast::f64;
  (void)TT;
  int rc = ast_grid_check(g);
  if (rc != AST_OK) return rc;
  if (!s || !u || !s2 || !u2 || !T || !T_do)
    return ast_fail(AST_ERR_ARG, "ast_step_f64: null state or time pointer");
  if (s == s2 || u == u2 || s == u2 || u == s2)
    return ast_fail(AST_ERR_ARG, "ast_step_f64: output planes must not alias the input planes");
  const long long ntiles = (long long)((g->N + TX - 1) / TX) * ((g->M + TY - 1) / TY);
  if (ntiles > 0x7fffffffLL) return ast_fail(AST_ERR_SIZE, "ast_step_f64: grid too large");
  hipStream_t st = (hipStream_t)stream;
  const double2* S = (const double2*)s;
  const double2* U = (const double2*)u;
  double2* S2 = (double2*)s2;
  double2* U2 = (double2*)u2;
  const bool axi = g->geometry == AST_AXISYMMETRIC;
  switch (method) {
    case AST_HLL:
      return axi ? launch<AST_HLL, AST_AXISYMMETRIC>(*g, S, U, S2, U2, T, T_do, stats, st)
                 : launch<AST_HLL, AST_PLANAR>(*g, S, U, S2, U2, T, T_do, stats, st);
    case AST_HLLC:
      return axi ? launch<AST_HLLC, AST_AXISYMMETRIC>(*g, S, U, S2, U2, T, T_do, stats, st)
                 : launch<AST_HLLC, AST_PLANAR>(*g, S, U, S2, U2, T, T_do, stats, st);
    case AST_EXACT:
      return axi ? launch<AST_EXACT, AST_AXISYMMETRIC>(*g, S, U, S2, U2, T, T_do, stats, st)
                 : launch<AST_EXACT, AST_PLANAR>(*g, S, U, S2, U2, T, T_do, stats, st);
    case AST_HLLC_A:
      return axi ? launch<AST_HLLC_A, AST_AXISYMMETRIC>(*g, S, U, S2, U2, T, T_do, stats, st)
                 : launch<AST_HLLC_A, AST_PLANAR>(*g, S, U, S2, U2, T, T_do, stats, st);
    default:
      return ast_fail(AST_ERR_ARG,
                      "ast_step_f64: method %d is not AST_HLL/AST_HLLC/AST_EXACT/AST_HLLC_A",
                      method);
  }
}
