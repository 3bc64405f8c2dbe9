// step_f64.hip -- the FP64 step, the drop-in for the reference's Kernel_TVD
// (kernel.cu:2513-3073; axisymmetric update KorolkovSolver.cu:13-32) on
// gfx950: face-centric wave strips in IEEE double.
//
// A wave owns 64 consecutive columns (60 output columns and 2 halo columns on
// each side) and walks the rows [m0, m1) of its strip top to bottom.  Per row
// each lane
//  * takes its x-neighbours' states from the adjacent lanes (DPP wave shifts),
//    forms the minmod face values of its cell, and solves its WEST face once;
//    its east flux is the west flux of the next lane (another shift);
//  * solves the NORTH face of its cell from the rows m-1 .. m+2 it holds in
//    registers; the south flux is the north flux of the row before;
//  * updates its cell from the four fluxes (conservative update, the
//    r-weighted source in axisymmetric geometry, the reference's floors).
// So every face of the grid is solved once (the reference solves each
// interior face twice, once per cell, kernel.cu:2799-2857), with no LDS and no
// barrier inside the strip.  Slopes are the minmod of neighbour differences,
// c +- minmod(c - a, b - c)/2, which is the reference's coordinate-difference
// form (linear/linear2, kernel.cu:35-50) on a uniform grid without its
// divisions.  Ghost cells follow the rules of ast_grid_t (fetch_cell).
//
// The state stays in the reference's layout (two planes of double pairs);
// loads and stores are 16-byte accesses of consecutive columns.  dt: each
// wave keeps the largest |wave speed| of the faces it solved for each
// direction, turns them into dt once, and workgroup minima lower *T with one
// 64-bit atomicMin on the IEEE bits (positive doubles order like their
// bits): the exact minimum over all faces, independent of the order.
#include "common.hpp"
#include "flux64.hpp"

namespace ast {
namespace s64 {

using namespace d64;

constexpr int WOUT = 60;  // output columns per wave
constexpr int WPB = 4;    // waves (stacked strips) per workgroup
constexpr int NT = 64 * WPB;

struct Q {
  double r, p, vx, vy;
};

AST_DEV Q q_sub(const Q& a, const Q& b) { return Q{a.r - b.r, a.p - b.p, a.vx - b.vx, a.vy - b.vy}; }

// lane i <- lane i - 1 (lane 0 keeps its own value) and lane i <- lane i + 1
// (lane 63 keeps its own): wave shifts by DPP on the two halves of a double.
AST_DEV double from_left(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false),
                          __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false));
}
AST_DEV double from_right(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false),
                          __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false));
}
AST_DEV Q from_left(const Q& q) {
  return Q{from_left(q.r), from_left(q.p), from_left(q.vx), from_left(q.vy)};
}
AST_DEV Q from_right(const Q& q) {
  return Q{from_right(q.r), from_right(q.p), from_right(q.vx), from_right(q.vy)};
}
AST_DEV FluxD from_right(const FluxD& f) {
  return FluxD{from_right(f.m), from_right(f.e), from_right(f.un), from_right(f.ut)};
}
AST_DEV bool from_left(bool b) {
  const int v = b ? 1 : 0;
  return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false) != 0;
}

// minmod(a, b) (kernel.cu:22-33): the one of a, b nearer zero when they have
// the same sign, else 0 -- the median of a, b and 0.
AST_DEV double minmod(double a, double b) { return fmax(fmin(a, b), fmin(fmax(a, b), 0.0)); }

// Face values c + s and c - s with s = minmod(c - a, b - c)/2 (dl = c - a,
// dr = b - c).  A density or pressure face value <= 0 falls back to the cell
// value (kernel.cu:2706-2788); that can only happen for a cell state that is
// itself not positive.
AST_DEV void face_values(const Q& c, const Q& dl, const Q& dr, Q& hi, Q& lo) {
  const double sr = 0.5 * minmod(dl.r, dr.r), sp = 0.5 * minmod(dl.p, dr.p);
  const double sx = 0.5 * minmod(dl.vx, dr.vx), sy = 0.5 * minmod(dl.vy, dr.vy);
  hi = Q{c.r + sr, c.p + sp, c.vx + sx, c.vy + sy};
  lo = Q{c.r - sr, c.p - sp, c.vx - sx, c.vy - sy};
  if (hi.r <= 0.0) hi.r = c.r;
  if (hi.p <= 0.0) hi.p = c.p;
  if (lo.r <= 0.0) lo.r = c.r;
  if (lo.p <= 0.0) lo.p = c.p;
}

// Cell (n, m) with the ghost rules of ast_grid_t; positions that no output
// cell reads (corners, columns past the halo of a partial band) take the
// nearest grid cell, so every lane holds a finite, positive-where-the-grid-is
// state.
AST_DEV Q load_q(const ast_grid_t& g, const double2* __restrict__ s,
                 const double2* __restrict__ u, int n, int m, double pulse) {
  const bool xin = (n >= 0) & (n < g.N), yin = (m >= 0) & (m < g.M);
  double2 a, b;
  if (xin & yin) {
    const size_t k = (size_t)m * g.N + n;
    a = s[k];
    b = u[k];
  } else if ((xin | yin) & (n >= -2) & (n <= g.N + 1) & (m >= -2) & (m <= g.M + 1)) {
    fetch_cell(g, s, u, n, m, a, b, pulse);
  } else {
    const int cn = n < 0 ? 0 : (n >= g.N ? g.N - 1 : n);
    const int cm = m < 0 ? 0 : (m >= g.M ? g.M - 1 : m);
    const size_t k = (size_t)cm * g.N + cn;
    a = s[k];
    b = u[k];
  }
  return Q{a.x, a.y, b.x, b.y};
}

// Exact-solver counters of one lane (the STATS kernels), ast_stats_t order.
struct Counters {
  uint32_t v[10];
  AST_DEV void clear() {
    for (int k = 0; k < 10; k++) v[k] = 0;
  }
  AST_DEV void add(const ExactInfo& e, bool counted) {
    if (!counted || e.pat == -3) return;
    v[0] += 1;
    if (e.pat == 2) { v[1] += 1; v[2] += e.it; }
    if (e.pat == 1) { v[3] += 1; v[4] += e.it; }
    if (e.pat == 0) v[5] += 1;
    if (e.pat == -1) v[6] += 1;
    if (e.capped) v[7] += 1;
    if (e.guessed && e.pat == 2) v[9] += 1;
  }
};

// Adds a workgroup's counters to *stats.
AST_DEV void block_counters_to(const Counters& c, uint32_t* red, ast_stats_t* stats) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 0; k < 10; k++) {
    const uint32_t w = wave_sum(c.v[k]);
    if (lane == 0) red[wave * 10 + k] = w;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    unsigned long long t = 0;
    for (int w = 0; w < WPB; w++) t += red[w * 10 + threadIdx.x];
    if (t) atomicAdd(reinterpret_cast<unsigned long long*>(stats) + threadIdx.x, t);
  }
}

// The face between states l (left/below) and r (right/above) for a normal
// along x (NX) or y: high order -- the method on the reconstructed values lv,
// rv -- or, outside the high-order window, HLL on the cell states
// (kernel.cu:2812, 2838-2839, 2856).  Returns the flux in the face frame
// (un: momentum along the normal, ut: along the other axis) and the face's
// largest |wave speed|.
template <int METHOD, bool WINDOW, bool NX>
AST_DEV FluxD face(const Q& lv, const Q& rv, const Q& l, const Q& r, bool ho, const GasD& a,
                   double& UU, ExactInfo& info) {
  info.pat = -3;  // no exact solve
  if (WINDOW && !ho) {
    const PrimD L = NX ? PrimD{l.r, l.p, l.vx, l.vy} : PrimD{l.r, l.p, l.vy, l.vx};
    const PrimD R = NX ? PrimD{r.r, r.p, r.vx, r.vy} : PrimD{r.r, r.p, r.vy, r.vx};
    return hll(L, R, a, UU);
  }
  const PrimD L = NX ? PrimD{lv.r, lv.p, lv.vx, lv.vy} : PrimD{lv.r, lv.p, lv.vy, lv.vx};
  const PrimD R = NX ? PrimD{rv.r, rv.p, rv.vx, rv.vy} : PrimD{rv.r, rv.p, rv.vy, rv.vx};
  return face_flux<METHOD>(L, R, a, UU, info);
}

struct StepArgs {
  ast_grid_t g;
  GasD gas;
  const double2* s;
  const double2* u;
  double2* s2;
  double2* u2;
  double* T;
  const double* T_do;
  const double* TT;  // simulation time (read only with a pulsed side; else null)
  ast_stats_t* stats;
  // neutral-H source planes of add_MK (kernel.cu:2495-2498), null when absent:
  // nn2 3 doubles per cell (x, y used), nn3 one
  const double* nn2;
  const double* nn3;
  double kmom, ken;
  int rows;    // rows per strip
  int nbands;  // bands of WOUT columns
  int nwg;     // workgroups
};

template <int METHOD, int GEOM, bool WINDOW, bool STATS>
__global__ __launch_bounds__(NT) void strip_kernel(StepArgs A) {
  __shared__ double red[WPB];
  __shared__ uint32_t redc[WPB * 10];
  const ast_grid_t& g = A.g;
  const GasD& gas = A.gas;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wg = xcd_tile(blockIdx.x, A.nwg);
  const int band = wg % A.nbands;
  const int m0 = ((wg / A.nbands) * WPB + wave) * A.rows;
  const int m1 = min(m0 + A.rows, g.M);
  const int n = band * WOUT + lane - 2;
  const bool out_col = (lane >= 2) & (lane < 2 + WOUT) & (n < g.N);
  // the west face of this lane's column is a face of an output cell
  const bool west_face = (lane >= 2) & (lane <= 2 + WOUT) & (n <= g.N);
  const double x = g.x_min + n * g.dx;
  const bool ho_x = high_order(g, x - g.dx / 2.0);  // west face (kernel.cu:2843)
  const bool ho_y = high_order(g, x);               // y faces: the cell x (kernel.cu:2816)
  const double T_do = *A.T_do;
  // pulsed Dirichlet sides: the density factor at the time before the step
  const double pulse = A.TT != nullptr ? pulse_factor(g, *A.TT) : 1.0;
  const double kx = T_do / g.dx, ky = T_do / g.dy;
  const bool frz = g.frozen_ax > 0.0;
  // frozen grid cells are copied (kernel.cu:2317-2324); a face votes for dt
  // when one of its cells is a grid cell that is not frozen (frozen cells
  // return before their faces)
  auto in_grid = [&](int nn, int mm) { return nn >= 0 && nn < g.N && mm >= 0 && mm < g.M; };
  auto frozen = [&](int nn, int mm) {
    return frz && in_grid(nn, mm) && is_frozen(g, g.x_min + nn * g.dx, g.y_min + mm * g.dy);
  };
  auto voter = [&](int nn, int mm) { return in_grid(nn, mm) && !frozen(nn, mm); };
  double UUx = 0.0, UUy = 0.0;
  Counters cnt;
  if constexpr (STATS) cnt.clear();

  if (m0 < g.M && band * WOUT < g.N) {  // wave-uniform
    // prologue: rows m0-2 .. m0+2; the south face of row m0
    const Q qa = load_q(g, A.s, A.u, n, m0 - 2, pulse);
    const Q qb = load_q(g, A.s, A.u, n, m0 - 1, pulse);
    Q q0 = load_q(g, A.s, A.u, n, m0, pulse);
    Q q1 = load_q(g, A.s, A.u, n, m0 + 1, pulse);
    Q q2 = load_q(g, A.s, A.u, n, m0 + 2, pulse);
    Q d1 = q_sub(q1, q0);  // q(m+1) - q(m)
    Q vn;                  // north face value of row m
    FluxD Gs;              // flux through the south face of row m
    {
      const Q db = q_sub(q0, qb);
      Q nb, sb, s0;
      face_values(qb, q_sub(qb, qa), db, nb, sb);
      face_values(q0, db, d1, vn, s0);
      double uu;
      ExactInfo ei;
      Gs = face<METHOD, WINDOW, false>(nb, s0, qb, q0, ho_y, gas, uu, ei);
      if (out_col && (voter(n, m0 - 1) || voter(n, m0))) UUy = fmax(UUy, uu);
      if constexpr (STATS) cnt.add(ei, out_col && m0 == 0);
    }
    bool fr_m = frozen(n, m0);
    for (int m = m0; m < m1; m++) {
      const Q q3 = load_q(g, A.s, A.u, n, m + 3, pulse);  // for the next row
      // y: face values of row m+1; the north face of row m
      const Q d2 = q_sub(q2, q1);
      Q vn1, vs1;
      face_values(q1, d1, d2, vn1, vs1);
      double uuy;
      ExactInfo eiy;
      const FluxD Gn = face<METHOD, WINDOW, false>(vn, vs1, q0, q1, ho_y, gas, uuy, eiy);
      // x: face values of row m; the west face (left value from lane - 1)
      const Q ql = from_left(q0);
      const Q qr = from_right(q0);
      Q ve, vw;
      face_values(q0, q_sub(q0, ql), q_sub(qr, q0), ve, vw);
      double uux;
      ExactInfo eix;
      const FluxD Fw = face<METHOD, WINDOW, true>(from_left(ve), vw, ql, q0, ho_x, gas, uux, eix);
      const FluxD Fe = from_right(Fw);
      // votes
      const bool vote_here = voter(n, m);
      const bool vote_up = voter(n, m + 1);
      const bool vote_left = from_left(vote_here);
      if (out_col && (vote_here || vote_up)) UUy = fmax(UUy, uuy);
      if (west_face && (vote_here || vote_left)) UUx = fmax(UUx, uux);
      if constexpr (STATS) {
        cnt.add(eiy, out_col);
        cnt.add(eix, west_face);
      }
      // conservative update (kernel.cu:3027-3032, 3059-3073): dU = -T_do/dx
      // (F_e - F_w) - T_do/dy (G_n - G_s)
      double dr = -kx * (Fe.m - Fw.m) - ky * (Gn.m - Gs.m);
      double dE = -kx * (Fe.e - Fw.e) - ky * (Gn.e - Gs.e);
      double dmx = -kx * (Fe.un - Fw.un) - ky * (Gn.ut - Gs.ut);
      double dmy = -kx * (Fe.ut - Fw.ut) - ky * (Gn.un - Gs.un);
      const double ke = 0.5 * q0.r * (q0.vx * q0.vx + q0.vy * q0.vy);
      const double mx = q0.r * q0.vx, my = q0.r * q0.vy;
      if constexpr (GEOM == AST_AXISYMMETRIC) {
        // KorolkovSolver.cu:13-32: -(T_do/y) v {rho, rho u, rho v, gamma p/(g-1) + KE}
        const double ty = T_do / (g.y_min + m * g.dy);
        const double tr = ty * my;
        dr -= tr;
        dmx -= tr * q0.vx;
        dmy -= tr * q0.vy;
        dE -= ty * q0.vy * (gas.g * gas.ig1 * q0.p + ke);
      }
      if (A.nn2 != nullptr || A.nn3 != nullptr) {  // wave-uniform
        // neutral-H sources (kernel.cu:2495-2498): momentum T_do (rho kmom)
        // nn2.{x,y}, energy T_do ken nn3; output cells only (in bounds)
        const size_t k = out_col ? (size_t)m * g.N + n : 0;
        if (A.nn2 != nullptr) {
          const double tk = T_do * (q0.r * A.kmom);
          dmx += tk * (out_col ? A.nn2[3 * k] : 0.0);
          dmy += tk * (out_col ? A.nn2[3 * k + 1] : 0.0);
        }
        if (A.nn3 != nullptr) dE += (T_do * A.ken) * (out_col ? A.nn3[k] : 0.0);
      }
      // density floor (kernel.cu:3060-3064): a non-positive density keeps
      // the old one, and the kinetic energy of the new state uses it
      const bool rfloor = q0.r + dr <= 0.0;
      const double drf = rfloor ? 0.0 : dr;
      const double r2 = q0.r + drf;
      const double vx2 = (mx + dmx) / r2;
      const double vy2 = (my + dmy) / r2;
      // p' = (E' - rho'|u'|^2/2)(g-1) with the O(E) terms cancelled exactly:
      // KE' - KE = (m.dm + |dm|^2/2 - KE dr)/rho'
      const double dke = (mx * dmx + my * dmy + 0.5 * (dmx * dmx + dmy * dmy) - ke * drf) / r2;
      double p2 = q0.p + gas.g1 * (dE - dke);
      if (p2 <= 0.0) p2 = 0.000001;  // kernel.cu:3070-3073
      if (out_col) {
        const size_t k = (size_t)m * g.N + n;
        if (fr_m) {  // frozen: copied (kernel.cu:2317-2324)
          A.s2[k] = make_double2(q0.r, q0.p);
          A.u2[k] = make_double2(q0.vx, q0.vy);
        } else {
          A.s2[k] = make_double2(r2, p2);
          A.u2[k] = make_double2(vx2, vy2);
        }
      }
      // slide the window one row
      fr_m = frozen(n, m + 1);
      Gs = Gn;
      vn = vn1;
      d1 = d2;
      q0 = q1;
      q1 = q2;
      q2 = q3;
    }
  }
  // per-face dt = krit*rad/UU (monotone in UU), capped at 1000 (kernel.cu:2530)
  UUx = wave_max(UUx);
  UUy = wave_max(UUy);
  const bool any_out = m0 < g.M && band * WOUT < g.N;
  double tmin = any_out ? 1000.0 : INFINITY;
  if (UUx > 0.0) tmin = fmin(tmin, g.krit * g.dx / UUx);
  if (UUy > 0.0) tmin = fmin(tmin, g.krit * g.dy / UUy);
  block_min_uniform<NT>(tmin, red, A.T, 0);
  if constexpr (STATS) block_counters_to(cnt, redc, A.stats);
}

// rows per strip of the FP64 step (ast_step_f64_rows)
constexpr int kDefaultRows = 6;

template <int METHOD, int GEOM, bool WINDOW>
static int launch3(const StepArgs& a, bool stats, hipStream_t st) {
  if (stats && METHOD == AST_EXACT)
    hipLaunchKernelGGL((strip_kernel<METHOD, GEOM, WINDOW, true>), dim3(a.nwg), dim3(NT), 0, st, a);
  else
    hipLaunchKernelGGL((strip_kernel<METHOD, GEOM, WINDOW, false>), dim3(a.nwg), dim3(NT), 0, st, a);
  AST_LAUNCH_CHECK("ast_step_f64 launch");
  return AST_OK;
}

template <int METHOD>
static int launch(const StepArgs& a, bool stats, bool window, hipStream_t st) {
  const bool axi = a.g.geometry == AST_AXISYMMETRIC;
  if (axi)
    return window ? launch3<METHOD, AST_AXISYMMETRIC, true>(a, stats, st)
                  : launch3<METHOD, AST_AXISYMMETRIC, false>(a, stats, st);
  return window ? launch3<METHOD, AST_PLANAR, true>(a, stats, st)
                : launch3<METHOD, AST_PLANAR, false>(a, stats, st);
}

}  // namespace s64
}  // namespace ast

// The step with a given strip height (rows > 0; 0 = the default) and
// optional neutral-H source planes (src may be null).
int ast_step_f64_rows(const ast_grid_t* g, const double* s, const double* u, double* s2,
                      double* u2, double* T, const double* T_do, const double* TT, int method,
                      ast_stats_t* stats, int rows, const ast_sources_t* src, void* stream) {
  using namespace ast::s64;
  int rc = ast_grid_check(g);
  if (rc != AST_OK) return rc;
  if (!s || !u || !s2 || !u2 || !T || !T_do)
    return ast_fail(AST_ERR_ARG, "ast_step_f64: null state or time pointer");
  if (s == s2 || u == u2 || s == u2 || u == s2)
    return ast_fail(AST_ERR_ARG, "ast_step_f64: output planes must not alias the input planes");
  if (rows == 0) rows = kDefaultRows;
  if (rows < 1 || rows > 4096) return ast_fail(AST_ERR_ARG, "ast_step_f64: %d rows per strip", rows);
  StepArgs a;
  a.g = *g;
  a.gas = ast::d64::make_gas(g->gamma);
  a.s = (const double2*)s;
  a.u = (const double2*)u;
  a.s2 = (double2*)s2;
  a.u2 = (double2*)u2;
  a.T = T;
  a.T_do = T_do;
  bool pulsed = false;
  for (int k = 0; k < 4; k++) pulsed |= g->bc[k] == AST_BC_DIRICHLET_PULSED;
  if (pulsed && !TT)
    return ast_fail(AST_ERR_ARG, "ast_step_f64: a pulsed Dirichlet side needs the time TT");
  a.TT = pulsed ? TT : nullptr;
  a.stats = stats;
  a.nn2 = src ? src->nn2 : nullptr;
  a.nn3 = src ? src->nn3 : nullptr;
  a.kmom = src ? src->kmom : 0.0;
  a.ken = src ? src->ken : 0.0;
  if (a.nn2 == s || a.nn2 == u || a.nn3 == s || a.nn3 == u)
    return ast_fail(AST_ERR_ARG, "ast_step_f64: a source plane aliases a state plane");
  a.rows = rows;
  a.nbands = (g->N + WOUT - 1) / WOUT;
  const long long strips = (g->M + rows - 1) / rows;
  const long long nwg = (long long)a.nbands * ((strips + WPB - 1) / WPB);
  if (nwg > 0x7fffffffLL) return ast_fail(AST_ERR_SIZE, "ast_step_f64: grid too large");
  a.nwg = (int)nwg;
  const bool window = !(g->ho_lo == -INFINITY && g->ho_hi == INFINITY);
  const bool st = stats != nullptr;
  hipStream_t hs = (hipStream_t)stream;
  switch (method) {
    case AST_HLL: return launch<AST_HLL>(a, st, window, hs);
    case AST_HLLC: return launch<AST_HLLC>(a, st, window, hs);
    case AST_EXACT: return launch<AST_EXACT>(a, st, window, hs);
    case AST_HLLC_A: return launch<AST_HLLC_A>(a, st, window, hs);
    default:
      return ast_fail(AST_ERR_ARG,
                      "ast_step_f64: method %d is not AST_HLL/AST_HLLC/AST_EXACT/AST_HLLC_A",
                      method);
  }
}

extern "C" int ast_step_f64(const ast_grid_t* g, const double* s, const double* u, double* s2,
                            double* u2, double* T, const double* T_do, double* TT, int method,
                            ast_stats_t* stats, void* stream) {
  // TT is read only with a pulsed Dirichlet side (unused by the reference's
  // live body)
  return ast_step_f64_rows(g, s, u, s2, u2, T, T_do, TT, method, stats, 0, nullptr, stream);
}

extern "C" int ast_step_f64_src(const ast_grid_t* g, const double* s, const double* u,
                                double* s2, double* u2, double* T, const double* T_do, double* TT,
                                int method, const ast_sources_t* src, ast_stats_t* stats,
                                void* stream) {
  return ast_step_f64_rows(g, s, u, s2, u2, T, T_do, TT, method, stats, 0, src, stream);
}
