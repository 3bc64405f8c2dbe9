// strip_f32.hpp -- the "wave strip" of the FP32 step, shared by the
// single-grid step (step_f32.hip) and the batched persistent launch
// (batch.hip).  See step_f32.hip for the scheme.
#pragma once
#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "exact_f32.hpp"
#include "fast_flux.hpp"

namespace ast {
namespace f32 {

constexpr int WOUT = 60;  // output columns per wave
constexpr int WPB = 4;    // waves per workgroup (stacked strips)
constexpr int NT = 64 * WPB;

// Buffer resource word 3 for raw (stride 0) buffers on the gfx9 family:
// DATA_FORMAT = 32 bit, no swizzle, no thread-id offset.
constexpr int kBufWord3 = 0x00020000;
// A byte offset beyond every range used here: a store at it (plus a row
// offset below 2^31) is dropped by the buffer range check (range sizes stay
// below 2^31, see kMaxStripBytes).
constexpr unsigned kOffDrop = 0x80000000u;
// Largest byte range one buffer resource of a strip covers.
constexpr long long kMaxStripBytes = 0x7fff0000LL;

struct Q {
  float r, p, vx, vy;
};

// lane i <- lane i-1; lane 0 reads 0.  Lane 0 is a halo lane: its results
// never reach an output lane (the zero fill lets the shift fold into the
// consuming instruction as a DPP operand).
AST_DEV float shr1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
// lane i <- lane i+1; lane 63 reads 0 (halo lane, as above)
AST_DEV float shl1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
// lane i <- lane i-1; lane 0 keeps its own value.  For the left state of a
// face: lane 0's (unused) face then stays an ordinary physical face, which
// matters to the iterative exact solver (a zero state is a vacuum problem
// that would hold the whole wave in its slowest path).
AST_DEV float shr1_keep(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138,
                                                    0xf, 0xf, false));
}
AST_DEV Q shr1(const Q& q) { return Q{shr1(q.r), shr1(q.p), shr1(q.vx), shr1(q.vy)}; }
AST_DEV Q shr1_keep(const Q& q) {
  return Q{shr1_keep(q.r), shr1_keep(q.p), shr1_keep(q.vx), shr1_keep(q.vy)};
}
AST_DEV Q sub(const Q& a, const Q& b) { return Q{a.r - b.r, a.p - b.p, a.vx - b.vx, a.vy - b.vy}; }

// minmod(a, b) (kernel.cu:22-33): the median of {a, b, 0} -- one v_med3_f32.
AST_DEV float minmod_f(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 0.f); }

// Face values at +-1/2 of a cell from its neighbour differences dl = c - a,
// dr = b - c.  The reference's positivity fallback (kernel.cu:2706-2788:
// rho, p <= 0 -> the cell value) never fires for positive cell states: with
// 0 <= |mm| <= |dl|, |dr| of the same sign, c + mm/2 and c - mm/2 lie between
// c/2 and the neighbours' means (also after rounding), so the fast path
// requires rho, p > 0 in every cell instead of testing each face value.
// c +- slope/2 as one fma each (0.5*slope is exact, so this rounds as
// c +- (0.5*slope) does).
AST_DEV void face_values(const Q& c, const Q& dl, const Q& dr, Q& plus, Q& minus) {
  const float mr = minmod_f(dl.r, dr.r), mp = minmod_f(dl.p, dr.p);
  const float mx = minmod_f(dl.vx, dr.vx), my = minmod_f(dl.vy, dr.vy);
  plus = Q{fmaf(0.5f, mr, c.r), fmaf(0.5f, mp, c.p), fmaf(0.5f, mx, c.vx), fmaf(0.5f, my, c.vy)};
  minus = Q{fmaf(-0.5f, mr, c.r), fmaf(-0.5f, mp, c.p), fmaf(-0.5f, mx, c.vx),
            fmaf(-0.5f, my, c.vy)};
}

// Cell (n, m) in FP32 with the ghost rules; corner/out-of-halo positions are
// clamped into the grid (never read by an output cell, kept finite).
AST_DEV Q load(const ast_grid_t& g, const float2* __restrict__ s, const float2* __restrict__ u,
               int n, int m, bool cols_in, double pulse) {
  float2 a, b;
  if (cols_in && m >= 0 && m < g.M) {
    const size_t k = (size_t)m * g.N + n;
    a = s[k];
    b = u[k];
    return Q{a.x, a.y, b.x, b.y};
  }
  const bool xin = (n >= 0) & (n < g.N), yin = (m >= 0) & (m < g.M);
  if ((!xin & !yin) | (n < -2) | (n > g.N + 1) | (m < -2) | (m > g.M + 1)) {
    const int cn = n < 0 ? 0 : (n >= g.N ? g.N - 1 : n);
    const int cm = m < 0 ? 0 : (m >= g.M ? g.M - 1 : m);
    const size_t k = (size_t)cm * g.N + cn;
    a = s[k];
    b = u[k];
  } else {
    fetch_cell(g, s, u, n, m, a, b, pulse);
  }
  return Q{a.x, a.y, b.x, b.y};
}

// ---- raw buffer access ----------------------------------------------------
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

AST_DEV __amdgpu_buffer_rsrc_t make_rsrc(const void* base, long long bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, kBufWord3);
}
AST_DEV float2 bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0);
  return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}
AST_DEV void bstore(float a, float b, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  u32x2 v;
  v.x = __float_as_uint(a);
  v.y = __float_as_uint(b);
  __builtin_amdgcn_raw_buffer_store_b64(v, r, (int)voff, (int)soff, 0);
}

struct Consts {
  float g, ig1, g1, gig1;  // gamma, 1/(gamma-1), gamma-1, gamma/(gamma-1)
  float fx0, iax2, iay2;   // frozen ellipse
  int frozen;
  int pulsed;  // a side is AST_BC_DIRICHLET_PULSED
};

static inline Consts make_consts(const ast_grid_t& g) {
  Consts c;
  c.g = (float)g.gamma;
  c.g1 = (float)(g.gamma - 1.0);
  c.ig1 = (float)(1.0 / (g.gamma - 1.0));
  c.gig1 = (float)(g.gamma / (g.gamma - 1.0));
  c.frozen = g.frozen_ax > 0.0;
  c.fx0 = (float)g.frozen_x0;
  c.iax2 = c.frozen ? (float)(1.0 / (g.frozen_ax * g.frozen_ax)) : 0.f;
  c.iay2 = c.frozen ? (float)(1.0 / (g.frozen_ay * g.frozen_ay)) : 0.f;
  c.pulsed = 0;
  for (int k = 0; k < 4; k++) c.pulsed |= g.bc[k] == AST_BC_DIRICHLET_PULSED;
  return c;
}

// Strip heights / band counts of one grid.
struct StripShape {
  int nbands, krows, nsg;
};
static inline StripShape strip_shape(const ast_grid_t& g, int krows) {
  StripShape sh;
  sh.krows = krows;
  sh.nbands = (g.N + WOUT - 1) / WOUT;
  const int nstrips = (g.M + krows - 1) / krows;
  sh.nsg = (nstrips + WPB - 1) / WPB;
  return sh;
}
// The tallest strip of a tiling (strip_rows): tile > 0 rows, or for S = -tile
// strips per band the end strips (M/S rounded to a multiple of 4) or the
// others (ceil of what is left over them).
static inline int tallest_strip(int M, int tile) {
  if (tile > 0) return tile;
  const int S = -tile, e = std::max(4, (M / S + 2) / 4 * 4), mid = M - 2 * e;
  if (S >= 3 && mid >= S - 2) return std::max(e, (mid + S - 3) / (S - 2));
  return (M + S - 1) / S;
}
// Do the buffer ranges of a strip fit one resource?  (rows krows + 4 of N
// float2 cells)
static inline bool strip_fits(const ast_grid_t& g, int krows) {
  return (long long)(krows + 4) * g.N * 8 <= kMaxStripBytes;
}
// The fast path reconstructs without the positivity fallback, which needs
// positive densities and pressures in every cell and ghost: a Dirichlet side
// must hold a positive state.
static inline int check_positive_bc(const ast_grid_t& g) {
  for (int k = 0; k < 4; k++) {
    const bool dir = g.bc[k] == AST_BC_DIRICHLET || g.bc[k] == AST_BC_DIRICHLET_PULSED;
    if (dir && !(g.bc_state[k][0] > 0.0 && g.bc_state[k][1] > 0.0))
      return ast_fail(AST_ERR_ARG, "FP32 step: Dirichlet side %d needs rho > 0 and p > 0", k);
  }
  return AST_OK;
}

// Kernel-internal method ids: the exact solver with the learned newton start
// compiled in (AST_EXACT kernels leave it out), and the plain exact solver
// that reports strong two-shock faces (the other half of a gated launch).
constexpr int kExactLearned = 16;
constexpr int kExactDetect = 17;
// the exact solver, in any of its kernel variants
constexpr bool is_exact(int method) {
  return method == AST_EXACT || method == kExactLearned || method == kExactDetect;
}

template <int METHOD>
AST_DEV FaceF face_flux(const PrimF& L, const PrimF& R, const Consts& c, float& UU,
                        const ExactCtx& ex, FaceInfo& fi) {
  if constexpr (METHOD == AST_HLL) return hll_frame_f32(L, R, c.g, c.ig1, UU);
  else if constexpr (METHOD == AST_HLLC) return hllc_frame_f32<false>(L, R, c.g, c.ig1, UU);
  else if constexpr (METHOD == AST_HLLC_A) return hllc_frame_f32<true>(L, R, c.g, c.ig1, UU);
  else return exact_frame_f32<METHOD == kExactLearned, METHOD == kExactDetect>(L, R, ex, UU, fi);
}

// Every face of the step is the physical flux of one state (gt = 0) unless
// HLL may serve it (the method, or the faces outside a high-order window).
template <int METHOD, bool WINDOW>
constexpr bool state_faces() {
  return METHOD != AST_HLL && !WINDOW;
}

// The flux of the face of lane + 1 (the east face of this lane's cell).
template <bool STATE>
AST_DEV FaceF shl1(const FaceF& f) {
  FaceF r;
  r.m = shl1(f.m);
  r.wn = shl1(f.wn);
  r.wt = shl1(f.wt);
  r.gn = shl1(f.gn);
  r.gt = STATE ? 0.f : shl1(f.gt);
  r.ge = shl1(f.ge);
  return r;
}

// A face's flux seen from a cell moving with (cn, ct) (face-frame
// components): the momentum fluxes gn - m dn, gt - m dt and the energy flux
// ge - dn gn - dt gt + m |d|^2/2, d = (cn, ct) - w.  The mass flux is
// frame-free.
struct CellFlux {
  float n, t, e;
};
template <bool STATE>
AST_DEV CellFlux in_cell_frame(const FaceF& f, float cn, float ct) {
  const float dn = cn - f.wn, dt = ct - f.wt;
  const float hk = 0.5f * f.m * fmaf(dn, dn, dt * dt);
  CellFlux r;
  r.n = fmaf(-f.m, dn, f.gn);
  if constexpr (STATE) {
    r.t = -f.m * dt;
    r.e = fmaf(-dn, f.gn, f.ge) + hk;
  } else {
    r.t = fmaf(-f.m, dt, f.gt);
    r.e = fmaf(-dt, f.gt, fmaf(-dn, f.gn, f.ge)) + hk;
  }
  return r;
}

// One face in the x (NORMAL_X) or y direction: rotate to the face frame,
// flux, and back.  With WINDOW, faces outside the high-order window use
// HLL on the cell states (kernel.cu:2812, 2838-2839, 2856).
// The exact solver's FaceInfo goes to st.add() after the call, in converged
// code (faces diverge inside with a window).
template <int METHOD, bool WINDOW, bool NORMAL_X, class ST>
AST_DEV FaceF face(const Q& l, const Q& r, const Q& lc, const Q& rc, bool ho, const Consts& c,
                   float& UU, const ExactCtx& ex, ST& st) {
  PrimF L, R;
  FaceInfo fi;
  fi.pat = kNoFace;
  fi.it = 0;
  fi.capped = fi.island = fi.learned = fi.strong = false;
  FaceF f;
  if constexpr (WINDOW) {  // faces diverge here: no wave-level learned guess
    if (!ho) {
      L = NORMAL_X ? PrimF{lc.r, lc.p, lc.vx, lc.vy} : PrimF{lc.r, lc.p, lc.vy, lc.vx};
      R = NORMAL_X ? PrimF{rc.r, rc.p, rc.vx, rc.vy} : PrimF{rc.r, rc.p, rc.vy, rc.vx};
      f = hll_frame_f32(L, R, c.g, c.ig1, UU);
    } else {
      ExactCtx e2 = ex;
      e2.mlp = nullptr;
      L = NORMAL_X ? PrimF{l.r, l.p, l.vx, l.vy} : PrimF{l.r, l.p, l.vy, l.vx};
      R = NORMAL_X ? PrimF{r.r, r.p, r.vx, r.vy} : PrimF{r.r, r.p, r.vy, r.vx};
      f = face_flux<METHOD>(L, R, c, UU, e2, fi);
    }
  } else {
    L = NORMAL_X ? PrimF{l.r, l.p, l.vx, l.vy} : PrimF{l.r, l.p, l.vy, l.vx};
    R = NORMAL_X ? PrimF{r.r, r.p, r.vx, r.vy} : PrimF{r.r, r.p, r.vy, r.vx};
    f = face_flux<METHOD>(L, R, c, UU, ex, fi);
  }
  if constexpr (is_exact(METHOD)) st.add(fi);
  return f;
}

AST_DEV bool frozen_at(const Consts& c, float x, float y) {
  if (!c.frozen) return false;
  const float d = x - c.fx0;
  return d * d * c.iax2 + y * y * c.iay2 < 1.f;
}

// Can any cell of columns [n0, n1] x rows [m0, m1] lie in the frozen
// ellipse?  (wave-uniform; false lets the wave skip the per-cell tests)
AST_DEV bool frozen_possible(const ast_grid_t& g, int n0, int n1, int m0, int m1) {
  if (!(g.frozen_ax > 0.0)) return false;
  const double x0 = g.x_min + n0 * g.dx, x1 = g.x_min + n1 * g.dx;
  const double y0 = g.y_min + m0 * g.dy, y1 = g.y_min + m1 * g.dy;
  const double ax = g.frozen_ax * 1.01, ay = g.frozen_ay * 1.01;
  return x1 >= g.frozen_x0 - ax && x0 <= g.frozen_x0 + ax && y1 >= -ay && y0 <= ay;
}

// add_MK's neutral-hydrogen source planes, either of a pair may be null:
// the reference's FP64 planes (nn2 three doubles per cell, x and y used; nn3
// one), or FP32 copies of them (xy: nn2's x and y; e: nn3), each value the
// FP32 rounding of the plane's -- the step rounds every plane value to FP32
// before it uses it, so the copies give the same bits in 12 B per cell
// instead of 32
struct SrcF {
  const double* nn2;
  const double* nn3;
  const float2* xy;
  const float* e;
  float kmom, ken;
};

// source kinds of the step kernels: none, the FP64 planes, the FP32 copies
constexpr int kSrcNone = 0, kSrcF64 = 1, kSrcF32 = 2;

// one cell's plane values, {nn2.x, nn2.y, nn3} (0 where a plane is absent),
// held in the planes' own type until used
template <int SRC>
struct SrcV {
  using T = typename std::conditional<SRC == kSrcF64, double, float>::type;
  T x, y, e;
};

// the plane values of cell (n, m), clamped into the grid: lanes and rows
// outside it read a cell inside it (their results are not stored)
template <int SRC>
AST_DEV SrcV<SRC> src_load(const SrcF& s, int n, int m, int N, int M) {
  const size_t k = (size_t)min(max(m, 0), M - 1) * N + min(max(n, 0), N - 1);
  SrcV<SRC> v{0, 0, 0};
  if constexpr (SRC == kSrcF64) {
    if (s.nn2) {  // wave-uniform
      v.x = s.nn2[3 * k];
      v.y = s.nn2[3 * k + 1];
    }
    if (s.nn3) v.e = s.nn3[k];
  } else if constexpr (SRC == kSrcF32) {
    if (s.xy) {
      const float2 q = s.xy[k];
      v.x = q.x;
      v.y = q.y;
    }
    if (s.e) v.e = s.e[k];
  }
  return v;
}

struct StripArgs {
  SrcF src;
  int n, m0, m1;
  bool cols_in, out_col, xvote, ho_x, ho_y;
  // fast strips of the first / last band: the ghost lanes (gmask) apply
  // their side's rule to the cell they load (col_off): vx -> med3(gsign * vx,
  // glo, ghi) (mirror, clamped outflow), or the Dirichlet state gq (gdir);
  // all wave-uniform
  unsigned long long gmask;
  bool gdir;
  float gsign, glo, ghi;
  float gq[4];
  // density factor of pulsed Dirichlet ghosts (pulse_factor), read by the
  // boundary variant (the fast variants of the PULSE kernels take the pulsed
  // densities from a PulsePatch)
  double pulse;
  // fast edge strips (kFastE): the rules of the y-ghost rows, per side (0:
  // y lo, 1: y hi), all wave-uniform: a ghost row loads the mirror row
  // (yrefl) or the edge row, then vy -> med3(ysign * vy, ylo, yhi), or the
  // Dirichlet state yq (ydir)
  bool yrefl[2], ydir[2];
  float ysign[2], ylo[2], yhi[2];
  float yq[2][4];
  float ax, ay;   // -T_do/dx, -T_do/dy
  float tdof, dxf, dyf, x, ymin;
  unsigned col_off;  // byte offset of the lane's column in a row (interior strips)
  unsigned st_off;   // the same for the stores, kOffDrop on lanes that write nothing
  unsigned row_bytes;
};

template <int J>
struct Row {
  static constexpr int w0 = J & 3, w1 = (J + 1) & 3, w2 = (J + 2) & 3, w3 = (J + 3) & 3;
  static constexpr int e0 = J & 1, e1 = (J + 1) & 1;
};

// One wave's strip.  FAST (wave-uniform): the strip has 4 to 64 rows and its
// columns are interior or its band's ghost columns follow a per-lane rule,
// so rows are raw buffer loads at fixed per-lane offsets (kFast, kFastX: no
// ghost row, no frozen cell either, rows in groups of four and a tail of 1-3
// rows, the exact solver a multiple of 4 rows; kFastE: ghost rows by row
// rules, frozen tests per row, a multiple of 4 rows);
// otherwise load() applies the boundary rules, the frozen tests run per row,
// and the rows of the last group of four that lie past the strip vote for no
// dt (their stores fall outside the store range).
//
// Rows are processed four per trip with the window held in w[4] (row r in
// slot (r - m0) & 3): the slide of the window is a change of names, not a
// copy.  Row m + 3 is fetched while row m is computed; no branch separates
// the four rows of a trip, so the fetch stays a row ahead of its use.  The
// stores of the lanes that own no output go to an offset beyond the store
// range and are dropped by the range check, so every row issues the same
// memory operations and the load of row m + 3 never waits for a store.
// Strip paths: kEdge -- load() with the ghost and clamping rules per cell;
// kFast -- raw buffer loads of interior columns; kFastX -- the same for a
// band that holds x-ghost columns, each lane loading the cell its ghost rule
// names and applying the rule's per-lane transform; kFastE -- a strip that
// holds y-ghost rows or may meet frozen cells: the x-ghost lanes of kFastX,
// the row a ghost row loads chosen in scalar code and the y rule applied to
// the whole row, the frozen tests of kEdge.
constexpr int kEdge = 0, kFast = 1, kFastX = 2, kFastE = 3;

// The ghost densities of pulsed Dirichlet sides in the fast edge variants
// (kernels compiled for grids with a pulsed side), applied to each loaded
// row: rows m < prows take the pulsed density on this lane if it is a ghost
// column of a pulsed x side (gl), and in the y ghost rows of a pulsed y side
// (plo, phi > 0; -1 where the side is not pulsed).  Each select is between
// a member and a value (two members would fold into a load through a
// selected address, which keeps the struct in scratch memory).  NoPulse, the
// other kernels' patch, leaves the rows as loaded.
struct NoPulse {
  AST_DEV Q operator()(const Q& q, int) const { return q; }
};
struct PulsePatch {
  bool gl;
  int prows, M;
  float px, plo, phi;
  AST_DEV Q operator()(Q q, int m) const {
    const bool below = m < prows;
    q.r = (gl && below) ? px : q.r;
    q.r = (m < 0 && below && plo > 0.f) ? plo : q.r;
    q.r = (m >= M && below && phi > 0.f) ? phi : q.r;
    return q;
  }
};

template <int METHOD, int GEOM, bool WINDOW, int PATH, int SRC, class ST, class PF = NoPulse>
AST_DEV void run_strip(const ast_grid_t& g, const Consts& c, const Consts& cw, const ExactCtx& ex,
                       const float2* __restrict__ s, const float2* __restrict__ u,
                       float2* __restrict__ s2, float2* __restrict__ u2, const StripArgs& a,
                       float& UUx, float& UUy, ST& st, const PF& pf = PF{}) {
  constexpr bool FAST = PATH != kEdge;
  const int n = a.n, m0 = a.m0, m1 = a.m1, N = g.N, M = g.M;
  const unsigned rb = a.row_bytes;
  // loads: FAST strips read rows m0-2 .. m1+1 through one resource per plane
  // (kFastE: the rows of the grid among them, which hold every row a ghost
  // row names; a strip of kFastE has at least 4 rows, so M >= 4)
  const int rlo = PATH == kFastE ? max(m0 - 2, 0) : m0 - 2;
  const int rhi = PATH == kFastE ? min(m1 + 1, M - 1) : m1 + 1;
  const size_t b0 = FAST ? (size_t)rlo * N : 0;
  const long long lbytes = FAST ? (long long)(rhi - rlo + 1) * rb : 0;
  const __amdgpu_buffer_rsrc_t rs = make_rsrc(s + b0, lbytes), ru = make_rsrc(u + b0, lbytes);
  auto LD = [&](int m) -> Q {
    if constexpr (FAST) {
      // kernel.cu:2564-2569, 2609-2614, 2661-2698 (y sides): the row a ghost
      // row loads (wave-uniform)
      const bool ylo_row = PATH == kFastE && m < 0, yhi_row = PATH == kFastE && m >= M;
      int ms = m;
      if (ylo_row) ms = a.yrefl[0] ? -1 - m : 0;
      if (yhi_row) ms = a.yrefl[1] ? 2 * M - 1 - m : M - 1;
      const unsigned so = (unsigned)(ms - rlo) * rb;
      const float2 p = bload(rs, a.col_off, so), q = bload(ru, a.col_off, so);
      if constexpr (PATH == kFastX || PATH == kFastE) {  // kernel.cu:2549-2580, 2639-2654 (x sides)
        const bool gl = (a.gmask >> (threadIdx.x & 63)) & 1ull;
        const bool gd = gl & a.gdir;  // (no branch: the loads stay a row ahead)
        const float vx = gl ? __builtin_amdgcn_fmed3f(a.gsign * q.x, a.glo, a.ghi) : q.x;
        Q r{gd ? a.gq[0] : p.x, gd ? a.gq[1] : p.y, gd ? a.gq[2] : vx, gd ? a.gq[3] : q.y};
        if constexpr (PATH == kFastE) {  // the y rule, the same for the whole row
          const bool gy = ylo_row || yhi_row;
          const float sg = yhi_row ? a.ysign[1] : a.ysign[0];
          const float lo = yhi_row ? a.ylo[1] : a.ylo[0];
          const float hi = yhi_row ? a.yhi[1] : a.yhi[0];
          const bool yd = gy && (yhi_row ? a.ydir[1] : a.ydir[0]);
          const float vy = gy ? __builtin_amdgcn_fmed3f(sg * r.vy, lo, hi) : r.vy;
          r = Q{yd ? (yhi_row ? a.yq[1][0] : a.yq[0][0]) : r.r,
                yd ? (yhi_row ? a.yq[1][1] : a.yq[0][1]) : r.p,
                yd ? (yhi_row ? a.yq[1][2] : a.yq[0][2]) : r.vx,
                yd ? (yhi_row ? a.yq[1][3] : a.yq[0][3]) : vy};
        }
        return r;
      }
      return Q{p.x, p.y, q.x, q.y};
    } else {
      return load(g, s, u, n, m, a.cols_in, a.pulse);
    }
  };
  // last row a prefetch may name (the boundary variant's load() clamps any
  // row into the grid)
  const int mlast = FAST ? m1 + 1 : m1 + 5;
  // stores: rows m0 .. m1-1
  const size_t sb = (size_t)m0 * N;
  const long long sbytes = (long long)(m1 - m0) * rb;
  const __amdgpu_buffer_rsrc_t ws = make_rsrc(s2 + sb, sbytes), wu = make_rsrc(u2 + sb, sbytes);

  const bool frz = (PATH == kEdge || PATH == kFastE) && cw.frozen;  // wave-uniform

  // prologue: rows m0-2 .. m0+1, the south face of row m0
  Q w[4], nv[2], dd[2];
  FaceF G[2];
  bool fr[2] = {false, false};
  {
    const Q qa = pf(LD(m0 - 2), m0 - 2), qb = pf(LD(m0 - 1), m0 - 1);
    w[0] = pf(LD(m0), m0);
    w[1] = pf(LD(m0 + 1), m0 + 1);
    w[2] = pf(LD(m0 + 2), m0 + 2);
    const Q d0 = sub(qb, qa), d1 = sub(w[0], qb);
    dd[0] = sub(w[1], w[0]);
    Q nvm1, svm1, sv0;
    face_values(qb, d0, d1, nvm1, svm1);
    face_values(w[0], d1, dd[0], nv[0], sv0);
    float uu;
    G[0] = face<METHOD, WINDOW, false>(nvm1, sv0, qb, w[0], a.ho_y, c, uu, ex, st);
    if (frz) {
      fr[0] = frozen_at(cw, a.x, a.ymin + m0 * a.dyf);
      const bool frm1 = frozen_at(cw, a.x, a.ymin + (m0 - 1) * a.dyf);
      if (!(fr[0] && frm1)) UUy = fmaxf(UUy, uu);
    } else {
      UUy = fmaxf(UUy, uu);
    }
  }

  // results of the previous row, stored after this row's fetch is issued
  // (memory operations retire in order: a store issued before the fetch of
  // row m+3 would make the wait for row m+2 wait for the store too).  The
  // first pending store goes to the drop offset.
  // FAST axisymmetric strips: T_do/y of row m0 + k in lane k (the strip has
  // at most 64 rows), read back per row as a scalar -- the per-row division
  // by the wave-uniform radius costs one lane read instead of a conversion,
  // an fma, a reciprocal and a product
  float tyl = 0.f;
  if constexpr (FAST && GEOM == AST_AXISYMMETRIC)
    tyl = a.tdof * frcp(a.ymin + (m0 + (int)(threadIdx.x & 63)) * a.dyf);
  Q pend = w[0];
  unsigned pend_vo = kOffDrop;
  // source planes: row m's values are fetched with row m-1's state fetch,
  // so that they arrive in order with it (memory operations retire in
  // order: a fetch issued where the values are used would make that use
  // wait for the state fetch issued before it)
  SrcV<SRC> scur{0, 0, 0};
  if constexpr (SRC != kSrcNone) scur = src_load<SRC>(a.src, n, m0, N, M);

  auto row = [&](auto tag, int m) {
    using R = decltype(tag);
    const Q& q0 = w[R::w0];
    const Q& q1 = w[R::w1];
    const Q& q2 = w[R::w2];
    w[R::w3] = pf(LD(min(m + 3, mlast)), min(m + 3, mlast));
    SrcV<SRC> snext{0, 0, 0};
    // (the last row of the strip re-reads its own values, which are cached,
    // rather than the next strip's)
    if constexpr (SRC != kSrcNone) snext = src_load<SRC>(a.src, n, min(m + 1, m1 - 1), N, M);
    bstore(pend.r, pend.p, ws, pend_vo, 0);
    bstore(pend.vx, pend.vy, wu, pend_vo, 0);
    // keep the fetch of row m+3 here, a row ahead of its use (the
    // scheduler would otherwise sink it towards its first use)
    if constexpr (FAST) __builtin_amdgcn_sched_barrier(0);
    // y: face values of row m+1, the north face of row m
    const Q d2 = sub(q2, q1);
    Q nv1, sv1;
    face_values(q1, dd[R::e0], d2, nv1, sv1);
    dd[R::e1] = d2;
    const float y = a.ymin + m * a.dyf;
    const bool live = FAST || m < m1;  // wave-uniform
    float uu;
    const FaceF Gn = face<METHOD, WINDOW, false>(nv[R::e0], sv1, q0, q1, a.ho_y, c, uu, ex, st);
    nv[R::e1] = nv1;
    // x: face values of row m, its west face (left state = east value of
    // lane-1); the east face flux is the west face flux of lane+1.  The
    // zero-filled shifts leave lane 0's slopes and lane 63's east flux
    // meaningless (finite, positive face values): both are halo lanes.
    const Q ql = shr1(q0);
    const Q dl = sub(q0, ql);
    const Q dr = Q{shl1(q0.r) - q0.r, shl1(q0.p) - q0.p, shl1(q0.vx) - q0.vx, shl1(q0.vy) - q0.vy};
    Q ev, wv;
    face_values(q0, dl, dr, ev, wv);
    float ux;
    const FaceF Fw = face<METHOD, WINDOW, true>(shr1_keep(ev), wv, ql, q0, a.ho_x, c, ux, ex, st);
    if (frz) {
      const bool f0 = fr[R::e0];
      const bool f1 = frozen_at(cw, a.x, y + a.dyf);
      const bool fl = frozen_at(cw, a.x - a.dxf, y);
      if (live && !(f0 && f1)) UUy = fmaxf(UUy, uu);
      if (live && !(f0 && fl)) UUx = fmaxf(UUx, ux);
      fr[R::e1] = f1;
    } else {
      UUy = fmaxf(UUy, live ? uu : 0.f);
      UUx = fmaxf(UUx, live ? ux : 0.f);
    }
    const FaceF& Gs = G[R::e0];
    // conservative update, kernel.cu:3027-3032, 3059-3073, written in the
    // frame of the cell (velocity u = (vx, vy)): with a = -T_do/dx, -T_do/dy
    // and each face's flux seen from the cell (in_cell_frame),
    //   D_rho = a (m_E - m_W) + ...             (rho' = rho + D_rho)
    //   D_1   = a (g_E - g_W) + ...  = dm - u D_rho
    //   D_2   = a (e_E - e_W) + ...  = dE - u.dm + |u|^2 D_rho / 2
    // and, exactly, u' = u + D_1/rho', rho' e' = rho e + D_2 - |D_1|^2/(2 rho').
    // No term is of the size of the kinetic energy, so the internal energy
    // -- O(1/M^2) of the total at Mach M -- keeps FP32's relative accuracy.
    constexpr bool ST1 = state_faces<METHOD, WINDOW>();
    const FaceF Fe = shl1<ST1>(Fw);
    const CellFlux cE = in_cell_frame<ST1>(Fe, q0.vx, q0.vy);
    const CellFlux cW = in_cell_frame<ST1>(Fw, q0.vx, q0.vy);
    const CellFlux cN = in_cell_frame<ST1>(Gn, q0.vy, q0.vx);
    const CellFlux cS = in_cell_frame<ST1>(Gs, q0.vy, q0.vx);
    float dr_ = a.ax * (Fe.m - Fw.m) + a.ay * (Gn.m - Gs.m);
    float de = a.ax * (cE.e - cW.e) + a.ay * (cN.e - cS.e);
    float d1x = a.ax * (cE.n - cW.n) + a.ay * (cN.t - cS.t);
    float d1y = a.ax * (cE.t - cW.t) + a.ay * (cN.n - cS.n);
    G[R::e1] = Gn;
    if constexpr (GEOM == AST_AXISYMMETRIC) {
      // KorolkovSolver.cu:13-32: -(T_do/y) v {rho, rho u, rho v, E + p}; in
      // the cell frame the momentum part vanishes and the energy part is
      // -(T_do/y) v (rho e + p) = -(T_do/y) v gamma p/(gamma-1)
      float ty;
      if constexpr (FAST) ty = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tyl), m - m0));
      else ty = a.tdof * frcp(y);
      const float f = ty * q0.vy;
      dr_ = fmaf(-f, q0.r, dr_);
      de = fmaf(-f, c.gig1 * q0.p, de);
    }
    if constexpr (SRC != kSrcNone) {
      // add_MK's neutral-hydrogen terms (kernel.cu:2495-2498), before the
      // floors: momentum T_do (rho kmom) nn2.{x, y}, energy T_do ken nn3;
      // seen from the cell the energy change loses u.(the momentum change)
      const float tk = a.tdof * (q0.r * a.src.kmom);
      const float mx = tk * (float)scur.x;
      const float my = tk * (float)scur.y;
      const float se = (a.tdof * a.src.ken) * (float)scur.e;
      scur = snext;
      d1x += mx;
      d1y += my;
      de += se - fmaf(q0.vx, mx, q0.vy * my);
    }
    // kernel.cu:3060-3064: a non-positive density keeps the old one, while
    // the momentum change stays (the reference divides rho u + dm by the old
    // density): then D_1 gains u D_rho and D_2 loses |u|^2 D_rho/2.  Rows
    // where no lane floors (nearly all) skip those terms (wave-uniform).
    const bool rfloor = q0.r + dr_ <= 0.f;
    float r2, vx2, vy2, d2f;
    if (__builtin_amdgcn_ballot_w64(rfloor) == 0) {
      r2 = q0.r + dr_;
      const float ir = frcp(r2);
      vx2 = fmaf(d1x, ir, q0.vx);
      vy2 = fmaf(d1y, ir, q0.vy);
      d2f = de - 0.5f * fmaf(d1x, d1x, d1y * d1y) * ir;
    } else {
      const float kept = rfloor ? dr_ : 0.f;
      r2 = q0.r + (dr_ - kept);
      const float ir = frcp(r2);
      const float e1x = fmaf(q0.vx, kept, d1x), e1y = fmaf(q0.vy, kept, d1y);
      vx2 = fmaf(e1x, ir, q0.vx);
      vy2 = fmaf(e1y, ir, q0.vy);
      d2f = de - 0.5f * (fmaf(q0.vx, q0.vx, q0.vy * q0.vy) * kept +
                         fmaf(e1x, e1x, e1y * e1y) * ir);
    }
    float p2 = fmaf(c.g1, d2f, q0.p);
    p2 = (p2 <= 0.f) ? 0.000001f : p2;  // kernel.cu:3070-3073
    if (frz && fr[R::e0]) {  // kernel.cu:2317-2324
      r2 = q0.r;
      p2 = q0.p;
      vx2 = q0.vx;
      vy2 = q0.vy;
    }
    // the row offset goes into the VGPR offset, which the range check
    // covers (the SGPR offset is not range checked): rows past the strip
    // and non-output lanes fall outside the store range
    pend = Q{r2, p2, vx2, vy2};
    pend_vo = a.st_off + (unsigned)(m - m0) * rb;
  };

  // kFast, kFastX: whole groups of four, then the 1-3 rows left
  // (wave-uniform).  Whole groups only: kFastE strips (a tail there costs
  // the kernel registers) and the exact solver's strips (a tail would inline
  // the solver twice more into a kernel whose code size already matters);
  // the boundary variant runs its last group whole (rows past m1 drop out).
  constexpr bool TAIL = (PATH == kFast || PATH == kFastX) && !is_exact(METHOD);
  const int mq = TAIL ? m0 + ((m1 - m0) & ~3) : m1;
  for (int m = m0; m < mq; m += 4) {
    if constexpr (FAST && !is_exact(METHOD)) {
      // wave priority falls with the fraction of the strip done: a SIMD
      // issues for its oldest wave first, so without this its waves finish
      // one after another and the last runs alone; with it the waves that
      // are behind catch up and the SIMD's waves end together
      const int q = (4 * (m - m0)) / (m1 - m0);  // wave-uniform, 0..3
      if (q == 0) __builtin_amdgcn_s_setprio(3);
      else if (q == 1) __builtin_amdgcn_s_setprio(2);
      else if (q == 2) __builtin_amdgcn_s_setprio(1);
      else __builtin_amdgcn_s_setprio(0);
    }
    row(Row<0>{}, m);
    row(Row<1>{}, m + 1);
    row(Row<2>{}, m + 2);
    row(Row<3>{}, m + 3);
  }
  if constexpr (TAIL) {
    const int rest = m1 - mq;
    if (rest > 0) row(Row<0>{}, mq);
    if (rest > 1) row(Row<1>{}, mq + 1);
    if (rest > 2) row(Row<2>{}, mq + 2);
  }
  bstore(pend.r, pend.p, ws, pend_vo, 0);
  bstore(pend.vx, pend.vy, wu, pend_vo, 0);
}

AST_DEV float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// The band and rows [m0, m1) of wave `wave` of workgroup `wg` under a strip
// tiling (wave-uniform).  tile > 0: strips of `tile` rows, WPB of them
// stacked per workgroup, band = wg mod nbands.  tile = -S: every band cut
// into S strips of about M/S rows, the waves of all workgroups numbered band
// by band (a workgroup may span two bands); this lets the launch hold
// exactly as many waves as the SIMDs take at once.  The first and last
// strip of a band (they hold the ghost rows) get M/S rounded to a multiple
// of 4, the others floor or ceil of what is left over them.  Waves past the
// last band get no rows (m0 = m1 = M).
AST_DEV void strip_rows(int M, int nbands, int tile, int wg, int wave, int& band, int& m0,
                        int& m1) {
  if (tile > 0) {
    band = wg % nbands;
    m0 = ((wg / nbands) * WPB + wave) * tile;
    m1 = min(m0 + tile, M);
  } else {
    const int S = -tile, gw = wg * WPB + wave;
    band = gw / S;
    const int k = gw - band * S;
    const int e = max(4, (M / S + 2) / 4 * 4), mid = M - 2 * e;
    if (S >= 3 && mid >= S - 2) {
      const int h = mid / (S - 2), r = mid - h * (S - 2), j = k - 1;
      m0 = k == 0 ? 0 : (k == S - 1 ? M - e : e + j * h + min(j, r));
      m1 = k == 0 ? e : (k == S - 1 ? M : m0 + h + (j < r ? 1 : 0));
    } else {
      const int h = M / S, r = M - h * S;
      m0 = k * h + min(k, r);
      m1 = m0 + h + (k < r ? 1 : 0);
    }
    if (band >= nbands) {
      band = nbands - 1;
      m0 = m1 = M;
    }
  }
}

// The strip of one wave given its band and rows (strip_rows): fills
// StripArgs and runs the fast or the boundary variant.  Returns the wave's
// dt candidate (wave-uniform).
template <int METHOD, int GEOM, bool WINDOW, int SRC = kSrcNone, bool PULSE = false, class ST>
AST_DEV double strip_of_wave(const ast_grid_t& g, const Consts& c, const ExactCtx& ex,
                             const float2* __restrict__ s, const float2* __restrict__ u,
                             float2* __restrict__ s2, float2* __restrict__ u2, double tdo,
                             int band, int m0, int m1, ST& st,
                             const SrcF& src = SrcF{nullptr, nullptr, nullptr, nullptr, 0.f, 0.f},
                             double pulse = 1.0) {
  const int lane = threadIdx.x & 63;
  StripArgs a;
  a.src = src;
  a.m0 = m0;
  a.m1 = m1;
  a.n = band * WOUT + lane - 2;
  a.cols_in = (band * WOUT - 2 >= 0) && (band * WOUT + 61 < g.N);  // wave-uniform
  a.out_col = (lane >= 2) & (lane < 2 + WOUT) & (a.n < g.N);
  a.xvote = (lane >= 2) & (lane <= 2 + WOUT) & (a.n <= g.N);
  a.ax = (float)(-tdo / g.dx);
  a.ay = (float)(-tdo / g.dy);
  a.tdof = (float)tdo;
  a.dxf = (float)g.dx;
  a.dyf = (float)g.dy;
  a.x = (float)(g.x_min + a.n * g.dx);
  a.ymin = (float)g.y_min;
  a.row_bytes = (unsigned)g.N * 8u;
  a.col_off = (unsigned)(a.n < 0 ? 0 : a.n) * 8u;
  // the x-ghost rule of fetch_cell for the ghost lanes: only the first or the
  // last band holds ghost columns, never both (more than one band)
  const bool xlo_band = band * WOUT - 2 < 0, xhi_band = band * WOUT + 61 >= g.N;
  const bool xfast = (xlo_band != xhi_band);  // wave-uniform
  const int side = xlo_band ? AST_XLO : AST_XHI;
  const int gbc = g.bc[side];
  // lanes of columns -2, -1 (first band) or N, N+1 (last band)
  const int l0 = xlo_band ? 0 : g.N - band * WOUT + 2;
  a.gmask = xfast ? (3ull << l0) : 0ull;
  a.gdir = gbc == AST_BC_DIRICHLET;
  a.gsign = gbc == AST_BC_REFLECT ? -1.f : 1.f;
  const float vc = (float)g.bc_vclamp[side];
  a.glo = (gbc == AST_BC_OUTFLOW_CLAMPED && side == AST_XHI) ? vc : -3.40282347e38f;
  a.ghi = (gbc == AST_BC_OUTFLOW_CLAMPED && side == AST_XLO) ? vc : 3.40282347e38f;
  for (int k = 0; k < 4; k++) a.gq[k] = (float)g.bc_state[side][k];
  a.pulse = pulse;
  if (xfast) {  // the cell each lane loads: ghost rule, else clamped into the grid
    int src = a.n < 0 ? 0 : (a.n >= g.N ? g.N - 1 : a.n);
    const bool ghost = (a.n < 0 && a.n >= -2) || (a.n >= g.N && a.n <= g.N + 1);
    if (ghost && gbc == AST_BC_REFLECT) src = a.n < 0 ? -1 - a.n : 2 * g.N - 1 - a.n;
    a.col_off = (unsigned)src * 8u;
  }
  a.st_off = a.out_col ? (unsigned)a.n * 8u : kOffDrop;
  a.ho_x = high_order(g, g.x_min + a.n * g.dx - 0.5 * g.dx);  // west face window
  a.ho_y = high_order(g, g.x_min + a.n * g.dx);               // y faces: cell x
  float UUx = 0.f, UUy = 0.f;
  Consts cw = c;  // frozen tests only where the wave can meet the region
  cw.frozen = c.frozen &&
              frozen_possible(g, band * WOUT - 3, band * WOUT + 62, a.m0 - 2, a.m1 + 1);
  if (a.m0 < g.M) {
    // wave-uniform
    const bool fast_rows = (a.cols_in || xfast) && a.m1 - a.m0 >= 4 && a.m1 - a.m0 <= 64 &&
                           (!is_exact(METHOD) || ((a.m1 - a.m0) & 3) == 0);
    const bool rows_in = a.m0 - 2 >= 0 && a.m1 + 1 < g.M;
    if constexpr (PULSE) {
      // a grid with a pulsed side: the strips that hold its ghosts run the
      // fast variants with the pulsed densities applied to each loaded row
      PulsePatch pp;
      pp.prows = -2;  // the first row that is not pulsed_row (monotone in m)
      for (int hi = g.M + 2; pp.prows < hi;) {
        const int mid = pp.prows + (hi - pp.prows) / 2;
        if (pulsed_row(g, mid)) pp.prows = mid + 1;
        else hi = mid;
      }
      pp.M = g.M;
      pp.px = pp.plo = pp.phi = -1.f;
      if (gbc == AST_BC_DIRICHLET_PULSED)
        pp.px = (float)((xlo_band ? g.bc_state[AST_XLO][0] : g.bc_state[AST_XHI][0]) * pulse);
      if (g.bc[AST_YLO] == AST_BC_DIRICHLET_PULSED) pp.plo = (float)(g.bc_state[AST_YLO][0] * pulse);
      if (g.bc[AST_YHI] == AST_BC_DIRICHLET_PULSED) pp.phi = (float)(g.bc_state[AST_YHI][0] * pulse);
      pp.gl = xfast && pp.px > 0.f && ((a.gmask >> lane) & 1ull);
      a.gdir = a.gdir || gbc == AST_BC_DIRICHLET_PULSED;
      if (fast_rows && rows_in && !cw.frozen) {
        if (a.cols_in)
          run_strip<METHOD, GEOM, WINDOW, kFast, SRC>(g, c, cw, ex, s, u, s2, u2, a, UUx, UUy, st);
        else
          run_strip<METHOD, GEOM, WINDOW, kFastX, SRC>(g, c, cw, ex, s, u, s2, u2, a, UUx, UUy, st,
                                                       pp);
      } else if (fast_rows && ((a.m1 - a.m0) & 3) == 0) {
        for (int k = 0; k < 2; k++) {
          const int ys = k ? AST_YHI : AST_YLO;
          const int ybc = g.bc[ys];
          const float yc = (float)g.bc_vclamp[ys];
          a.yrefl[k] = ybc == AST_BC_REFLECT;
          a.ydir[k] = ybc == AST_BC_DIRICHLET || ybc == AST_BC_DIRICHLET_PULSED;
          a.ysign[k] = ybc == AST_BC_REFLECT ? -1.f : 1.f;
          a.ylo[k] = (ybc == AST_BC_OUTFLOW_CLAMPED && k == 1) ? yc : -3.40282347e38f;
          a.yhi[k] = (ybc == AST_BC_OUTFLOW_CLAMPED && k == 0) ? yc : 3.40282347e38f;
          for (int j = 0; j < 4; j++) a.yq[k][j] = (float)g.bc_state[ys][j];
        }
        run_strip<METHOD, GEOM, WINDOW, kFastE, SRC>(g, c, cw, ex, s, u, s2, u2, a, UUx, UUy, st,
                                                     pp);
      } else {
        run_strip<METHOD, GEOM, WINDOW, kEdge, SRC>(g, c, cw, ex, s, u, s2, u2, a, UUx, UUy, st);
      }
    } else {
      // kernels without the PULSE variants (the exact solver's): strips
      // holding ghosts of a pulsed side take the boundary variant (fetch_cell
      // applies the pulse; the results are those of the fast variants, bit
      // for bit)
      const bool xpulse = c.pulsed && xfast && gbc == AST_BC_DIRICHLET_PULSED;
      const bool ypulse = c.pulsed && (g.bc[AST_YLO] == AST_BC_DIRICHLET_PULSED ||
                                       g.bc[AST_YHI] == AST_BC_DIRICHLET_PULSED);
      if (fast_rows && rows_in && !cw.frozen && !xpulse) {
        if (a.cols_in)
          run_strip<METHOD, GEOM, WINDOW, kFast, SRC>(g, c, cw, ex, s, u, s2, u2, a, UUx, UUy, st);
        else
          run_strip<METHOD, GEOM, WINDOW, kFastX, SRC>(g, c, cw, ex, s, u, s2, u2, a, UUx, UUy, st);
      } else if (fast_rows && ((a.m1 - a.m0) & 3) == 0 && !xpulse && !ypulse) {
        // the y-ghost rules of fetch_cell as row rules (only the first or last
        // strips of a band hold ghost rows)
        for (int k = 0; k < 2; k++) {
          const int ys = k ? AST_YHI : AST_YLO;
          const int ybc = g.bc[ys];
          const float yc = (float)g.bc_vclamp[ys];
          a.yrefl[k] = ybc == AST_BC_REFLECT;
          a.ydir[k] = ybc == AST_BC_DIRICHLET;
          a.ysign[k] = ybc == AST_BC_REFLECT ? -1.f : 1.f;
          a.ylo[k] = (ybc == AST_BC_OUTFLOW_CLAMPED && k == 1) ? yc : -3.40282347e38f;
          a.yhi[k] = (ybc == AST_BC_OUTFLOW_CLAMPED && k == 0) ? yc : 3.40282347e38f;
          for (int j = 0; j < 4; j++) a.yq[k][j] = (float)g.bc_state[ys][j];
        }
        run_strip<METHOD, GEOM, WINDOW, kFastE, SRC>(g, c, cw, ex, s, u, s2, u2, a, UUx, UUy, st);
      } else {
        run_strip<METHOD, GEOM, WINDOW, kEdge, SRC>(g, c, cw, ex, s, u, s2, u2, a, UUx, UUy, st);
      }
    }
  }
  // only output lanes vote (halo lanes may hold garbage speeds)
  UUx = wave_max(a.xvote ? UUx : 0.f);
  UUy = wave_max(a.out_col ? UUy : 0.f);
  // per-face dt = krit*rad/UU (monotone in UU), capped at 1000 (kernel.cu:2530)
  const bool any_out = (band * WOUT < g.N) && a.m0 < g.M;  // wave-uniform
  double tmin = any_out ? 1000.0 : INFINITY;
  if (UUx > 0.f) tmin = fmin(tmin, g.krit * g.dx / (double)UUx);
  if (UUy > 0.f) tmin = fmin(tmin, g.krit * g.dy / (double)UUy);
  return tmin;
}

}  // namespace f32
}  // namespace ast
