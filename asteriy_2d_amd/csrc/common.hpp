// common.hpp -- shared device/host helpers of the HIP kernels: error
// reporting of the C-ABI, the XCD-aware tile order, the ghost-cell rules,
// and the deterministic block reductions (dt min).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/asteriy.h"
#include "dev.hpp"
#include "time_slots.h"

// ---- host side: status + message of the last failure (api.cpp) ----------
int ast_fail(int code, const char* fmt, ...);
int ast_hip_fail(hipError_t e, const char* where);
#define AST_HIP_CHECK(call)                                   \
  do {                                                        \
    hipError_t _e = (call);                                   \
    if (_e != hipSuccess) return ast_hip_fail(_e, #call);     \
  } while (0)
#define AST_LAUNCH_CHECK(where)                               \
  do {                                                        \
    hipError_t _e = hipGetLastError();                        \
    if (_e != hipSuccess) return ast_hip_fail(_e, where);     \
  } while (0)

namespace ast {

// The hardware hands workgroup i to XCD (i mod 8).  Remap so that each XCD
// receives one contiguous run of tiles: neighbouring tiles share halo rows
// and then meet in the same XCD's L2.
AST_DEV int xcd_tile(int bid, int ntiles) {
  const int nx = 8;
  int q = ntiles / nx, r = ntiles % nx;
  int xcd = bid % nx, local = bid / nx;
  return xcd * q + (xcd < r ? xcd : r) + local;
}

// Density factor of AST_BC_DIRICHLET_PULSED sides at simulation time TT
// (add2's inflow perturbation, kernel.cu:2022-2025): 1 + a sin(w TT) while
// TT < t_lim, else 1.
AST_DEV double pulse_factor(const ast_grid_t& g, double TT) {
  return TT < g.bc_pulse[3] ? 1.0 + g.bc_pulse[0] * sin(TT * g.bc_pulse[1]) : 1.0;
}

// Does row m (ghost rows included) have its centre below the pulsed sides'
// y_lim?  The sum is formed unfused, as the host forms it.
__host__ __device__ __forceinline__ bool pulsed_row(const ast_grid_t& g, int m) {
#pragma clang fp contract(off)
  return g.y_min + m * g.dy < g.bc_pulse[2];
}

// Cell (n, m) of a state plane with the ghost rules of Kernel_TVD
// (kernel.cu:2549-2698) and add_MK (kernel.cu:2328-2377).  Indices may lie
// up to 2 cells outside the grid on at most one axis; anything else
// (corners, beyond the halo of a partial tile) is never read by a cell and
// returns zeros.  pulse: the density factor of pulsed Dirichlet ghosts whose
// centre lies below y_lim (pulse_factor).
template <typename P>
AST_DEV void fetch_cell(const ast_grid_t& g, const P* __restrict__ s, const P* __restrict__ u,
                        int n, int m, P& so, P& uo, double pulse = 1.0) {
  const int N = g.N, M = g.M;
  const bool xin = (n >= 0) & (n < N), yin = (m >= 0) & (m < M);
  if (xin & yin) {
    so = s[(size_t)m * N + n];
    uo = u[(size_t)m * N + n];
    return;
  }
  so.x = so.y = uo.x = uo.y = 0;
  if ((!xin & !yin) | (n < -2) | (n > N + 1) | (m < -2) | (m > M + 1)) return;
  int side, in_n = n, in_m = m;
  if (n < 0) { side = AST_XLO; in_n = -1 - n; }
  else if (n >= N) { side = AST_XHI; in_n = 2 * N - 1 - n; }
  else if (m < 0) { side = AST_YLO; in_m = -1 - m; }
  else { side = AST_YHI; in_m = 2 * M - 1 - m; }
  const int bc = g.bc[side];
  if (bc == AST_BC_DIRICHLET || bc == AST_BC_DIRICHLET_PULSED) {
    so.x = g.bc_state[side][0];
    so.y = g.bc_state[side][1];
    uo.x = g.bc_state[side][2];
    uo.y = g.bc_state[side][3];
    if (bc == AST_BC_DIRICHLET_PULSED && pulsed_row(g, m))
      so.x = (decltype(so.x))(g.bc_state[side][0] * pulse);
    return;
  }
  if (bc == AST_BC_REFLECT) {
    size_t k = (size_t)in_m * N + in_n;
    so = s[k];
    uo = u[k];
    if (side <= AST_XHI) uo.x = -uo.x;
    else uo.y = -uo.y;
    return;
  }
  int en = n < 0 ? 0 : (n >= N ? N - 1 : n);
  int em = m < 0 ? 0 : (m >= M ? M - 1 : m);
  size_t k = (size_t)em * N + en;
  so = s[k];
  uo = u[k];
  if (bc == AST_BC_OUTFLOW_CLAMPED) {
    // kernel.cu:2347-2350 caps the lo-side normal velocity from above; the
    // hi side is the mirror image.
    const auto c = (decltype(uo.x))g.bc_vclamp[side];
    if (side == AST_XLO) { if (uo.x > c) uo.x = c; }
    else if (side == AST_XHI) { if (uo.x < c) uo.x = c; }
    else if (side == AST_YLO) { if (uo.y > c) uo.y = c; }
    else { if (uo.y < c) uo.y = c; }
  }
}

// Frozen region of add_MK (kernel.cu:2317-2324) as an ellipse.
AST_DEV bool is_frozen(const ast_grid_t& g, double x, double y) {
  if (!(g.frozen_ax > 0.0)) return false;
  double dxr = x - g.frozen_x0;
  double d = (dxr * dxr) / (g.frozen_ax * g.frozen_ax) + (y * y) / (g.frozen_ay * g.frozen_ay);
  return d < 1.0;
}

AST_DEV bool high_order(const ast_grid_t& g, double xf) { return xf > g.ho_lo && xf < g.ho_hi; }

// ---- wave64 / block reductions ------------------------------------------
AST_DEV double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
AST_DEV double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
AST_DEV uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Lowers *T to the block minimum of tmin.  Positive IEEE doubles order like
// their bit patterns, so one 64-bit atomicMin on the bits gives the exact,
// order-independent minimum (the reference's store, kernel.cu:3013-3017, is
// racy).  `red` is LDS of at least nwaves doubles.
template <int NT>
AST_DEV void block_min_to(double tmin, double* red, double* T) {
  constexpr int NW = NT / 64;
  tmin = wave_min(tmin);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = tmin;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
#pragma unroll
    for (int i = 1; i < NW; i++) t = fmin(t, red[i]);
    if (t < INFINITY)
      atomicMin(reinterpret_cast<unsigned long long*>(T), (unsigned long long)__double_as_longlong(t));
  }
}

// block_min_to for a tmin that is already uniform across each wave.  With
// nslots > 0, T is a set of dt slots (time_slots.h) and workgroup b lowers
// slot (b mod nslots), so that no single address takes every workgroup's
// atomic; the time advance folds the slots.
template <int NT>
AST_DEV void block_min_uniform(double tmin, double* red, double* T, int nslots) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = tmin;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
#pragma unroll
    for (int i = 1; i < NW; i++) t = fmin(t, red[i]);
    double* dst = nslots > 0 ? T + kSlotStride * (int)(blockIdx.x % (unsigned)nslots) : T;
    if (t < INFINITY)
      atomicMin(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)__double_as_longlong(t));
  }
}

// The time arguments of one step launch (time_slots.h): prev, the slot set
// of the previous step (null: dt is *T_do); cur, the set this step lowers
// (null: *T itself); rst, a set to clear (may be null).
struct KTime {
  const double* prev;
  double* cur;
  double* rst;
  int nslots;
};

// dt of this step, in two parts.  step_dt_load (loads only): with a previous
// slot set, dt is the minimum of its slots (every wave folds them itself: no
// workgroup waits for another), else *T_do.  step_dt_commit (the writes):
// with a previous set, wave 0 of workgroup 0 performs the previous step's
// time advance, funk_time of kernel.cu:1525-1536 (T_do = that minimum,
// TT += T_do, i += 1; T stays at its reset value); it clears the set `rst`
// for the step after.
AST_DEV double step_dt_load(const double* T_do, const KTime& kt) {
  const int lane = threadIdx.x & 63;
  if (kt.prev != nullptr) return wave_min(lane < kt.nslots ? kt.prev[kSlotStride * lane] : INFINITY);
  return *T_do;
}
AST_DEV void step_dt_commit(double* T, const KTime& kt, double t) {
  const int lane = threadIdx.x & 63;
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    if (kt.rst != nullptr) {
      if (lane < kt.nslots) kt.rst[kSlotStride * lane] = kSlotEmpty;
      if (lane == 0) reinterpret_cast<uint32_t*>(kt.rst + 1)[0] = 0u;
    }
    if (kt.prev != nullptr && lane == 0) {
      T[1] = t;
      T[2] = T[2] + t;
      int* i = reinterpret_cast<int*>(T + 3);
      *i = *i + 1;
    }
  }
}

}  // namespace ast
