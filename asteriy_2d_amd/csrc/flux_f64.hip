// flux_f64.hip -- batched FP64 face-flux operators (HLL, HLLC, exact,
// HLLC_Aleksashov) over SoA face arrays: the reference's per-face operators
// (kernel.cu:1399, Gas_Dinamic.cu:1814, kernel.cu:128, kernel.cu:767)
// applied to n independent faces with a common normal (n1, n2), computed by
// the face-frame solvers of flux64.hpp.
//
// Memory: 8 input planes and 5 output planes, each lane one face per
// iteration of a grid-stride loop; every load/store is a coalesced 8-byte
// access of consecutive faces (104 B/face of algorithmic traffic).
#include "common.hpp"
#include "flux64.hpp"

namespace ast {
namespace f64 {

using namespace d64;

constexpr int FNT = 256;

// Exact-solver counters of one lane, ast_stats_t order.
struct LaneStats {
  uint32_t v[10];
  AST_DEV void clear() {
    for (int k = 0; k < 10; k++) v[k] = 0;
  }
  AST_DEV void add(const ExactInfo& e) {
    v[0] += 1;
    if (e.pat == 2) { v[1] += 1; v[2] += e.it; }
    if (e.pat == 1) { v[3] += 1; v[4] += e.it; }
    if (e.pat == 0) v[5] += 1;
    if (e.pat == -1) v[6] += 1;
    if (e.capped) v[7] += 1;
    if (e.guessed && e.pat == 2) v[9] += 1;
  }
};

AST_DEV void block_stats_to(const LaneStats& c, uint32_t* red, ast_stats_t* stats) {
  constexpr int NW = FNT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 0; k < 10; k++) {
    const uint32_t w = wave_sum(c.v[k]);
    if (lane == 0) red[wave * 10 + k] = w;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    unsigned long long t = 0;
    for (int w = 0; w < NW; w++) t += red[w * 10 + threadIdx.x];
    if (t) atomicAdd(reinterpret_cast<unsigned long long*>(stats) + threadIdx.x, t);
  }
}

// face i of the SoA batch in the frame of the normal (n1, n2)
AST_DEV void load_face(const double* __restrict__ in, size_t n, size_t i, double n1, double n2,
                       PrimD& L, PrimD& R) {
  const double vxl = in[2 * n + i], vyl = in[3 * n + i];
  const double vxr = in[6 * n + i], vyr = in[7 * n + i];
  L = PrimD{in[i], in[n + i], vxl * n1 + vyl * n2, -vxl * n2 + vyl * n1};
  R = PrimD{in[4 * n + i], in[5 * n + i], vxr * n1 + vyr * n2, -vxr * n2 + vyr * n1};
}

// the flux back in lab axes, and the operator's return value: dt =
// krit*rad/UU (HLL also caps it at 1e7, kernel.cu:1442-1443)
template <int METHOD>
AST_DEV void store_face(double* __restrict__ out, size_t n, size_t i, double n1, double n2,
                        const FluxD& f, double krad, double UU) {
  out[i] = f.m;
  out[n + i] = f.e;
  out[2 * n + i] = f.un * n1 - f.ut * n2;
  out[3 * n + i] = f.un * n2 + f.ut * n1;
  const double t = krad / UU;
  out[4 * n + i] = METHOD == AST_HLL ? min2(1e7, t) : t;
}

template <int METHOD>
__global__ __launch_bounds__(FNT) void flux_kernel(size_t n, const double* __restrict__ in,
                                                   double n1, double n2, double krad, GasD gas,
                                                   double* __restrict__ out,
                                                   ast_stats_t* stats) {
  __shared__ uint32_t redi[(FNT / 64) * 10];
  LaneStats st;
  st.clear();
  const size_t stride = (size_t)gridDim.x * FNT;
  for (size_t i = (size_t)blockIdx.x * FNT + threadIdx.x; i < n; i += stride) {
    PrimD L, R;
    load_face(in, n, i, n1, n2, L, R);
    double UU;
    ExactInfo ei;
    const FluxD f = face_flux<METHOD>(L, R, gas, UU, ei);
    if (METHOD == AST_EXACT) st.add(ei);
    store_face<METHOD>(out, n, i, n1, n2, f, krad, UU);
  }
  if (METHOD == AST_EXACT && stats != nullptr) block_stats_to(st, redi, stats);
}

// The exact solver on pattern-sorted faces: each workgroup classifies a tile
// of STILE faces (pattern_key), orders them by pattern with a counting sort
// in LDS and solves them in that order, so that a wave's lanes mostly take
// the same wave pattern and iteration loop (the faces of one wave otherwise
// run the branches of every pattern one after another).  Each face's result
// depends only on its own data.
constexpr int SORT_PER = 4;
constexpr int STILE = FNT * SORT_PER;

__global__ __launch_bounds__(FNT) void exact_sorted_kernel(size_t n, const double* __restrict__ in,
                                                           double n1, double n2, double krad,
                                                           GasD gas, double* __restrict__ out,
                                                           ast_stats_t* stats) {
  __shared__ uint32_t redi[(FNT / 64) * 10];
  __shared__ unsigned short perm[STILE];
  __shared__ int cnt[kPatternKeys + 1], cur[kPatternKeys + 1];
  LaneStats st;
  st.clear();
  const size_t ntiles = (n + STILE - 1) / STILE;
  for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {  // workgroup-uniform
    const size_t base = tile * STILE;
    const int nv = (int)min((size_t)STILE, n - base);
    if (threadIdx.x <= kPatternKeys) cnt[threadIdx.x] = 0;
    __syncthreads();
    int key[SORT_PER];
#pragma unroll
    for (int j = 0; j < SORT_PER; j++) {
      const int li = j * FNT + threadIdx.x;
      key[j] = kPatternKeys;  // past the end of the batch
      if (li < nv) {
        PrimD L, R;
        load_face(in, n, base + li, n1, n2, L, R);
        key[j] = pattern_key(L, R, gas);
      }
      atomicAdd(&cnt[key[j]], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int s = 0;
      for (int k = 0; k <= kPatternKeys; k++) {
        cur[k] = s;
        s += cnt[k];
      }
      cnt[0] = cur[kPatternKeys];  // the tile's faces of the batch
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SORT_PER; j++) {
      const int slot = atomicAdd(&cur[key[j]], 1);
      perm[slot] = (unsigned short)(j * FNT + threadIdx.x);
    }
    __syncthreads();
    const int nfaces = cnt[0];
#pragma unroll 1
    for (int j = 0; j < SORT_PER; j++) {
      const int pos = j * FNT + threadIdx.x;
      if (pos < nfaces) {
        const size_t i = base + perm[pos];
        PrimD L, R;
        load_face(in, n, i, n1, n2, L, R);
        double UU;
        ExactInfo ei;
        const FluxD f = exact(L, R, gas, UU, ei);
        st.add(ei);
        store_face<AST_EXACT>(out, n, i, n1, n2, f, krad, UU);
      }
    }
    __syncthreads();  // perm and the counters are reused by the next tile
  }
  if (stats != nullptr) block_stats_to(st, redi, stats);
}

// The exact solver with per-face newton starting points and diagnostics.
__global__ __launch_bounds__(FNT) void guess_kernel(size_t n, const double* __restrict__ in,
                                                    double n1, double n2, double krad, GasD gas,
                                                    const float* __restrict__ guess,
                                                    double* __restrict__ out, ast_stats_t* stats,
                                                    int* __restrict__ info) {
  __shared__ uint32_t redi[(FNT / 64) * 10];
  LaneStats st;
  st.clear();
  const size_t stride = (size_t)gridDim.x * FNT;
  for (size_t i = (size_t)blockIdx.x * FNT + threadIdx.x; i < n; i += stride) {
    PrimD L, R;
    load_face(in, n, i, n1, n2, L, R);
    const double gv = guess ? (double)guess[i] : -1.0;
    double UU;
    ExactInfo ei;
    const FluxD f = exact(L, R, gas, UU, ei, gv);
    st.add(ei);
    store_face<AST_EXACT>(out, n, i, n1, n2, f, krad, UU);
    if (info) {
      info[2 * i] = ei.it;
      info[2 * i + 1] = (ei.pat == -9 ? -1 : ei.pat + 1) | (ei.capped ? 8 : 0);
    }
  }
  if (stats != nullptr) block_stats_to(st, redi, stats);
}

// HLLC with a passive scalar: 10 input planes (the 8 of load_face, Q_L,
// Q_R) and 6 output planes (the 5 of store_face, PQ); 128 B per face.
template <bool PER_SIDE>
__global__ __launch_bounds__(FNT) void flux_q_kernel(size_t n, const double* __restrict__ in,
                                                     double n1, double n2, double krad, GasD gas,
                                                     double* __restrict__ out) {
  const size_t stride = (size_t)gridDim.x * FNT;
  for (size_t i = (size_t)blockIdx.x * FNT + threadIdx.x; i < n; i += stride) {
    PrimD L, R;
    load_face(in, n, i, n1, n2, L, R);
    double UU, PQ;
    const FluxD f = hllc<PER_SIDE, true>(L, R, gas, UU, in[8 * n + i], in[9 * n + i], &PQ);
    store_face<AST_HLLC>(out, n, i, n1, n2, f, krad, UU);
    out[5 * n + i] = PQ;
  }
}

static unsigned flux_blocks(size_t n) {
  // enough workgroups to fill 256 CUs several times over; the loop covers
  // the rest
  size_t blocks = (n + FNT - 1) / FNT;
  if (blocks > 256 * 64) blocks = 256 * 64;
  return (unsigned)blocks;
}

template <int METHOD>
static int launch_flux(size_t n, const double* in, double n1, double n2, double krad,
                       const GasD& gas, double* out, ast_stats_t* stats, hipStream_t stream) {
  if constexpr (METHOD == AST_EXACT) {
    size_t blocks = (n + STILE - 1) / STILE;
    if (blocks > 256 * 16) blocks = 256 * 16;
    hipLaunchKernelGGL(exact_sorted_kernel, dim3((unsigned)blocks), dim3(FNT), 0, stream, n, in,
                       n1, n2, krad, gas, out, stats);
    AST_LAUNCH_CHECK("ast_flux_f64 launch");
    return AST_OK;
  } else {
    hipLaunchKernelGGL(flux_kernel<METHOD>, dim3(flux_blocks(n)), dim3(FNT), 0, stream, n, in,
                       n1, n2, krad, gas, out, stats);
    AST_LAUNCH_CHECK("ast_flux_f64 launch");
    return AST_OK;
  }
}

}  // namespace f64
}  // namespace ast

extern "C" int ast_flux_f64(int method, size_t n, const double* in, double n1, double n2,
                            double rad, double gamma, double krit, double* out,
                            ast_stats_t* stats, void* stream) {
  using namespace ast::f64;
  if (n == 0) return AST_OK;
  if (!in || !out) return ast_fail(AST_ERR_ARG, "ast_flux_f64: null in/out");
  if (!(gamma > 1.0)) return ast_fail(AST_ERR_ARG, "ast_flux_f64: gamma must be > 1");
  const GasD gas = ast::d64::make_gas(gamma);
  const double krad = krit * rad;
  hipStream_t st = (hipStream_t)stream;
  switch (method) {
    case AST_HLL: return launch_flux<AST_HLL>(n, in, n1, n2, krad, gas, out, stats, st);
    case AST_HLLC: return launch_flux<AST_HLLC>(n, in, n1, n2, krad, gas, out, stats, st);
    case AST_EXACT: return launch_flux<AST_EXACT>(n, in, n1, n2, krad, gas, out, stats, st);
    case AST_HLLC_A: return launch_flux<AST_HLLC_A>(n, in, n1, n2, krad, gas, out, stats, st);
    default: return ast_fail(AST_ERR_ARG, "ast_flux_f64: unknown method %d", method);
  }
}

extern "C" int ast_flux_f64_guess(size_t n, const double* in, double n1, double n2, double rad,
                                  double gamma, double krit, const float* guess, double* out,
                                  ast_stats_t* stats, int* info, void* stream) {
  using namespace ast::f64;
  if (n == 0) return AST_OK;
  if (!in || !out) return ast_fail(AST_ERR_ARG, "ast_flux_f64_guess: null in/out");
  if (!(gamma > 1.0)) return ast_fail(AST_ERR_ARG, "ast_flux_f64_guess: gamma must be > 1");
  const GasD gas = ast::d64::make_gas(gamma);
  hipLaunchKernelGGL(guess_kernel, dim3(flux_blocks(n)), dim3(FNT), 0, (hipStream_t)stream, n, in,
                     n1, n2, krit * rad, gas, guess, out, stats, info);
  AST_LAUNCH_CHECK("ast_flux_f64_guess launch");
  return AST_OK;
}

extern "C" int ast_flux_q_f64(int method, size_t n, const double* in, double n1, double n2,
                              double rad, double gamma, double krit, double* out, void* stream) {
  using namespace ast::f64;
  if (n == 0) return AST_OK;
  if (!in || !out) return ast_fail(AST_ERR_ARG, "ast_flux_q_f64: null in/out");
  if (!(gamma > 1.0)) return ast_fail(AST_ERR_ARG, "ast_flux_q_f64: gamma must be > 1");
  const GasD gas = ast::d64::make_gas(gamma);
  hipStream_t st = (hipStream_t)stream;
  if (method == AST_HLLC)
    hipLaunchKernelGGL(flux_q_kernel<false>, dim3(flux_blocks(n)), dim3(FNT), 0, st, n, in, n1,
                       n2, krit * rad, gas, out);
  else if (method == AST_HLLCQ || method == AST_HLLCQ_A)
    hipLaunchKernelGGL(flux_q_kernel<true>, dim3(flux_blocks(n)), dim3(FNT), 0, st, n, in, n1,
                       n2, krit * rad, gas, out);
  else
    return ast_fail(AST_ERR_ARG, "ast_flux_q_f64: method %d (AST_HLLC, AST_HLLCQ, AST_HLLCQ_A)",
                    method);
  AST_LAUNCH_CHECK("ast_flux_q_f64 launch");
  return AST_OK;
}
