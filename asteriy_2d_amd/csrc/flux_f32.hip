// flux_f32.hip -- batched FP32 face-flux operators over SoA face arrays:
// HLL, HLLC (both wave-speed estimates) and the exact solver (FP64 islands)
// of the fast path applied to
// n independent faces with a common normal (n1, n2).  52 B/face of
// algorithmic traffic (8 words in, 5 out).
#include <type_traits>

#include "common.hpp"
#include "exact_f32.hpp"
#include "fast_flux.hpp"

namespace ast {
namespace f32 {

constexpr int FNT = 256;

// STATS: the variant that counts the exact solver's work into *stats;
// LEARNED: the exact solver with the learned newton start compiled in
template <int METHOD, bool STATS, bool LEARNED>
__global__ __launch_bounds__(FNT) void flux_kernel(size_t n, const float* __restrict__ in,
                                                   float n1, float n2, float krad,
                                                   ExactCtx ex, float* __restrict__ out,
                                                   ast_stats_t* stats, int* __restrict__ info) {
  __shared__ uint32_t redi[(FNT / 64) * 10];
  typename std::conditional<STATS, StatsF32, NoStats>::type st;
  st.clear();
  // whole 64-face tiles per wave: every lane runs every iteration (EXEC full,
  // as the wave-level learned guess needs); lanes past n compute on a clamped
  // face and do not store
  const size_t stride = (size_t)gridDim.x * FNT;
  const size_t nt = (n + 63) / 64 * 64;
  const int lane = threadIdx.x & 63;
  const size_t base0 = (size_t)blockIdx.x * FNT + 64 * __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (size_t base = base0; base < nt; base += stride) {  // wave-uniform
    const size_t ii = base + lane;
    const bool valid = ii < n;
    const size_t i = valid ? ii : n - 1;
    const float rl = in[i], pl = in[n + i], vxl = in[2 * n + i], vyl = in[3 * n + i];
    const float rr = in[4 * n + i], pr = in[5 * n + i], vxr = in[6 * n + i], vyr = in[7 * n + i];
    // face frame: normal (n1, n2), tangent (-n2, n1)
    const PrimF L = {rl, pl, vxl * n1 + vyl * n2, -vxl * n2 + vyl * n1};
    const PrimF R = {rr, pr, vxr * n1 + vyr * n2, -vxr * n2 + vyr * n1};
    float UU;
    FluxF f;
    if constexpr (METHOD == AST_HLL) f = hll_f32(L, R, ex.g, ex.ig1, UU);
    else if constexpr (METHOD == AST_HLLC) f = hllc_f32<false>(L, R, ex.g, ex.ig1, UU);
    else if constexpr (METHOD == AST_HLLC_A) f = hllc_f32<true>(L, R, ex.g, ex.ig1, UU);
    else {
      FaceInfo fi;
      f = exact_f32<LEARNED>(L, R, ex, UU, fi);
      if (!valid) fi = FaceInfo{kNoFace, 0, false, false, false, false};
      st.add(fi);
      if (info && valid) {
        info[2 * i] = fi.it;
        info[2 * i + 1] = (fi.pat == kNoPattern ? -1 : fi.pat + 1) | (fi.capped ? 8 : 0) |
                          (fi.island ? 16 : 0);
      }
    }
    if (valid) {
      out[i] = f.m;
      out[n + i] = f.e;
      out[2 * n + i] = f.un * n1 - f.ut * n2;
      out[3 * n + i] = f.un * n2 + f.ut * n1;
      out[4 * n + i] = krad * frcp(UU);
    }
  }
  if constexpr (STATS) block_stats_f32_to<FNT>(st, redi, stats);
}

// ---- the exact solver on pattern-sorted faces ------------------------------
//
// Faces of one wave that take different wave patterns (and different
// iteration counts) run their branches one after another with most lanes
// masked.  Here each workgroup takes a tile of STILE faces, classifies them
// (pattern_key: vacuum, two rarefactions, shock + rarefaction, two shocks,
// strong two shocks, no pattern), orders the tile's faces by that key with a
// counting sort in LDS, and then solves them in that order, so that the 64
// lanes of a wave mostly hold faces of one pattern.  Results depend only on
// each face's own data: the order changes no bit.
//
// The tile's faces are kept in LDS from the classifying (coalesced) loads on:
// the solve reads its permuted faces there and leaves its fluxes there, and
// the tile's fluxes go out with coalesced stores (measured: re-reading the
// faces from global memory in sorted order, 32 scattered loads per lane and
// tile, took 60% of the kernel's time).
//
// With the learned newton start, a first kernel (kGuessPass) evaluates the
// network for the strong two-shock faces only -- they are contiguous after
// the sort, so its waves are full of faces that use the guess -- and leaves
// each guess in the face's dt slot of `out`; the solve kernel
// (kSolveGuessed) reads it there before it writes the dt.  The solver kernel
// thus carries no network registers, and the network's own kernel runs only
// on the faces it serves.
constexpr int SORT_NT = 256;
constexpr int SORT_PER = 2;                  // faces per thread per tile
constexpr int STILE = SORT_NT * SORT_PER;    // faces per tile
constexpr int kSolve = 0, kGuessPass = 1, kSolveGuessed = 2;

// The face's eight words (rhoL, pL, vxL, vyL, rhoR, pR, vxR, vyR) in the
// frame of the normal (n1, n2).
AST_DEV void face_frame(const float* w, float n1, float n2, PrimF& L, PrimF& R) {
  L = PrimF{w[0], w[1], w[2] * n1 + w[3] * n2, -w[2] * n2 + w[3] * n1};
  R = PrimF{w[4], w[5], w[6] * n1 + w[7] * n2, -w[6] * n2 + w[7] * n1};
}

template <bool STATS, int MODE>
__global__ __launch_bounds__(SORT_NT) void exact_sorted_kernel(size_t n,
                                                               const float* __restrict__ in,
                                                               float n1, float n2, float krad,
                                                               ExactCtx ex, float* out,
                                                               ast_stats_t* stats,
                                                               int* __restrict__ info) {
  constexpr bool STRONG = MODE != kSolve;  // strong faces form a group of their own
  __shared__ uint32_t redi[(SORT_NT / 64) * 10];
  __shared__ unsigned short perm[STILE];
  __shared__ int cnt[kPatternKeys + 1], cur[kPatternKeys + 1];
  __shared__ float stage[8][STILE];  // the tile's faces, then (5 planes) its fluxes
  typename std::conditional<STATS, StatsF32, NoStats>::type st;
  st.clear();
  const size_t ntiles = (n + STILE - 1) / STILE;
  for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {  // workgroup-uniform
    const size_t base = tile * STILE;
    const int nv = (int)min((size_t)STILE, n - base);
    if (threadIdx.x <= kPatternKeys) cnt[threadIdx.x] = 0;
    __syncthreads();
    // classify (coalesced loads, kept in LDS) and count
    int key[SORT_PER];
#pragma unroll
    for (int j = 0; j < SORT_PER; j++) {
      const int li = j * SORT_NT + threadIdx.x;
      key[j] = kPatternKeys;  // past the end of the batch
      if (li < nv) {
        float w[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          w[k] = in[k * n + base + li];
          stage[k][li] = w[k];
        }
        PrimF L, R;
        face_frame(w, n1, n2, L, R);
        key[j] = pattern_key(L, R, ex);
        if (!STRONG && key[j] == 4) key[j] = 3;  // no network: strong faces are plain ones
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
      cnt[0] = cur[4];             // first strong face
      cnt[1] = cur[5];             // one past the last strong face
      cnt[2] = cur[kPatternKeys];  // the tile's faces of the batch
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SORT_PER; j++) {
      const int slot = atomicAdd(&cur[key[j]], 1);
      perm[slot] = (unsigned short)(j * SORT_NT + threadIdx.x);
    }
    __syncthreads();
    const int s0 = cnt[0], s1 = cnt[1], nfaces = cnt[2];
    if constexpr (MODE == kGuessPass) {
      // the network on the strong faces, whole waves of them (EXEC full)
      const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
      for (int w0 = s0 + 64 * wave; w0 < s1; w0 += SORT_NT) {  // wave-uniform
        const int li = perm[min(w0 + lane, s1 - 1)];
        float w[8];
#pragma unroll
        for (int k = 0; k < 8; k++) w[k] = stage[k][li];
        PrimF L, R;
        face_frame(w, n1, n2, L, R);
        const bool sw = L.p > R.p;
        float fe[mlp::IN];
        const float lp = mlp::features<float>(sw ? R.r : L.r, sw ? R.p : L.p, sw ? -R.u : L.u,
                                              sw ? L.r : R.r, sw ? L.p : R.p, sw ? -L.u : R.u,
                                              ex.g, fe);
        const float y = mlp::forward_wave_bf16(ex.mlp, fe);
        if (w0 + lane < s1) out[4 * n + base + li] = __expf(y + lp);
      }
    } else {
      // solve in sorted order
#pragma unroll 1
      for (int j = 0; j < SORT_PER; j++) {  // uniform
        const int pos = j * SORT_NT + threadIdx.x;
        FaceInfo fi{kNoFace, 0, false, false, false, false};
        if (pos < nfaces) {
          const int li = perm[pos];
          const size_t i = base + li;
          float w[8];
#pragma unroll
          for (int k = 0; k < 8; k++) w[k] = stage[k][li];
          PrimF L, R;
          face_frame(w, n1, n2, L, R);
          const float g = (MODE == kSolveGuessed && pos >= s0 && pos < s1) ? out[4 * n + i] : -1.f;
          float UU;
          const FluxF f = exact_f32<false>(L, R, ex, UU, fi, g);
          // this lane alone holds face li: its words are read, its fluxes
          // take their place
          stage[0][li] = f.m;
          stage[1][li] = f.e;
          stage[2][li] = f.un * n1 - f.ut * n2;
          stage[3][li] = f.un * n2 + f.ut * n1;
          stage[4][li] = krad * frcp(UU);
          if (info) {
            info[2 * i] = fi.it;
            info[2 * i + 1] = (fi.pat == kNoPattern ? -1 : fi.pat + 1) | (fi.capped ? 8 : 0) |
                              (fi.island ? 16 : 0);
          }
        }
        st.add(fi);
      }
      __syncthreads();
      // the tile's fluxes, coalesced
#pragma unroll
      for (int j = 0; j < SORT_PER; j++) {
        const int li = j * SORT_NT + threadIdx.x;
        if (li < nv) {
#pragma unroll
          for (int k = 0; k < 5; k++) out[k * n + base + li] = stage[k][li];
        }
      }
    }
    __syncthreads();  // stage, perm and the counters are reused by the next tile
  }
  if constexpr (STATS) block_stats_f32_to<SORT_NT>(st, redi, stats);
}

template <int METHOD>
static int launch_unsorted(size_t n, const float* in, float n1, float n2, float krad,
                           const ExactCtx& ex, float* out, ast_stats_t* stats, int* info,
                           hipStream_t stream);

template <int METHOD>
static int launch(size_t n, const float* in, float n1, float n2, float krad, const ExactCtx& ex,
                  float* out, ast_stats_t* stats, int* info, hipStream_t stream) {
  if constexpr (METHOD == AST_EXACT) {
    // one tile per workgroup (measured on 2^24 faces: 512-face tiles, one per
    // workgroup, 346 us; 1024-face tiles 367-403 us; 4096 workgroups looping
    // over the tiles 389-427 us), a loop over tiles beyond 2^20 workgroups
    size_t blocks = (n + STILE - 1) / STILE;
    if (blocks > (size_t(1) << 20)) blocks = size_t(1) << 20;
    const dim3 gr((unsigned)blocks), bl(SORT_NT);
    const bool st = stats != nullptr;
    if (ex.mlp != nullptr) {
      hipLaunchKernelGGL((exact_sorted_kernel<false, kGuessPass>), gr, bl, 0, stream, n, in, n1,
                         n2, krad, ex, out, nullptr, nullptr);
      AST_LAUNCH_CHECK("ast_flux_f32 guess launch");
      if (st)
        hipLaunchKernelGGL((exact_sorted_kernel<true, kSolveGuessed>), gr, bl, 0, stream, n, in,
                           n1, n2, krad, ex, out, stats, info);
      else
        hipLaunchKernelGGL((exact_sorted_kernel<false, kSolveGuessed>), gr, bl, 0, stream, n, in,
                           n1, n2, krad, ex, out, stats, info);
    } else if (st) {
      hipLaunchKernelGGL((exact_sorted_kernel<true, kSolve>), gr, bl, 0, stream, n, in, n1, n2,
                         krad, ex, out, stats, info);
    } else {
      hipLaunchKernelGGL((exact_sorted_kernel<false, kSolve>), gr, bl, 0, stream, n, in, n1, n2,
                         krad, ex, out, stats, info);
    }
    AST_LAUNCH_CHECK("ast_flux_f32 launch");
    return AST_OK;
  } else {
    return launch_unsorted<METHOD>(n, in, n1, n2, krad, ex, out, stats, info, stream);
  }
}

template <int METHOD>
static int launch_unsorted(size_t n, const float* in, float n1, float n2, float krad,
                           const ExactCtx& ex, float* out, ast_stats_t* stats, int* info,
                           hipStream_t stream) {
  size_t blocks = (n + FNT - 1) / FNT;
  if (blocks > 256 * 64) blocks = 256 * 64;
  hipLaunchKernelGGL((flux_kernel<METHOD, false, false>), dim3((unsigned)blocks), dim3(FNT), 0,
                     stream, n, in, n1, n2, krad, ex, out, stats, info);
  AST_LAUNCH_CHECK("ast_flux_f32 launch");
  return AST_OK;
}

}  // namespace f32
}  // namespace ast

extern "C" int ast_flux_f32_info(int method, size_t n, const float* in, float n1, float n2,
                                 float rad, float gamma, float krit, const ast_mlp_t* mlp,
                                 float* out, ast_stats_t* stats, int* info, void* stream) {
  using namespace ast::f32;
  if (n == 0) return AST_OK;
  if (!in || !out) return ast_fail(AST_ERR_ARG, "ast_flux_f32: null in/out");
  if (!(gamma > 1.0f)) return ast_fail(AST_ERR_ARG, "ast_flux_f32: gamma must be > 1");
  const ast::ExactCtx ex = ast::make_exact_ctx((double)gamma, mlp);
  const float krad = krit * rad;
  hipStream_t st = (hipStream_t)stream;
  switch (method) {
    case AST_HLL: return launch<AST_HLL>(n, in, n1, n2, krad, ex, out, stats, nullptr, st);
    case AST_HLLC: return launch<AST_HLLC>(n, in, n1, n2, krad, ex, out, stats, nullptr, st);
    case AST_EXACT: return launch<AST_EXACT>(n, in, n1, n2, krad, ex, out, stats, info, st);
    case AST_HLLC_A: return launch<AST_HLLC_A>(n, in, n1, n2, krad, ex, out, stats, nullptr, st);
    default: return ast_fail(AST_ERR_ARG, "ast_flux_f32: unknown method %d", method);
  }
}

extern "C" int ast_flux_f32(int method, size_t n, const float* in, float n1, float n2, float rad,
                            float gamma, float krit, float* out, ast_stats_t* stats,
                            void* stream) {
  return ast_flux_f32_info(method, n, in, n1, n2, rad, gamma, krit, nullptr, out, stats, nullptr,
                           stream);
}
