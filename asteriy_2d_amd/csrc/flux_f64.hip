// flux_f64.hip -- batched FP64 face-flux operators (HLL, HLLC, exact, HLLC_A) over
// SoA face arrays: the reference's per-face operators (kernel.cu:1399,
// Gas_Dinamic.cu:1814, kernel.cu:128) applied to n independent faces.
// Built with -ffp-contract=off for bit parity with the reference arithmetic.
//
// Memory: 8 input planes and 5 output planes, each lane one face per
// iteration of a grid-stride loop; every load/store is a coalesced 8-byte
// access of consecutive faces (104 B/face of algorithmic traffic).
#include "common.hpp"

namespace ast {
namespace f64 {

constexpr int FNT = 256;

template <int METHOD>
__global__ __launch_bounds__(FNT) void flux_kernel(size_t n, const double* __restrict__ in,
                                                   double n1, double n2, double rad,
                                                   Phys<double> ph, double* __restrict__ out,
                                                   ast_stats_t* stats) {
  __shared__ uint32_t redi[(FNT / 64) * 10];
  FaceStats st;
  st.clear();
  const size_t stride = (size_t)gridDim.x * FNT;
  for (size_t i = (size_t)blockIdx.x * FNT + threadIdx.x; i < n; i += stride) {
    V2<double> Ls = {in[i], in[n + i]};
    V2<double> Lu = {in[2 * n + i], in[3 * n + i]};
    V2<double> Rs = {in[4 * n + i], in[5 * n + i]};
    V2<double> Ru = {in[6 * n + i], in[7 * n + i]};
    V2<double> Ps, Pu;
    double t;
    if constexpr (METHOD == AST_HLL) t = HLL(ph, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
    else if constexpr (METHOD == AST_HLLC) t = HLLC(ph, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
    else if constexpr (METHOD == AST_HLLC_A) t = HLLC_A(ph, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
    else t = GODUNOV(ph, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad, st);
    out[i] = Ps.x;
    out[n + i] = Ps.y;
    out[2 * n + i] = Pu.x;
    out[3 * n + i] = Pu.y;
    out[4 * n + i] = t;
  }
  if (METHOD == AST_EXACT && stats != nullptr) block_stats_to<FNT>(st, redi, stats);
}

// The exact solver with per-face newton starting points and diagnostics.
__global__ __launch_bounds__(FNT) void guess_kernel(size_t n, const double* __restrict__ in,
                                                    double n1, double n2, double rad,
                                                    Phys<double> ph, const float* __restrict__ guess,
                                                    double* __restrict__ out, ast_stats_t* stats,
                                                    int* __restrict__ info) {
  __shared__ uint32_t redi[(FNT / 64) * 10];
  FaceStats st;
  st.clear();
  const size_t stride = (size_t)gridDim.x * FNT;
  for (size_t i = (size_t)blockIdx.x * FNT + threadIdx.x; i < n; i += stride) {
    V2<double> Ls = {in[i], in[n + i]};
    V2<double> Lu = {in[2 * n + i], in[3 * n + i]};
    V2<double> Rs = {in[4 * n + i], in[5 * n + i]};
    V2<double> Ru = {in[6 * n + i], in[7 * n + i]};
    V2<double> Ps, Pu;
    const float gv = guess ? guess[i] : 0.f;
    const bool use = gv > 0.f && gv < 3.0e38f;
    FaceStats fs;
    fs.clear();
    const double t = GODUNOV(ph, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad, fs, use, (double)gv);
    out[i] = Ps.x;
    out[n + i] = Ps.y;
    out[2 * n + i] = Pu.x;
    out[3 * n + i] = Pu.y;
    out[4 * n + i] = t;
    if (info) {
      const int pat = fs.vacuum ? -1 : fs.raref ? 0 : fs.bisect ? 1 : fs.newton ? 2 : -2;
      info[2 * i] = (int)(fs.newton_it + fs.bisect_it);
      info[2 * i + 1] = (pat + 1) | (fs.capped ? 8 : 0);
    }
    fs.mlp = (use && fs.newton) ? 1 : 0;
    st.exact += fs.exact; st.newton += fs.newton; st.newton_it += fs.newton_it;
    st.bisect += fs.bisect; st.bisect_it += fs.bisect_it; st.raref += fs.raref;
    st.vacuum += fs.vacuum; st.capped += fs.capped; st.mlp += fs.mlp;
  }
  if (stats != nullptr) block_stats_to<FNT>(st, redi, stats);
}

template <int METHOD>
static int launch_flux(size_t n, const double* in, double n1, double n2, double rad,
                       const Phys<double>& ph, double* out, ast_stats_t* stats,
                       hipStream_t stream) {
  // enough workgroups to fill 256 CUs several times over; the loop covers
  // the rest
  size_t blocks = (n + FNT - 1) / FNT;
  if (blocks > 256 * 64) blocks = 256 * 64;
  hipLaunchKernelGGL(flux_kernel<METHOD>, dim3((unsigned)blocks), dim3(FNT), 0, stream, n, in,
                     n1, n2, rad, ph, out, stats);
  AST_LAUNCH_CHECK("ast_flux_f64 launch");
  return AST_OK;
}

}  // namespace f64
}  // namespace ast

// host-side twin of make_phys (Header.h:17-26)
static ast::Phys<double> host_phys(double gamma, double krit) {
  ast::Phys<double> p;
  p.ga = gamma;
  p.g1 = gamma - 1.0;
  p.g2 = gamma + 1.0;
  p.gp = (p.g2 / gamma) / 2.0;
  p.gm = (p.g1 / gamma) / 2.0;
  p.krit = krit;
  return p;
}

extern "C" int ast_flux_f64(int method, size_t n, const double* in, double n1, double n2,
                            double rad, double gamma, double krit, double* out,
                            ast_stats_t* stats, void* stream) {
  using namespace ast::f64;
  if (n == 0) return AST_OK;
  if (!in || !out) return ast_fail(AST_ERR_ARG, "ast_flux_f64: null in/out");
  if (!(gamma > 1.0)) return ast_fail(AST_ERR_ARG, "ast_flux_f64: gamma must be > 1");
  const ast::Phys<double> ph = host_phys(gamma, krit);
  hipStream_t st = (hipStream_t)stream;
  switch (method) {
    case AST_HLL: return launch_flux<AST_HLL>(n, in, n1, n2, rad, ph, out, stats, st);
    case AST_HLLC: return launch_flux<AST_HLLC>(n, in, n1, n2, rad, ph, out, stats, st);
    case AST_EXACT: return launch_flux<AST_EXACT>(n, in, n1, n2, rad, ph, out, stats, st);
    case AST_HLLC_A: return launch_flux<AST_HLLC_A>(n, in, n1, n2, rad, ph, out, stats, st);
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
  const ast::Phys<double> ph = host_phys(gamma, krit);
  size_t blocks = (n + FNT - 1) / FNT;
  if (blocks > 256 * 64) blocks = 256 * 64;
  hipLaunchKernelGGL(guess_kernel, dim3((unsigned)blocks), dim3(FNT), 0, (hipStream_t)stream, n,
                     in, n1, n2, rad, ph, guess, out, stats, info);
  AST_LAUNCH_CHECK("ast_flux_f64_guess launch");
  return AST_OK;
}
