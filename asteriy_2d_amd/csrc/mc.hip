// mc.hip -- the Monte-Carlo code's host helpers (Gas_Dinamic.cu:1468-1602)
// over device arrays, one particle (or one rate pair) per lane, from the
// __host__ __device__ functions of include/asteriy_mc.hpp.  Planes are SoA;
// every access is a coalesced 8-byte (4-byte for the integer outputs) load
// or store of consecutive items.
#include "common.hpp"

#include "../../include/asteriy_mc.hpp"

namespace ast {
namespace mck {

constexpr int MC_NT = 256;

static unsigned mc_blocks(size_t n) {
  size_t b = (n + MC_NT - 1) / MC_NT;
  if (b > (size_t(1) << 20)) b = size_t(1) << 20;
  return (unsigned)b;
}

__global__ __launch_bounds__(MC_NT) void rates_kernel(size_t n, const double* __restrict__ u,
                                                      const double* __restrict__ cp,
                                                      double* __restrict__ out) {
  const size_t stride = (size_t)gridDim.x * MC_NT;
  for (size_t i = (size_t)blockIdx.x * MC_NT + threadIdx.x; i < n; i += stride) {
    double v1, v2, v3;
    asteriy::mc::velosities(u[i], cp[i], v1, v2, v3);
    out[i] = v1;
    out[n + i] = v2;
    out[2 * n + i] = v3;
  }
}

__global__ __launch_bounds__(MC_NT) void cell_exit_kernel(asteriy::mc::McGrid g, size_t n,
                                                          const double* __restrict__ in,
                                                          double* __restrict__ t,
                                                          int* __restrict__ mode,
                                                          int* __restrict__ hit) {
  const size_t stride = (size_t)gridDim.x * MC_NT;
  for (size_t i = (size_t)blockIdx.x * MC_NT + threadIdx.x; i < n; i += stride) {
    int md;
    double tt;
    const bool h = asteriy::mc::peresechenie(g, in[i], in[n + i], in[2 * n + i], in[3 * n + i],
                                             in[4 * n + i], in[5 * n + i], in[6 * n + i],
                                             in[7 * n + i], md, tt);
    t[i] = tt;
    mode[i] = md;
    hit[i] = h ? 1 : 0;
  }
}

__global__ __launch_bounds__(MC_NT) void belong_kernel(asteriy::mc::McGrid g, size_t n,
                                                       const double* __restrict__ xy,
                                                       int* __restrict__ nm) {
  const size_t stride = (size_t)gridDim.x * MC_NT;
  for (size_t i = (size_t)blockIdx.x * MC_NT + threadIdx.x; i < n; i += stride) {
    int a, b;
    asteriy::mc::belong_point(g, xy[i], xy[n + i], a, b);
    nm[i] = a;
    nm[n + i] = b;
  }
}

}  // namespace mck
}  // namespace ast

extern "C" int ast_mc_rates_f64(size_t n, const double* u, const double* cp, double* out,
                                void* stream) {
  using namespace ast::mck;
  if (n == 0) return AST_OK;
  if (!u || !cp || !out) return ast_fail(AST_ERR_ARG, "ast_mc_rates_f64: null argument");
  hipLaunchKernelGGL(rates_kernel, dim3(mc_blocks(n)), dim3(MC_NT), 0, (hipStream_t)stream, n, u,
                     cp, out);
  AST_LAUNCH_CHECK("ast_mc_rates_f64 launch");
  return AST_OK;
}

extern "C" int ast_mc_cell_exit_f64(const ast_grid_t* g, size_t n, const double* in, double* t,
                                    int* mode, int* hit, void* stream) {
  using namespace ast::mck;
  if (!g) return ast_fail(AST_ERR_ARG, "ast_mc_cell_exit_f64: null grid");
  if (!(g->dx > 0.0) || !(g->dy > 0.0))
    return ast_fail(AST_ERR_ARG, "ast_mc_cell_exit_f64: dx, dy must be > 0");
  if (n == 0) return AST_OK;
  if (!in || !t || !mode || !hit) return ast_fail(AST_ERR_ARG, "ast_mc_cell_exit_f64: null array");
  hipLaunchKernelGGL(cell_exit_kernel, dim3(mc_blocks(n)), dim3(MC_NT), 0, (hipStream_t)stream,
                     asteriy::mc::make_grid(*g), n, in, t, mode, hit);
  AST_LAUNCH_CHECK("ast_mc_cell_exit_f64 launch");
  return AST_OK;
}

extern "C" int ast_mc_belong_point_f64(const ast_grid_t* g, size_t n, const double* xy, int* nm,
                                       void* stream) {
  using namespace ast::mck;
  if (!g) return ast_fail(AST_ERR_ARG, "ast_mc_belong_point_f64: null grid");
  if (!(g->dx > 0.0) || !(g->dy > 0.0))
    return ast_fail(AST_ERR_ARG, "ast_mc_belong_point_f64: dx, dy must be > 0");
  if (n == 0) return AST_OK;
  if (!xy || !nm) return ast_fail(AST_ERR_ARG, "ast_mc_belong_point_f64: null array");
  hipLaunchKernelGGL(belong_kernel, dim3(mc_blocks(n)), dim3(MC_NT), 0, (hipStream_t)stream,
                     asteriy::mc::make_grid(*g), n, xy, nm);
  AST_LAUNCH_CHECK("ast_mc_belong_point_f64 launch");
  return AST_OK;
}
