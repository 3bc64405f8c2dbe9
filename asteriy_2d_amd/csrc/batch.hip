// batch.hip -- many independent FP32 instances in one persistent launch.
//
// A parameter sweep (uneven grids, own physics and dt per instance) is one
// work table of (instance, band, strip group) items, in instance order.  The
// step kernel runs a fixed grid of workgroups (the device's co-resident
// capacity); workgroup b takes items b, b + G, b + 2G, ... (remapped within
// each round so that one XCD gets a contiguous run of items, which keeps
// neighbouring strips of an instance in one L2), runs the same wave strip as
// the single-grid step and lowers its instance's T.  A second kernel advances
// every instance's time (funk_time, kernel.cu:1525-1536, per instance).  The
// two-step ping-pong of both kernels is captured in one HIP graph.  Strip
// heights are per grid shape and chosen by timing the batch itself.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <new>
#include <utility>
#include <vector>

#include "strip_f32.hpp"

namespace ast {
namespace f32 {

struct BatchInst {
  ast_grid_t g;
  Consts c;
  ExactCtx ex;
  float2* plane[4];  // s0, u0, s1, u1
  double* scal;      // T, T_do, TT, i
  int nbands, krows, nsg, first_item;
};

template <int METHOD, int GEOM>
AST_DEV void batch_item(const BatchInst& I, int local, int from, double* red) {
  NoStats st;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int band, m0, m1;
  strip_rows(I.g.M, I.nbands, I.krows, local, wave, band, m0, m1);
  const double tmin = strip_of_wave<METHOD, GEOM, false>(
      I.g, I.c, I.ex, I.plane[2 * from], I.plane[2 * from + 1], I.plane[2 * (1 - from)],
      I.plane[2 * (1 - from) + 1], I.scal[1], band, m0, m1, st);
  block_min_uniform<NT>(tmin, red, I.scal, 0);
}

template <int METHOD>
__global__ __launch_bounds__(NT) void batch_step_kernel(const BatchInst* __restrict__ inst,
                                                        const int* __restrict__ item_inst,
                                                        int total, int from) {
  __shared__ double red[WPB];
  for (int base = 0; base < total; base += gridDim.x) {
    const int in_round = min((int)gridDim.x, total - base);
    if ((int)blockIdx.x >= in_round) break;  // workgroup-uniform
    const int item = base + xcd_tile(blockIdx.x, in_round);
    const int k = __builtin_amdgcn_readfirstlane(item_inst[item]);  // instance of the item
    const BatchInst& I = inst[k];
    const int local = item - I.first_item;
    if (I.g.geometry == AST_AXISYMMETRIC) batch_item<METHOD, AST_AXISYMMETRIC>(I, local, from, red);
    else batch_item<METHOD, AST_PLANAR>(I, local, from, red);
    __syncthreads();  // red[] is reused by the next item
  }
}

__global__ void batch_time_kernel(const BatchInst* __restrict__ inst, int count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) {  // funk_time per instance
    double* T = inst[k].scal;
    T[1] = T[0];
    T[2] = T[2] + T[1];
    T[0] = 10000000;
    int* i = (int*)(T + 3);
    *i = *i + 1;
  }
}

}  // namespace f32
}  // namespace ast

using namespace ast;
using namespace ast::f32;

struct ast_batch {
  int count, method, cur;
  std::vector<BatchInst> h;
  BatchInst* d;
  int* item_inst;  // instance of each work item
  int item_cap;
  float* pool;   // all planes
  double* scal;  // count x 4
  int total_items, grid_wgs;
  hipStream_t stream;
  hipGraph_t graph;
  hipGraphExec_t exec;
  bool graph_ok;
};

static void batch_drop_graph(ast_batch_t* B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
  B->graph_ok = false;
}

// lays out the work table from the per-instance strip heights and uploads
// the descriptors
static int batch_layout(ast_batch_t* B) {
  int item = 0;
  for (auto& I : B->h) {
    const StripShape sh = strip_shape(I.g, I.krows);
    I.nbands = sh.nbands;
    I.nsg = sh.nsg;
    I.first_item = item;
    item += sh.nbands * sh.nsg;
  }
  B->total_items = item;
  std::vector<int> map(item);
  for (int k = 0; k < B->count; k++) {
    const int n = B->h[k].nbands * B->h[k].nsg;
    for (int q = 0; q < n; q++) map[B->h[k].first_item + q] = k;
  }
  if (item > B->item_cap) {
    if (B->item_inst) AST_HIP_CHECK(hipFree(B->item_inst));
    B->item_inst = nullptr;
    AST_HIP_CHECK(hipMalloc((void**)&B->item_inst, sizeof(int) * item));
    B->item_cap = item;
  }
  AST_HIP_CHECK(hipMemcpy(B->item_inst, map.data(), sizeof(int) * item, hipMemcpyHostToDevice));
  AST_HIP_CHECK(hipMemcpy(B->d, B->h.data(), sizeof(BatchInst) * B->count, hipMemcpyHostToDevice));
  batch_drop_graph(B);
  return AST_OK;
}

template <int METHOD>
static int launch_batch(ast_batch_t* B, int from) {
  const int g = std::min(B->grid_wgs, B->total_items);
  hipLaunchKernelGGL(batch_step_kernel<METHOD>, dim3(g), dim3(NT), 0, B->stream, B->d,
                     B->item_inst, B->total_items, from);
  AST_LAUNCH_CHECK("ast_batch step launch");
  return AST_OK;
}

static int batch_step(ast_batch_t* B, int from, bool advance_time) {
  int rc;
  switch (B->method) {
    case AST_HLL: rc = launch_batch<AST_HLL>(B, from); break;
    case AST_HLLC: rc = launch_batch<AST_HLLC>(B, from); break;
    case AST_HLLC_A: rc = launch_batch<AST_HLLC_A>(B, from); break;
    default: rc = launch_batch<AST_EXACT>(B, from); break;
  }
  if (rc != AST_OK || !advance_time) return rc;
  hipLaunchKernelGGL(batch_time_kernel, dim3((B->count + 63) / 64), dim3(64), 0, B->stream, B->d,
                     B->count);
  AST_LAUNCH_CHECK("ast_batch time launch");
  return AST_OK;
}

extern "C" int ast_batch_create(int count, const ast_grid_t* grids, int method, ast_batch_t** out) {
  if (!out || !grids || count < 1) return ast_fail(AST_ERR_ARG, "ast_batch_create: bad argument");
  *out = nullptr;
  if (method < AST_HLL || method > AST_HLLC_A)
    return ast_fail(AST_ERR_ARG, "ast_batch_create: method %d", method);
  size_t floats = 0;
  for (int k = 0; k < count; k++) {
    int rc = ast_grid_check(&grids[k]);
    if (rc != AST_OK) return rc;
    if (!(grids[k].ho_lo == -INFINITY && grids[k].ho_hi == INFINITY))
      return ast_fail(AST_ERR_ARG, "ast_batch_create: instance %d: the batch reconstructs every face "
                      "(ho_lo = -inf, ho_hi = +inf)", k);
    if ((rc = check_positive_bc(grids[k])) != AST_OK) return rc;
    for (int j = 0; j < 4; j++)
      if (grids[k].bc[j] == AST_BC_DIRICHLET_PULSED)
        return ast_fail(AST_ERR_ARG, "ast_batch_create: instance %d: the batch has no simulation "
                        "time for the pulsed inflow of side %d", k, j);
    if (!strip_fits(grids[k], 16))
      return ast_fail(AST_ERR_SIZE, "ast_batch_create: instance %d: %d columns exceed one buffer range",
                      k, grids[k].N);
    floats += (size_t)grids[k].N * grids[k].M * 2 * 4;
  }
  if (ast_device_count() < 1) return ast_fail(AST_ERR_NOGPU, "ast_batch_create: no HIP device");
  ast_batch_t* B = new (std::nothrow) ast_batch_t();
  if (!B) return ast_fail(AST_ERR_SIZE, "ast_batch_create: out of host memory");
  B->count = count;
  B->method = method;
  B->h.resize(count);
  hipError_t e = hipMalloc((void**)&B->pool, floats * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&B->scal, sizeof(double) * 4 * count);
  if (e == hipSuccess) e = hipMalloc((void**)&B->d, sizeof(BatchInst) * count);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    if (B->pool) (void)hipFree(B->pool);
    if (B->scal) (void)hipFree(B->scal);
    if (B->d) (void)hipFree(B->d);
    delete B;
    return ast_hip_fail(e, "ast_batch_create: allocation");
  }
  std::vector<double> init(4 * count, 0.0);
  size_t off = 0;
  for (int k = 0; k < count; k++) {
    BatchInst& I = B->h[k];
    I.g = grids[k];
    I.c = make_consts(grids[k]);
    I.ex = make_exact_ctx(grids[k].gamma, nullptr);
    const size_t pl = (size_t)grids[k].N * grids[k].M * 2;
    for (int q = 0; q < 4; q++) {
      I.plane[q] = (float2*)(B->pool + off);
      off += pl;
    }
    I.scal = B->scal + 4 * k;
    I.krows = 8;
    init[4 * k + 0] = 1e7;  // kernel.cu:3321-3324
    init[4 * k + 1] = 1e-9;
    init[4 * k + 2] = 0.0;
    init[4 * k + 3] = 0.0;
  }
  AST_HIP_CHECK(hipMemcpy(B->scal, init.data(), sizeof(double) * init.size(), hipMemcpyHostToDevice));
  // co-resident capacity: workgroups per CU x CUs
  int dev = 0, cus = 256, per_cu = 1;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
  hipError_t oe;
  switch (method) {
    case AST_HLL: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_step_kernel<AST_HLL>, NT, 0); break;
    case AST_HLLC: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_step_kernel<AST_HLLC>, NT, 0); break;
    case AST_HLLC_A: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_step_kernel<AST_HLLC_A>, NT, 0); break;
    default: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_step_kernel<AST_EXACT>, NT, 0); break;
  }
  if (oe != hipSuccess || per_cu < 1) per_cu = 1;
  B->grid_wgs = cus * per_cu;
  int rc = batch_layout(B);
  if (rc != AST_OK) {
    ast_batch_destroy(B);
    return rc;
  }
  *out = B;
  return AST_OK;
}

extern "C" int ast_batch_destroy(ast_batch_t* B) {
  if (!B) return AST_OK;
  (void)hipStreamSynchronize(B->stream);
  batch_drop_graph(B);
  (void)hipFree(B->pool);
  (void)hipFree(B->scal);
  (void)hipFree(B->d);
  if (B->item_inst) (void)hipFree(B->item_inst);
  (void)hipStreamDestroy(B->stream);
  delete B;
  return AST_OK;
}

static int batch_check(ast_batch_t* B, int k, const char* fn) {
  if (!B) return ast_fail(AST_ERR_ARG, "%s: null batch", fn);
  if (k < 0 || k >= B->count) return ast_fail(AST_ERR_ARG, "%s: instance %d out of range", fn, k);
  return AST_OK;
}

extern "C" int ast_batch_set_state(ast_batch_t* B, int k, const double* s, const double* u) {
  int rc = batch_check(B, k, "ast_batch_set_state");
  if (rc != AST_OK) return rc;
  if (!s || !u) return ast_fail(AST_ERR_ARG, "ast_batch_set_state: null state");
  AST_HIP_CHECK(hipStreamSynchronize(B->stream));
  const BatchInst& I = B->h[k];
  const size_t n = (size_t)I.g.N * I.g.M * 2;
  std::vector<float> a(n), b(n);
  for (size_t q = 0; q < n; q++) {
    a[q] = (float)s[q];
    b[q] = (float)u[q];
  }
  // the FP32 strip reconstructs without the positivity fallback: every cell
  // needs rho > 0 and p > 0 (as ast_sim_set_state checks)
  for (size_t q = 0; q < n; q++)
    if (!(a[q] > 0.f))
      return ast_fail(AST_ERR_ARG, "ast_batch_set_state: instance %d cell %zu has %s %g (the FP32 "
                      "step needs rho > 0 and p > 0)", k, q / 2, (q & 1) ? "p" : "rho", (double)a[q]);
  AST_HIP_CHECK(hipMemcpy(I.plane[2 * B->cur], a.data(), n * sizeof(float), hipMemcpyHostToDevice));
  AST_HIP_CHECK(hipMemcpy(I.plane[2 * B->cur + 1], b.data(), n * sizeof(float), hipMemcpyHostToDevice));
  return AST_OK;
}

extern "C" int ast_batch_get_state(ast_batch_t* B, int k, double* s, double* u) {
  int rc = batch_check(B, k, "ast_batch_get_state");
  if (rc != AST_OK) return rc;
  if (!s || !u) return ast_fail(AST_ERR_ARG, "ast_batch_get_state: null state");
  AST_HIP_CHECK(hipStreamSynchronize(B->stream));
  const BatchInst& I = B->h[k];
  const size_t n = (size_t)I.g.N * I.g.M * 2;
  std::vector<float> a(n), b(n);
  AST_HIP_CHECK(hipMemcpy(a.data(), I.plane[2 * B->cur], n * sizeof(float), hipMemcpyDeviceToHost));
  AST_HIP_CHECK(hipMemcpy(b.data(), I.plane[2 * B->cur + 1], n * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t q = 0; q < n; q++) {
    s[q] = a[q];
    u[q] = b[q];
  }
  return AST_OK;
}

extern "C" int ast_batch_get_time(ast_batch_t* B, int k, double* T, double* T_do, double* TT,
                                  int* i) {
  int rc = batch_check(B, k, "ast_batch_get_time");
  if (rc != AST_OK) return rc;
  double v[4];
  AST_HIP_CHECK(hipStreamSynchronize(B->stream));
  AST_HIP_CHECK(hipMemcpy(v, B->h[k].scal, sizeof v, hipMemcpyDeviceToHost));
  if (T) *T = v[0];
  if (T_do) *T_do = v[1];
  if (TT) *TT = v[2];
  if (i) memcpy(i, &v[3], sizeof *i);
  return AST_OK;
}

extern "C" int ast_batch_set_time(ast_batch_t* B, int k, double T, double T_do, double TT, int i) {
  int rc = batch_check(B, k, "ast_batch_set_time");
  if (rc != AST_OK) return rc;
  double v[4] = {T, T_do, TT, 0.0};
  memcpy(&v[3], &i, sizeof i);
  AST_HIP_CHECK(hipStreamSynchronize(B->stream));
  AST_HIP_CHECK(hipMemcpy(B->h[k].scal, v, sizeof v, hipMemcpyHostToDevice));
  return AST_OK;
}

// Per grid shape, tries each candidate strip height while the other shapes
// keep theirs, timing `reps` launches of the whole batch step kernel (no
// time advance: the launches read the current planes and write the spare
// ones); keeps the fastest.  Time scalars are restored.
extern "C" int ast_batch_autotune(ast_batch_t* B, int reps) {
  if (!B) return ast_fail(AST_ERR_ARG, "ast_batch_autotune: null batch");
  if (reps < 1) reps = 5;
  static const int cand[] = {4, 8, 12, 16};  // multiples of 4 take the fast strip path
  std::vector<double> saved(4 * B->count);
  AST_HIP_CHECK(hipStreamSynchronize(B->stream));
  AST_HIP_CHECK(hipMemcpy(saved.data(), B->scal, sizeof(double) * saved.size(), hipMemcpyDeviceToHost));
  std::map<std::pair<int, int>, std::vector<int>> shapes;
  for (int k = 0; k < B->count; k++) shapes[{B->h[k].g.N, B->h[k].g.M}].push_back(k);
  hipEvent_t e0, e1;
  AST_HIP_CHECK(hipEventCreate(&e0));
  AST_HIP_CHECK(hipEventCreate(&e1));
  for (auto& sh : shapes) {
    int best = B->h[sh.second[0]].krows;
    float best_t = 1e30f;
    for (int kr : cand) {
      for (int k : sh.second) B->h[k].krows = kr;
      int rc = batch_layout(B);
      if (rc != AST_OK) return rc;
      for (int r = -1; r < reps; r++) {
        if (r == 0) AST_HIP_CHECK(hipEventRecord(e0, B->stream));
        if ((rc = batch_step(B, B->cur, false)) != AST_OK) return rc;
      }
      AST_HIP_CHECK(hipEventRecord(e1, B->stream));
      AST_HIP_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      AST_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best_t) {
        best_t = ms;
        best = kr;
      }
    }
    for (int k : sh.second) B->h[k].krows = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int rc = batch_layout(B);
  if (rc != AST_OK) return rc;
  AST_HIP_CHECK(hipMemcpy(B->scal, saved.data(), sizeof(double) * saved.size(), hipMemcpyHostToDevice));
  return AST_OK;
}

static int batch_build_graph(ast_batch_t* B) {
  batch_drop_graph(B);
  AST_HIP_CHECK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
  int rc = batch_step(B, 0, true);
  if (rc == AST_OK) rc = batch_step(B, 1, true);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(B->stream, &gr);
  if (rc != AST_OK) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess) return ast_hip_fail(e, "ast_batch_run: hipStreamEndCapture");
  B->graph = gr;
  AST_HIP_CHECK(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
  B->graph_ok = true;
  return AST_OK;
}

extern "C" int ast_batch_run(ast_batch_t* B, int steps, int use_graph) {
  if (!B) return ast_fail(AST_ERR_ARG, "ast_batch_run: null batch");
  if (steps < 0) return ast_fail(AST_ERR_ARG, "ast_batch_run: steps %d", steps);
  int rc;
  if (!use_graph) {
    for (int k = 0; k < steps; k++) {
      if ((rc = batch_step(B, B->cur, true)) != AST_OK) return rc;
      B->cur ^= 1;
    }
    return AST_OK;
  }
  if (B->cur == 1 && steps > 0) {
    if ((rc = batch_step(B, 1, true)) != AST_OK) return rc;
    B->cur = 0;
    steps--;
  }
  if (steps >= 2 && !B->graph_ok)
    if ((rc = batch_build_graph(B)) != AST_OK) return rc;
  for (int k = 0; k + 1 < steps; k += 2) AST_HIP_CHECK(hipGraphLaunch(B->exec, B->stream));
  if (steps & 1) {
    if ((rc = batch_step(B, 0, true)) != AST_OK) return rc;
    B->cur = 1;
  }
  return AST_OK;
}

extern "C" int ast_batch_sync(ast_batch_t* B) {
  if (!B) return ast_fail(AST_ERR_ARG, "ast_batch_sync: null batch");
  AST_HIP_CHECK(hipStreamSynchronize(B->stream));
  return AST_OK;
}

extern "C" void* ast_batch_stream(ast_batch_t* B) { return B ? (void*)B->stream : nullptr; }

extern "C" long long ast_batch_cells(const ast_batch_t* B) {
  if (!B) return 0;
  long long c = 0;
  for (const auto& I : B->h) c += (long long)I.g.N * I.g.M;
  return c;
}
