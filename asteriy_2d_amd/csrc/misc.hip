// misc.hip -- small kernels around the step: the time advance (funk_time),
// FP64 <-> FP32 state conversion, and compensated FP64 conservation sums.
#include <mutex>

#include "common.hpp"

namespace ast {

// funk_time, kernel.cu:1525-1536 (the progress print is left out).
__global__ void funk_time_kernel(double* T, double* T_do, double* TT, int* i) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *T_do = *T;
    *TT = *TT + *T_do;
    *T = 10000000;
    if (i) *i = *i + 1;
  }
}

// The time advance that closes a run of steps on the runtime's time block
// (time_slots.h): funk_time with T = the minimum of the slots of set `last`
// (the last step's), and every set cleared for the next run.  With `flags`
// (pinned host memory, 2 words), the run's strong-face flag goes to
// flags[run & 1] and is cleared, and the run is counted.  One wave; lane k
// folds and clears slot k.
__global__ void funk_time_fold_kernel(double* T, int last, int nslots, uint32_t* flags) {
  const int k = threadIdx.x;
  double* L = T + kSlot0 + kSetDoubles * last;
  double t = k < nslots ? L[kSlotStride * k] : INFINITY;
  t = wave_min(t);
  if (k == 0) {
    t = fmin(t, *T);  // a step launched without slots lowers T itself
    T[1] = t;
    T[2] = T[2] + t;
    T[0] = kSlotEmpty;
    int* i = reinterpret_cast<int*>(T + 3);
    *i = *i + 1;
    uint32_t* w = reinterpret_cast<uint32_t*>(T + kGateWord);  // {strong, run}
    if (flags != nullptr) {
      const uint32_t run = w[1];
      __atomic_store_n(&flags[run & 1u], w[0], __ATOMIC_RELAXED);
      __threadfence_system();
      w[0] = 0u;
      w[1] = run + 1u;
    }
  }
  for (int j = 0; j < kSets; j++) {
    double* S = T + kSlot0 + kSetDoubles * j;
    if (k < nslots) S[kSlotStride * k] = kSlotEmpty;
    if (k == 0) reinterpret_cast<uint32_t*>(S + 1)[0] = 0u;
  }
}

__global__ void f64_to_f32_kernel(size_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    b[k] = (float)a[k];
}
__global__ void f32_to_f64_kernel(size_t n, const float* __restrict__ a, double* __restrict__ b) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    b[k] = (double)a[k];
}

// Neumaier compensated accumulation.
struct KSum {
  double s, c;
  AST_DEV void add(double x) {
    double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  AST_DEV void merge(const KSum& o) {
    add(o.s);
    c += o.c;
  }
};

constexpr int SUM_BLOCKS = 256, SUM_NT = 256;

// Scratch of ast_conservation_sums_noalloc: the per-workgroup partial sums
// of pass 1 (4 doubles per workgroup, 8 KiB).  Each device holds its own
// copy until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns on that device (scratch_on_device).
constexpr int kScratchDoubles = 4 * SUM_BLOCKS;
static __device__ __align__(256) double scratch[kScratchDoubles];

// Pass 1: block b sums rows [b*M/B, (b+1)*M/B) in a fixed order and writes
// its (sum, compensation) pairs for mass and energy to work[4*b..4*b+3].
template <typename P>
__global__ __launch_bounds__(SUM_NT) void sums_pass1(ast_grid_t g, const P* __restrict__ s,
                                                     const P* __restrict__ u, double* work) {
  __shared__ double sh[4][SUM_NT];
  const int b = blockIdx.x;
  const long long m0 = (long long)g.M * b / SUM_BLOCKS, m1 = (long long)g.M * (b + 1) / SUM_BLOCKS;
  const double dV = g.dx * g.dy, g1 = g.gamma - 1.0;
  KSum ms = {0, 0}, es = {0, 0};
  for (long long m = m0; m < m1; m++) {
    const double w = (g.geometry == AST_AXISYMMETRIC) ? (g.y_min + m * g.dy) * dV : dV;
    for (int n = threadIdx.x; n < g.N; n += SUM_NT) {
      const size_t k = (size_t)m * g.N + n;
      const double r = s[k].x, p = s[k].y, vx = u[k].x, vy = u[k].y;
      ms.add(r * w);
      es.add((p / g1 + 0.5 * r * (vx * vx + vy * vy)) * w);
    }
  }
  sh[0][threadIdx.x] = ms.s;
  sh[1][threadIdx.x] = ms.c;
  sh[2][threadIdx.x] = es.s;
  sh[3][threadIdx.x] = es.c;
  __syncthreads();
  if (threadIdx.x == 0) {
    KSum a = {0, 0}, e = {0, 0};
    for (int t = 0; t < SUM_NT; t++) {
      a.merge(KSum{sh[0][t], sh[1][t]});
      e.merge(KSum{sh[2][t], sh[3][t]});
    }
    work[4 * b + 0] = a.s;
    work[4 * b + 1] = a.c;
    work[4 * b + 2] = e.s;
    work[4 * b + 3] = e.c;
  }
}

__global__ void sums_pass2(const double* work, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    KSum a = {0, 0}, e = {0, 0};
    for (int b = 0; b < SUM_BLOCKS; b++) {
      a.merge(KSum{work[4 * b + 0], work[4 * b + 1]});
      e.merge(KSum{work[4 * b + 2], work[4 * b + 3]});
    }
    out[0] = a.s + a.c;
    out[1] = e.s + e.c;
  }
}

}  // namespace ast

// The time advance closing a run whose last step lowered slot set `last`.
int ast_funk_time_fold(double* T, int last, int nslots, uint32_t* flags, void* stream) {
  if (!T || last < 0 || nslots < 1 || nslots > ast::kMaxSlots)
    return ast_fail(AST_ERR_ARG, "ast_funk_time_fold: bad argument");
  hipLaunchKernelGGL(ast::funk_time_fold_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, T,
                     last % ast::kSets, nslots, flags);
  AST_LAUNCH_CHECK("ast_funk_time_fold launch");
  return AST_OK;
}

extern "C" int ast_funk_time(double* T, double* T_do, double* TT, int* i, void* stream) {
  if (!T || !T_do || !TT) return ast_fail(AST_ERR_ARG, "ast_funk_time: null pointer");
  hipLaunchKernelGGL(ast::funk_time_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, T, T_do,
                     TT, i);
  AST_LAUNCH_CHECK("ast_funk_time launch");
  return AST_OK;
}

extern "C" int ast_convert_f64_to_f32(size_t n, const double* src, float* dst, void* stream) {
  if (n == 0) return AST_OK;
  if (!src || !dst) return ast_fail(AST_ERR_ARG, "ast_convert_f64_to_f32: null pointer");
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(ast::f64_to_f32_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, n, src, dst);
  AST_LAUNCH_CHECK("ast_convert_f64_to_f32 launch");
  return AST_OK;
}

extern "C" int ast_convert_f32_to_f64(size_t n, const float* src, double* dst, void* stream) {
  if (n == 0) return AST_OK;
  if (!src || !dst) return ast_fail(AST_ERR_ARG, "ast_convert_f32_to_f64: null pointer");
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(ast::f32_to_f64_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, n, src, dst);
  AST_LAUNCH_CHECK("ast_convert_f32_to_f64 launch");
  return AST_OK;
}

extern "C" size_t ast_conservation_work_doubles(void) { return 2 + 4 * ast::SUM_BLOCKS; }

// The device address of `scratch` on the current device, looked up once per
// device (the array's host-side name is not a device address).
static int scratch_on_device(double** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double* addr[kMaxDevices] = {};
  int dev = 0;
  AST_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return ast_fail(AST_ERR_SIZE, "ast_conservation_sums_noalloc: device %d beyond %d", dev,
                    kMaxDevices);
  std::lock_guard<std::mutex> lock(mu);
  if (addr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(ast::scratch));
    if (e != hipSuccess || p == nullptr)
      return ast_hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                          "ast_conservation_sums_noalloc: hipGetSymbolAddress(scratch)");
    addr[dev] = static_cast<double*>(p);
  }
  *out = addr[dev];
  return AST_OK;
}

extern "C" int ast_conservation_sums_noalloc(const ast_grid_t* g, const void* s, const void* u,
                                             int fp32, double* out, void* stream) {
  int rc = ast_grid_check(g);
  if (rc != AST_OK) return rc;
  if (!s || !u || !out)
    return ast_fail(AST_ERR_ARG, "ast_conservation_sums_noalloc: null pointer");
  // pass 1 writes 4 partials per workgroup, whatever the grid size
  static_assert(4 * ast::SUM_BLOCKS <= ast::kScratchDoubles, "scratch holds every partial");
  double* work = nullptr;
  if ((rc = scratch_on_device(&work)) != AST_OK) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (fp32)
    hipLaunchKernelGGL(ast::sums_pass1<float2>, dim3(ast::SUM_BLOCKS), dim3(ast::SUM_NT), 0, st,
                       *g, (const float2*)s, (const float2*)u, work);
  else
    hipLaunchKernelGGL(ast::sums_pass1<double2>, dim3(ast::SUM_BLOCKS), dim3(ast::SUM_NT), 0, st,
                       *g, (const double2*)s, (const double2*)u, work);
  AST_LAUNCH_CHECK("ast_conservation_sums_noalloc pass 1");
  hipLaunchKernelGGL(ast::sums_pass2, dim3(1), dim3(64), 0, st, work, out);
  AST_LAUNCH_CHECK("ast_conservation_sums_noalloc pass 2");
  return AST_OK;
}

extern "C" int ast_conservation_sums(const ast_grid_t* g, const void* s, const void* u, int fp32,
                                     double* out, void* stream) {
  int rc = ast_grid_check(g);
  if (rc != AST_OK) return rc;
  if (!s || !u || !out) return ast_fail(AST_ERR_ARG, "ast_conservation_sums: null pointer");
  hipStream_t st = (hipStream_t)stream;
  double* work = out + 2;
  if (fp32)
    hipLaunchKernelGGL(ast::sums_pass1<float2>, dim3(ast::SUM_BLOCKS), dim3(ast::SUM_NT), 0, st,
                       *g, (const float2*)s, (const float2*)u, work);
  else
    hipLaunchKernelGGL(ast::sums_pass1<double2>, dim3(ast::SUM_BLOCKS), dim3(ast::SUM_NT), 0, st,
                       *g, (const double2*)s, (const double2*)u, work);
  AST_LAUNCH_CHECK("ast_conservation_sums pass 1");
  hipLaunchKernelGGL(ast::sums_pass2, dim3(1), dim3(64), 0, st, work, out);
  AST_LAUNCH_CHECK("ast_conservation_sums pass 2");
  return AST_OK;
}
