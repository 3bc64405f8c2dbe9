// sim.cpp -- the host runtime around the step: the reference's time loop
// (kernel.cu:3868-3906: Kernel_TVD, funk_time, ping-pong by swapping the
// plane arguments) as a native object.
//
// The reference launches, checks and device-synchronises after every kernel
// (4 launches + 4 blocking syncs per 2 steps).  Here a run of kChunk steps
// (ping-pong A->B->A...) is captured once into a HIP graph on the
// simulation's own stream and replayed; dt never leaves the device (T, T_do,
// TT, i live in one device block, as in kernel.cu:3290-3293).  State is held
// in FP64 (the reference layout: each step is Kernel_TVD + funk_time) or FP32
// planes (each step is one kernel that also performs the previous step's time
// advance, time_slots.h; one small kernel closes the run).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <fstream>
#include <new>
#include <vector>

#include "../../include/asteriy.h"
#include "step_sources.h"
#include "time_slots.h"

int ast_fail(int code, const char* fmt, ...);
int ast_hip_fail(hipError_t e, const char* where);
int ast_step_f32_rows(const ast_grid_t* g, const float* s, const float* u, float* s2, float* u2,
                      double* T, const double* T_do, const double* TT, int method,
                      const ast_mlp_t* mlp, ast_stats_t* stats, int krows,
                      const ast::StepTime* tm, const ast::StepSources* src, void* stream);
int ast_funk_time_fold(double* T, int last, int nslots, uint32_t* flags, void* stream);
int ast_step_f64_rows(const ast_grid_t* g, const double* s, const double* u, double* s2,
                      double* u2, double* T, const double* T_do, const double* TT, int method,
                      ast_stats_t* stats, int rows, const ast_sources_t* src, void* stream);
extern "C" int ast_get_tile_rows(void);

#define CK(call)                                              \
  do {                                                        \
    hipError_t _e = (call);                                   \
    if (_e != hipSuccess) return ast_hip_fail(_e, #call);     \
  } while (0)

struct ast_sim {
  ast_grid_t g;
  int precision;  // 0: FP64 planes, 1: FP32 planes
  int method;
  int krows;   // FP32 strip tiling (rows per strip, or -S strips per band)
  int rows64;  // FP64 rows per strip (0: the kernel's default)
  size_t cells;
  void* plane[4];  // s0, u0, s1, u1
  int cur;         // 0: state in (s0,u0); 1: in (s1,u1)
  double* scal;    // time block: T, T_do, TT (+ int i in the 4th word), then
                   // the FP32 step's dt slots (time_slots.h)
  ast_stats_t* stats;
  bool collect;  // count the exact solver's work into stats (counting kernels)
  const ast_mlp_t* mlp;
  hipStream_t stream;
  hipGraph_t graph[3];      // a captured run of kChunk steps, per run kind
  hipGraphExec_t exec[3];
  bool graph_ok[3];
  // learned mode (FP32 exact solver with a learned guess): each run of steps
  // uses one solver kernel -- the one with the learned start, the plain one
  // that flags strong two-shock faces, or the plain one alone (kRunQuiet);
  // the kernel that closes run k writes whether it met strong faces to
  // hflags[k & 1] (pinned host memory), and run k + 2 takes that as its
  // mode, after waiting for run k's completion event -- the host never
  // waits for the run in flight
  uint32_t* hflags;         // pinned host, 2 words (fine-grained: kernel stores reach it)
  uint32_t* dflags;         // its device address
  hipEvent_t run_ev[2];     // completion of run k in run_ev[k & 1]
  long long runs;           // runs launched in learned mode
  int mode_init[2];         // modes of runs 0 and 1
  int mode_force;           // -1: automatic; 0 / 1: every run plain / learned
  // neutral-H source planes, device copies owned by the simulation (null
  // when not set): FP64 simulations keep the reference's planes (src.nn2,
  // src.nn3), FP32 ones their FP32 roundings (src32.xy, src32.e), which the
  // FP32 step reads with the same results
  ast_sources_t src;
  ast::StepSources src32;
};
typedef struct ast_sim ast_sim_t;

static size_t elem(const ast_sim_t* s) { return s->precision ? sizeof(float) : sizeof(double); }

// steps per captured graph (even: a replay ends on the plane it starts from)
constexpr int kChunk = 20;
// run kinds of the learned mode: plain solver that flags strong faces,
// learned start (flags them too), plain solver without the flag (its runs
// record "none"); a plain run flags only every kDetectEvery-th run, since
// the flag costs the plain kernel ~4% and strong faces are rare
constexpr int kRunPlain = 0, kRunLearned = 1, kRunQuiet = 2;
constexpr int kDetectEvery = 4;

static bool has_sources(const ast_sim_t* S) {
  return S->src.nn2 || S->src.nn3 || S->src32.xy || S->src32.e;
}
static const ast_sources_t* sources64(const ast_sim_t* S) {
  return S->src.nn2 || S->src.nn3 ? &S->src : nullptr;
}
static const ast::StepSources* sources32(const ast_sim_t* S) {
  return S->src32.xy || S->src32.e ? &S->src32 : nullptr;
}

// the learned mode applies: FP32 exact solver with a learned guess, no face
// window (the learned kernels need full waves) and no source planes
static bool learned_mode(const ast_sim_t* S) {
  return S->precision == 1 && S->method == AST_EXACT && S->mlp != nullptr &&
         S->g.ho_lo == -INFINITY && S->g.ho_hi == INFINITY && !has_sources(S);
}

static ast::StepTime step_time(ast_sim_t* S, int j, int mode) {
  double* T = S->scal;
  ast::StepTime tm;
  tm.nslots = ast::kMaxSlots;
  tm.prev = j > 0 ? ast::slot_set(T, j - 1) : nullptr;
  tm.cur = ast::slot_set(T, j);
  tm.rst = ast::slot_set(T, j + 1);
  tm.T = T;
  tm.strong = (mode == kRunPlain || mode == kRunLearned)
                  ? reinterpret_cast<uint32_t*>(T + ast::kGateWord)
                  : nullptr;
  tm.learned = mode == kRunLearned;
  return tm;
}

// n steps from plane `from`.  FP32: step j lowers slot set j (mod 3), takes
// its dt from set j-1 and advances the time; a fold of the last set closes
// the run, so between runs the time block reads as after funk_time.  mode:
// the run kind in the learned mode (-1: not in it).
static int launch_steps(ast_sim_t* S, int from, int n, int mode) {
  double* T = S->scal;
  for (int j = 0; j < n; j++) {
    const int f = from ^ (j & 1);
    void* s = S->plane[2 * f];
    void* u = S->plane[2 * f + 1];
    void* s2 = S->plane[2 * (1 - f)];
    void* u2 = S->plane[2 * (1 - f) + 1];
    int rc;
    if (S->precision == 0) {
      rc = ast_step_f64_rows(&S->g, (const double*)s, (const double*)u, (double*)s2,
                             (double*)u2, T, T + 1, T + 2, S->method,
                             S->collect ? S->stats : nullptr,
                             S->rows64, sources64(S), S->stream);
      if (rc == AST_OK) rc = ast_funk_time(T, T + 1, T + 2, (int*)(T + 3), S->stream);
    } else {
      const ast::StepTime tm = step_time(S, j, mode);
      rc = ast_step_f32_rows(&S->g, (const float*)s, (const float*)u, (float*)s2, (float*)u2, T,
                             T + 1, T + 2, S->method, mode == kRunQuiet ? nullptr : S->mlp,
                             S->collect ? S->stats : nullptr, S->krows, &tm, sources32(S),
                             S->stream);
    }
    if (rc != AST_OK) return rc;
  }
  if (S->precision == 1 && n > 0)
    return ast_funk_time_fold(T, n - 1, ast::kMaxSlots, mode >= 0 ? S->dflags : nullptr,
                              S->stream);
  return AST_OK;
}

// The slot sets as a run leaves them (every slot at funk_time's reset value).
static int clear_slot_sets(ast_sim_t* S) {
  std::vector<double> v(ast::kTimeBlockDoubles - ast::kSlot0, 0.0);
  for (int j = 0; j < ast::kSets; j++)
    for (int k = 0; k < ast::kMaxSlots; k++) v[j * ast::kSetDoubles + ast::kSlotStride * k] = ast::kSlotEmpty;
  CK(hipMemcpy(S->scal + ast::kSlot0, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  return AST_OK;
}

static void drop_graph(ast_sim_t* S) {
  for (int m = 0; m < 3; m++) {
    if (S->exec[m]) (void)hipGraphExecDestroy(S->exec[m]);
    if (S->graph[m]) (void)hipGraphDestroy(S->graph[m]);
    S->exec[m] = nullptr;
    S->graph[m] = nullptr;
    S->graph_ok[m] = false;
  }
}

extern "C" int ast_sim_create(const ast_grid_t* g, int precision, int method, ast_sim_t** out) {
  if (!out) return ast_fail(AST_ERR_ARG, "ast_sim_create: null out");
  *out = nullptr;
  int rc = ast_grid_check(g);
  if (rc != AST_OK) return rc;
  if (precision != 0 && precision != 1)
    return ast_fail(AST_ERR_ARG, "ast_sim_create: precision %d (0 = FP64, 1 = FP32)", precision);
  if (method < AST_HLL || method > AST_HLLC_A)
    return ast_fail(AST_ERR_ARG, "ast_sim_create: method %d", method);
  if (ast_device_count() < 1) return ast_fail(AST_ERR_NOGPU, "ast_sim_create: no HIP device");
  ast_sim_t* S = new (std::nothrow) ast_sim_t();
  if (!S) return ast_fail(AST_ERR_SIZE, "ast_sim_create: out of host memory");
  S->g = *g;
  S->precision = precision;
  S->method = method;
  S->krows = ast_get_tile_rows();
  S->cells = (size_t)g->N * g->M;
  const size_t bytes = S->cells * 2 * elem(S);
  hipError_t e = hipSuccess;
  for (int i = 0; i < 4 && e == hipSuccess; i++) e = hipMalloc(&S->plane[i], bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&S->scal, ast::kTimeBlockDoubles * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&S->stats, sizeof(ast_stats_t));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void**)&S->hflags, 2 * sizeof(uint32_t),
                                           hipHostMallocMapped | hipHostMallocCoherent);
  if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&S->dflags, S->hflags, 0);
  for (int k = 0; k < 2 && e == hipSuccess; k++)
    e = hipEventCreateWithFlags(&S->run_ev[k], hipEventDisableTiming);
  if (e != hipSuccess) {
    ast_sim_t* tmp = S;
    for (int i = 0; i < 4; i++) if (tmp->plane[i]) (void)hipFree(tmp->plane[i]);
    if (tmp->scal) (void)hipFree(tmp->scal);
    if (tmp->stats) (void)hipFree(tmp->stats);
    if (tmp->hflags) (void)hipHostFree(tmp->hflags);
    for (int k = 0; k < 2; k++) if (tmp->run_ev[k]) (void)hipEventDestroy(tmp->run_ev[k]);
    if (tmp->stream) (void)hipStreamDestroy(tmp->stream);
    delete tmp;
    return ast_hip_fail(e, "ast_sim_create: allocation");
  }
  S->hflags[0] = S->hflags[1] = 0u;
  S->mode_force = -1;
  // kernel.cu:3321-3324: T = 1e7, T_do = 1e-9, TT = 0, i = 0; dt slots 1e7
  double init[ast::kSlot0] = {1e7, 1e-9};
  CK(hipMemcpy(S->scal, init, sizeof init, hipMemcpyHostToDevice));
  if ((rc = clear_slot_sets(S)) != AST_OK) return rc;
  CK(hipMemset(S->stats, 0, sizeof(ast_stats_t)));
  *out = S;
  return AST_OK;
}

extern "C" int ast_sim_destroy(ast_sim_t* S) {
  if (!S) return AST_OK;
  (void)hipStreamSynchronize(S->stream);
  drop_graph(S);
  for (int i = 0; i < 4; i++) (void)hipFree(S->plane[i]);
  (void)hipFree(S->scal);
  (void)hipFree(S->stats);
  (void)hipHostFree(S->hflags);
  (void)hipFree(const_cast<double*>(S->src.nn2));
  (void)hipFree(const_cast<double*>(S->src.nn3));
  (void)hipFree(const_cast<float*>(S->src32.xy));
  (void)hipFree(const_cast<float*>(S->src32.e));
  for (int k = 0; k < 2; k++) (void)hipEventDestroy(S->run_ev[k]);
  (void)hipStreamDestroy(S->stream);
  delete S;
  return AST_OK;
}

// Neutral-H source planes from host memory (the reference fills nn1..nn3 on
// the host and copies them once, kernel.cu:3337-3339, 3453-3455): nn2 N*M
// triples, nn3 N*M doubles, either may be NULL; both NULL removes them.  An
// FP32 simulation keeps the planes' FP32 roundings (nn2's x and y, nn3): its
// step rounds each plane value to FP32 before using it, so they give the
// same results in 12 B per cell instead of 32.  The learned start is not
// used while planes are set.
extern "C" int ast_sim_set_sources(ast_sim_t* S, const double* nn2, const double* nn3,
                                   double kmom, double ken) {
  if (!S) return ast_fail(AST_ERR_ARG, "ast_sim_set_sources: null sim");
  CK(hipStreamSynchronize(S->stream));
  drop_graph(S);  // captured runs hold the old planes
  CK(hipFree(const_cast<double*>(S->src.nn2)));
  CK(hipFree(const_cast<double*>(S->src.nn3)));
  CK(hipFree(const_cast<float*>(S->src32.xy)));
  CK(hipFree(const_cast<float*>(S->src32.e)));
  S->src = ast_sources_t{nullptr, nullptr, kmom, ken};
  S->src32 = ast::StepSources{nullptr, nullptr, nullptr, nullptr, kmom, ken};
  const size_t n = S->cells;
  if (S->precision == 1) {
    std::vector<float> h;
    if (nn2) {
      h.resize(2 * n);
      for (size_t k = 0; k < n; k++) {
        h[2 * k] = (float)nn2[3 * k];
        h[2 * k + 1] = (float)nn2[3 * k + 1];
      }
      float* d = nullptr;
      CK(hipMalloc(&d, 2 * n * sizeof(float)));
      S->src32.xy = d;
      CK(hipMemcpy(d, h.data(), 2 * n * sizeof(float), hipMemcpyHostToDevice));
    }
    if (nn3) {
      h.resize(n);
      for (size_t k = 0; k < n; k++) h[k] = (float)nn3[k];
      float* d = nullptr;
      CK(hipMalloc(&d, n * sizeof(float)));
      S->src32.e = d;
      CK(hipMemcpy(d, h.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
    return AST_OK;
  }
  if (nn2) {
    double* d = nullptr;
    CK(hipMalloc(&d, n * 3 * sizeof(double)));
    S->src.nn2 = d;
    CK(hipMemcpy(d, nn2, n * 3 * sizeof(double), hipMemcpyHostToDevice));
  }
  if (nn3) {
    double* d = nullptr;
    CK(hipMalloc(&d, n * sizeof(double)));
    S->src.nn3 = d;
    CK(hipMemcpy(d, nn3, n * sizeof(double), hipMemcpyHostToDevice));
  }
  return AST_OK;
}

// Host state in the reference layout: s, u are N*M {x, y} pairs of doubles.
extern "C" int ast_sim_set_state(ast_sim_t* S, const double* s, const double* u) {
  if (!S || !s || !u) return ast_fail(AST_ERR_ARG, "ast_sim_set_state: null pointer");
  const size_t n = S->cells * 2;
  CK(hipStreamSynchronize(S->stream));
  S->cur = 0;
  if (S->precision == 0) {
    CK(hipMemcpy(S->plane[0], s, n * sizeof(double), hipMemcpyHostToDevice));
    CK(hipMemcpy(S->plane[1], u, n * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> a(n), b(n);
    for (size_t k = 0; k < n; k++) {
      a[k] = (float)s[k];
      b[k] = (float)u[k];
    }
    // the FP32 step needs positive densities and pressures
    for (size_t k = 0; k < n; k++)
      if (!(a[k] > 0.f))
        return ast_fail(AST_ERR_ARG, "ast_sim_set_state: cell %zu has %s %g (the FP32 step needs "
                        "rho > 0 and p > 0)", k / 2, (k & 1) ? "p" : "rho", (double)a[k]);
    CK(hipMemcpy(S->plane[0], a.data(), n * sizeof(float), hipMemcpyHostToDevice));
    CK(hipMemcpy(S->plane[1], b.data(), n * sizeof(float), hipMemcpyHostToDevice));
  }
  return AST_OK;
}

extern "C" int ast_sim_get_state(ast_sim_t* S, double* s, double* u) {
  if (!S || !s || !u) return ast_fail(AST_ERR_ARG, "ast_sim_get_state: null pointer");
  const size_t n = S->cells * 2;
  CK(hipStreamSynchronize(S->stream));
  void* ps = S->plane[2 * S->cur];
  void* pu = S->plane[2 * S->cur + 1];
  if (S->precision == 0) {
    CK(hipMemcpy(s, ps, n * sizeof(double), hipMemcpyDeviceToHost));
    CK(hipMemcpy(u, pu, n * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> a(n), b(n);
    CK(hipMemcpy(a.data(), ps, n * sizeof(float), hipMemcpyDeviceToHost));
    CK(hipMemcpy(b.data(), pu, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < n; k++) {
      s[k] = a[k];
      u[k] = b[k];
    }
  }
  return AST_OK;
}

// Device planes holding the current state (for callers that keep data on
// the GPU, e.g. error norms against another run).
extern "C" int ast_sim_device_state(ast_sim_t* S, void** s, void** u) {
  if (!S || !s || !u) return ast_fail(AST_ERR_ARG, "ast_sim_device_state: null pointer");
  *s = S->plane[2 * S->cur];
  *u = S->plane[2 * S->cur + 1];
  return AST_OK;
}

extern "C" int ast_sim_set_time(ast_sim_t* S, double T, double T_do, double TT, int i) {
  if (!S) return ast_fail(AST_ERR_ARG, "ast_sim_set_time: null sim");
  double v[4] = {T, T_do, TT, 0.0};
  memcpy(&v[3], &i, sizeof i);
  CK(hipStreamSynchronize(S->stream));
  CK(hipMemcpy(S->scal, v, sizeof v, hipMemcpyHostToDevice));
  return AST_OK;
}

extern "C" int ast_sim_get_time(ast_sim_t* S, double* T, double* T_do, double* TT, int* i) {
  if (!S) return ast_fail(AST_ERR_ARG, "ast_sim_get_time: null sim");
  double v[4];
  CK(hipStreamSynchronize(S->stream));
  CK(hipMemcpy(v, S->scal, sizeof v, hipMemcpyDeviceToHost));
  if (T) *T = v[0];
  if (T_do) *T_do = v[1];
  if (TT) *TT = v[2];
  if (i) memcpy(i, &v[3], sizeof *i);
  return AST_OK;
}

extern "C" int ast_sim_get_stats(ast_sim_t* S, ast_stats_t* out, int reset) {
  if (!S || !out) return ast_fail(AST_ERR_ARG, "ast_sim_get_stats: null pointer");
  CK(hipStreamSynchronize(S->stream));
  CK(hipMemcpy(out, S->stats, sizeof *out, hipMemcpyDeviceToHost));
  if (reset) CK(hipMemset(S->stats, 0, sizeof(ast_stats_t)));
  return AST_OK;
}

extern "C" int ast_sim_collect_stats(ast_sim_t* S, int on) {
  if (!S) return ast_fail(AST_ERR_ARG, "ast_sim_collect_stats: null sim");
  CK(hipStreamSynchronize(S->stream));
  if ((on != 0) != S->collect) {
    S->collect = on != 0;
    drop_graph(S);  // the captured step kernels change
  }
  return AST_OK;
}

extern "C" int ast_sim_set_mlp(ast_sim_t* S, const ast_mlp_t* mlp) {
  if (!S) return ast_fail(AST_ERR_ARG, "ast_sim_set_mlp: null sim");
  if (mlp && S->precision != 1)
    return ast_fail(AST_ERR_ARG, "ast_sim_set_mlp: the learned guess serves the FP32 step");
  CK(hipStreamSynchronize(S->stream));
  S->mlp = mlp;
  drop_graph(S);
  return AST_OK;
}

// Strip tiling of the FP32 step (its tile shape; see the autotuner): rows
// per wave strip, or -S for S strips of (nearly) equal height per band.
extern "C" int ast_sim_set_tile_rows(ast_sim_t* S, int krows) {
  if (!S) return ast_fail(AST_ERR_ARG, "ast_sim_set_tile_rows: null sim");
  if (S->precision == 0 ? (krows < 1 || krows > 4096)
                        : (krows == 0 || krows > 1024 || (krows < 0 && -(long long)krows > S->g.M)))
    return ast_fail(AST_ERR_ARG, "ast_sim_set_tile_rows: %d", krows);
  CK(hipStreamSynchronize(S->stream));
  if (S->precision == 0) S->rows64 = krows;
  else S->krows = krows;
  drop_graph(S);
  return AST_OK;
}

extern "C" void* ast_sim_stream(ast_sim_t* S) { return S ? (void*)S->stream : nullptr; }

// Tile autotuner of the FP32 step: times `reps` launches of the step kernel
// for each candidate strip tiling (rows per wave strip, or -S: S strips of
// (nearly) equal height per band) on the simulation's own
// state and stream, keeps the fastest, and restores the time scalars (the
// repeated launches read the current planes and only write the spare ones).
// best_rows / best_us (may be NULL) report the choice.  The FP64 step has a
// fixed tile; for it this only reports the step time.
extern "C" int ast_sim_autotune(ast_sim_t* S, const int* candidates, int ncand, int reps,
                                int* best_rows, float* best_us) {
  if (!S) return ast_fail(AST_ERR_ARG, "ast_sim_autotune: null sim");
  // default: strips of 4-16 rows, and the tilings that give every SIMD 2 to
  // 5 waves of (nearly) equal height, S strips per band with S * bands
  // waves (a wave strip has 60 output columns; a CU has 4 SIMDs); the FP64
  // step takes strips of 4 to 32 rows
  int dflt[8] = {4, 8, 12, 16};
  static const int dflt64[6] = {4, 6, 8, 12, 16, 32};
  if ((!candidates || ncand <= 0) && S->precision == 0) {
    candidates = dflt64;
    ncand = 6;
  }
  if (!candidates || ncand <= 0) {
    int nd = 4, dev = 0, cus = 0;
    CK(hipGetDevice(&dev));
    CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int nbands = (S->g.N + 59) / 60;
    for (int waves = 2; waves <= 5; waves++) {
      const int ns = waves * 4 * cus / nbands;
      if (ns >= 1 && S->g.M / ns >= 4) dflt[nd++] = -ns;
    }
    candidates = dflt;
    ncand = nd;
  }
  if (reps < 1) reps = 10;
  double saved[ast::kGateWord + 1];  // T, T_do, TT, i and the learned-mode words
  CK(hipStreamSynchronize(S->stream));
  CK(hipMemcpy(saved, S->scal, sizeof saved, hipMemcpyDeviceToHost));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const int from = S->cur;
  void* sp = S->plane[2 * from];
  void* up = S->plane[2 * from + 1];
  void* s2 = S->plane[2 * (1 - from)];
  void* u2 = S->plane[2 * (1 - from) + 1];
  int best = S->precision == 1 ? S->krows : S->rows64;
  float best_t = 1e30f;
  // the first step of a run (dt from T_do), in the kernel most runs take:
  // with the learned start attached, the plain solver without the strong-face
  // flag unless the learned start is forced on
  const int kind = learned_mode(S) ? (S->mode_force == 1 ? kRunLearned : kRunQuiet) : -1;
  const ast::StepTime tm = step_time(S, 0, kind);
  auto step = [&](int kr) {
    return S->precision == 1
               ? ast_step_f32_rows(&S->g, (const float*)sp, (const float*)up, (float*)s2,
                                   (float*)u2, S->scal, S->scal + 1, S->scal + 2, S->method,
                                   kind == kRunQuiet ? nullptr : S->mlp, nullptr, kr, &tm,
                                   sources32(S), S->stream)
               : ast_step_f64_rows(&S->g, (const double*)sp, (const double*)up, (double*)s2,
                                   (double*)u2, S->scal, S->scal + 1, S->scal + 2, S->method,
                                   nullptr, kr, sources64(S), S->stream);
  };
  auto valid = [&](int kr) {
    if (S->precision == 0) return kr >= 1 && kr <= 4096;
    return !(kr == 0 || kr > 1024 || (kr < 0 && -kr > S->g.M));
  };
  const int n = ncand;
  // warm-up: the clocks of an idle device ramp over the first tens of
  // milliseconds of work, which would slow the first candidates; untimed
  // launches of the first valid candidate until 20 ms have passed
  for (int c = 0; c < n; c++) {
    const int kr = candidates[c];
    if (!valid(kr)) continue;
    float ms = 0.f;
    CK(hipEventRecord(e0, S->stream));
    for (int launched = 0; ms < 20.f && launched < 4000; launched += 10) {
      for (int r = 0; r < 10; r++) {
        const int rc = step(kr);
        if (rc != AST_OK) return rc;
      }
      CK(hipEventRecord(e1, S->stream));
      CK(hipEventSynchronize(e1));
      CK(hipEventElapsedTime(&ms, e0, e1));
    }
    break;
  }
  for (int c = 0; c < n; c++) {
    const int kr = candidates[c];
    if (!valid(kr)) continue;
    for (int r = -2; r < reps; r++) {  // two untimed launches first
      if (r == 0) CK(hipEventRecord(e0, S->stream));
      const int rc = step(kr);
      if (rc != AST_OK) return rc;
    }
    CK(hipEventRecord(e1, S->stream));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const float us = 1000.f * ms / reps;
    if (us < best_t) {
      best_t = us;
      best = kr;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CK(hipMemcpy(S->scal, saved, sizeof saved, hipMemcpyHostToDevice));
  int rc = clear_slot_sets(S);
  if (rc != AST_OK) return rc;
  if (S->precision == 1 && best != S->krows) {
    S->krows = best;
    drop_graph(S);
  }
  if (S->precision == 0 && best != S->rows64) {
    S->rows64 = best;
    drop_graph(S);
  }
  if (best_rows) *best_rows = best;
  if (best_us) *best_us = best_t;
  return AST_OK;
}

static int build_graph(ast_sim_t* S, int mode) {
  const int m = mode < 0 ? 0 : mode;
  if (S->exec[m]) (void)hipGraphExecDestroy(S->exec[m]);
  if (S->graph[m]) (void)hipGraphDestroy(S->graph[m]);
  S->exec[m] = nullptr;
  S->graph[m] = nullptr;
  S->graph_ok[m] = false;
  CK(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal));
  int rc = launch_steps(S, 0, kChunk, mode);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(S->stream, &gr);
  if (rc != AST_OK) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess) return ast_hip_fail(e, "ast_sim_run: hipStreamEndCapture");
  S->graph[m] = gr;
  CK(hipGraphInstantiate(&S->exec[m], S->graph[m], nullptr, nullptr, 0));
  S->graph_ok[m] = true;
  return AST_OK;
}

// The kind of the next run (learned_mode): its own initial value for the
// first two runs, then the learned start if the run two before met strong
// faces (read after that run's completion event; the run in between stays
// queued), else the plain solver, flagging strong faces every kDetectEvery
// runs.  Forced plain runs do not flag.
static int next_run_mode(ast_sim_t* S, int* mode) {
  if (S->mode_force >= 0) {
    *mode = S->mode_force == 1 ? kRunLearned : kRunQuiet;
    return AST_OK;
  }
  const long long k = S->runs;
  if (k < 2) {
    *mode = S->mode_init[k] ? kRunLearned : kRunPlain;
    return AST_OK;
  }
  CK(hipEventSynchronize(S->run_ev[k & 1]));
  const bool strong = __atomic_load_n(&S->hflags[k & 1], __ATOMIC_ACQUIRE) != 0u;
  *mode = strong ? kRunLearned : (k % kDetectEvery == 0 ? kRunPlain : kRunQuiet);
  return AST_OK;
}

// One run of n steps from plane `from` (graph: a replay of the captured
// kChunk steps of its mode, from plane 0).
static int one_run(ast_sim_t* S, int from, int n, bool graph) {
  const bool lm = learned_mode(S);
  int mode = -1, rc;
  if (lm && (rc = next_run_mode(S, &mode)) != AST_OK) return rc;
  if (graph) {
    const int m = mode < 0 ? 0 : mode;
    if (!S->graph_ok[m] && (rc = build_graph(S, mode)) != AST_OK) return rc;
    CK(hipGraphLaunch(S->exec[m], S->stream));
  } else if ((rc = launch_steps(S, from, n, mode)) != AST_OK) {
    return rc;
  }
  if (lm) {
    CK(hipEventRecord(S->run_ev[S->runs & 1], S->stream));
    S->runs++;
  }
  return AST_OK;
}

// The learned modes of the next two runs (waits for every launched run).
extern "C" int ast_sim_learned_modes(ast_sim_t* S, int* modes) {
  if (!S || !modes) return ast_fail(AST_ERR_ARG, "ast_sim_learned_modes: null argument");
  CK(hipStreamSynchronize(S->stream));
  const long long k = S->runs;
  for (int d = 0; d < 2; d++)
    modes[d] = S->mode_force >= 0
                   ? S->mode_force
                   : (k + d < 2 ? S->mode_init[k + d] : (S->hflags[(k + d) & 1] ? 1 : 0));
  return AST_OK;
}

extern "C" int ast_sim_set_learned_mode(ast_sim_t* S, int mode) {
  if (!S || mode < -1 || mode > 1)
    return ast_fail(AST_ERR_ARG, "ast_sim_set_learned_mode: mode %d (-1, 0 or 1)", mode);
  S->mode_force = mode;
  return AST_OK;
}

// Advances `steps` steps asynchronously on the simulation stream.  With
// use_graph, runs of kChunk steps replay the captured graph (which starts
// from plane 0, so an odd position first takes one eager step); the rest
// are launched directly.  Both give the same result bit for bit.  In the
// learned mode the eager path runs the same runs (1 step from an odd
// position, then kChunk steps at a time, then the rest), so that each run
// takes the same mode in both.
extern "C" int ast_sim_run(ast_sim_t* S, int steps, int use_graph) {
  if (!S) return ast_fail(AST_ERR_ARG, "ast_sim_run: null sim");
  if (steps < 0) return ast_fail(AST_ERR_ARG, "ast_sim_run: steps %d", steps);
  int rc;
  if (!use_graph && !learned_mode(S)) {
    if ((rc = launch_steps(S, S->cur, steps, -1)) != AST_OK) return rc;
    S->cur ^= steps & 1;
    return AST_OK;
  }
  if (S->cur == 1 && steps > 0) {
    if ((rc = one_run(S, 1, 1, false)) != AST_OK) return rc;
    S->cur = 0;
    steps--;
  }
  for (; steps >= kChunk; steps -= kChunk)
    if ((rc = one_run(S, 0, kChunk, use_graph != 0)) != AST_OK) return rc;
  if (steps > 0) {
    if ((rc = one_run(S, 0, steps, false)) != AST_OK) return rc;
    S->cur = steps & 1;
  }
  return AST_OK;
}

extern "C" int ast_sim_sync(ast_sim_t* S) {
  if (!S) return ast_fail(AST_ERR_ARG, "ast_sim_sync: null sim");
  CK(hipStreamSynchronize(S->stream));
  return AST_OK;
}

// ---- on-disk formats ------------------------------------------------------

// The reference's binary state file (B19.bin read at kernel.cu:3414-3436,
// B31.bin written at kernel.cu:4111-4148): per cell in k order, 6 doubles
// x, y, rho, p, vx, vy.  It holds no time: a resume from it restarts with
// T_do = 1e-9 as the reference does.
extern "C" int ast_sim_write_b31(ast_sim_t* S, const char* path) {
  if (!S || !path) return ast_fail(AST_ERR_ARG, "ast_sim_write_b31: null argument");
  std::vector<double> s(S->cells * 2), u(S->cells * 2);
  int rc = ast_sim_get_state(S, s.data(), u.data());
  if (rc != AST_OK) return rc;
  FILE* f = fopen(path, "wb");
  if (!f) return ast_fail(AST_ERR_ARG, "ast_sim_write_b31: cannot open %s", path);
  for (size_t k = 0; k < S->cells; k++) {
    const int n = (int)(k % S->g.N), m = (int)(k / S->g.N);
    double rec[6] = {S->g.x_min + n * S->g.dx, S->g.y_min + m * S->g.dy, s[2 * k], s[2 * k + 1],
                     u[2 * k], u[2 * k + 1]};
    if (fwrite(rec, sizeof rec, 1, f) != 1) {
      fclose(f);
      return ast_fail(AST_ERR_ARG, "ast_sim_write_b31: write failed");
    }
  }
  fclose(f);
  return AST_OK;
}

// Tecplot ASCII file of the state in the form of the reference's final
// output (kernel.cu:4127-4129, 4170-4215): every stride-th node in x and y
// (the reference's Nmin = 2), FEPOINT with quadrilateral elements (1-based
// corner lists), variables X, Y, Ro, P, Vx, Vy, Mach, T = p/rho and Zav (the
// vorticity dvy/dx - dvx/dy by central differences, 0 on the boundary rows
// and columns), written with the default iostream formatting as there.
extern "C" int ast_sim_write_tecplot(ast_sim_t* S, const char* path, int stride) {
  if (!S || !path || stride < 1) return ast_fail(AST_ERR_ARG, "ast_sim_write_tecplot: bad argument");
  const int N = S->g.N, M = S->g.M;
  std::vector<double> s(S->cells * 2), u(S->cells * 2);
  int rc = ast_sim_get_state(S, s.data(), u.data());
  if (rc != AST_OK) return rc;
  std::ofstream out(path);
  if (!out) return ast_fail(AST_ERR_ARG, "ast_sim_write_tecplot: cannot open %s", path);
  const int nn = (N + stride - 1) / stride, mm = (M + stride - 1) / stride;
  out << "TITLE = \"HP\"  VARIABLES = \"X\", \"Y\", \"Ro\", \"P\", \"Vx\", \"Vy\", \"Mach\", "
         "\"T\", \"Zav\", ZONE T = \"HP\", N = "
      << nn * mm << " , E = " << (nn - 1) * (mm - 1) << ", F = FEPOINT, ET = quadrilateral\n";
  const double g = S->g.gamma, dx = S->g.dx, dy = S->g.dy;
  for (int m = 0; m < M; m += stride)
    for (int n = 0; n < N; n += stride) {
      const size_t k = (size_t)m * N + n;
      const double rho = s[2 * k], p = s[2 * k + 1], vx = u[2 * k], vy = u[2 * k + 1];
      double zav = 0.0;
      if (n > 0 && m > 0 && n < N - 1 && m < M - 1)
        zav = (u[2 * (k + 1) + 1] - u[2 * (k - 1) + 1]) / (2 * dx) -
              (u[2 * (k + N)] - u[2 * (k - N)]) / (2 * dy);
      double mach = 0.0, temp = 0.0;
      if (rho > 0.0) {
        mach = std::sqrt((vx * vx + vy * vy) / (g * p / rho));
        temp = p / rho;
      }
      out << S->g.x_min + n * dx << " " << S->g.y_min + m * dy << " " << rho << " " << p << " " << vx
          << " " << vy << " " << mach << " " << temp << " " << zav << "\n";
    }
  for (int m = 0; m + 1 < mm; m++)
    for (int n = 0; n + 1 < nn; n++)
      out << m * nn + n + 1 << " " << m * nn + n + 2 << " " << (m + 1) * nn + n + 2 << " "
          << (m + 1) * nn + n + 1 << "\n";
  out.close();
  if (!out) return ast_fail(AST_ERR_ARG, "ast_sim_write_tecplot: write failed");
  return AST_OK;
}

// The row next to the axis (m = 0) as text, the reference's "_param_y=0"
// file (kernel.cu:4151-4164): x y rho p vx vy 1/c per line.
extern "C" int ast_sim_write_axis(ast_sim_t* S, const char* path) {
  if (!S || !path) return ast_fail(AST_ERR_ARG, "ast_sim_write_axis: null argument");
  std::vector<double> s(S->cells * 2), u(S->cells * 2);
  int rc = ast_sim_get_state(S, s.data(), u.data());
  if (rc != AST_OK) return rc;
  std::ofstream out(path);
  if (!out) return ast_fail(AST_ERR_ARG, "ast_sim_write_axis: cannot open %s", path);
  for (int n = 0; n < S->g.N; n++) {
    const double rho = s[2 * n], p = s[2 * n + 1];
    const double ss = rho > 0 ? 1.0 / std::sqrt(S->g.gamma * p / rho) : 0.0;
    out << S->g.x_min + n * S->g.dx << " " << S->g.y_min << " " << rho << " " << p << " "
        << u[2 * n] << " " << u[2 * n + 1] << " " << ss << "\n";
  }
  out.close();
  if (!out) return ast_fail(AST_ERR_ARG, "ast_sim_write_axis: write failed");
  return AST_OK;
}

extern "C" int ast_sim_read_b31(ast_sim_t* S, const char* path) {
  if (!S || !path) return ast_fail(AST_ERR_ARG, "ast_sim_read_b31: null argument");
  FILE* f = fopen(path, "rb");
  if (!f) return ast_fail(AST_ERR_ARG, "ast_sim_read_b31: cannot open %s", path);
  std::vector<double> s(S->cells * 2), u(S->cells * 2);
  for (size_t k = 0; k < S->cells; k++) {
    double rec[6];
    if (fread(rec, sizeof rec, 1, f) != 1) {
      fclose(f);
      return ast_fail(AST_ERR_ARG, "ast_sim_read_b31: %s is shorter than N*M records", path);
    }
    s[2 * k] = rec[2];
    s[2 * k + 1] = rec[3];
    u[2 * k] = rec[4];
    u[2 * k + 1] = rec[5];
  }
  fclose(f);
  return ast_sim_set_state(S, s.data(), u.data());
}

// Checkpoint with the time state the reference does not save: header
// {magic, version, N, M, precision, i, T, T_do, TT} + raw device planes.
// A run resumed from it continues bit for bit.
struct CkptHeader {
  char magic[8];
  int32_t version, N, M, precision, i;
  uint32_t gate_mode;  // learned modes of the next two runs (bits 0 and 1)
  double T, T_do, TT;
};

extern "C" int ast_sim_save(ast_sim_t* S, const char* path) {
  if (!S || !path) return ast_fail(AST_ERR_ARG, "ast_sim_save: null argument");
  CkptHeader h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, "ASTCKPT", 8);
  h.version = 1;
  h.N = S->g.N;
  h.M = S->g.M;
  h.precision = S->precision;
  int rc = ast_sim_get_time(S, &h.T, &h.T_do, &h.TT, &h.i);
  if (rc != AST_OK) return rc;
  // every launched run has completed (ast_sim_get_time synchronised): the
  // modes of the next two runs are known
  {
    int m[2];
    if ((rc = ast_sim_learned_modes(S, m)) != AST_OK) return rc;
    h.gate_mode = (uint32_t)(m[0] | (m[1] << 1));
  }
  const size_t bytes = S->cells * 2 * elem(S);
  std::vector<char> a(bytes), b(bytes);
  CK(hipMemcpy(a.data(), S->plane[2 * S->cur], bytes, hipMemcpyDeviceToHost));
  CK(hipMemcpy(b.data(), S->plane[2 * S->cur + 1], bytes, hipMemcpyDeviceToHost));
  FILE* f = fopen(path, "wb");
  if (!f) return ast_fail(AST_ERR_ARG, "ast_sim_save: cannot open %s", path);
  bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(a.data(), 1, bytes, f) == bytes &&
            fwrite(b.data(), 1, bytes, f) == bytes;
  fclose(f);
  return ok ? AST_OK : ast_fail(AST_ERR_ARG, "ast_sim_save: write failed");
}

extern "C" int ast_sim_load(ast_sim_t* S, const char* path) {
  if (!S || !path) return ast_fail(AST_ERR_ARG, "ast_sim_load: null argument");
  FILE* f = fopen(path, "rb");
  if (!f) return ast_fail(AST_ERR_ARG, "ast_sim_load: cannot open %s", path);
  CkptHeader h;
  const size_t bytes = S->cells * 2 * elem(S);
  std::vector<char> a(bytes), b(bytes);
  bool ok = fread(&h, sizeof h, 1, f) == 1;
  if (ok && (memcmp(h.magic, "ASTCKPT", 8) != 0 || h.N != S->g.N || h.M != S->g.M ||
             h.precision != S->precision)) {
    fclose(f);
    return ast_fail(AST_ERR_ARG, "ast_sim_load: %s does not match this simulation", path);
  }
  ok = ok && fread(a.data(), 1, bytes, f) == bytes && fread(b.data(), 1, bytes, f) == bytes;
  fclose(f);
  if (!ok) return ast_fail(AST_ERR_ARG, "ast_sim_load: %s is truncated", path);
  if (S->precision == 1) {  // the FP32 step needs rho > 0 and p > 0 in every cell
    const float* sv = reinterpret_cast<const float*>(a.data());
    for (size_t k = 0; k < S->cells * 2; k++)
      if (!(sv[k] > 0.f))
        return ast_fail(AST_ERR_ARG, "ast_sim_load: %s: cell %zu has %s %g (the FP32 step needs "
                        "rho > 0 and p > 0)", path, k / 2, (k & 1) ? "p" : "rho", (double)sv[k]);
  }
  CK(hipStreamSynchronize(S->stream));
  S->cur = 0;
  CK(hipMemcpy(S->plane[0], a.data(), bytes, hipMemcpyHostToDevice));
  CK(hipMemcpy(S->plane[1], b.data(), bytes, hipMemcpyHostToDevice));
  const uint32_t words[2] = {0u, 0u};  // {strong, run}
  CK(hipMemcpy(S->scal + ast::kGateWord, words, sizeof words, hipMemcpyHostToDevice));
  S->mode_init[0] = (int)(h.gate_mode & 1u);
  S->mode_init[1] = (int)((h.gate_mode >> 1) & 1u);
  S->runs = 0;
  S->hflags[0] = S->hflags[1] = 0u;
  return ast_sim_set_time(S, h.T, h.T_do, h.TT, h.i);
}
