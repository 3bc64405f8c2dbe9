// driver.cpp -- asteriy_run: the reference's host program (main(),
// kernel.cu:3253-4224) as a native executable over the C-ABI.
//
// It sets up one of the reference's problems, runs the step/funk_time loop
// through the graph-replaying runtime (ast_sim_*), prints progress in the
// reference's form ("i = .., TT = .., dt = ..", kernel.cu:1531-1534, and the
// wall time of a block of steps, kernel.cu:3965-3968) and writes the final
// state in the reference's binary format (B31.bin, kernel.cu:4111-4148) and,
// on request, its text outputs (Tecplot file of every 2nd node, the axis row,
// kernel.cu:4127-4215).
//
//   asteriy_run --problem wind|sod|live [--N n --M m] [--method hll|hllc|exact|hllc_a]
//               [--precision f64|f32] [--steps k] [--print-every k]
//               [--out B31.bin] [--tecplot file] [--axis file]
//               [--checkpoint file] [--resume file] [--mlp params.f32] [--autotune]
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "../../include/asteriy.h"

namespace {

struct Opts {
  std::string problem = "wind", method = "hllc", precision = "f64", out, ckpt, resume, mlp, tec,
      axis;
  int N = 0, M = 0, steps = 200, every = 5000;
  bool autotune = false;
};

int die(const char* what) {
  fprintf(stderr, "asteriy_run: %s: %s\n", what, ast_last_error());
  return 1;
}

void usage() {
  printf(
      "usage: asteriy_run --problem wind|sod|live [--N n --M m] [--method hll|hllc|exact|hllc_a]\n"
      "                   [--precision f64|f32] [--steps k] [--print-every k] [--out B31.bin]\n"
      "                   [--tecplot file] [--axis file] [--checkpoint file] [--resume file]\n"
      "                   [--mlp params.f32] [--autotune]\n"
      "  wind : stellar wind / counter-flow, z in [-400,400], r in [0,400], gamma 5/3, CFL 0.3\n"
      "  sod  : axisymmetric Sod tube, z in [0,1], r in [0,0.5], gamma 1.4, CFL 0.4\n"
      "  live : the configuration the reference's main() runs (Header.h), x in [0,80],\n"
      "         r in [0,1], HLLC only on faces with x in (10,40), CFL 0.05\n");
}

// reference spacing convention, Header.h:11-16
void extent(ast_grid_t& g, int N, int M, double x_min, double x_max, double y_max) {
  g.N = N;
  g.M = M;
  g.x_min = x_min;
  g.dx = (x_max - x_min) / (N - 1);
  g.dy = y_max / M;
  g.y_min = y_max / (2.0 * M);
}

// Builds the grid and the initial state (s = {rho,p}, u = {vx,vy} per cell).
bool setup(const Opts& o, ast_grid_t& g, std::vector<double>& s, std::vector<double>& u) {
  memset(&g, 0, sizeof g);
  g.ho_lo = -INFINITY;
  g.ho_hi = INFINITY;
  int N = o.N, M = o.M;
  if (o.problem == "wind") {
    if (!N) N = 2048;
    if (!M) M = 1024;
    extent(g, N, M, -400.0, 400.0, 400.0);
    g.gamma = 5.0 / 3.0;
    g.krit = 0.3;
    g.geometry = AST_AXISYMMETRIC;
    // rho_w = 1, v_w = 1, Mach 10 at R = 1; counter-flow rho 1e-3, speed 1,
    // Mach 2 along -z (kernel.cu:2328-2337, 3355-3373); bubble radius 20
    const double gam = g.gamma, pw = 1.0 / (gam * 100.0), rinf = 1e-3,
                 pinf = rinf / (gam * 4.0);
    g.bc[AST_XLO] = AST_BC_OUTFLOW;
    g.bc[AST_XHI] = AST_BC_DIRICHLET;
    g.bc[AST_YLO] = AST_BC_REFLECT;
    g.bc[AST_YHI] = AST_BC_OUTFLOW;
    g.bc_state[AST_XHI][0] = rinf;
    g.bc_state[AST_XHI][1] = pinf;
    g.bc_state[AST_XHI][2] = -1.0;
    g.frozen_ax = g.frozen_ay = 1.0;
    s.resize(2 * (size_t)N * M);
    u.resize(2 * (size_t)N * M);
    for (int m = 0; m < M; m++)
      for (int n = 0; n < N; n++) {
        const size_t k = (size_t)m * N + n;
        const double x = g.x_min + n * g.dx, y = g.y_min + m * g.dy;
        const double R = fmax(sqrt(x * x + y * y), 1e-12);
        const bool in = R < 20.0;
        s[2 * k] = in ? 1.0 / (R * R) : rinf;
        s[2 * k + 1] = in ? pw * pow(R, -2.0 * gam) : pinf;
        u[2 * k] = in ? x / R : -1.0;
        u[2 * k + 1] = in ? y / R : 0.0;
      }
    return true;
  }
  if (o.problem == "sod") {
    if (!N) N = 256;
    if (!M) M = 128;
    extent(g, N, M, 0.0, 1.0, 0.5);
    g.gamma = 1.4;
    g.krit = 0.4;
    g.geometry = AST_AXISYMMETRIC;
    g.bc[AST_XLO] = g.bc[AST_XHI] = g.bc[AST_YHI] = AST_BC_OUTFLOW;
    g.bc[AST_YLO] = AST_BC_REFLECT;
    s.resize(2 * (size_t)N * M);
    u.assign(2 * (size_t)N * M, 0.0);
    for (int m = 0; m < M; m++)
      for (int n = 0; n < N; n++) {
        const size_t k = (size_t)m * N + n;
        const bool left = g.x_min + n * g.dx < 0.5;
        s[2 * k] = left ? 1.0 : 0.125;
        s[2 * k + 1] = left ? 1.0 : 0.1;
      }
    return true;
  }
  if (o.problem == "live") {
    if (!N) N = 32000;
    if (!M) M = 400;
    extent(g, N, M, 0.0, 80.0, 1.0);  // Header.h:8-16
    g.gamma = 5.0 / 3.0;
    g.krit = 0.05;                    // Header.h:47
    g.geometry = AST_PLANAR;
    g.bc[AST_XLO] = AST_BC_DIRICHLET;  // kernel.cu:2577-2580
    g.bc[AST_XHI] = AST_BC_OUTFLOW;    // kernel.cu:2549-2552
    g.bc[AST_YLO] = g.bc[AST_YHI] = AST_BC_REFLECT;
    g.bc_state[AST_XLO][0] = 1.0;
    g.bc_state[AST_XLO][1] = 7.0;  // p_H, Header.h:6
    g.bc_state[AST_XLO][2] = 1.0;
    g.ho_lo = 10.000001;  // kernel.cu:2801, 2815, 2842
    g.ho_hi = 40.000001;
    s.resize(2 * (size_t)N * M);
    u.resize(2 * (size_t)N * M);
    for (size_t k = 0; k < (size_t)N * M; k++) {  // kernel.cu:3375-3378
      s[2 * k] = 1.0;
      s[2 * k + 1] = 7.0;
      u[2 * k] = 1.0;
      u[2 * k + 1] = 0.0;
    }
    return true;
  }
  return false;
}

}  // namespace

int main(int argc, char** argv) {
  Opts o;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto next = [&](void) -> const char* {
      if (i + 1 >= argc) {
        fprintf(stderr, "asteriy_run: %s needs a value\n", a.c_str());
        exit(2);
      }
      return argv[++i];
    };
    if (a == "--help" || a == "-h") { usage(); return 0; }
    else if (a == "--problem") o.problem = next();
    else if (a == "--method") o.method = next();
    else if (a == "--precision") o.precision = next();
    else if (a == "--N") o.N = atoi(next());
    else if (a == "--M") o.M = atoi(next());
    else if (a == "--steps") o.steps = atoi(next());
    else if (a == "--print-every") o.every = atoi(next());
    else if (a == "--out") o.out = next();
    else if (a == "--tecplot") o.tec = next();
    else if (a == "--axis") o.axis = next();
    else if (a == "--checkpoint") o.ckpt = next();
    else if (a == "--resume") o.resume = next();
    else if (a == "--mlp") o.mlp = next();
    else if (a == "--autotune") o.autotune = true;
    else { fprintf(stderr, "asteriy_run: unknown option %s\n", a.c_str()); usage(); return 2; }
  }
  const int method = o.method == "hll" ? AST_HLL : o.method == "hllc" ? AST_HLLC
                     : o.method == "exact" ? AST_EXACT : o.method == "hllc_a" ? AST_HLLC_A : -1;
  const int prec = o.precision == "f64" ? 0 : o.precision == "f32" ? 1 : -1;
  if (method < 0 || prec < 0 || o.steps < 0 || o.every < 1) { usage(); return 2; }
  ast_grid_t g;
  std::vector<double> s, u;
  if (!setup(o, g, s, u)) { usage(); return 2; }
  if (ast_device_count() < 1) { fprintf(stderr, "asteriy_run: no HIP device\n"); return 3; }

  ast_sim_t* sim = nullptr;
  if (ast_sim_create(&g, prec, method, &sim) != AST_OK) return die("ast_sim_create");
  if (ast_sim_set_state(sim, s.data(), u.data()) != AST_OK) return die("ast_sim_set_state");
  if (!o.resume.empty() && ast_sim_load(sim, o.resume.c_str()) != AST_OK) return die("resume");
  void* mlp = nullptr;
  if (!o.mlp.empty()) {
    std::vector<float> p(AST_MLP_NPARAM);
    FILE* f = fopen(o.mlp.c_str(), "rb");
    if (!f || fread(p.data(), sizeof(float), p.size(), f) != p.size()) {
      fprintf(stderr, "asteriy_run: cannot read %d floats from %s\n", AST_MLP_NPARAM, o.mlp.c_str());
      return 2;
    }
    fclose(f);
    if (hipMalloc(&mlp, ast_mlp_bytes()) != hipSuccess) return die("hipMalloc");
    if (ast_mlp_set_params((ast_mlp_t*)mlp, p.data(), nullptr) != AST_OK) return die("mlp");
    if (ast_sim_set_mlp(sim, (ast_mlp_t*)mlp) != AST_OK) return die("ast_sim_set_mlp");
  }
  if (o.autotune && prec == 1) {
    int rows = 0;
    float us = 0.f;
    if (ast_sim_autotune(sim, nullptr, 0, 10, &rows, &us) != AST_OK) return die("autotune");
    printf("autotune: %d rows per wave strip, %.1f us per step\n", rows, us);
  }
  printf("%s %dx%d, %s, %s, %d steps\n", o.problem.c_str(), g.N, g.M, o.method.c_str(),
         o.precision.c_str(), o.steps);
  const double cells = (double)g.N * g.M;
  auto t0 = std::chrono::steady_clock::now();
  int done = 0;
  while (done < o.steps) {
    const int chunk = std::min(o.every, o.steps - done);
    if (ast_sim_run(sim, chunk, 1) != AST_OK) return die("ast_sim_run");
    if (ast_sim_sync(sim) != AST_OK) return die("ast_sim_sync");
    done += chunk;
    double T, Td, TT;
    int it;
    if (ast_sim_get_time(sim, &T, &Td, &TT, &it) != AST_OK) return die("ast_sim_get_time");
    printf("i = %d,  TT = %lf ,  dt = %lf \n", it, TT, Td);  // kernel.cu:1533
  }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  printf("%d step - Time:  %.2f sec  (%.3f Gcell-updates/s)\n", o.steps, sec,
         o.steps * cells / sec / 1e9);
  if (!o.out.empty() && ast_sim_write_b31(sim, o.out.c_str()) != AST_OK) return die("write B31");
  if (!o.tec.empty() && ast_sim_write_tecplot(sim, o.tec.c_str(), 2) != AST_OK) return die("write Tecplot");
  if (!o.axis.empty() && ast_sim_write_axis(sim, o.axis.c_str()) != AST_OK) return die("write axis");
  if (!o.ckpt.empty() && ast_sim_save(sim, o.ckpt.c_str()) != AST_OK) return die("checkpoint");
  ast_sim_destroy(sim);
  if (mlp) (void)hipFree(mlp);
  return 0;
}
