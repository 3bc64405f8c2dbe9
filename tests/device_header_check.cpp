// device_header_check.cpp -- the per-face wrappers of include/asteriy_device.hpp
// compiled for the host (built and run by tests/test_device_header.py, no
// GPU): reads n faces (8 SoA planes of doubles: rhoL, pL, vxL, vyL, rhoR, pR,
// vxR, vyR) from argv[1], applies the operator argv[3] (0 HLL, 1
// HLLC_2d_Korolkov, 2 Godunov_Solver_Alexashov, 3 HLLC_Aleksashov) with the
// normal (argv[4], argv[5]), rad argv[6], gamma argv[7], krit argv[8], and
// writes 5 planes (Ps.x, Ps.y, Pu.x, Pu.y, return value) to argv[2].
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "../include/asteriy_device.hpp"

int main(int argc, char** argv) {
  if (argc != 9) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 3;
  std::vector<double> in;
  double v;
  while (fread(&v, sizeof v, 1, f) == 1) in.push_back(v);
  fclose(f);
  const size_t n = in.size() / 8;
  const int method = atoi(argv[3]);
  const double n1 = atof(argv[4]), n2 = atof(argv[5]), rad = atof(argv[6]);
  const asteriy::Gas gas = asteriy::make_gas(atof(argv[7]), atof(argv[8]));
  std::vector<double> out(5 * n);
  for (size_t i = 0; i < n; i++) {
    const double2 Ls = {in[i], in[n + i]}, Lu = {in[2 * n + i], in[3 * n + i]};
    const double2 Rs = {in[4 * n + i], in[5 * n + i]}, Ru = {in[6 * n + i], in[7 * n + i]};
    double2 Ps, Pu;
    double t;
    switch (method) {
      case AST_HLL: t = asteriy::HLL(gas, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad); break;
      case AST_HLLC: t = asteriy::HLLC_2d_Korolkov(gas, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad); break;
      case AST_EXACT:
        t = asteriy::Godunov_Solver_Alexashov(gas, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
        break;
      case AST_HLLC_A: t = asteriy::HLLC_Aleksashov(gas, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad); break;
      default: return 4;
    }
    out[i] = Ps.x;
    out[n + i] = Ps.y;
    out[2 * n + i] = Pu.x;
    out[3 * n + i] = Pu.y;
    out[4 * n + i] = t;
  }
  f = fopen(argv[2], "wb");
  if (!f) return 5;
  fwrite(out.data(), sizeof(double), out.size(), f);
  fclose(f);
  return 0;
}
