// mc_header_check.cpp -- include/asteriy_mc.hpp compiled for the host (built
// and run by tests/test_mc_helpers.py, no GPU).  Mode "rates": reads pairs
// (u, cp) from argv[2] (2 planes of doubles), writes Velosity_1/2/3 (3
// planes) to argv[3].  Mode "exit": reads 8 planes (x0, y0, x, y, z, Vx, Vy,
// Vz) and writes t, mode, hit (3 planes of doubles) with the grid spacing
// dx = argv[4], dy = argv[5].  Mode "belong": reads 2 planes (x, r), writes
// n, m (2 planes of doubles) with x_min argv[4], dx argv[5], y_min argv[6],
// dy argv[7].
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../include/asteriy_mc.hpp"

static std::vector<double> read_all(const char* path) {
  std::vector<double> v;
  FILE* f = fopen(path, "rb");
  if (!f) exit(3);
  double d;
  while (fread(&d, sizeof d, 1, f) == 1) v.push_back(d);
  fclose(f);
  return v;
}

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  const std::vector<double> in = read_all(argv[2]);
  std::vector<double> out;
  if (!strcmp(argv[1], "rates")) {
    const size_t n = in.size() / 2;
    out.resize(3 * n);
    for (size_t i = 0; i < n; i++) {
      out[i] = asteriy::mc::velosity_1(in[i], in[n + i]);
      out[n + i] = asteriy::mc::velosity_2(in[i], in[n + i]);
      out[2 * n + i] = asteriy::mc::velosity_3(in[i], in[n + i]);
      double v1, v2, v3;  // the fused form: the same bits
      asteriy::mc::velosities(in[i], in[n + i], v1, v2, v3);
      if (!(v1 == out[i] || (v1 != v1 && out[i] != out[i])) ||
          !(v2 == out[n + i] || (v2 != v2 && out[n + i] != out[n + i])) ||
          !(v3 == out[2 * n + i] || (v3 != v3 && out[2 * n + i] != out[2 * n + i])))
        return 6;
    }
  } else if (!strcmp(argv[1], "exit") && argc == 6) {
    const size_t n = in.size() / 8;
    const asteriy::mc::McGrid g{0.0, atof(argv[4]), 0.0, atof(argv[5])};
    out.resize(3 * n);
    for (size_t i = 0; i < n; i++) {
      int mode;
      double t;
      const bool hit = asteriy::mc::peresechenie(g, in[i], in[n + i], in[2 * n + i], in[3 * n + i],
                                                 in[4 * n + i], in[5 * n + i], in[6 * n + i],
                                                 in[7 * n + i], mode, t);
      out[i] = t;
      out[n + i] = mode;
      out[2 * n + i] = hit ? 1.0 : 0.0;
    }
  } else if (!strcmp(argv[1], "belong") && argc == 8) {
    const size_t n = in.size() / 2;
    const asteriy::mc::McGrid g{atof(argv[4]), atof(argv[5]), atof(argv[6]), atof(argv[7])};
    out.resize(2 * n);
    for (size_t i = 0; i < n; i++) {
      int a, b;
      asteriy::mc::belong_point(g, in[i], in[n + i], a, b);
      out[i] = a;
      out[n + i] = b;
    }
  } else {
    return 4;
  }
  FILE* f = fopen(argv[3], "wb");
  if (!f) return 5;
  fwrite(out.data(), sizeof(double), out.size(), f);
  fclose(f);
  return 0;
}
