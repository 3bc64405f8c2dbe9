// device_header_kernel.hip -- include/asteriy_device.hpp in a caller's HIP
// kernel, compiled for gfx950 by tests/test_device_header.py (no GPU needed):
// one face per lane through each of the four wrappers.
#include <hip/hip_runtime.h>

#include "../include/asteriy_device.hpp"

__global__ void caller_kernel(const double2* s, const double2* u, double2* ps, double2* pu,
                              double* t, int n, asteriy::Gas gas, int method) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n) return;
  double2 Ps, Pu;
  double d;
  if (method == AST_HLL) d = asteriy::HLL(gas, s[i], u[i], s[i + 1], u[i + 1], 1.0, 0.0, Ps, Pu, 0.1);
  else if (method == AST_HLLC)
    d = asteriy::HLLC_2d_Korolkov(gas, s[i], u[i], s[i + 1], u[i + 1], 1.0, 0.0, Ps, Pu, 0.1);
  else if (method == AST_HLLC_A)
    d = asteriy::HLLC_Aleksashov(gas, s[i], u[i], s[i + 1], u[i + 1], 1.0, 0.0, Ps, Pu, 0.1);
  else
    d = asteriy::Godunov_Solver_Alexashov(gas, s[i], u[i], s[i + 1], u[i + 1], 1.0, 0.0, Ps, Pu, 0.1);
  ps[i] = Ps;
  pu[i] = Pu;
  t[i] = d;
}
