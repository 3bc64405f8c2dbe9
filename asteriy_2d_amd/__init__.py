"""asteriy_2d_amd: MI355X (gfx950) implementation of Asteriy_2d's explicit
(z,r) finite-volume step -- HIP kernels behind a C-ABI (include/asteriy.h),
with a thin Python host mirror of the reference's operator interface."""
from .grid import (AXISYMMETRIC, BC_DIRICHLET, BC_DIRICHLET_PULSED, BC_OUTFLOW,
                   BC_OUTFLOW_CLAMPED, BC_REFLECT, EXACT, HLL, HLLC, HLLC_A, HLLCQ, HLLCQ_A,
                   HLLC_MHD, HLLD, PLANAR, PULSE_ADD2, Grid, reference_live_grid)

__all__ = ["Grid", "reference_live_grid", "HLL", "HLLC", "EXACT", "HLLC_A", "HLLCQ", "HLLCQ_A",
           "HLLC_MHD", "HLLD",
           "PLANAR", "AXISYMMETRIC", "BC_OUTFLOW", "BC_REFLECT", "BC_DIRICHLET",
           "BC_OUTFLOW_CLAMPED", "BC_DIRICHLET_PULSED", "PULSE_ADD2"]
