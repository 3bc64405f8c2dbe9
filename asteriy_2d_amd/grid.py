"""Grid/physics description shared by the host API and the C-ABI.

`Grid` mirrors `ast_grid_t` of include/asteriy.h, which replaces the
compile-time configuration of the reference (Header.h:6-63): cell centres are
x = x_min + n*dx and y = y_min + m*dy with the reference's spacing convention
dx = (x_max - x_min)/(N - 1), dy = y_max/M, y_min = dy/2 (Header.h:11-16).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Sequence

HLL, HLLC, EXACT, HLLC_A = 0, 1, 2, 3
HLLCQ, HLLCQ_A = 4, 5  # passive-scalar fluxes (ast_flux_q_*)
HLLC_MHD, HLLD = 6, 7  # MHD fluxes with a passive scalar (ast_flux_mhd_f64)
METHODS = {"hll": HLL, "hllc": HLLC, "exact": EXACT, "hllc_a": HLLC_A}
PLANAR, AXISYMMETRIC = 0, 1
BC_OUTFLOW, BC_REFLECT, BC_DIRICHLET, BC_OUTFLOW_CLAMPED, BC_DIRICHLET_PULSED = 0, 1, 2, 3, 4
XLO, XHI, YLO, YHI = 0, 1, 2, 3
# the inflow perturbation of add2 (kernel.cu:2022-2025): 1% density, period 1,
# rows y < 0.5, until TT = 3
PULSE_ADD2 = (0.01, 2.0 * math.pi, 0.5, 3.0)
ITER_CAP = 1100


class CGrid(ctypes.Structure):
    """ctypes layout of ast_grid_t (include/asteriy.h)."""

    _fields_ = [
        ("N", ctypes.c_int32),
        ("M", ctypes.c_int32),
        ("geometry", ctypes.c_int32),
        ("bc", ctypes.c_int32 * 4),
        ("x_min", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("y_min", ctypes.c_double),
        ("dy", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("krit", ctypes.c_double),
        ("bc_state", (ctypes.c_double * 4) * 4),
        ("bc_vclamp", ctypes.c_double * 4),
        ("ho_lo", ctypes.c_double),
        ("ho_hi", ctypes.c_double),
        ("frozen_x0", ctypes.c_double),
        ("frozen_ax", ctypes.c_double),
        ("frozen_ay", ctypes.c_double),
        ("bc_pulse", ctypes.c_double * 4),
    ]


class CStats(ctypes.Structure):
    """ctypes layout of ast_stats_t."""

    _fields_ = [(name, ctypes.c_ulonglong) for name in (
        "exact_faces", "newton_faces", "newton_iters", "bisect_faces",
        "bisect_iters", "rarefaction_faces", "vacuum_faces", "capped",
        "fp64_faces", "mlp_faces")]


STATS_FIELDS = [f[0] for f in CStats._fields_]


@dataclass
class Grid:
    """Runtime grid + physics (Header.h macros as data)."""

    N: int
    M: int
    x_min: float
    dx: float
    y_min: float
    dy: float
    gamma: float = 5.0 / 3.0
    krit: float = 0.05
    geometry: int = PLANAR
    bc: Sequence[int] = (BC_DIRICHLET, BC_OUTFLOW, BC_REFLECT, BC_REFLECT)
    bc_state: Sequence[Sequence[float]] = field(
        default_factory=lambda: [[1.0, 7.0, 1.0, 0.0]] + [[0.0] * 4] * 3)
    bc_vclamp: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    ho_lo: float = -math.inf
    ho_hi: float = math.inf
    frozen_x0: float = 0.0
    frozen_ax: float = 0.0
    frozen_ay: float = 0.0
    # BC_DIRICHLET_PULSED sides: density factor 1 + a sin(w TT) for rows with
    # y < y_lim while TT < t_lim (a, w, y_lim, t_lim); add2's inflow
    # perturbation (kernel.cu:2022-2025) is PULSE_ADD2
    bc_pulse: Sequence[float] = (0.0, 0.0, 0.0, 0.0)

    @classmethod
    def from_extent(cls, N: int, M: int, x_min: float, x_max: float,
                    y_max: float, **kw) -> "Grid":
        """Reference spacing convention, Header.h:11-16."""
        dx = (x_max - x_min) / (N - 1)
        dy = y_max / M
        return cls(N=N, M=M, x_min=x_min, dx=dx, y_min=y_max / (2.0 * M), dy=dy, **kw)

    @property
    def cells(self) -> int:
        return self.N * self.M

    def x(self, n):
        return self.x_min + n * self.dx

    def y(self, m):
        return self.y_min + m * self.dy

    def to_c(self) -> CGrid:
        g = CGrid()
        g.N, g.M, g.geometry = self.N, self.M, self.geometry
        for i in range(4):
            g.bc[i] = int(self.bc[i])
            g.bc_vclamp[i] = float(self.bc_vclamp[i])
            for j in range(4):
                g.bc_state[i][j] = float(self.bc_state[i][j])
        g.x_min, g.dx, g.y_min, g.dy = self.x_min, self.dx, self.y_min, self.dy
        g.gamma, g.krit = self.gamma, self.krit
        g.ho_lo, g.ho_hi = self.ho_lo, self.ho_hi
        g.frozen_x0, g.frozen_ax, g.frozen_ay = self.frozen_x0, self.frozen_ax, self.frozen_ay
        for i in range(4):
            g.bc_pulse[i] = float(self.bc_pulse[i])
        return g

    def replace(self, **kw) -> "Grid":
        d = dict(self.__dict__)
        d.update(kw)
        return Grid(**d)


def reference_live_grid(N: int = 32000, M: int = 400) -> Grid:
    """The configuration `main()` runs (Header.h:6-16, 47; kernel.cu:2513-2857):
    x in [0,80], r in [0,1], Dirichlet inflow {rho=1, p=p_H=7, u=(1,0)} at
    n=0, outflow at n=N-1, mirrors at m=0 and m=M-1, HLLC on reconstructed
    faces only for face x in (10,40), CFL 0.05."""
    return Grid.from_extent(
        N, M, 0.0, 80.0, 1.0, gamma=5.0 / 3.0, krit=0.05, geometry=PLANAR,
        bc=(BC_DIRICHLET, BC_OUTFLOW, BC_REFLECT, BC_REFLECT),
        bc_state=[[1.0, 7.0, 1.0, 0.0], [0.0] * 4, [0.0] * 4, [0.0] * 4],
        ho_lo=10.000001, ho_hi=40.000001)
