"""Problem set-ups of the benchmark configurations (host, numpy).

* Sod tube in (z, r): 256x128 over z in [0,1], r in [0,0.5], gamma 1.4,
  rho=1, p=1 for z<0.5 else rho=0.125, p=0.1, at rest; reflective axis,
  zero-gradient outflow elsewhere; HLL + minmod, CFL 0.4.
* Stellar wind / counter-flow interaction (the astrosphere problem of the
  reference): z in [-400,400], r in [0,400], gamma 5/3.  A spherical wind
  leaves a frozen sphere R=1 (add_MK's frozen region, kernel.cu:2317-2324)
  with rho = rho_w/R^2, radial speed v_w and Mach M_w at R=1, p =
  rho_w v_w^2/(gamma M_w^2) R^(-2 gamma) (the commented wind initial state of
  kernel.cu:3355-3373); a uniform counter-flow rho=1e-3 enters at z=+400
  moving along -z with Mach M_inf (the Dirichlet inflow of add_MK,
  kernel.cu:2328-2337); outflow at z=-400 and r=400; reflective axis.  The
  builder fixes rho_w = 1, v_w = 1, counter-flow speed 1 and an initial wind
  bubble of radius 20 (inside the ram-pressure balance radius ~31.6).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .grid import (AXISYMMETRIC, BC_DIRICHLET, BC_OUTFLOW, BC_REFLECT, HLL, HLLC, PLANAR,
                   Grid)


@dataclass
class Problem:
    name: str
    grid: Grid
    s: np.ndarray  # (M, N, 2) rho, p
    u: np.ndarray  # (M, N, 2) vx, vy
    method: int
    steps: int


def sod(N: int = 256, M: int = 128, geometry: int = AXISYMMETRIC) -> Problem:
    g = Grid.from_extent(N, M, 0.0, 1.0, 0.5, gamma=1.4, krit=0.4, geometry=geometry,
                         bc=(BC_OUTFLOW, BC_OUTFLOW, BC_REFLECT, BC_OUTFLOW))
    x = g.x(np.arange(N))[None, :].repeat(M, 0)
    s = np.empty((M, N, 2))
    s[..., 0] = np.where(x < 0.5, 1.0, 0.125)
    s[..., 1] = np.where(x < 0.5, 1.0, 0.1)
    u = np.zeros((M, N, 2))
    return Problem("sod", g, s, u, HLL, 200)


@dataclass
class WindParams:
    rho_w: float = 1.0
    v_w: float = 1.0
    mach_w: float = 10.0
    rho_inf: float = 1e-3
    v_inf: float = 1.0
    mach_inf: float = 2.0
    r_frozen: float = 1.0
    r_bubble: float = 20.0


def wind_grid(N: int, M: int, wp: WindParams = WindParams(), krit: float = 0.3,
              gamma: float = 5.0 / 3.0) -> Grid:
    p_inf = wp.rho_inf * wp.v_inf ** 2 / (gamma * wp.mach_inf ** 2)
    return Grid.from_extent(
        N, M, -400.0, 400.0, 400.0, gamma=gamma, krit=krit, geometry=AXISYMMETRIC,
        bc=(BC_OUTFLOW, BC_DIRICHLET, BC_REFLECT, BC_OUTFLOW),
        bc_state=[[0.0] * 4, [wp.rho_inf, p_inf, -wp.v_inf, 0.0], [0.0] * 4, [0.0] * 4],
        frozen_x0=0.0, frozen_ax=wp.r_frozen, frozen_ay=wp.r_frozen)


def wind_state(g: Grid, wp: WindParams = WindParams()):
    gamma = g.gamma
    n = np.arange(g.N)
    m = np.arange(g.M)
    x = g.x(n)[None, :]
    y = g.y(m)[:, None]
    R = np.sqrt(x * x + y * y)
    p_w = wp.rho_w * wp.v_w ** 2 / (gamma * wp.mach_w ** 2)
    p_inf = wp.rho_inf * wp.v_inf ** 2 / (gamma * wp.mach_inf ** 2)
    inside = R < wp.r_bubble
    Rc = np.maximum(R, 1e-12)
    s = np.empty((g.M, g.N, 2))
    u = np.empty((g.M, g.N, 2))
    s[..., 0] = np.where(inside, wp.rho_w / Rc ** 2, wp.rho_inf)
    s[..., 1] = np.where(inside, p_w * Rc ** (-2.0 * gamma), p_inf)
    u[..., 0] = np.where(inside, wp.v_w * x / Rc, -wp.v_inf)
    u[..., 1] = np.where(inside, wp.v_w * y / Rc, 0.0)
    return s, u


def wind(N: int = 2048, M: int = 1024, method: int = HLLC, steps: int = 5000,
         wp: WindParams = WindParams(), krit: float = 0.3) -> Problem:
    g = wind_grid(N, M, wp, krit)
    s, u = wind_state(g, wp)
    return Problem(f"wind{N}x{M}", g, s, u, method, steps)


def korolkov_initial_state(g: Grid):
    """InitialConditions (KorolkovSolver.cu:35-53): inside the radius 100 of
    the origin (x^2 + y^2 < 10000) rho = 8, p = 2, u = (-90, 0); elsewhere
    rho = 1, p = 1, u = (-1, 0).  The reference defines it beside its own
    Initialization (KorolkovSolver.cu:56) and calls neither from the live
    main; the state is given for runs of that setup."""
    x = g.x(np.arange(g.N))[None, :]
    y = g.y(np.arange(g.M))[:, None]
    inside = np.broadcast_to(x * x + y * y < 10000, (g.M, g.N))
    s = np.empty((g.M, g.N, 2))
    u = np.zeros((g.M, g.N, 2))
    s[..., 0] = np.where(inside, 8.0, 1.0)
    s[..., 1] = np.where(inside, 2.0, 1.0)
    u[..., 0] = np.where(inside, -90.0, -1.0)
    return s, u


def random_state(g: Grid, seed: int = 0, smooth: bool = False):
    """A random but physical state (positive rho, p) for parity tests."""
    rng = np.random.default_rng(seed)
    s = np.empty((g.M, g.N, 2))
    u = np.empty((g.M, g.N, 2))
    s[..., 0] = np.exp(rng.uniform(-1.0, 1.0, (g.M, g.N)))
    s[..., 1] = np.exp(rng.uniform(-1.0, 1.0, (g.M, g.N)))
    u[...] = rng.uniform(-1.0, 1.0, (g.M, g.N, 2))
    if smooth:
        x = np.linspace(0, 2 * math.pi, g.N)[None, :]
        y = np.linspace(0, 2 * math.pi, g.M)[:, None]
        s[..., 0] = 1.0 + 0.3 * np.sin(x) * np.cos(y)
        s[..., 1] = 1.0 + 0.2 * np.cos(x + y)
        u[..., 0] = 0.3 * np.sin(y)
        u[..., 1] = 0.2 * np.cos(x)
    return s, u


__all__ = ["Problem", "sod", "wind", "wind_grid", "wind_state", "WindParams", "random_state",
           "PLANAR", "AXISYMMETRIC"]
