"""Seeded face-pair generator of the Riemann-flux microbenchmark.

2^24 independent pairs by default: rhoL, rhoR and pL, pR log-uniform in
[1e-3, 1e3]; normal velocities uniform in +-5*max(cL, cR); tangential
velocities uniform in +-1; a fraction (2%) of near-vacuum pairs whose sides
diverge strongly: uR - uL uniform in [0.9, 1.5] x 2(cL + cR)/(gamma - 1)
(the vacuum threshold is 1.0 x), split symmetrically.  The normal is (1, 0),
so the normal velocity is vx and the tangential one vy.
"""
from __future__ import annotations

import numpy as np

SEED = 20261202


def faces(n: int = 1 << 24, gamma: float = 5.0 / 3.0, seed: int = SEED,
          vacuum_frac: float = 0.02, dtype=np.float64) -> np.ndarray:
    """Returns an (8, n) SoA array: rhoL, pL, vxL, vyL, rhoR, pR, vxR, vyR."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(1e-3), np.log(1e3)
    rl = np.exp(rng.uniform(lo, hi, n))
    rr = np.exp(rng.uniform(lo, hi, n))
    pl = np.exp(rng.uniform(lo, hi, n))
    pr = np.exp(rng.uniform(lo, hi, n))
    cl = np.sqrt(gamma * pl / rl)
    cr = np.sqrt(gamma * pr / rr)
    cm = np.maximum(cl, cr)
    ul = rng.uniform(-5.0, 5.0, n) * cm
    ur = rng.uniform(-5.0, 5.0, n) * cm
    vl = rng.uniform(-1.0, 1.0, n)
    vr = rng.uniform(-1.0, 1.0, n)
    vac = rng.random(n) < vacuum_frac
    du = rng.uniform(0.9, 1.5, n) * 2.0 * (cl + cr) / (gamma - 1.0)
    ul = np.where(vac, -0.5 * du, ul)
    ur = np.where(vac, 0.5 * du, ur)
    out = np.stack([rl, pl, ul, vl, rr, pr, ur, vr]).astype(dtype)
    return np.ascontiguousarray(out)


__all__ = ["faces", "SEED"]
