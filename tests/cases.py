"""Host-side test cases with closed-form answers, shared by the oracle tests
(CPU) and the GPU tests."""
import math

import numpy as np

from asteriy_2d_amd import grid as G


def floor_case(T_do, w=0.6, a=2.0, gamma=5.0 / 3.0, geometry=G.PLANAR, first_order=True):
    """A 5x2 planar grid whose middle column (n = 2) sits at rest between
    columns moving away from it at speed a, every cell with rho = p = 1 and
    tangential velocity w, all faces first order (HLL on the cell states,
    kernel.cu:2812, 2856).  The rows are equal and the y ghosts copy the edge
    rows, so the y fluxes cancel.  Returns (grid, s, u, expected middle cell
    (rho', p', vx', vy') by the closed form below).  first_order=False
    reconstructs every face instead (no closed form then: the expected
    values are those of the first-order case, for orientation only)."""
    ho = (0.0, 0.0) if first_order else (-math.inf, math.inf)
    g = G.Grid.from_extent(5, 2, 0.0, 0.4, 0.2, gamma=gamma, krit=0.3, geometry=geometry,
                           bc=(G.BC_OUTFLOW,) * 4, ho_lo=ho[0], ho_hi=ho[1])
    s = np.ones((2, 5, 2))
    u = np.zeros((2, 5, 2))
    u[:, :2, 0], u[:, 3:, 0] = -a, a
    u[..., 1] = w
    # HLL with Davis speeds at the east face of the middle cell: SL = -c,
    # SR = a + c; its mass flux c a/(a + 2c), tangential momentum c a w/(a + 2c),
    # energy c ((E_R + p) a - (a + c) a^2/2)/(a + 2c) with E_R = p/(g-1) +
    # (a^2 + w^2)/2 (the w^2 parts of E_R - E_L cancel); the west face is its
    # mirror image (mass and energy leave through it as well, the normal
    # momentum flux is even and cancels)
    c = math.sqrt(gamma)
    E0 = 1.0 / (gamma - 1) + 0.5 * w * w
    ER = 1.0 / (gamma - 1) + 0.5 * (a * a + w * w)
    k = T_do / g.dx
    dr = -2 * k * c * a / (a + 2 * c)
    dE = -2 * k * c * ((ER + 1.0) * a - (a + c) * a * a / 2) / (a + 2 * c)
    dmy = w * dr
    r2 = 1.0 + dr
    if r2 <= 0:
        r2 = 1.0                    # kernel.cu:3060-3064: keep the old density
    vy = (w + dmy) / r2
    p2 = (E0 + dE - 0.5 * r2 * vy * vy) * (gamma - 1)
    if p2 <= 0:
        p2 = 1e-6                   # kernel.cu:3070-3073
    return g, s, u, (r2, p2, 0.0, vy)


# T_do of the three regimes of floor_case: no floor, the density floor alone,
# both floors
FLOOR_REGIMES = [(0.005, False, False), (0.1, True, False), (0.3, True, True)]


def uniform_axisymmetric(T_do=0.01, rho=0.8, p=1.3, ux=0.45, uy=-0.7):
    """A uniform state with v != 0 on an axisymmetric grid with zero-gradient
    ghosts on every side: every face carries the same flux, so one step
    changes each cell only by the r-source -(T_do/y) v {rho, rho u, rho v,
    gamma p/(g-1) + rho|u|^2/2} at the cell-centre radius y
    (KorolkovSolver.cu:13-32).  Worked through: rho' = rho (1 - f), u' = u,
    v' = v, p' = p (1 - gamma f) with f = T_do v / y.  Returns (grid, s, u,
    expected s2, expected u2)."""
    g = G.Grid.from_extent(12, 9, -1.0, 1.0, 0.9, gamma=5.0 / 3.0, krit=0.3,
                           geometry=G.AXISYMMETRIC, bc=(G.BC_OUTFLOW,) * 4)
    s = np.empty((g.M, g.N, 2))
    s[..., 0], s[..., 1] = rho, p
    u = np.empty_like(s)
    u[..., 0], u[..., 1] = ux, uy
    y = g.y(np.arange(g.M))[:, None] * np.ones((1, g.N))
    f = T_do * uy / y
    s2 = np.stack([rho * (1 - f), p * (1 - g.gamma * f)], -1)
    u2 = u.copy()
    return g, s, u, s2, u2


# the momentum coefficient of the reference's neutral-H source, n_H/Kn
# (Header.h:84-85, kernel.cu:2495-2496)
K_MOM_REF = 3.0 / 0.419231


def neutral_sources(M, N, seed=7, scale=0.3, scale3=None):
    """Random neutral-H moment planes: nn2 (M,N,3) with a z component of a
    very different size (the step must not read it) and nn3 (M,N); entries
    uniform in +-scale (+-scale3 for nn3)."""
    rng = np.random.default_rng(seed)
    nn2 = rng.uniform(-scale, scale, (M, N, 3))
    nn2[..., 2] = rng.uniform(1e3, 2e3, (M, N))
    s3 = scale if scale3 is None else scale3
    nn3 = rng.uniform(-s3, s3, (M, N))
    return nn2, nn3


def frozen_mask(g):
    """Cells inside the grid's frozen ellipse (copied by the step,
    kernel.cu:2317-2324)."""
    if not g.frozen_ax > 0:
        return np.zeros((g.M, g.N), bool)
    x = g.x(np.arange(g.N))[None, :]
    y = g.y(np.arange(g.M))[:, None]
    return ((x - g.frozen_x0) ** 2 / g.frozen_ax ** 2 + y ** 2 / g.frozen_ay ** 2) < 1.0


def sources_case(N=40, M=24, seed=3):
    """The wind problem with a frozen sphere that covers cells of a small
    grid, and random source planes: (problem, nn2, nn3, frozen mask).  The
    planes are sized so that with the reference's kmom = n_H/Kn and ken =
    0.02 a step of this coarse grid (dt up to ~4) changes a velocity by a few
    per cent of the inflow speed and a pressure by a tenth of the inflow
    pressure at most -- strong enough to matter, not so strong that the
    reference's own update (E' - rho'|u'|^2/2) cancels to round-off."""
    from asteriy_2d_amd import problems
    p = problems.wind(N, M, method=G.HLLC, wp=problems.WindParams(r_frozen=60.0))
    nn2, nn3 = neutral_sources(p.grid.M, p.grid.N, seed, scale=2e-3, scale3=1e-4)
    return p, nn2, nn3, frozen_mask(p.grid)


def uniform_with_sources(geometry, T_do=0.01, rho=0.8, p=1.3, ux=0.45, uy=-0.7,
                         kmom=K_MOM_REF, ken=0.7, seed=7):
    """A uniform state on a grid with zero-gradient ghosts on every side (all
    fluxes cancel) under the neutral-H sources of add_MK (kernel.cu:2495-2498):
    momentum gains T_do rho kmom nn2.{x,y}, energy T_do ken nn3, and in
    axisymmetric geometry the r-source acts as in uniform_axisymmetric.  With
    f = T_do v / y (0 when planar): rho' = rho (1 - f),
    u' = u + T_do kmom nn2.x / (1 - f), v' = v + T_do kmom nn2.y / (1 - f),
    p' = (E (1 - f) - f p + T_do ken nn3 - rho'|u'|^2/2)(g - 1).  Returns
    (grid, s, u, nn2, nn3, expected s2, expected u2)."""
    g = G.Grid.from_extent(12, 9, -1.0, 1.0, 0.9, gamma=5.0 / 3.0, krit=0.3,
                           geometry=geometry, bc=(G.BC_OUTFLOW,) * 4)
    nn2, nn3 = neutral_sources(g.M, g.N, seed)
    s = np.empty((g.M, g.N, 2))
    s[..., 0], s[..., 1] = rho, p
    u = np.empty_like(s)
    u[..., 0], u[..., 1] = ux, uy
    if geometry == G.AXISYMMETRIC:
        y = g.y(np.arange(g.M))[:, None] * np.ones((1, g.N))
        f = T_do * uy / y
    else:
        f = np.zeros((g.M, g.N))
    gm = g.gamma
    r2 = rho * (1 - f)
    vx2 = ux + T_do * kmom * nn2[..., 0] / (1 - f)
    vy2 = uy + T_do * kmom * nn2[..., 1] / (1 - f)
    E = p / (gm - 1) + 0.5 * rho * (ux * ux + uy * uy)
    E2 = E * (1 - f) - f * p + T_do * ken * nn3
    p2 = (E2 - 0.5 * r2 * (vx2 * vx2 + vy2 * vy2)) * (gm - 1)
    return g, s, u, nn2, nn3, np.stack([r2, p2], -1), np.stack([vx2, vy2], -1)


# ---- comparison of an FP64 result with the oracle ---------------------------
#
# The FP64 GPU path and the oracle compute the same step with different (but
# algebraically equal) operation orders: the GPU solves each face once, the
# oracle (like the reference) twice with coordinate-difference slopes.  They
# agree to round-off; the bar is 1e-12 relative per step, measured per cell
# against the natural scale of each quantity: rho against its value, each
# velocity component against the local signal speed |u| + c, and p against
# the energy it is recovered from, (gamma-1) E = p + (gamma-1) rho |u|^2/2.
# (The reference forms p' = (E' - rho'|u'|^2/2)(gamma-1): at Mach M the two
# terms cancel to 1 part in 1 + gamma(gamma-1)M^2/2, so its own rounding
# moves p by that many ulps -- ~3000 at the Mach-70 edge of the wind bubble,
# 7e-13 of p; the GPU step forms p' without that cancellation.)

PARITY_REL = 1e-12


def step_mismatch(s2, u2, rs2, ru2, gamma, rel=PARITY_REL):
    """Number of cells (and the worst relative difference) where an FP64 step
    result (s2, u2) differs from the oracle's (rs2, ru2) by more than rel."""
    s2, u2, rs2, ru2 = (np.asarray(a, dtype=np.float64) for a in (s2, u2, rs2, ru2))
    c = np.sqrt(gamma * np.abs(rs2[..., 1]) / np.abs(rs2[..., 0]))
    v2 = ru2[..., 0] ** 2 + ru2[..., 1] ** 2
    vs = np.sqrt(v2) + c
    pe = np.abs(rs2[..., 1]) + (gamma - 1) * 0.5 * np.abs(rs2[..., 0]) * v2
    d = np.stack([np.abs(s2[..., 0] - rs2[..., 0]) / np.abs(rs2[..., 0]),
                  np.abs(s2[..., 1] - rs2[..., 1]) / pe,
                  np.abs(u2[..., 0] - ru2[..., 0]) / vs,
                  np.abs(u2[..., 1] - ru2[..., 1]) / vs], -1)
    d = np.where(np.isnan(d), np.inf, d)
    return int((d > rel).any(-1).sum()), float(d.max())


def flux_scales(F, gamma):
    """Natural scale of each output plane of a face batch F (8, n): the mass
    flux rho (|u| + c), the energy flux (E + p)(|u| + c), the momentum fluxes
    rho (|u| + c)^2 + p, with the larger density, pressure, speed and sound
    speed of the two sides."""
    r = np.maximum(F[0], F[4])
    p = np.maximum(F[1], F[5])
    c = np.sqrt(gamma * np.maximum(F[1] / F[0], F[5] / F[4]))
    v = np.abs(F[[2, 3, 6, 7]]).max(0) + c
    E = p / (gamma - 1) + 0.5 * r * v * v
    mom = r * v * v + p
    return np.stack([r * v, (E + p) * v, mom, mom])


def q_scale(Fq, gamma):
    """Natural scale of the passive scalar's flux on faces Fq (10, n):
    max(|Q_L|, |Q_R|) times the largest speed |u| + c of the two sides."""
    c = np.sqrt(gamma * np.maximum(Fq[1] / Fq[0], Fq[5] / Fq[4]))
    v = np.abs(Fq[[2, 3, 6, 7]]).max(0) + c
    return np.maximum(np.abs(Fq[8]), np.abs(Fq[9])) * v + 1e-300


def flux_mismatch(out, ref, F, gamma, rel=PARITY_REL):
    """(count, worst) of flux values (4 planes) differing from the oracle by
    more than rel of their natural scale, plus dt (plane 4) relative."""
    sc = flux_scales(F, gamma)
    d = np.abs(out[:4] - ref[:4]) / sc
    both_nan = np.isnan(out[:4]) & np.isnan(ref[:4])
    d = np.where(both_nan, 0.0, np.where(np.isnan(d), np.inf, d))
    dt = np.abs(out[4] - ref[4]) / np.abs(ref[4])
    dt = np.where((out[4] == ref[4]), 0.0, np.where(np.isnan(dt), np.inf, dt))
    return int((d > rel).sum() + (dt > rel).sum()), float(max(d.max(), dt.max()))


def pulsed_case(N=48, M=16, pulse=(0.2, 2.0 * math.pi * 100.0, 0.45, 0.004)):
    """The reference's live configuration shrunk to N x M (Dirichlet inflow
    {1, 7, 1, 0} at x-lo, outflow at x-hi, mirrors on y, HLLC faces inside
    x in (10, 40)) with the x-lo side pulsed: density factor 1 + a sin(w TT)
    in the rows with y < y_lim (about half of them) while TT < t_lim.  The
    amplitude and frequency are raised over add2's (0.01, 2 pi) and t_lim
    lowered from 3 so that the pulse is visible and ends within a few steps
    (dt is about 7e-4 here).  State: the inflow state with a
    density bump downstream.  Returns (grid, s, u)."""
    g = G.reference_live_grid(N, M).replace(
        bc=(G.BC_DIRICHLET_PULSED, G.BC_OUTFLOW, G.BC_REFLECT, G.BC_REFLECT), bc_pulse=pulse)
    s = np.empty((M, N, 2))
    u = np.zeros((M, N, 2))
    x = g.x(np.arange(N))[None, :]
    s[..., 0] = 1.0 + 0.5 * np.exp(-((x - 20.0) / 8.0) ** 2)
    s[..., 1] = 7.0
    u[..., 0] = 1.0
    return g, s, u


def pulsed_equivalent(g, TT):
    """The two plain-Dirichlet grids a pulsed step at time TT must match row
    by row: (grid for the rows with y < y_lim, grid for the others) -- the
    first with the inflow density times 1 + a sin(TT w) while TT < t_lim
    (add2, kernel.cu:2022-2025), the second unchanged."""
    a, w, ylim, tlim = g.bc_pulse
    f = (1.0 + a * math.sin(TT * w)) if TT < tlim else 1.0
    st = [list(r) for r in g.bc_state]
    lo = [list(r) for r in g.bc_state]
    lo[G.XLO][0] = st[G.XLO][0] * f
    bc = (G.BC_DIRICHLET,) + tuple(g.bc[1:])
    rows_lo = g.y(np.arange(g.M)) < ylim
    return g.replace(bc=bc, bc_state=lo), g.replace(bc=bc, bc_state=st), rows_lo


# ---- MHD faces (HLLDQ_Korolkov, Gas_Dinamic.cu:874) ---------------------------

SPI4 = 3.544907701811032  # spi4, Header.h:69: the solver works with b = B / SPI4
AXES = [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0),
        (-1.0, 0.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, -1.0)]


def mhd_faces(n, seed, axis=0, continuous_bn=True):
    """(18, n) MHD face pairs (rho, Q, p, v1..v3, Bx..Bz left, then right):
    densities and pressures over 3 and 3.5 decades, plasma beta from ~1e-3 to
    ~1e3, Mach numbers up to ~5; with continuous_bn the field component along
    `axis` is the same on both sides, as in a divergence-free field.  The
    last rows hold the solver's special cases: a zero normal field (the
    Alfven fan switched off), |b_n| just below and above 0.001 (the switch),
    fields along the normal with beta < 1 (t = rho (S-u)(S-SM) - b_n^2 near
    zero: the degenerate star states), and a field-free pair."""
    rng = np.random.default_rng(seed)
    F = np.empty((18, n))
    for k in (0, 9):
        F[k] = np.exp(rng.uniform(-3.0, 3.0, n))
        F[k + 1] = rng.uniform(-2.0, 2.0, n) * F[k]
        F[k + 2] = np.exp(rng.uniform(-4.0, 3.0, n))
        c = np.sqrt(5.0 / 3.0 * F[k + 2] / F[k])
        F[k + 3:k + 6] = rng.normal(0.0, 1.5, (3, n)) * c
        F[k + 6:k + 9] = (rng.normal(0.0, 1.0, (3, n)) * np.sqrt(F[k + 2])
                          * np.exp(rng.uniform(-1.5, 1.5, n)) * SPI4)
    if continuous_bn:
        F[15 + axis] = F[6 + axis]
    m = max(n // 16, 1)
    tail = n - 6 * m
    if tail > 0:
        rows = np.arange(tail, n)
        seg = np.array_split(rows, 6)
        for k in (0, 9):
            F[k + 6 + axis, seg[0]] = 0.0                                   # b_n = 0
            F[k + 6 + axis, seg[1]] = 0.0009 * SPI4 * np.sign(rng.normal(size=len(seg[1])))
            F[k + 6 + axis, seg[2]] = 0.0011 * SPI4 * np.sign(rng.normal(size=len(seg[2])))
            for j in range(3):                                              # field along n
                if j != axis:
                    F[k + 6 + j, seg[3]] = 0.0
            F[k + 6 + axis, seg[3]] = np.sqrt(3.0 * F[k + 2, seg[3]]) * SPI4
            F[k + 6:k + 9, seg[4]] = 0.0                                    # no field
        for s in seg[1:4]:
            F[15 + axis, s] = F[6 + axis, s]
        # stationary contacts: one state, the right density tripled, u_n = 0
        s = seg[5]
        F[9:18, s] = F[0:9, s]
        F[9, s] *= 3.0
        F[3 + axis, s] = F[12 + axis, s] = 0.0
    return F


def mhd_state(W, nvec, gamma):
    """Primitive pieces of (9, n) states along the normal nvec."""
    rho, Q, p = W[0], W[1], W[2]
    v = W[3:6]
    b = W[6:9] / SPI4
    nv = np.asarray(nvec, dtype=np.float64)[:, None]
    un = (v * nv).sum(0)
    bn = (b * nv).sum(0)
    bb = (b * b).sum(0)
    pT = p + 0.5 * bb
    e = p / (gamma - 1.0) + 0.5 * rho * (v * v).sum(0) + 0.5 * bb
    return rho, Q, p, v, b, un, bn, bb, pT, e


def mhd_phys_flux(W, nvec, gamma):
    """The ideal-MHD flux of states W (9, n) through the normal nvec in the
    reference's output order (mass, momentum x/y/z, field x/y/z, energy),
    and the scalar's flux, written in vector form from the equations."""
    rho, Q, p, v, b, un, bn, bb, pT, e = mhd_state(W, nvec, gamma)
    nv = np.asarray(nvec, dtype=np.float64)[:, None]
    out = np.empty((8,) + rho.shape)
    out[0] = rho * un
    out[1:4] = rho * v * un + pT * nv - bn * b
    out[4:7] = SPI4 * (un * b - bn * v)
    out[7] = (e + pT) * un - bn * (v * b).sum(0)
    return out, Q * un


def mhd_fast_speed(W, nvec, gamma):
    """Fast magnetosonic speed along nvec from its closed form."""
    rho, Q, p, v, b, un, bn, bb, pT, e = mhd_state(W, nvec, gamma)
    a2 = gamma * p / rho
    s = a2 + bb / rho
    return np.sqrt(0.5 * (s + np.sqrt(np.maximum(s * s - 4.0 * a2 * bn * bn / rho, 0.0))))


def mhd_scale(F, gamma):
    """Per-face natural scale of the MHD fluxes: the largest of each side's
    (rho + rho|v|^2 + e + pT + |b|^2 + SPI4 |b| |v|) (|v| + c_f) and |Q| (|v| + c_f)."""
    out = np.zeros(F.shape[1])
    for k in (0, 9):
        W = F[k:k + 9]
        rho, Q, p, v, b, un, bn, bb, pT, e = mhd_state(W, (1.0, 0.0, 0.0), gamma)
        vm = np.sqrt((v * v).sum(0))
        cf = np.sqrt((gamma * p + bb) / rho)
        sp = vm + cf
        out = np.maximum(out, (rho + rho * vm * vm + e + pT + bb + SPI4 * np.sqrt(bb) * vm
                               + np.abs(Q)) * sp)
    return out


def mhd_rotational(rng, left_going):
    """An isolated rotational (Alfven) discontinuity with the normal along x:
    rho, p, u_n, b_n and |b_t| continuous, b_t rotated by an angle, and
    [v_t] = +-sign(b_n) [b_t] / sqrt(rho) (from rho (u - S)[v_t] = b_n [b_t]
    at S = u -+ |b_n| / sqrt(rho)); u_n puts x/t = 0 between the
    discontinuity and the contact, so the exact flux is that of the state
    on the far side."""
    rho, p = np.exp(rng.uniform(-1, 1)), np.exp(rng.uniform(-1, 1))
    bn = rng.choice([-1.0, 1.0]) * rng.uniform(0.3, 1.5)
    bt = rng.normal(0, 1, 2)
    phi = rng.uniform(0.3, 3.0)
    c, s = np.cos(phi), np.sin(phi)
    btR = np.array([c * bt[0] - s * bt[1], s * bt[0] + c * bt[1]])
    ca = abs(bn) / np.sqrt(rho)
    sgn = np.sign(bn) * (1.0 if left_going else -1.0)
    vt = rng.normal(0, 1, 2)
    vtR = vt + sgn * (btR - bt) / np.sqrt(rho)
    u_n = (0.5 if left_going else -0.5) * ca
    Q = rng.uniform(-1, 1)
    L = np.array([rho, Q, p, u_n, vt[0], vt[1], bn * SPI4, bt[0] * SPI4, bt[1] * SPI4])
    R = np.array([rho, Q, p, u_n, vtR[0], vtR[1], bn * SPI4, btR[0] * SPI4, btR[1] * SPI4])
    return L, R


# ---- the Monte-Carlo helpers (Gas_Dinamic.cu:1468-1602) -----------------------

def mc_rate_pairs(n, seed):
    """(2, n) pairs (u, cp): speeds over five decades, a share below the
    u = 0.001 series switch and at it, thermal speeds over three decades."""
    rng = np.random.default_rng(seed)
    cp = np.exp(rng.uniform(-3.0, 3.0, n))
    u = cp * np.exp(rng.uniform(-6.0, 2.5, n))
    k = n // 8
    u[:k] = rng.uniform(0.0, 0.001, k)
    u[k:2 * k] = 0.001 * (1.0 + rng.uniform(-1e-9, 1e-9, k))
    u[2 * k] = 0.0
    return np.ascontiguousarray(np.stack([u, cp]))


def mc_particles(grid, n, seed):
    """(8, n) particles (x0, y0, x, y, z, Vx, Vy, Vz) inside random cells of
    the grid (the axis cells included), with random velocities, some purely
    axial, some purely radial, some at rest."""
    rng = np.random.default_rng(seed)
    ci = rng.integers(0, grid.N, n)
    cj = rng.integers(0, grid.M, n)
    x0 = grid.x_min + ci * grid.dx
    y0 = grid.y_min + cj * grid.dy
    x = x0 + rng.uniform(-0.45, 0.45, n) * grid.dx
    r = y0 + rng.uniform(-0.45, 0.45, n) * grid.dy
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    V = rng.normal(0.0, 1.0, (3, n))
    k = n // 10
    V[1:, :k] = 0.0
    V[0, k:2 * k] = 0.0
    V[:, 2 * k:2 * k + 3] = 0.0
    return np.ascontiguousarray(np.stack([x0, y0, x, r * np.cos(phi), r * np.sin(phi),
                                          V[0], V[1], V[2]]))
