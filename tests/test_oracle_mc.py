"""CPU checks that pin the oracle's host helpers of the Monte-Carlo
neutral-hydrogen code (Gas_Dinamic.cu:1468-1602).

The rates Velosity_1/2/3 close Maxwellian averages over the relative
velocity w = u - v of a particle (speed u) and a Maxwellian of thermal
speed cp (f ~ exp(-|v|^2/cp^2)): <|w|>, the momentum-exchange integral
Int |w| w_par exp(-|v|^2/cp^2) d^3v, and <|w|^3>/<|w|^2>; they are checked
against those integrals by quadrature, with their series below u = 0.001.
The reference's sqrtpi is 1.77245385 (Header.h:65, nine digits): the closed
forms carry its ~3e-10 relative error, and Velosity_2's closed form, whose
leading terms cancel at small u, jumps by ~1e-5 at its u = 0.001 switch
(checked: a restatement with the exact sqrt(pi) fails it).  Peresechenie is
checked against a brute march of the ray through the cell, peresich by
substitution, Belong_point and minplus by cases."""
import math

import numpy as np
import pytest
from scipy import integrate

from asteriy_2d_amd import grid as G

SQRTPI_REF = 1.77245385


def _maxwell(F, u, cp, norm=True, epsrel=1e-12):
    """Int F(|w|, w_par) exp(-|v|^2/cp^2) d^3v (w = u - v, w_par along u),
    over (|v|, cos angle to u); divided by pi^1.5 cp^3 when norm."""
    def f(mu, v):
        w = math.sqrt(max(v * v + u * u - 2.0 * v * u * mu, 0.0))
        return F(w, u - v * mu) * math.exp(-v * v / (cp * cp)) * 2.0 * math.pi * v * v
    val, _ = integrate.dblquad(f, 0.0, 12.0 * cp, -1.0, 1.0, epsabs=0.0, epsrel=epsrel)
    return val / (math.pi ** 1.5 * cp ** 3) if norm else val


def _rates(u, cp, epsrel=1e-12):
    v1 = _maxwell(lambda w, wp: w, u, cp, epsrel=epsrel)
    v2 = _maxwell(lambda w, wp: w * wp, u, cp, norm=False, epsrel=epsrel)
    v3 = _maxwell(lambda w, wp: w ** 3, u, cp, epsrel=epsrel) / (1.5 * cp * cp + u * u)
    return v1, v2, v3


@pytest.mark.parametrize("u,cp", [(0.1, 1.0), (0.7, 1.3), (2.0, 0.8), (5.0, 1.1), (0.35, 0.2)])
def test_velosity_rates_are_the_maxwellian_averages(ora, u, cp):
    want = _rates(u, cp)
    for k in (1, 2, 3):
        got = ora.velosity(k, u, cp)
        assert abs(got - want[k - 1]) <= 2e-8 * abs(want[k - 1]), (k, got, want[k - 1])


def _closed_forms_40_digits(u, cp):
    """The three rates' closed forms with the exact pi, in 40-digit decimal
    arithmetic (erf by its Taylor series), for small u/cp."""
    import decimal
    D = decimal.Decimal
    with decimal.localcontext() as ctx:
        ctx.prec = 40
        pi = D("3.141592653589793238462643383279502884197")
        spi = pi.sqrt()
        u, cp = D(repr(u)), D(repr(cp))
        x = u / cp
        s, term, k = D(0), x, 0
        while abs(term) > D(10) ** -45:
            s += term / (2 * k + 1)
            k += 1
            term = -term * x * x / k
        erf = 2 / spi * s
        e = (-x * x).exp()
        v1 = cp / spi * e + (u + cp * cp / (2 * u)) * erf
        v2 = cp ** 3 * pi * (e * cp * u * 2 * (cp * cp + 2 * u * u) +
                             spi * (4 * u ** 4 + 4 * cp * cp * u * u - cp ** 4) * erf) / (4 * u * u)
        v3 = (e * cp * (5 * cp * cp + 2 * u * u) / (spi * (3 * cp * cp + 2 * u * u)) +
              (4 * u ** 4 + 12 * cp * cp * u * u + 3 * cp ** 4) * erf /
              (2 * u * (3 * cp * cp + 2 * u * u)))
        return float(v1), float(v2), float(v3)


@pytest.mark.parametrize("u", [0.0, 1e-4, 6e-4, 0.00099])
def test_velosity_series_below_the_switch(ora, u):
    """u < 0.001: the series are the closed forms' Taylor expansions: against
    the closed forms at 40 digits to 1e-14 (Velosity_1 and _3 scaled by the
    truncated sqrtpi they divide by); at u = 0 the closed values 2cp/sqrtpi,
    0 and 8cp/(3 sqrtpi) (the reference's sqrtpi)."""
    cp = 0.9
    if u == 0.0:
        assert ora.velosity(1, 0.0, cp) == 2.0 * cp / SQRTPI_REF
        assert ora.velosity(2, 0.0, cp) == 0.0
        assert ora.velosity(3, 0.0, cp) == 8.0 * cp / (3.0 * SQRTPI_REF)
        return
    want = _closed_forms_40_digits(u, cp)
    scale = (SQRTPI_REF / math.sqrt(math.pi), 1.0, SQRTPI_REF / math.sqrt(math.pi))
    for k in (1, 2, 3):
        got = ora.velosity(k, u, cp) * scale[k - 1]
        assert abs(got - want[k - 1]) <= 1e-14 * abs(want[k - 1]), (k, got, want[k - 1])


def test_velosity_switch_carries_the_reference_sqrtpi(ora):
    """Velosity_1 and _3 are continuous at u = 0.001 to the truncated
    constant's 1e-9; Velosity_2's cancelling closed form is off there by
    ~1e-5 relative -- the reference's nine-digit sqrtpi, kept."""
    cp = 1.0
    lo, hi = 0.001 * (1 - 1e-12), 0.001
    for k in (1, 3):
        a, b = ora.velosity(k, lo, cp), ora.velosity(k, hi, cp)
        assert abs(a - b) <= 1e-9 * abs(a)
    a, b = ora.velosity(2, lo, cp), ora.velosity(2, hi, cp)
    assert 1e-6 < abs(a - b) / abs(a) < 1e-3


def test_minplus_cases(ora):
    assert ora.minplus(-1.0, -2.0) == -1.0
    assert ora.minplus(0.0, 3.0) == 3.0
    assert ora.minplus(2.0, -1.0) == 2.0
    assert ora.minplus(2.0, 3.0) == 2.0
    assert ora.minplus(3.0, 2.0) == 2.0
    assert ora.minplus(1e-12, 1e-12) == -1.0


def test_peresich_roots(ora):
    rng = np.random.default_rng(71)
    for _ in range(500):
        y, z = rng.normal(0, 1, 2)
        Vy, Vz = rng.normal(0, 1, 2)
        R = abs(rng.normal(0, 1.5))
        t1, t2 = ora.peresich(y, z, Vy, Vz, R)
        # closest approach of the line to the axis
        tc = -(y * Vy + z * Vz) / (Vy * Vy + Vz * Vz)
        dmin = math.hypot(y + tc * Vy, z + tc * Vz)
        if dmin > R * (1 + 1e-12):
            assert (t1, t2) == (-1.0, -1.0)
            continue
        assert t1 >= t2
        for t in (t1, t2):
            r = math.hypot(y + t * Vy, z + t * Vz)
            assert abs(r - R) <= 1e-9 * max(R, 1.0), (r, R)


def _grid():
    return G.Grid.from_extent(40, 20, -2.0, 2.0, 1.0)


def _exit_march(g, x0, y0, x, y, z, Vx, Vy, Vz):
    """First time the ray leaves the cell [x0 -+ dx/2] x [y0 -+ dy/2] in r,
    by a fine march and bisection, and the side it leaves by."""
    def inside(t):
        xx, r = x + t * Vx, math.hypot(y + t * Vy, z + t * Vz)
        return abs(xx - x0) <= g.dx / 2 and y0 - g.dy / 2 <= r <= y0 + g.dy / 2
    speed = math.sqrt(Vx * Vx + Vy * Vy + Vz * Vz)
    h = min(g.dx, g.dy) / speed / 4000.0
    t = 0.0
    while inside(t + h):
        t += h
    a, b = t, t + h
    for _ in range(80):
        mid = 0.5 * (a + b)
        a, b = (mid, b) if inside(mid) else (a, mid)
    te = b
    xx, r = x + te * Vx, math.hypot(y + te * Vy, z + te * Vz)
    d = [abs(xx - (x0 + g.dx / 2)), abs(xx - (x0 - g.dx / 2)),
         abs(r - (y0 + g.dy / 2)), abs(r - (y0 - g.dy / 2))]
    return te, 1 + int(np.argmin(d))


def test_peresechenie_matches_a_marched_ray(ora):
    """Particles inside random cells (the axis cell included) with random
    velocities: the exit time and side equal those of a brute march."""
    g = _grid()
    rng = np.random.default_rng(72)
    checked = 0
    for _ in range(300):
        n, m = int(rng.integers(0, g.N)), int(rng.integers(0, g.M))
        x0, y0 = g.x_min + n * g.dx, g.y_min + m * g.dy
        x = x0 + rng.uniform(-0.45, 0.45) * g.dx
        r = y0 + rng.uniform(-0.45, 0.45) * g.dy
        phi = rng.uniform(0, 2 * math.pi)
        y, z = r * math.cos(phi), r * math.sin(phi)
        Vx, Vy, Vz = rng.normal(0, 1, 3)
        hit, mode, t = ora.peresechenie(g, x0, y0, x, y, z, Vx, Vy, Vz)
        te, side = _exit_march(g, x0, y0, x, y, z, Vx, Vy, Vz)
        assert hit
        assert mode == side, (mode, side, t, te)
        assert abs(t - te) <= 1e-9 * te + 1e-12
        checked += 1
    assert checked == 300


def test_peresechenie_edge_cases(ora):
    g = _grid()
    x0, y0 = g.x_min + 5 * g.dx, g.y_min + 3 * g.dy
    # purely axial motion: only the x faces
    hit, mode, t = ora.peresechenie(g, x0, y0, x0, y0, 0.0, 2.0, 0.0, 0.0)
    assert hit and mode == 1 and abs(t - g.dx / 4) < 1e-15
    hit, mode, t = ora.peresechenie(g, x0, y0, x0, y0, 0.0, -2.0, 0.0, 0.0)
    assert hit and mode == 2
    # at rest: no face
    hit, mode, t = ora.peresechenie(g, x0, y0, x0, y0, 0.0, 0.0, 0.0, 0.0)
    assert not hit and mode == 0


def test_belong_point_cases(ora):
    g = _grid()
    rng = np.random.default_rng(73)
    for _ in range(1000):
        x = rng.uniform(g.x_min - 0.49 * g.dx, g.x_min + (g.N - 0.51) * g.dx)
        y = rng.uniform(0.0, g.M * g.dy * 0.999)
        n, m = ora.belong_point(g, x, y)
        assert abs(x - (g.x_min + n * g.dx)) <= 0.5 * g.dx * (1 + 1e-12)
        assert abs(y - (g.y_min + m * g.dy)) <= 0.5 * g.dy * (1 + 1e-12)
