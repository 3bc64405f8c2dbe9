"""GPU checks of the passive-scalar flux operators (ast_flux_q_f64):
HLLCQ_Korolkov_2D (Gas_Dinamic.cu:249), HLLCQ_Aleksashov (Gas_Dinamic.cu:355)
and the scalar of the live HLLC_2d_Korolkov_b_s (Gas_Dinamic.cu:1604), each
against its own oracle restatement to 1e-12 of the natural scales, under
several normals; the scalar moves with the mass (PQ = phi * mass flux for
Q = phi rho); NaN faces give NaN; bad methods are refused."""
import numpy as np
import pytest

from asteriy_2d_amd import facegen
from asteriy_2d_amd import grid as G

from cases import PARITY_REL, flux_mismatch, q_scale

pytestmark = pytest.mark.gpu

Q_METHODS = [G.HLLC, G.HLLCQ, G.HLLCQ_A]


def _faces_q(n, seed, phi=None):
    F = facegen.faces(n, seed=seed, vacuum_frac=0.0)
    rng = np.random.default_rng(seed + 100)
    if phi is None:
        q = rng.uniform(-2.0, 2.0, (2, n))
        QL, QR = q[0] * F[0], q[1] * F[4]
    else:
        QL, QR = phi * F[0], phi * F[4]
    return np.ascontiguousarray(np.concatenate([F, QL[None], QR[None]]))


def _run(method, Fq, n1, n2, rad, dev):
    import torch
    from asteriy_2d_amd import solver
    out = solver.flux_q_batch(method, torch.from_numpy(Fq).to(dev), n1, n2, rad)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("method", Q_METHODS)
@pytest.mark.parametrize("normal", [(1.0, 0.0), (0.0, -1.0), (-0.6, 0.8)])
def test_flux_q_matches_oracle(dev, ora, method, normal):
    n = 1 << 15
    print(f"flux_q method={method} normal={normal} n={n}", flush=True)
    Fq = _faces_q(n, seed=31)
    out = _run(method, Fq, *normal, 0.25, dev)
    ref = ora.flux_q_batch(method, Fq, *normal, 0.25)
    bad, worst = flux_mismatch(out[:5], ref[:5], Fq[:8], 5.0 / 3.0)
    assert bad == 0, (bad, worst)
    dq = np.abs(out[5] - ref[5]) / q_scale(Fq, 5.0 / 3.0)
    print(f"  worst relative difference {max(worst, dq.max()):.2e}")
    assert dq.max() <= PARITY_REL, dq.max()


@pytest.mark.parametrize("method", Q_METHODS)
def test_flux_q_scalar_moves_with_the_mass(dev, method):
    print(f"flux_q phi-consistency method={method}", flush=True)
    for phi in (1.0, -0.37):
        Fq = _faces_q(1 << 14, seed=32, phi=phi)
        out = _run(method, Fq, 0.6, 0.8, 0.5, dev)
        d = np.abs(out[5] - phi * out[0])
        assert (d / q_scale(Fq, 5.0 / 3.0)).max() <= 1e-13


def test_flux_q_nan_faces_and_errors(dev, ora):
    """Faces with a NaN input: the fluxes and PQ are NaN exactly where the
    oracle's are (a face whose upwind side is finite keeps its finite flux)
    and equal it elsewhere; dt is not compared there (the reference's min and
    max skip a NaN wave speed, the GPU's propagate it).  No hang, no error.
    A method without a scalar is refused with a status code."""
    import torch
    from asteriy_2d_amd import native, solver
    print("flux_q NaN faces and errors", flush=True)
    Fq = _faces_q(256, seed=33)
    Fq[0, 5] = np.nan
    Fq[7, 9] = np.nan
    Fq[9, 11] = np.nan
    for m in Q_METHODS:
        out = _run(m, Fq, 1.0, 0.0, 0.5, dev)
        ref = ora.flux_q_batch(m, Fq, 1.0, 0.0, 0.5)
        planes = [0, 1, 2, 3, 5]
        assert np.array_equal(np.isnan(out[planes]), np.isnan(ref[planes]))
        assert np.isnan(out[planes, 9]).all()
        ok = np.ones(256, bool)
        ok[[5, 9, 11]] = False
        assert np.isfinite(out[:, ok]).all()
        bad, worst = flux_mismatch(out[:5, ok], ref[:5, ok], Fq[:8, ok], 5.0 / 3.0)
        assert bad == 0, worst
        assert (np.abs(out[5, ok] - ref[5, ok]) / q_scale(Fq[:, ok], 5.0 / 3.0)).max() <= PARITY_REL
    with pytest.raises(native.AsteriyError):
        solver.flux_q_batch(G.EXACT, torch.from_numpy(Fq).to(dev), 1.0, 0.0, 0.5)
