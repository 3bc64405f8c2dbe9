"""GPU checks of the MHD flux operators (ast_flux_mhd_f64): HLLDQ_Korolkov
(Gas_Dinamic.cu:874-1415) with metod 2 (HLLC for MHD) and metod 3 (HLLD),
against the oracle's restatement to 1e-12 of the faces' natural scales under
the six axis normals and an oblique one (the reference's frame rule then
keeps a tangent pair that is not orthogonal to the normal), on continuous
and jumping normal fields, the solver's switch cases (b_n = 0, |b_n| at the
0.001 switch, degenerate star states, no field), a ragged large batch, NaN
faces; HLLD resolves isolated rotational discontinuities exactly on the GPU
as well."""
import numpy as np
import pytest

from asteriy_2d_amd import grid as G

from cases import (AXES, PARITY_REL, SPI4, mhd_faces, mhd_phys_flux, mhd_rotational,
                   mhd_scale)

pytestmark = pytest.mark.gpu

METHODS = [G.HLLC_MHD, G.HLLD]
GAMMA = 5.0 / 3.0


def _run(method, F, nvec, rad, dev):
    import torch
    from asteriy_2d_amd import solver
    out = solver.flux_mhd_batch(method, torch.from_numpy(np.ascontiguousarray(F)).to(dev),
                                *nvec, rad)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _mismatch(out, ref, F):
    sc = mhd_scale(F, GAMMA)
    flux = np.abs(out[:9] - ref[:9]) / sc
    dt = np.abs(out[9] - ref[9]) / np.abs(ref[9])
    return flux.max(), dt.max()


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("continuous", [True, False])
def test_flux_mhd_matches_oracle(ora, dev, method, continuous):
    for k, nvec in enumerate(AXES + [(0.6, 0.8, 0.0)]):
        ax = k % 3 if k < 6 else 0
        F = mhd_faces(1 << 14, seed=40 + k, axis=ax, continuous_bn=continuous)
        print(f"ast_flux_mhd_f64 method {method} normal {nvec} continuous {continuous}",
              flush=True)
        out = _run(method, F, nvec, 0.7, dev)
        ref = ora.flux_mhd_batch(method, F, nvec, 0.7)
        fl, dt = _mismatch(out, ref, F)
        assert fl < PARITY_REL, (nvec, fl)
        assert dt < 1e-14, (nvec, dt)


def test_flux_mhd_ragged_large_batch(ora, dev):
    """2^20 + 3 faces (the grid-stride loop's tail) against the oracle."""
    F = mhd_faces((1 << 20) + 3, seed=48, axis=1)
    print("ast_flux_mhd_f64 HLLD n = 2^20 + 3", flush=True)
    out = _run(G.HLLD, F, (0.0, 1.0, 0.0), 0.3, dev)
    ref = ora.flux_mhd_batch(G.HLLD, F, (0.0, 1.0, 0.0), 0.3)
    fl, dt = _mismatch(out, ref, F)
    assert fl < PARITY_REL and dt < 1e-14


@pytest.mark.parametrize("method", METHODS)
def test_flux_mhd_nan_faces(ora, dev, method):
    """A NaN in a face gives the oracle's NaN pattern, and only in that face."""
    F = mhd_faces(256, seed=49, axis=0)
    F[2, 5] = np.nan
    F[13, 77] = np.nan
    F[0, 200] = np.nan
    out = _run(method, F, (1.0, 0.0, 0.0), 1.0, dev)
    ref = ora.flux_mhd_batch(method, F, (1.0, 0.0, 0.0), 1.0)
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    ok = np.ones(F.shape[1], dtype=bool)
    ok[[5, 77, 200]] = False
    assert np.isfinite(out[:, ok]).all()
    fl, dt = _mismatch(out[:, ok], ref[:, ok], F[:, ok])
    assert fl < PARITY_REL and dt < 1e-14


def test_hlld_rotational_discontinuity_on_gpu(dev):
    """Isolated rotational discontinuities: the GPU HLLD flux is the exact
    flux of the state on the far side of the discontinuity."""
    rng = np.random.default_rng(50)
    cols = []
    far = []
    for left_going in (True, False):
        for _ in range(64):
            L, R = mhd_rotational(rng, left_going)
            cols.append(np.concatenate([L, R]))
            far.append(R if left_going else L)
    F = np.ascontiguousarray(np.array(cols).T)
    out = _run(G.HLLD, F, (1.0, 0.0, 0.0), 1.0, dev)
    want, wq = mhd_phys_flux(np.array(far).T, (1.0, 0.0, 0.0), GAMMA)
    sc = mhd_scale(F, GAMMA)
    assert (np.abs(out[:8] - want) / sc).max() < 1e-13
    assert (np.abs(out[8] - wq) / sc).max() < 1e-13
