"""GPU checks of the Monte-Carlo helpers over device arrays
(ast_mc_rates_f64, ast_mc_cell_exit_f64, ast_mc_belong_point_f64;
Gas_Dinamic.cu:1468-1602) against the oracle: the rates to 1e-14 relative
(the device's exp and erf against the host's: a few ulp; Velosity_2's
closed form, whose leading terms cancel for small u/cp, to 1e-14 of its
terms' size, pi cp^6/(2u)), the cell exits' sides and hits exactly and their
times to 1e-14, the cells of points exactly."""
import numpy as np
import pytest

from asteriy_2d_amd import grid as G

from cases import mc_particles, mc_rate_pairs

pytestmark = pytest.mark.gpu


def test_mc_rates_match_oracle(ora, dev):
    import torch
    from asteriy_2d_amd import solver
    P = mc_rate_pairs(1 << 16, seed=91)
    print("ast_mc_rates_f64 n = 65536", flush=True)
    out = solver.mc_rates(torch.from_numpy(P[0].copy()).to(dev),
                          torch.from_numpy(P[1].copy()).to(dev)).cpu().numpy()
    u, cp = P
    for k in (1, 2, 3):
        ref = np.array([ora.velosity(k, a, c) for a, c in P.T])
        scale = np.abs(ref)
        if k == 2:
            scale = np.maximum(scale, np.pi * cp ** 6 / (2.0 * np.maximum(u, 1e-300)))
        bad = np.abs(out[k - 1] - ref) > 1e-14 * scale
        assert not bad.any(), (k, int(bad.sum()), P[:, bad][:, :3])


def test_mc_cell_exit_matches_oracle(ora, dev):
    import torch
    from asteriy_2d_amd import solver
    g = G.Grid.from_extent(256, 128, -40.0, 40.0, 40.0)
    P = mc_particles(g, 20000, seed=92)
    print("ast_mc_cell_exit_f64 n = 20000", flush=True)
    t, mode, hit = (a.cpu().numpy() for a in solver.mc_cell_exit(g, torch.from_numpy(P).to(dev)))
    for i in range(P.shape[1]):
        h, m, tt = ora.peresechenie(g, *P[:, i])
        assert (int(hit[i]), int(mode[i])) == (int(h), m), i
        assert abs(t[i] - tt) <= 1e-14 * abs(tt), (i, t[i], tt)


def test_mc_belong_point_matches_oracle(ora, dev):
    import torch
    from asteriy_2d_amd import solver
    g = G.Grid.from_extent(256, 128, -40.0, 40.0, 40.0)
    rng = np.random.default_rng(93)
    xy = np.ascontiguousarray(np.stack([rng.uniform(-40.0, 40.0, 20000),
                                        rng.uniform(0.0, 39.99, 20000)]))
    print("ast_mc_belong_point_f64 n = 20000", flush=True)
    nm = solver.mc_belong_point(g, torch.from_numpy(xy).to(dev)).cpu().numpy()
    for i in range(xy.shape[1]):
        assert tuple(int(v) for v in nm[:, i]) == ora.belong_point(g, *xy[:, i]), i
