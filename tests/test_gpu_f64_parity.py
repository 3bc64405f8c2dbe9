"""GPU parity of the FP64 path against the CPU oracle (through the C-ABI).

The FP64 step solves each face once and reconstructs from neighbour
differences; the oracle, like the reference, solves every face from both of
its cells with coordinate-difference slopes.  The two agree to round-off: the
bar is 1e-12 relative per step (tests/cases.py: rho and p against their
value, velocities against |u| + c; fluxes against their natural scale), and
the exact solver's wave patterns and iteration counts are equal.
"""
import numpy as np
import pytest

from asteriy_2d_amd import grid as G
from asteriy_2d_amd import facegen, problems

from cases import PARITY_REL, flux_mismatch, step_mismatch

pytestmark = pytest.mark.gpu

REL = PARITY_REL


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("method", [G.HLL, G.HLLC, G.EXACT, G.HLLC_A])
def test_flux_f64_matches_oracle(dev, ora, method):
    import torch
    from asteriy_2d_amd import solver
    n = 1 << 16
    faces = facegen.faces(n, seed=7)
    print(f"flux_f64 method={method} n={n}", flush=True)
    out = solver.flux_batch(method, _t(faces, dev), 1.0, 0.0, 0.25).cpu().numpy()
    torch.cuda.synchronize()
    ref = ora.flux_batch(method, faces, 1.0, 0.0, 0.25)
    bad, worst = flux_mismatch(out, ref, faces, 5.0 / 3.0)
    print(f"  worst relative difference {worst:.2e}")
    assert bad == 0, (bad, worst)


@pytest.mark.parametrize("normal", [(1.0, 0.0), (0.0, -1.0), (-1.0, 0.0), (0.0, 1.0)])
def test_flux_f64_normals(dev, ora, normal):
    from asteriy_2d_amd import solver
    faces = facegen.faces(4096, seed=11)
    for method in (G.HLL, G.HLLC, G.EXACT, G.HLLC_A):
        out = solver.flux_batch(method, _t(faces, dev), *normal, 0.5).cpu().numpy()
        ref = ora.flux_batch(method, faces, *normal, 0.5)
        bad, worst = flux_mismatch(out, ref, faces, 5.0 / 3.0)
        assert bad == 0, (method, bad, worst)


def _gpu_step(grid, s, u, T, T_do, method, dev, stats=None):
    import torch
    from asteriy_2d_amd import solver
    S, U = _t(s, dev), _t(u, dev)
    S2, U2 = torch.empty_like(S), torch.empty_like(U)
    Tt = torch.tensor([T], dtype=torch.float64, device=dev)
    Td = torch.tensor([T_do], dtype=torch.float64, device=dev)
    TT = torch.zeros(1, dtype=torch.float64, device=dev)
    solver.Kernel_TVD(grid, S, U, S2, U2, Tt, Td, TT, method, stats=stats)
    torch.cuda.synchronize()
    return S2.cpu().numpy(), U2.cpu().numpy(), float(Tt.item())


def _grids():
    base = G.Grid.from_extent(37, 21, -1.0, 2.0, 1.5, gamma=5.0 / 3.0, krit=0.3)
    yield "planar-ref-bcs", base.replace(geometry=G.PLANAR)
    yield "axi-ref-bcs", base.replace(geometry=G.AXISYMMETRIC)
    yield "axi-outflow", base.replace(geometry=G.AXISYMMETRIC,
                                      bc=(G.BC_OUTFLOW, G.BC_OUTFLOW, G.BC_REFLECT, G.BC_OUTFLOW))
    yield "clamped+dirichlet", base.replace(
        geometry=G.AXISYMMETRIC, bc=(G.BC_OUTFLOW_CLAMPED, G.BC_DIRICHLET, G.BC_REFLECT,
                                     G.BC_OUTFLOW_CLAMPED),
        bc_state=[[0.0] * 4, [0.5, 0.3, -0.7, 0.1], [0.0] * 4, [0.0] * 4],
        bc_vclamp=(-0.2, 0.0, 0.0, 0.15))
    yield "window+frozen", base.replace(geometry=G.AXISYMMETRIC, ho_lo=-0.3, ho_hi=1.1,
                                        frozen_x0=0.4, frozen_ax=0.5, frozen_ay=0.4)
    yield "tiny-2x2", G.Grid.from_extent(2, 2, 0.0, 1.0, 1.0, krit=0.2,
                                         bc=(G.BC_REFLECT, G.BC_REFLECT, G.BC_REFLECT,
                                             G.BC_REFLECT))


@pytest.mark.parametrize("method", [G.HLL, G.HLLC, G.EXACT, G.HLLC_A])
@pytest.mark.parametrize("name,grid", list(_grids()), ids=[n for n, _ in _grids()])
def test_step_f64_matches_oracle(dev, ora, name, grid, method):
    print(f"step_f64 {name} method={method} N={grid.N} M={grid.M}", flush=True)
    s, u = problems.random_state(grid, seed=3)
    T_do = 0.01
    s2, u2, T = _gpu_step(grid, s, u, 1e7, T_do, method, dev)
    rs2, ru2, rT = ora.step(grid, s, u, 1e7, T_do, method)
    bad, worst = step_mismatch(s2, u2, rs2, ru2, grid.gamma)
    print(f"  worst relative difference {worst:.2e}")
    assert bad == 0, (bad, worst)
    assert abs(T - rT) <= REL * rT


def test_reference_live_configuration(dev, ora):
    """The configuration main() runs: Dirichlet inflow at n=0, HLLC only on
    reconstructed faces with x in (10,40), first-order HLL elsewhere
    (kernel.cu:2799-2857), on a 64x8 slice of its grid."""
    g = G.reference_live_grid(N=640, M=8)
    s, u = problems.random_state(g, seed=5)
    for method in (G.HLLC,):
        s2, u2, T = _gpu_step(g, s, u, 1e7, 1e-4, method, dev)
        rs2, ru2, rT = ora.step(g, s, u, 1e7, 1e-4, method)
        assert step_mismatch(s2, u2, rs2, ru2, g.gamma)[0] == 0
        assert abs(T - rT) <= REL * rT


def test_sod_200_steps_every_step(dev, ora):
    """Sod (256x128, HLL, CFL 0.4, axisymmetric, lagged dt from 1e-9): each
    GPU step compared with the oracle step started from the same input."""
    import torch
    from asteriy_2d_amd import solver
    p = problems.sod()
    g = p.grid
    S, U = _t(p.s, dev), _t(p.u, dev)
    S2, U2 = torch.empty_like(S), torch.empty_like(U)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev)
    Td = torch.tensor([1e-9], dtype=torch.float64, device=dev)
    TT = torch.zeros(1, dtype=torch.float64, device=dev)
    it = torch.zeros(1, dtype=torch.int32, device=dev)
    for k in range(p.steps):
        s_in, u_in = S.cpu().numpy(), U.cpu().numpy()
        T_do = float(Td.item())
        solver.Kernel_TVD(g, S, U, S2, U2, T, Td, TT, p.method)
        T_after = float(T.item())
        rs2, ru2, rT = ora.step(g, s_in, u_in, 1e7, T_do, p.method)
        bad, worst = step_mismatch(S2.cpu().numpy(), U2.cpu().numpy(), rs2, ru2, g.gamma)
        assert bad == 0, (k, bad, worst)
        assert abs(T_after - rT) <= REL * rT, f"step {k}"
        solver.funk_time(T, Td, TT, it)
        S, S2, U, U2 = S2, S, U2, U
    assert int(it.item()) == p.steps
    # the tube has moved: the solution is not the initial state
    assert float(TT.item()) > 0.05


def test_step_deterministic(dev):
    import torch
    p = problems.wind(256, 128)
    outs = []
    for _ in range(2):
        s2, u2, T = _gpu_step(p.grid, p.s, p.u, 1e7, 0.01, G.HLLC, dev)
        outs.append((s2, u2, T))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2]


def test_exact_stats_counted(dev, ora):
    import torch
    from asteriy_2d_amd import native, solver
    faces = facegen.faces(8192, seed=2)
    st = solver.new_stats(dev)
    solver.flux_batch(G.EXACT, _t(faces, dev), stats=st)
    d = native.stats_from_tensor(st)
    _, it = ora.flux_batch(G.EXACT, faces, want_iters=True)
    assert d["exact_faces"] == 8192
    assert d["newton_faces"] == int((it[:, 1] == 2).sum())
    assert d["bisect_faces"] == int((it[:, 1] == 1).sum())
    assert d["vacuum_faces"] == int((it[:, 1] == -1).sum())
    assert d["newton_iters"] == int(it[it[:, 1] == 2, 0].sum())


def test_funk_time(dev, ora):
    import torch
    from asteriy_2d_amd import solver
    T = torch.tensor([0.125], dtype=torch.float64, device=dev)
    Td = torch.tensor([1e-9], dtype=torch.float64, device=dev)
    TT = torch.tensor([2.0], dtype=torch.float64, device=dev)
    i = torch.tensor([41], dtype=torch.int32, device=dev)
    solver.funk_time(T, Td, TT, i)
    assert (T.item(), Td.item(), TT.item(), i.item()) == ora.funk_time(0.125, 1e-9, 2.0, 41)


def test_errors_are_status_codes(dev):
    import torch
    from asteriy_2d_amd import native, solver
    g = problems.sod(16, 8).grid.replace(N=1)
    S = torch.zeros(16, dtype=torch.float64, device=dev)
    T = torch.zeros(1, dtype=torch.float64, device=dev)
    with pytest.raises(native.AsteriyError):
        solver.Kernel_TVD(g, S, S, S.clone(), S.clone(), T, T, T, G.HLL)
