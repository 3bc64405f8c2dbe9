"""GPU checks of the closed-form cases that pin the oracle (tests/cases.py):
the update floors (kernel.cu:3060-3073), the axisymmetric source with v != 0
(KorolkovSolver.cu:13-32), and the exact solver's bounded loops on faces the
reference would loop on forever (kernel.cu:647-675, 710-740) -- for the FP64
step, the FP32 step and the batched FP32 launch, each through the C-ABI."""
import numpy as np
import pytest

from asteriy_2d_amd import grid as G

from cases import FLOOR_REGIMES, floor_case, uniform_axisymmetric

pytestmark = pytest.mark.gpu


def _step(grid, s, u, T_do, method, dev, dtype):
    import torch
    from asteriy_2d_amd import solver
    S = torch.from_numpy(np.ascontiguousarray(s)).to(dev, dtype)
    U = torch.from_numpy(np.ascontiguousarray(u)).to(dev, dtype)
    S2, U2 = torch.empty_like(S), torch.empty_like(U)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev)
    Td = torch.tensor([T_do], dtype=torch.float64, device=dev)
    TT = torch.zeros(1, dtype=torch.float64, device=dev)
    solver.Kernel_TVD(grid, S, U, S2, U2, T, Td, TT, method)
    torch.cuda.synchronize()
    return S2.double().cpu().numpy(), U2.double().cpu().numpy()


def _batch_step(grid, s, u, T_do, method):
    from asteriy_2d_amd import batch as B
    with B.Batch([grid], method, [(s, u)]) as bt:
        bt.set_time(0, 1e7, T_do, 0.0, 0)
        bt.run(1, graph=False)
        bt.sync()
        return bt.state(0)


@pytest.mark.parametrize("path", ["f64", "f32"])
@pytest.mark.parametrize("T_do,rho_floor,p_floor", FLOOR_REGIMES)
def test_floors_gpu(dev, ora, path, T_do, rho_floor, p_floor):
    """The middle cell of floor_case against its closed form: no floor, the
    density floor alone (the new kinetic energy then uses the kept density),
    both floors.  FP64 to 1e-13 and equal to the oracle to 1e-12; FP32 to
    2e-6 of the closed form with the floors firing exactly where they do."""
    import torch
    print(f"floors path={path} T_do={T_do}", flush=True)
    g, s, u, want = floor_case(T_do)
    if path == "f64":
        s2, u2 = _step(g, s, u, T_do, G.HLLC, dev, torch.float64)
        rs2, ru2, _ = ora.step(g, s, u, 1e7, T_do, G.HLLC)
        assert np.abs(s2 - rs2).max() <= 1e-12 * np.abs(rs2).max()
        assert np.abs(u2 - ru2).max() <= 1e-12 * (np.abs(ru2).max() + 1)
        tol = 1e-13
    else:
        s2, u2 = _step(g, s, u, T_do, G.HLLC, dev, torch.float32)
        tol = 2e-6
    got = (s2[0, 2, 0], s2[0, 2, 1], u2[0, 2, 0], u2[0, 2, 1])
    p_floor_value = 1e-6 if path == "f64" else float(np.float32(1e-6))
    assert (got[0] == 1.0) == rho_floor, got
    assert (got[1] == p_floor_value) == p_floor, got
    for k in range(4):
        assert abs(got[k] - want[k]) <= tol * (abs(want[k]) + 1), (k, got, want)


@pytest.mark.parametrize("T_do", [r[0] for r in FLOOR_REGIMES])
def test_floors_reconstructed_batch_and_step(dev, ora, T_do):
    """floor_case with every face reconstructed (the batched launch takes no
    face window): the batch equals the FP32 step bit for bit, both follow the
    oracle to FP32 accuracy and floor the same cells the oracle floors."""
    import torch
    print(f"floors reconstructed T_do={T_do}", flush=True)
    g, s, u, _ = floor_case(T_do, first_order=False)
    rs2, ru2, _ = ora.step(g, s, u, 1e7, T_do, G.HLLC)
    s2, u2 = _step(g, s, u, T_do, G.HLLC, dev, torch.float32)
    b2, bu2 = _batch_step(g, s, u, T_do, G.HLLC)
    assert np.array_equal(s2, b2) and np.array_equal(u2, bu2)
    assert np.array_equal(s2[..., 0] == 1.0, rs2[..., 0] == 1.0)
    assert np.array_equal(s2[..., 1] == float(np.float32(1e-6)), rs2[..., 1] == 1e-6)
    assert np.abs(s2 - rs2).max() <= 2e-6 * (np.abs(rs2).max() + 1)
    assert np.abs(u2 - ru2).max() <= 2e-6 * (np.abs(ru2).max() + 1)


@pytest.mark.parametrize("path", ["f64", "f32", "batch"])
@pytest.mark.parametrize("method", [G.HLL, G.HLLC, G.EXACT, G.HLLC_A])
def test_axisymmetric_source_gpu(dev, path, method):
    """uniform_axisymmetric: a uniform state with v != 0 changes only by the
    r-source, rho' = rho (1 - f), p' = p (1 - gamma f), u' = u, with
    f = T_do v / y at the cell-centre radius."""
    import torch
    print(f"r-source path={path} method={method}", flush=True)
    T_do = 0.01
    g, s, u, ws, wu = uniform_axisymmetric(T_do)
    if path == "batch":
        s2, u2 = _batch_step(g, s, u, T_do, method)
    else:
        s2, u2 = _step(g, s, u, T_do, method, dev,
                       torch.float64 if path == "f64" else torch.float32)
    tol = 1e-14 if path == "f64" else 1e-6
    assert np.abs(s2 - ws).max() <= tol * 2
    assert np.abs(u2 - wu).max() <= tol


def _faces(rows):
    """(8, n) SoA of faces given as (rhoL, pL, vxL, vyL, rhoR, pR, vxR, vyR)."""
    return np.ascontiguousarray(np.array(rows, dtype=np.float64).T)


def test_exact_solver_loops_are_bounded(dev, ora):
    """Zero pressure on the low side at rest: a shock + rarefaction face whose
    bracket function is NaN everywhere, on which the reference's bisection
    never stops.  The oracle caps it at AST_ITER_CAP iterations and flags it;
    the FP64 and FP32 flux kernels return with the face counted as capped (or,
    FP32, ended by its bracket collapsing), the other faces of the batch
    unaffected."""
    import torch
    from asteriy_2d_amd import native, solver
    _, _, _, info = ora.godunov_info([1.0, 0.0], [0.0, 0.0], [1.0, 1.0], [0.0, 0.0])
    assert info.capped == 1
    F = _faces([(1.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0),
                (1.0, 1.0, 0.3, 0.1, 0.5, 0.4, -0.2, 0.0),
                (0.8, 2.0, -0.5, 0.0, 1.2, 0.3, 0.6, 0.2)])
    ref = ora.flux_batch(G.EXACT, F)
    st = solver.new_stats(dev)
    out = solver.flux_batch(G.EXACT, torch.from_numpy(F).to(dev), stats=st).cpu().numpy()
    d = native.stats_from_tensor(st)
    assert d["capped"] == 1 and d["exact_faces"] == 3
    assert np.abs(out[:, 1:] - ref[:, 1:]).max() <= 1e-12 * np.abs(ref[:, 1:]).max()
    st = solver.new_stats(dev)
    out32 = solver.flux_batch(G.EXACT, torch.from_numpy(F).to(dev).float(), stats=st)
    torch.cuda.synchronize()
    out32 = out32.double().cpu().numpy()
    assert np.abs(out32[:, 1:] - ref[:, 1:]).max() <= 1e-5 * np.abs(ref[:, 1:]).max()


def test_nan_input_face_terminates(dev, ora):
    """A NaN state matches no wave pattern: every flux kernel returns (NaN
    fluxes for that face), no iteration runs, and the other faces are
    unaffected."""
    import torch
    from asteriy_2d_amd import native, solver
    nan = float("nan")
    F = _faces([(nan, 1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0),
                (1.0, 1.0, 0.3, 0.1, 0.5, 0.4, -0.2, 0.0)])
    ref = ora.flux_batch(G.EXACT, F)
    assert np.isnan(ref[:4, 0]).all()
    for dt in (torch.float64, torch.float32):
        for m in (G.HLL, G.HLLC, G.EXACT, G.HLLC_A):
            st = solver.new_stats(dev)
            out = solver.flux_batch(m, torch.from_numpy(F).to(dev, dt), stats=st)
            torch.cuda.synchronize()
            out = out.double().cpu().numpy()
            r = ora.flux_batch(m, F)
            tol = 1e-12 if dt == torch.float64 else 1e-5
            assert np.abs(out[:, 1] - r[:, 1]).max() <= tol * np.abs(r[:, 1]).max()
            if m == G.EXACT:
                assert np.isnan(out[:4, 0]).all()
                d = native.stats_from_tensor(st)
                assert d["newton_iters"] + d["bisect_iters"] <= 64 and d["capped"] == 0
