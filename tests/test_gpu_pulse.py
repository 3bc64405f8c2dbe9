"""GPU checks of the pulsed Dirichlet inflow (AST_BC_DIRICHLET_PULSED, add2's
perturbation, kernel.cu:2022-2025) in the FP64 step and the runtime:

* a pulsed step at time TT against the oracle's, and row by row against
  plain Dirichlet steps with the inflow density scaled or not (1e-12);
* the runtime (TT on the device, advanced by funk_time): every step against
  the oracle step from the same input and time, across t_lim; 25 steps of
  graph replay against 25 oracle steps;
* the FP32 paths refuse the pulsed side."""
import numpy as np
import pytest

from asteriy_2d_amd import grid as G

from cases import PARITY_REL, pulsed_case, pulsed_equivalent, step_mismatch

pytestmark = pytest.mark.gpu


def _step(g, s, u, T_do, TT, dev):
    import torch
    from asteriy_2d_amd import solver
    S = torch.from_numpy(np.ascontiguousarray(s)).to(dev)
    U = torch.from_numpy(np.ascontiguousarray(u)).to(dev)
    S2, U2 = torch.empty_like(S), torch.empty_like(U)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev)
    Td = torch.tensor([T_do], dtype=torch.float64, device=dev)
    TTt = torch.tensor([TT], dtype=torch.float64, device=dev)
    solver.Kernel_TVD(g, S, U, S2, U2, T, Td, TTt, G.HLLC)
    torch.cuda.synchronize()
    return S2.cpu().numpy(), U2.cpu().numpy(), float(T.item())


@pytest.mark.parametrize("TT", [0.0, 0.0013, 0.0031, 0.0039, 0.02])
def test_pulsed_step_gpu(dev, ora, TT):
    print(f"pulsed step TT={TT}", flush=True)
    g, s, u = pulsed_case()
    ps, pu, pT = _step(g, s, u, 0.01, TT, dev)
    rs, ru, rT = ora.step(g, s, u, 1e7, 0.01, G.HLLC, TT=TT)
    bad, w = step_mismatch(ps, pu, rs, ru, g.gamma)
    assert bad == 0, w
    assert abs(pT - rT) <= PARITY_REL * rT
    g_lo, g_hi, rows = pulsed_equivalent(g, TT)
    ls, lu, _ = _step(g_lo, s, u, 0.01, 0.0, dev)
    hs, hu, _ = _step(g_hi, s, u, 0.01, 0.0, dev)
    assert step_mismatch(ps[rows], pu[rows], ls[rows], lu[rows], g.gamma)[0] == 0
    assert step_mismatch(ps[~rows], pu[~rows], hs[~rows], hu[~rows], g.gamma)[0] == 0
    assert np.array_equal(ps[~rows], hs[~rows]) and np.array_equal(pu[~rows], hu[~rows])


def test_pulsed_runtime_every_step(dev, ora):
    """12 steps (t_lim = 0.004 is passed after 5 steps): each runtime step from
    the oracle's input and time against the oracle step; the pulse moves the
    result while it is on (against an unpulsed step)."""
    from asteriy_2d_amd.sim import Simulation
    print("pulsed runtime every step", flush=True)
    g, s, u = pulsed_case()
    ref = ora.Sim(g, s, u, G.HLLC)
    on = 0
    with Simulation(g, G.HLLC, "f64") as sim:
        for k in range(12):
            sim.set_state(ref.s, ref.u)
            sim.set_time(1e7, ref.T_do, ref.TT, ref.i)
            TT_in = ref.TT
            sim.run(1, graph=False)
            gs, gu = sim.state()
            t = sim.time()
            ref.advance(1)
            bad, w = step_mismatch(gs, gu, ref.s, ref.u, g.gamma)
            assert bad == 0, (k, w)
            assert abs(t["TT"] - ref.TT) <= PARITY_REL * ref.TT, k
            if TT_in < g.bc_pulse[3]:
                on += 1
        assert 2 <= on < 12
    with Simulation(g, G.HLLC, "f64", s, u) as sim:
        sim.run(25)
        gs, gu = sim.state()
    r25 = ora.Sim(g, s, u, G.HLLC).advance(25)
    bad, w = step_mismatch(gs, gu, r25.s, r25.u, g.gamma, rel=1e-10)
    assert bad == 0, w
    plain = ora.Sim(g.replace(bc=(G.BC_DIRICHLET,) + tuple(g.bc[1:])), s, u, G.HLLC).advance(25)
    assert not np.allclose(r25.s, plain.s, rtol=1e-8, atol=0)


def test_pulsed_refused_by_fp32_paths(dev):
    import torch
    from asteriy_2d_amd import batch as B
    from asteriy_2d_amd import native, solver
    from asteriy_2d_amd.sim import Simulation
    print("pulsed refused by FP32 paths", flush=True)
    g, s, u = pulsed_case()
    with pytest.raises(native.AsteriyError):
        Simulation(g, G.HLLC, "f32", s, u)
    S = torch.from_numpy(s).to(dev, torch.float32)
    U = torch.from_numpy(u).to(dev, torch.float32)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev)
    Td = torch.tensor([1e-9], dtype=torch.float64, device=dev)
    with pytest.raises(native.AsteriyError):
        solver.Kernel_TVD(g, S, U, torch.empty_like(S), torch.empty_like(U), T, Td, T, G.HLLC)
    with pytest.raises(native.AsteriyError):
        B.Batch([g], G.HLLC, [(s, u)])
