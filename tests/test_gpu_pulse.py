"""GPU checks of the pulsed Dirichlet inflow (AST_BC_DIRICHLET_PULSED, add2's
perturbation, kernel.cu:2022-2025) in the FP64 step and the runtime:

* a pulsed step at time TT against the oracle's, and row by row against
  plain Dirichlet steps with the inflow density scaled or not (1e-12);
* the runtime (TT on the device, advanced by funk_time): every step against
  the oracle step from the same input and time, across t_lim; 25 steps of
  graph replay against 25 oracle steps;
* the FP32 step: a pulsed x side row by row, and a pulsed y side whole, bit
  for bit against plain FP32 Dirichlet steps with the density scaled or
  not, on the fast and the boundary strip paths (three tilings, with and
  without a high-order window); against the FP64 pulsed step;
* the FP32 runtime (each step takes the time the previous one used plus its
  own dt): bit for bit the stand-alone FP32 step with funk_time, over graph
  and eager runs across t_lim; against the FP64 runtime;
* the FP32 runtime with the pulse and neutral-H planes together: bit for bit
  the stand-alone FP32 step with both;
* the exact solver with the learned start forced on (its kernels take the
  boundary strip path for the pulsed band) against the plain exact solver;
* the batch refuses the pulsed side (it has no simulation time)."""
import numpy as np
import pytest

from asteriy_2d_amd import grid as G

from cases import (K_MOM_REF, PARITY_REL, neutral_sources, pulsed_case, pulsed_equivalent,
                   step_mismatch)

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


def test_pulsed_refused_by_batch(dev):
    from asteriy_2d_amd import batch as B
    from asteriy_2d_amd import native
    print("pulsed refused by the batch", flush=True)
    g, s, u = pulsed_case()
    with pytest.raises(native.AsteriyError):
        B.Batch([g], G.HLLC, [(s, u)])


def _step32(g, s, u, T_do, TT, dev, tile=None, method=G.HLLC):
    """One FP32 Kernel_TVD at strip tiling `tile` (None: the library's)."""
    import torch
    from asteriy_2d_amd import native, solver
    L = native.load()
    old = L.ast_get_tile_rows()
    if tile is not None:
        L.ast_set_tile_rows(int(tile))
    try:
        S = torch.from_numpy(np.ascontiguousarray(s)).to(dev, torch.float32)
        U = torch.from_numpy(np.ascontiguousarray(u)).to(dev, torch.float32)
        S2, U2 = torch.empty_like(S), torch.empty_like(U)
        T = torch.tensor([1e7], dtype=torch.float64, device=dev)
        Td = torch.tensor([T_do], dtype=torch.float64, device=dev)
        TTt = torch.tensor([TT], dtype=torch.float64, device=dev)
        solver.Kernel_TVD(g, S, U, S2, U2, T, Td, TTt, method)
        torch.cuda.synchronize()
    finally:
        L.ast_set_tile_rows(old)
    return S2.double().cpu().numpy(), U2.double().cpu().numpy(), float(T.item())


def _wide_case(window):
    """pulsed_case on 200 x 64 cells: four bands, so that the first band's
    strips take the fast x-ghost path (kFastX, and kFastE at the y edges)."""
    g, s, u = pulsed_case(N=200, M=64)
    if not window:
        g = g.replace(ho_lo=-np.inf, ho_hi=np.inf)
    return g, s, u


@pytest.mark.parametrize("method", [G.HLLC, G.EXACT])
@pytest.mark.parametrize("tile", [8, 6, -5])
@pytest.mark.parametrize("window", [True, False])
@pytest.mark.parametrize("TT", [0.0013, 0.0031, 0.02])
def test_pulsed_step_fp32_rows(dev, method, tile, window, TT):
    """A pulsed x side at time TT: rows below y_lim equal the plain FP32 step
    with the inflow density scaled by 1 + a sin(w TT) (1 from t_lim on), the
    other rows the unscaled one, bit for bit; 8-row strips take the fast
    paths, 6-row ones the boundary path at the y edges (the exact solver's:
    everywhere but in whole groups of four rows)."""
    print(f"FP32 pulsed x side method={method} tile={tile} window={window} TT={TT}", flush=True)
    g, s, u = _wide_case(window)
    ps, pu, pT = _step32(g, s, u, 1e-3, TT, dev, tile, method)
    g_lo, g_hi, rows = pulsed_equivalent(g, TT)
    ls, lu, _ = _step32(g_lo, s, u, 1e-3, 0.0, dev, tile, method)
    hs, hu, hT = _step32(g_hi, s, u, 1e-3, 0.0, dev, tile, method)
    assert 0 < rows.sum() < g.M
    assert np.array_equal(ps[rows], ls[rows]) and np.array_equal(pu[rows], lu[rows])
    assert np.array_equal(ps[~rows], hs[~rows]) and np.array_equal(pu[~rows], hu[~rows])
    if TT < g.bc_pulse[3]:
        assert not np.array_equal(ps[rows], hs[rows])


@pytest.mark.parametrize("tile", [8, 6])
@pytest.mark.parametrize("TT", [0.0013, 0.02])
def test_pulsed_step_fp32_y_side(dev, tile, TT):
    """Pulsed Dirichlet sides on y: every y-lo ghost row lies below y_lim
    (scaled density), no y-hi ghost row does; the step equals the plain FP32
    step with the y-lo density scaled, bit for bit."""
    print(f"FP32 pulsed y sides tile={tile} TT={TT}", flush=True)
    g0, s, u = _wide_case(False)
    inflow = [1.0, 7.0, 0.0, 0.5]
    st = [list(r) for r in g0.bc_state]
    st[G.YLO] = list(inflow)
    st[G.YHI] = [1.2, 6.0, 0.0, -0.3]
    g = g0.replace(bc=(G.BC_DIRICHLET, G.BC_OUTFLOW, G.BC_DIRICHLET_PULSED,
                       G.BC_DIRICHLET_PULSED), bc_state=st)
    a, w, ylim, tlim = g.bc_pulse
    f = (1.0 + a * np.sin(TT * w)) if TT < tlim else 1.0
    sl = [list(r) for r in st]
    sl[G.YLO][0] = st[G.YLO][0] * f
    gp = g.replace(bc=(G.BC_DIRICHLET, G.BC_OUTFLOW, G.BC_DIRICHLET, G.BC_DIRICHLET),
                   bc_state=sl)
    ps, pu, pT = _step32(g, s, u, 1e-3, TT, dev, tile)
    qs, qu, qT = _step32(gp, s, u, 1e-3, 0.0, dev, tile)
    assert np.array_equal(ps, qs) and np.array_equal(pu, qu) and pT == qT


@pytest.mark.parametrize("TT", [0.0, 0.0013, 0.0039, 0.02])
def test_pulsed_step_fp32_vs_fp64(dev, TT):
    """The FP32 pulsed step against the FP64 pulsed step on the same
    (FP32-rounded) input: relative L1 of rho and p below 1e-7."""
    print(f"FP32 pulsed step vs FP64 TT={TT}", flush=True)
    g, s, u = _wide_case(True)
    s = s.astype(np.float32).astype(np.float64)
    u = u.astype(np.float32).astype(np.float64)
    a32 = _step32(g, s, u, 1e-3, TT, dev)
    a64 = _step(g, s, u, 1e-3, TT, dev)
    for c in (0, 1):
        l1 = np.abs(a32[0][..., c] - a64[0][..., c]).sum() / np.abs(a64[0][..., c]).sum()
        print(f"  component {c}: relative L1 {l1:.2e}", flush=True)
        assert l1 < 1e-7, (c, l1)


@pytest.mark.parametrize("runs", [[(26, True)], [(3, False), (21, True), (2, False)]])
def test_pulsed_runtime_fp32_equals_step(dev, runs):
    """The FP32 runtime on the pulsed case (t_lim is passed after about 22
    of the 26 steps) equals the stand-alone FP32 step with funk_time after
    each, bit for bit, whatever the split of the steps into graph and eager
    runs."""
    import torch
    from asteriy_2d_amd import solver
    from asteriy_2d_amd.sim import Simulation
    print(f"FP32 pulsed runtime vs step runs={runs}", flush=True)
    g, s, u = _wide_case(True)
    steps = sum(n for n, _ in runs)
    with Simulation(g, G.HLLC, "f32", s, u) as sim:
        for n, graph in runs:
            sim.run(n, graph=graph)
        sa, ua = sim.state()
        ta = sim.time()
    S = torch.from_numpy(s).to(dev, torch.float32)
    U = torch.from_numpy(u).to(dev, torch.float32)
    S2, U2 = torch.empty_like(S), torch.empty_like(U)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev)
    Td = torch.tensor([1e-9], dtype=torch.float64, device=dev)
    TT = torch.zeros(1, dtype=torch.float64, device=dev)
    on = 0
    for _ in range(steps):
        on += float(TT.item()) < g.bc_pulse[3]
        solver.Kernel_TVD(g, S, U, S2, U2, T, Td, TT, G.HLLC)
        solver.funk_time(T, Td, TT)
        S, S2, U, U2 = S2, S, U2, U
    torch.cuda.synchronize()
    assert 2 <= on < steps
    assert np.array_equal(sa, S.double().cpu().numpy())
    assert np.array_equal(ua, U.double().cpu().numpy())
    assert ta["TT"] == float(TT.item()) and ta["i"] == steps


def test_pulsed_runtime_fp32_vs_fp64(dev):
    """40 steps of the FP32 runtime against the FP64 runtime on the pulsed
    case: relative L1 of rho below 1e-6; and the pulse's own effect (pulsed
    minus unpulsed run, a few cells next to the inflow) the same in both
    precisions to 1e-3 of its size."""
    from asteriy_2d_amd.sim import Simulation
    print("FP32 pulsed runtime vs FP64", flush=True)
    g, s, u = _wide_case(True)
    gp = g.replace(bc=(G.BC_DIRICHLET,) + tuple(g.bc[1:]))
    res = {}
    for prec in ("f32", "f64"):
        for name, gg in (("pulsed", g), ("plain", gp)):
            with Simulation(gg, G.HLLC, prec, s, u) as sim:
                sim.run(40)
                res[prec, name] = sim.state()[0][..., 0]
    l1 = (np.abs(res["f32", "pulsed"] - res["f64", "pulsed"]).sum()
          / np.abs(res["f64", "pulsed"]).sum())
    e32 = res["f32", "pulsed"] - res["f32", "plain"]
    e64 = res["f64", "pulsed"] - res["f64", "plain"]
    rel = np.abs(e32 - e64).sum() / np.abs(e64).sum()
    print(f"  relative L1 FP32 vs FP64 {l1:.2e}; pulse effect {np.abs(e64).sum():.3e}, "
          f"FP32 vs FP64 {rel:.2e}", flush=True)
    assert l1 < 1e-6
    assert np.abs(e64).sum() > 0 and rel < 1e-3


def test_pulsed_fp32_with_sources_runtime_equals_step(dev):
    """The pulsed case with neutral-H planes: 26 steps of the FP32 runtime
    (graph and eager runs, across t_lim) equal the stand-alone FP32 step with
    both the time and the planes, and funk_time after each, bit for bit."""
    import torch
    from asteriy_2d_amd import solver
    from asteriy_2d_amd.sim import Simulation
    print("FP32 pulsed runtime with sources vs step", flush=True)
    g, s, u = _wide_case(True)
    nn2, nn3 = neutral_sources(g.M, g.N, seed=11, scale=2e-3, scale3=1e-4)
    with Simulation(g, G.HLLC, "f32", s, u) as sim:
        sim.set_sources(nn2, nn3, K_MOM_REF, 0.02)
        sim.run(3, graph=False)
        sim.run(23)
        sa, ua = sim.state()
        ta = sim.time()
    src = solver.NeutralSources(torch.from_numpy(nn2).to(dev), torch.from_numpy(nn3).to(dev),
                                K_MOM_REF, 0.02)
    S = torch.from_numpy(s).to(dev, torch.float32)
    U = torch.from_numpy(u).to(dev, torch.float32)
    S2, U2 = torch.empty_like(S), torch.empty_like(U)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev)
    Td = torch.tensor([1e-9], dtype=torch.float64, device=dev)
    TT = torch.zeros(1, dtype=torch.float64, device=dev)
    for _ in range(26):
        solver.Kernel_TVD(g, S, U, S2, U2, T, Td, TT, G.HLLC, sources=src)
        solver.funk_time(T, Td, TT)
        S, S2, U, U2 = S2, S, U2, U
    torch.cuda.synchronize()
    assert np.array_equal(sa, S.double().cpu().numpy())
    assert np.array_equal(ua, U.double().cpu().numpy())
    assert ta["TT"] == float(TT.item()) and ta["i"] == 26


def test_pulsed_fp32_exact_learned_start(dev):
    """The pulsed case without a window, exact solver: 26 steps with the
    learned-start kernels forced on (they take the boundary strip path for
    the pulsed band) agree with the plain exact solver to 1e-6 (relative L1
    of rho; measured: bit for bit, as no face here is a strong two-shock
    face, the only ones the guess serves) and the pulse acts."""
    from asteriy_2d_amd import learned
    from asteriy_2d_amd.sim import Simulation
    print("FP32 pulsed exact with the learned start", flush=True)
    g, s, u = _wide_case(False)
    model = learned.LearnedGuess.load(device=dev)
    res = {}
    for name, gg, mlp in (("learned", g, model), ("plain", g, None),
                          ("unpulsed", g.replace(bc=(G.BC_DIRICHLET,) + tuple(g.bc[1:])), None)):
        with Simulation(gg, G.EXACT, "f32", s, u) as sim:
            if mlp is not None:
                sim.set_mlp(mlp)
                sim.set_learned_mode(1)
            sim.run(26)
            res[name] = sim.state()[0][..., 0]
    assert np.all(np.isfinite(res["learned"]))
    l1 = np.abs(res["learned"] - res["plain"]).sum() / np.abs(res["plain"]).sum()
    d = np.abs(res["plain"] - res["unpulsed"]).sum() / np.abs(res["unpulsed"]).sum()
    print(f"  learned vs plain {l1:.2e}; pulse effect {d:.2e}", flush=True)
    assert l1 < 1e-6 and d > 1e-7
