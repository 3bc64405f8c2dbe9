"""GPU checks of the neutral-H source terms of add_MK (kernel.cu:2495-2498)
behind the FP64 step (ast_step_f64_src) and the runtime (ast_sim_set_sources):

* the uniform-state closed form of tests/cases.py (planar and axisymmetric,
  the reference's energy coefficient 0 and a non-zero one), to 1e-13;
* the wind problem with a frozen sphere and random moment planes: every step
  against the oracle step from the same input (1e-12 per step), through the
  stand-alone entry point and through the runtime's graph replay;
* zero planes give the source-free step bit for bit;
* the FP32 step with the same planes (ast_step_f32_src): the closed form to
  FP32 accuracy, 8 steps of the wind problem against the FP64 step with the
  planes, zero planes bit for bit the plain FP32 step, and the FP32 runtime
  against the FP64 runtime with the planes;
* the FP32 runtime, which keeps the planes' FP32 roundings, against the
  stand-alone FP32 step reading the FP64 planes: bit for bit;
* at the headline size (2048 x 1024), 40 steps of the FP32 runtime with the
  planes against the FP64 runtime with them."""
import numpy as np
import pytest

from asteriy_2d_amd import grid as G

from cases import K_MOM_REF, PARITY_REL, sources_case, step_mismatch, uniform_with_sources

pytestmark = pytest.mark.gpu


def _dev_step(g, s, u, T_do, method, dev, src_np):
    """One ast_step_f64_src call; src_np = (nn2, nn3, kmom, ken) host arrays."""
    import torch
    from asteriy_2d_amd import solver
    S = torch.from_numpy(np.ascontiguousarray(s)).to(dev)
    U = torch.from_numpy(np.ascontiguousarray(u)).to(dev)
    S2, U2 = torch.empty_like(S), torch.empty_like(U)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev)
    Td = torch.tensor([T_do], dtype=torch.float64, device=dev)
    TT = torch.zeros(1, dtype=torch.float64, device=dev)
    src = None
    if src_np is not None:
        nn2, nn3, kmom, ken = src_np
        src = solver.NeutralSources(None if nn2 is None else torch.from_numpy(nn2).to(dev),
                                    None if nn3 is None else torch.from_numpy(nn3).to(dev),
                                    kmom, ken)
    solver.Kernel_TVD(g, S, U, S2, U2, T, Td, TT, method, sources=src)
    torch.cuda.synchronize()
    return S2.cpu().numpy(), U2.cpu().numpy(), float(T.item())


@pytest.mark.parametrize("geometry", [G.PLANAR, G.AXISYMMETRIC])
@pytest.mark.parametrize("ken", [0.0, 0.7])
def test_sources_closed_form_gpu(dev, geometry, ken):
    print(f"sources closed form geometry={geometry} ken={ken}", flush=True)
    T_do = 0.01
    g, s, u, nn2, nn3, s2w, u2w = uniform_with_sources(geometry, T_do=T_do, ken=ken)
    s2, u2, _ = _dev_step(g, s, u, T_do, G.HLLC, dev, (nn2, nn3, K_MOM_REF, ken))
    np.testing.assert_allclose(s2[..., 0], s2w[..., 0], rtol=1e-13)
    np.testing.assert_allclose(u2, u2w, rtol=1e-13, atol=1e-15)
    E = s[..., 1] / (g.gamma - 1)
    assert np.max(np.abs(s2[..., 1] - s2w[..., 1]) / E) < 1e-13


@pytest.mark.parametrize("method", [G.HLLC, G.HLL, G.HLLC_A])
@pytest.mark.parametrize("planes", ["both", "nn2", "nn3"])
def test_sources_every_step_vs_oracle(dev, ora, method, planes):
    """8 steps of the wind problem (frozen sphere, random planes, energy
    coefficient 0.02): each GPU step against the oracle step from the same
    input, 1e-12 per cell; dt likewise.  The sources enter after the fluxes,
    whatever the method; the exact solver is left out here because the
    nearly uniform inflow that the energy source perturbs by parts in 1e6
    puts faces at its bisection's stopping rule (|f| <= 1e-9, an absolute
    velocity), where inputs that differ in the last bit (the reference's
    coordinate-difference slopes) may stop one iteration apart: 1e-12 of the
    flux is then not a bound of the reference's own solver."""
    import torch
    from asteriy_2d_amd import solver
    print(f"sources every step method={method} planes={planes}", flush=True)
    p, nn2, nn3, fr = sources_case()
    g = p.grid
    nn2 = nn2 if planes in ("both", "nn2") else None
    nn3 = nn3 if planes in ("both", "nn3") else None
    ken = 0.02
    src = solver.NeutralSources(None if nn2 is None else torch.from_numpy(nn2).to(dev),
                                None if nn3 is None else torch.from_numpy(nn3).to(dev),
                                K_MOM_REF, ken)
    osrc = ora.Sources(nn2, nn3, K_MOM_REF, ken)
    S = torch.from_numpy(p.s).to(dev)
    U = torch.from_numpy(p.u).to(dev)
    S2, U2 = torch.empty_like(S), torch.empty_like(U)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev)
    Td = torch.tensor([2e-3], dtype=torch.float64, device=dev)
    TT = torch.zeros(1, dtype=torch.float64, device=dev)
    for k in range(8):
        s_in, u_in, T_do = S.cpu().numpy(), U.cpu().numpy(), float(Td.item())
        solver.Kernel_TVD(g, S, U, S2, U2, T, Td, TT, method, sources=src)
        rs2, ru2, rT = ora.step(g, s_in, u_in, 1e7, T_do, method, src=osrc)
        s2, u2 = S2.cpu().numpy(), U2.cpu().numpy()
        bad, w = step_mismatch(s2, u2, rs2, ru2, g.gamma)
        assert bad == 0, (k, bad, w)
        assert np.array_equal(s2[fr], s_in[fr]) and np.array_equal(u2[fr], u_in[fr])
        assert abs(float(T.item()) - rT) <= PARITY_REL * rT, k
        solver.funk_time(T, Td, TT)
        S, S2, U, U2 = S2, S, U2, U


def test_sources_runtime_graph_vs_oracle(dev, ora):
    """The runtime with source planes (graph replay of 20 steps + 5 direct)
    against 25 oracle steps of the reference protocol: the FP64 step is
    within 1e-12 per step, so 25 steps stay within 1e-10 of the oracle; the
    planes change the result (against a run without them); clearing them
    restores the source-free run bit for bit."""
    from asteriy_2d_amd.sim import Simulation
    print("sources runtime 40x24, 25 steps", flush=True)
    p, nn2, nn3, _ = sources_case()
    g, ken = p.grid, 0.02
    with Simulation(g, G.HLLC, "f64", p.s, p.u) as sim:
        sim.set_sources(nn2, nn3, K_MOM_REF, ken)
        sim.run(25)
        s, u = sim.state()
        t = sim.time()
    ref = ora.Sim(g, p.s, p.u, G.HLLC, src=ora.Sources(nn2, nn3, K_MOM_REF, ken)).advance(25)
    bad, w = step_mismatch(s, u, ref.s, ref.u, g.gamma, rel=1e-10)
    assert bad == 0, w
    assert abs(t["TT"] - ref.TT) <= 1e-10 * ref.TT
    with Simulation(g, G.HLLC, "f64", p.s, p.u) as a, Simulation(g, G.HLLC, "f64", p.s, p.u) as b:
        a.set_sources(nn2, nn3, K_MOM_REF, ken)
        a.run(5)
        a.set_sources(None, None)
        a.set_state(p.s, p.u)
        a.set_time(1e7, 1e-9, 0.0, 0)
        a.run(25)
        b.run(25)
        sa, ua = a.state()
        sb, ub = b.state()
        assert np.array_equal(sa, sb) and np.array_equal(ua, ub)
        assert not np.allclose(u, ub, rtol=1e-6, atol=0)


def test_zero_planes_give_the_plain_step(dev):
    """The reference initialises the moments to zero (kernel.cu:3337-3339):
    zero planes (any coefficients) give ast_step_f64's result bit for bit."""
    print("sources zero planes", flush=True)
    p, nn2, nn3, _ = sources_case()
    z2, z3 = np.zeros_like(nn2), np.zeros_like(nn3)
    a = _dev_step(p.grid, p.s, p.u, 2e-3, G.HLLC, dev, None)
    b = _dev_step(p.grid, p.s, p.u, 2e-3, G.HLLC, dev, (z2, z3, K_MOM_REF, 0.0))
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_sources_refused_where_not_served(dev):
    """Planes of the wrong size, and planes with the learned guess on the
    FP32 step, raise; nothing runs."""
    import torch
    from asteriy_2d_amd import learned, solver
    print("sources refused", flush=True)
    p, nn2, nn3, _ = sources_case()
    g = p.grid
    src = solver.NeutralSources(torch.from_numpy(nn2).to(dev), None)
    S = torch.from_numpy(p.s).to(dev, torch.float32)
    U = torch.from_numpy(p.u).to(dev, torch.float32)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev)
    Td = torch.tensor([1e-9], dtype=torch.float64, device=dev)
    with pytest.raises(ValueError):
        solver.Kernel_TVD(g, S, U, torch.empty_like(S), torch.empty_like(U), T, Td, T, G.EXACT,
                          sources=src, mlp=learned.LearnedGuess.load())
    small = solver.NeutralSources(torch.from_numpy(nn2[:-1]).to(dev).contiguous(), None)
    S64 = torch.from_numpy(p.s).to(dev)
    with pytest.raises(ValueError):
        solver.Kernel_TVD(g, S64, torch.from_numpy(p.u).to(dev), torch.empty_like(S64),
                          torch.empty_like(S64), T, Td, T, G.HLLC, sources=small)


def _dev_step32(g, s, u, T_do, method, dev, src_np):
    """One ast_step_f32_src call on the FP32-rounded state."""
    import torch
    from asteriy_2d_amd import solver
    S = torch.from_numpy(np.ascontiguousarray(s)).to(dev, torch.float32)
    U = torch.from_numpy(np.ascontiguousarray(u)).to(dev, torch.float32)
    S2, U2 = torch.empty_like(S), torch.empty_like(U)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev)
    Td = torch.tensor([T_do], dtype=torch.float64, device=dev)
    TT = torch.zeros(1, dtype=torch.float64, device=dev)
    src = None
    if src_np is not None:
        nn2, nn3, kmom, ken = src_np
        src = solver.NeutralSources(None if nn2 is None else torch.from_numpy(nn2).to(dev),
                                    None if nn3 is None else torch.from_numpy(nn3).to(dev),
                                    kmom, ken)
    solver.Kernel_TVD(g, S, U, S2, U2, T, Td, TT, method, sources=src)
    torch.cuda.synchronize()
    return S2.double().cpu().numpy(), U2.double().cpu().numpy(), float(T.item())


@pytest.mark.parametrize("geometry", [G.PLANAR, G.AXISYMMETRIC])
@pytest.mark.parametrize("ken", [0.0, 0.7])
def test_sources_closed_form_fp32(dev, geometry, ken):
    """The uniform-state closed form through the FP32 step: density to
    FP32's 1e-6, velocity to 1e-6 of its size, pressure to 1e-6 of E."""
    print(f"FP32 sources closed form geometry={geometry} ken={ken}", flush=True)
    T_do = 0.01
    g, s, u, nn2, nn3, s2w, u2w = uniform_with_sources(geometry, T_do=T_do, ken=ken)
    s2, u2, _ = _dev_step32(g, s, u, T_do, G.HLLC, dev, (nn2, nn3, K_MOM_REF, ken))
    np.testing.assert_allclose(s2[..., 0], s2w[..., 0], rtol=1e-6)
    assert np.max(np.abs(u2 - u2w)) < 1e-6 * np.max(np.abs(u2w))
    E = s[..., 1] / (g.gamma - 1)
    assert np.max(np.abs(s2[..., 1] - s2w[..., 1]) / E) < 1e-6


@pytest.mark.parametrize("method", [G.HLLC, G.HLL, G.HLLC_A, G.EXACT])
def test_sources_fp32_vs_fp64(dev, method):
    """8 steps of the wind problem with the planes: the FP32 step against the
    FP64 step with the same planes from the same (FP32-rounded) input each
    step, relative L1 of rho and p below 2e-7 per step; zero planes give the
    plain FP32 step bit for bit."""
    import torch
    print(f"FP32 sources vs FP64 method={method}", flush=True)
    p, nn2, nn3, _ = sources_case()
    g = p.grid
    s = p.s.astype(np.float32).astype(np.float64)
    u = p.u.astype(np.float32).astype(np.float64)
    T_do = 2e-3  # then each step's dt, as funk_time passes it on
    for k in range(8):
        a32 = _dev_step32(g, s, u, T_do, method, dev, (nn2, nn3, K_MOM_REF, 0.02))
        a64 = _dev_step(g, s, u, T_do, method, dev, (nn2, nn3, K_MOM_REF, 0.02))
        for c in (0, 1):
            l1 = np.abs(a32[0][..., c] - a64[0][..., c]).sum() / np.abs(a64[0][..., c]).sum()
            print(f"  step {k} component {c}: relative L1 {l1:.2e}", flush=True)
            assert l1 < 2e-7, (k, c, l1)
        s = a32[0]
        u = a32[1]
        T_do = a64[2]
    z2, z3 = np.zeros_like(nn2), np.zeros_like(nn3)
    b = _dev_step32(g, s, u, T_do, method, dev, (z2, z3, K_MOM_REF, 0.02))
    c = _dev_step32(g, s, u, T_do, method, dev, None)
    assert np.array_equal(b[0], c[0]) and np.array_equal(b[1], c[1])


def test_sources_fp32_runtime(dev):
    """The FP32 runtime with the planes (graph replay) against the FP64
    runtime with them: 40 steps, relative L1 of rho below 5e-7."""
    from asteriy_2d_amd.sim import Simulation
    print("FP32 runtime with sources", flush=True)
    p, nn2, nn3, _ = sources_case()
    res = []
    for prec in ("f32", "f64"):
        with Simulation(p.grid, G.HLLC, prec, p.s, p.u) as sim:
            sim.set_sources(nn2, nn3, K_MOM_REF, 0.02)
            sim.run(40).sync()
            res.append(sim.state())
    l1 = np.abs(res[0][0][..., 0] - res[1][0][..., 0]).sum() / np.abs(res[1][0][..., 0]).sum()
    print(f"  relative L1 of rho after 40 steps: {l1:.2e}", flush=True)
    assert l1 < 5e-7, l1


@pytest.mark.parametrize("method,tile,stats", [(G.HLLC, 0, False), (G.HLLC, 6, False),
                                               (G.HLLC, -9, False), (G.EXACT, 0, False),
                                               (G.EXACT, 0, True)])
def test_sources_fp32_runtime_equals_step(dev, method, tile, stats):
    """The FP32 runtime keeps the planes' FP32 roundings (12 B per cell); its
    steps (a graph replay of 20 steps and direct launches, several strip
    tilings, the exact solver's counting kernels) equal the stand-alone FP32
    step reading the FP64 planes, with funk_time after each, bit for bit."""
    import torch
    from asteriy_2d_amd import solver
    from asteriy_2d_amd.sim import Simulation
    print(f"FP32 runtime vs step with sources method={method} tile={tile} stats={stats}",
          flush=True)
    p, nn2, nn3, _ = sources_case()
    g, steps = p.grid, 26
    with Simulation(g, method, "f32", p.s, p.u) as sim:
        sim.set_sources(nn2, nn3, K_MOM_REF, 0.02)
        if tile:
            sim.set_tile_rows(tile)
        if stats:
            sim.collect_stats()
        sim.run(21, graph=True)
        sim.run(steps - 21, graph=False)
        sa, ua = sim.state()
        ta = sim.time()
    src = solver.NeutralSources(torch.from_numpy(nn2).to(dev), torch.from_numpy(nn3).to(dev),
                                K_MOM_REF, 0.02)
    S = torch.from_numpy(p.s).to(dev, torch.float32)
    U = torch.from_numpy(p.u).to(dev, torch.float32)
    S2, U2 = torch.empty_like(S), torch.empty_like(U)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev)
    Td = torch.tensor([1e-9], dtype=torch.float64, device=dev)
    TT = torch.zeros(1, dtype=torch.float64, device=dev)
    for _ in range(steps):
        solver.Kernel_TVD(g, S, U, S2, U2, T, Td, TT, method, sources=src)
        solver.funk_time(T, Td, TT)
        S, S2, U, U2 = S2, S, U2, U
    torch.cuda.synchronize()
    assert np.array_equal(sa, S.double().cpu().numpy())
    assert np.array_equal(ua, U.double().cpu().numpy())
    assert ta["TT"] == float(TT.item()) and ta["i"] == steps


def test_sources_fp32_full_size(dev):
    """The wind workload at 2048 x 1024 with random planes (kmom = n_H/Kn,
    ken = 0.02): 40 steps of the FP32 runtime against the FP64 runtime,
    relative L1 of rho and p below 1e-6, and the planes' own effect (against
    the runs without them) the same in both precisions to 1e-3 of its size."""
    from asteriy_2d_amd import problems
    from asteriy_2d_amd.sim import Simulation
    print("FP32 sources at 2048x1024", flush=True)
    p = problems.wind(2048, 1024, method=G.HLLC)
    g = p.grid
    rng = np.random.default_rng(5)
    nn2 = rng.uniform(-1e-3, 1e-3, (g.M, g.N, 3))
    nn3 = rng.uniform(-1e-4, 1e-4, (g.M, g.N))
    res = {}
    for prec in ("f32", "f64"):
        for planes in (True, False):
            with Simulation(g, G.HLLC, prec, p.s, p.u) as sim:
                if planes:
                    sim.set_sources(nn2, nn3, K_MOM_REF, 0.02)
                sim.run(40)
                res[prec, planes] = sim.state()[0]
    for c in (0, 1):
        a, b = res["f32", True][..., c], res["f64", True][..., c]
        l1 = np.abs(a - b).sum() / np.abs(b).sum()
        e32 = res["f32", True][..., c] - res["f32", False][..., c]
        e64 = res["f64", True][..., c] - res["f64", False][..., c]
        rel = np.abs(e32 - e64).sum() / np.abs(e64).sum()
        print(f"  component {c}: relative L1 {l1:.2e}; planes' effect {np.abs(e64).sum():.3e}, "
              f"FP32 vs FP64 {rel:.2e}", flush=True)
        assert l1 < 1e-6 and np.abs(e64).sum() > 0 and rel < 1e-3, (c, l1, rel)
