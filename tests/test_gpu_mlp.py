"""GPU tests of the learned p* guess: the solver's face log, FP32 and BF16
MFMA forward passes, the training gradient and Adam, each against the numpy
oracle (oracle/mlp_ref.py), and what the trained network gains."""
import os

import numpy as np
import pytest

from asteriy_2d_amd import facegen, learned, problems
from asteriy_2d_amd import grid as G

pytestmark = pytest.mark.gpu

GAMMA = 5.0 / 3.0


def _t(a, dev, dt=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return t.to(dt) if dt is not None else t


def test_face_log_matches_oracle(dev, ora):
    from asteriy_2d_amd import learned
    from oracle import mlp_ref as R
    F = facegen.faces(4096, seed=31)
    feats, target, mask = learned.log_faces(_t(F, dev), GAMMA)
    X, lp, st = R.features(F, GAMMA)
    # FP32 logs and ratios (velocity features reach |u/c_ref| ~ 1e3)
    assert (np.abs(feats.cpu().numpy().T - X) / (1 + np.abs(X))).max() < 2e-5
    ts = R.two_shock(*st, GAMMA)
    assert np.array_equal(mask.cpu().numpy().astype(bool), ts)
    _, it, pstar = ora.godunov_guess_batch(F, None, gamma=GAMMA)
    sel = it[:, 1] == 2
    assert np.array_equal(sel, ts)
    tgt = np.log(pstar[sel]) - lp[sel]
    assert np.abs(target.cpu().numpy()[sel] - tgt).max() < 2e-5


@pytest.mark.parametrize("n", [1, 63, 64, 1000, 4099])
def test_forward_f32_and_bf16_match_oracle(dev, n):
    import torch
    from asteriy_2d_amd import learned
    from oracle import mlp_ref as R
    rng = np.random.default_rng(n)
    p = learned.glorot_init(5)
    p[-1] = 0.25
    p[192:224] = rng.uniform(-0.2, 0.2, 32)  # non-zero b1
    model = learned.LearnedGuess(p, dev)
    X = rng.standard_normal((n, 6)).astype(np.float32)
    feats = _t(X.T, dev)
    y32 = model.forward(feats, 0).cpu().numpy()
    ybf = model.forward(feats, 1).cpu().numpy()
    yref, _ = R.forward(p.astype(np.float64), X.astype(np.float64))
    # FP32 masters on the vector ALU: FP32 rounding + the exp2/rcp tanh
    assert np.abs(y32 - yref).max() < 1e-4
    # BF16 weights and activations (8-bit mantissa), FP32 accumulation
    err = np.abs(ybf - yref)
    print(f"bf16 MFMA forward: mean |err| {err.mean():.2e}, max {err.max():.2e}")
    assert err.mean() < 1e-2 and err.max() < 6e-2


def test_gradient_matches_oracle(dev):
    from asteriy_2d_amd import learned
    from oracle import mlp_ref as R
    rng = np.random.default_rng(7)
    n = 3000
    p = learned.glorot_init(9)
    X = rng.standard_normal((n, 6)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    m = (rng.random(n) < 0.6).astype(np.uint8)
    model = learned.LearnedGuess(p, dev)
    loss, grad = model.loss_grad(_t(X.T, dev), _t(t, dev), _t(m, dev))
    lref, gref = R.loss_grad(p.astype(np.float64), X.astype(np.float64), t.astype(np.float64), m)
    g = grad.cpu().numpy()
    assert abs(loss.item() - lref) <= 1e-4 * lref
    assert np.abs(g - gref).max() <= 1e-4 * np.abs(gref).max()


def test_gradient_deterministic(dev):
    from asteriy_2d_amd import learned
    rng = np.random.default_rng(8)
    X = _t(rng.standard_normal((6, 5000)).astype(np.float32), dev)
    t = _t(rng.standard_normal(5000).astype(np.float32), dev)
    model = learned.LearnedGuess(learned.glorot_init(2), dev)
    a = model.loss_grad(X, t)[1].cpu().numpy()
    b = model.loss_grad(X, t)[1].cpu().numpy()
    assert np.array_equal(a, b)


def test_adam_matches_oracle(dev):
    import torch
    from asteriy_2d_amd import learned
    from oracle import mlp_ref as R
    rng = np.random.default_rng(3)
    p = learned.glorot_init(4)
    g = rng.standard_normal(p.size).astype(np.float32) * 0.01
    model = learned.LearnedGuess(p, dev)
    state = torch.zeros(2 * p.size, dtype=torch.float32, device=dev)
    model.adam(_t(g, dev), state, 1, 1e-3)
    model.adam(_t(g * 0.5, dev), state, 2, 1e-3)
    pr, m, v = R.adam(p.astype(np.float64), g.astype(np.float64), 0 * p, 0 * p, 1, 1e-3)
    pr, m, v = R.adam(pr, 0.5 * g.astype(np.float64), m, v, 2, 1e-3)
    assert np.abs(model.params() - pr).max() < 1e-6


def test_short_training_reduces_loss(dev):
    from asteriy_2d_amd import learned
    model, (vf, vt) = learned.train(n_faces=200_000, steps=150, batch=8192, lr=3e-3,
                                    log=lambda *a: None)
    fresh = learned.LearnedGuess(seed=1, device=dev)
    l0 = fresh.loss_grad(vf, vt)[0].item()
    l1 = model.loss_grad(vf, vt)[0].item()
    print(f"validation loss {l0:.4f} -> {l1:.4f}")
    assert l1 < 0.5 * l0


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..",
                                                    "asteriy_2d_amd", "data", "mlp_guess.npy")),
                    reason="trained parameters not present")
def test_trained_guess_gain(dev):
    """Held-out two-shock faces of the microbenchmark distribution: the FP64
    exact solver seeded by the trained network converges to the same p*
    (reference stopping rule) in fewer newton iterations than from the
    reference's PVRS start."""
    import torch
    from asteriy_2d_amd import learned
    model = learned.LearnedGuess.load(device=dev)
    F = _t(facegen.faces(1 << 18, seed=4242), dev)
    _, _, mask = learned.log_faces(F, GAMMA)
    sel = mask.bool()
    F2 = F[:, sel].contiguous()
    g = learned.guesses(model, F2, GAMMA, precision=1)
    out0, info0 = learned.solve_with_guess(F2, None, GAMMA)
    out1, info1 = learned.solve_with_guess(F2, g, GAMMA)
    it0 = info0[:, 0].double().mean().item()
    it1 = info1[:, 0].double().mean().item()
    print(f"two-shock faces {F2.shape[1]}: newton iterations PVRS {it0:.3f} -> learned {it1:.3f}")
    assert it1 < it0 - 0.5
    assert int((info1[:, 1] & 8).sum()) == 0  # never capped
    rel = ((out1 - out0).abs() / (out0.abs().max(0).values + 1e-300)).max().item()
    assert rel < 1e-8, rel


def _colliding(N=128, M=64):
    """Two streams colliding at z = 0.5 (Mach ~3.9 each): strong two-shock faces."""
    g = G.Grid.from_extent(N, M, 0.0, 1.0, 0.5, krit=0.3, geometry=G.PLANAR,
                           bc=(G.BC_OUTFLOW, G.BC_OUTFLOW, G.BC_REFLECT, G.BC_OUTFLOW))
    x = g.x(np.arange(N))[None, :]
    s = np.ones((M, N, 2))
    u = np.zeros((M, N, 2))
    u[..., 0] = np.where(x < 0.5, 5.0, -5.0)
    return g, s, u


def test_gated_learned_step_without_strong_faces_is_the_plain_step(dev):
    """No strong two-shock face in the wind workload: the gated simulation
    never switches to the learned kernel and steps bit for bit like the one
    without a network."""
    from asteriy_2d_amd.sim import Simulation
    p = problems.wind(256, 128)
    model = learned.LearnedGuess.load(device=dev)
    out = []
    for mlp in (None, model):
        with Simulation(p.grid, G.EXACT, "f32", p.s, p.u) as sim:
            if mlp is not None:
                sim.set_mlp(mlp)
            sim.collect_stats(True)
            sim.run(12)
            out.append((sim.state(), sim.time(), sim.stats()))
    (sa, ta, ca), (sb, tb, cb) = out
    assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1]) and ta == tb
    assert cb["mlp_faces"] == 0


def test_learned_mode_follows_strong_faces(dev):
    """Colliding streams: the first run of 20 steps meets strong two-shock
    faces, so the automatic choice (the flag of the run two before) gives the
    runs after it (0 plain, 1 learned); a run without strong faces switches
    the run two later back to the plain solver."""
    from asteriy_2d_amd.sim import Simulation
    g, s, u = _colliding()
    model = learned.LearnedGuess.load(device=dev)
    with Simulation(g, G.EXACT, "f32", s, u) as sim:
        sim.set_mlp(model)
        assert sim.learned_modes() == (0, 0)
        sim.run(20)
        assert sim.learned_modes() == (0, 1)
        sim.run(20)
        m = sim.learned_modes()
        assert m[0] == 1
        sim.run(60)
        assert sim.learned_modes() == (0, 0)  # the collision has passed


def test_plain_runs_flag_strong_faces_every_fourth_run(dev):
    """Plain runs flag strong faces only every fourth run (the flag costs the
    plain kernel), forced plain runs never.  The collision's strong faces
    last a few steps, so the colliding state is restored (set_state keeps
    the run count) before runs 4 and 5: run 4 (4 % 4 = 0) flags them and
    run 6 takes the learned start; run 5 meets them without flagging, so run
    7 does not."""
    from asteriy_2d_amd.sim import Simulation
    print("plain runs flag every fourth run", flush=True)
    g, s, u = _colliding()
    model = learned.LearnedGuess.load(device=dev)
    with Simulation(g, G.EXACT, "f32", s, u) as sim:
        sim.set_mlp(model)
        sim.set_learned_mode(0)
        for _ in range(4):
            sim.run(4)  # even: one run per call
        sim.set_learned_mode(-1)
        assert sim.learned_modes() == (0, 0)  # runs 4, 5: the forced runs flagged nothing
        sim.set_state(s, u)
        sim.run(4)                             # run 4 flags the collision
        assert sim.learned_modes() == (0, 1)  # run 5 from run 3, run 6 from run 4
        sim.set_state(s, u)
        sim.run(4)                             # run 5 meets it without the flag
        assert sim.learned_modes() == (1, 0)  # run 6 from run 4, run 7 from run 5


def test_gated_learned_runs_step_like_the_plain_solver(dev):
    """100 steps (five runs: two flagging, two quiet, one flagging) of the
    wind workload with the network attached equal the plain exact solver's
    bit for bit: no face is strong, and the flag changes no result."""
    from asteriy_2d_amd.sim import Simulation
    print("gated learned runs vs plain solver, 100 steps", flush=True)
    p = problems.wind(256, 128)
    model = learned.LearnedGuess.load(device=dev)
    out = []
    for mlp in (None, model):
        with Simulation(p.grid, G.EXACT, "f32", p.s, p.u) as sim:
            if mlp is not None:
                sim.set_mlp(mlp)
            sim.run(100)
            out.append((sim.state(), sim.time()))
    (sa, ta), (sb, tb) = out
    assert np.array_equal(sa[0], sb[0]) and np.array_equal(sa[1], sb[1]) and ta == tb


def test_learned_start_serves_strong_faces(dev):
    """Colliding streams with the learned start on for every run: its faces
    start newton from the network, the result agrees with the plain solver to
    the FP32 newton tolerance, and graph replay and eager launches (the same
    runs) agree bit for bit."""
    from asteriy_2d_amd.sim import Simulation
    g, s, u = _colliding()
    model = learned.LearnedGuess.load(device=dev)
    res = {}
    for tag, mlp, graph in (("plain", None, True), ("learned", model, True),
                            ("learned_eager", model, False)):
        with Simulation(g, G.EXACT, "f32", s, u) as sim:
            if mlp is not None:
                sim.set_mlp(mlp)
                sim.set_learned_mode(1)
            sim.collect_stats(True)
            sim.run(25, graph=graph)  # graph: one replay of 20 steps + 5 direct
            res[tag] = (sim.state(), sim.stats())
    (a, ca), (b, cb), (c, _) = res["plain"], res["learned"], res["learned_eager"]
    assert ca["mlp_faces"] == 0 and cb["mlp_faces"] > 0, (ca, cb)
    assert cb["newton_faces"] > 0
    for k in (0, 1):
        rel = np.abs(b[0][..., k] - a[0][..., k]).sum() / np.abs(a[0][..., k]).sum()
        assert rel < 1e-5, (k, rel)
    assert np.array_equal(b[0], c[0]) and np.array_equal(b[1], c[1])
