"""GPU tests of the batched persistent launch: every instance of a batch
steps exactly as the single-grid FP32 step would (bit for bit: the same
strip code, exact dt min, per-instance time), whatever the tile heights,
with or without graph capture."""
import numpy as np
import pytest

from asteriy_2d_amd import grid as G

pytestmark = pytest.mark.gpu


def _instances():
    from asteriy_2d_amd import batch
    return batch.sweep(count=5, seed=3, shapes=[(256, 128), (130, 70), (300, 64), (64, 200)])


@pytest.mark.parametrize("method", [G.HLLC, G.HLLC_A])
def test_batch_equals_single_grid(dev, method):
    from asteriy_2d_amd.batch import Batch
    from asteriy_2d_amd.sim import Simulation
    inst = _instances()
    with Batch([g for g, _, _, _ in inst], method, [(s, u) for _, s, u, _ in inst]) as b:
        b.run(23)
        got = [b.state(k) for k in range(len(inst))]
        times = [b.time(k) for k in range(len(inst))]
    for k, (g, s, u, _) in enumerate(inst):
        print(f"instance {k}: {g.N}x{g.M}", flush=True)
        with Simulation(g, method, "f32", s, u) as sim:
            sim.run(23)
            rs, ru = sim.state()
            t = sim.time()
        assert np.array_equal(got[k][0], rs) and np.array_equal(got[k][1], ru)
        assert times[k]["TT"] == t["TT"] and times[k]["i"] == 23


def test_batch_autotune_and_graph_do_not_change_results(dev):
    from asteriy_2d_amd.batch import Batch
    inst = _instances()
    grids = [g for g, _, _, _ in inst]
    states = [(s, u) for _, s, u, _ in inst]
    with Batch(grids, G.HLLC, states) as a:
        a.run(11, graph=False)
        ra = [a.state(k) for k in range(len(grids))]
    with Batch(grids, G.HLLC, states) as b:
        b.autotune(reps=2)
        b.run(11, graph=True)
        rb = [b.state(k) for k in range(len(grids))]
    for x, y in zip(ra, rb):
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


def test_batch_rejects_window(dev):
    from asteriy_2d_amd import native
    from asteriy_2d_amd.batch import Batch
    g = G.reference_live_grid(N=128, M=8)
    with pytest.raises(native.AsteriyError):
        Batch([g], G.HLLC)
