#!/usr/bin/env python3
"""Benchmark of the explicit (z,r) finite-volume step on one MI355X.

Workload: the stellar-wind / counter-flow interaction on a 2048x1024 (z,r)
grid, z in [-400,400], r in [0,400], gamma 5/3, HLLC + minmod, CFL 0.3
(asteriy_2d_amd/problems.py).  A "step" is one pass of the step kernel over
the grid followed by the time advance (the reference's Kernel_TVD + funk_time
pair), replayed from a captured HIP graph with the state resident in HBM.

Headline `value`: Gcell-updates/s of the FP32 fast step over the timed steps.
The same run also measures the plain FP64 step (the bit-faithful restatement
of the reference step) on the same workload and the L1 density/pressure error
of the FP32 run against it, and the exact-solver FP32 step.

Roofline: algorithmic bytes of the step kernel (FP32: 16 B read + 16 B
written per cell-update) over its average launch duration, timed with HIP
events on the stream the kernel runs on; peak 8 TB/s HBM.
CPU baseline: the oracle (line-faithful C restatement of the reference step,
FP64, OpenMP) on the host cores, on a bounded sample of full-grid steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_GRID, M_GRID = 2048, 1024


def _timed(sim, steps: int, stream):
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    sim.sync()
    e0.record(stream)
    sim.run(steps, graph=True)
    e1.record(stream)
    sim.sync()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def _kernel_ms(grid, precision, method, stream, reps=50):
    """Average duration of the step kernel alone: `reps` back-to-back launches
    on `stream` between two HIP events (the queue stays full, so the launch
    gaps are hidden)."""
    import numpy as np
    import torch

    from asteriy_2d_amd import problems, solver
    dt = torch.float32 if precision == "f32" else torch.float64
    s, u = problems.wind_state(grid)
    dev = torch.device("cuda:0")
    S = torch.from_numpy(s).to(dev, dt)
    U = torch.from_numpy(u).to(dev, dt)
    S2, U2 = torch.empty_like(S), torch.empty_like(U)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev)
    Td = torch.tensor([0.05], dtype=torch.float64, device=dev)
    TT = torch.zeros(1, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        for _ in range(3):
            solver.Kernel_TVD(grid, S, U, S2, U2, T, Td, TT, method, stream=stream)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            solver.Kernel_TVD(grid, S, U, S2, U2, T, Td, TT, method, stream=stream)
        e1.record(stream)
    torch.cuda.synchronize()
    del np
    return e0.elapsed_time(e1) / reps


def _l1(a, b):
    import numpy as np
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def _cpu_baseline(grid, s, u, method, budget_s: float):
    """Oracle steps on the full grid until ~budget_s of CPU time."""
    import numpy as np

    from oracle import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    O.build()
    T_do = 0.05
    steps = 0
    t0 = time.perf_counter()
    while True:
        O.step(grid, s, u, 1e7, T_do, method, nthreads=threads)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 50:
            break
    del np
    return {"value": round(grid.cells * steps / el / 1e9, 6), "unit": "Gcell-updates/s",
            "cores": threads, "kind": "port",
            "sample": f"{steps} full-grid FP64 HLLC steps of the {grid.N}x{grid.M} wind workload "
                      f"by the C oracle ({el:.1f} s, {threads} OpenMP threads)"}


def _traffic(name: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy kernel), recorded in profiles/."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-exact", action="store_true")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU slab decomposition is design-only)", file=sys.stderr)
        return 2

    import torch

    from asteriy_2d_amd import EXACT, HLLC, native, problems
    from asteriy_2d_amd.sim import Simulation

    native.load()
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 3
    p = problems.wind(N_GRID, M_GRID, method=HLLC, steps=a.steps)
    g = p.grid
    cells = g.cells
    K, W = a.steps, a.warmup

    # -- FP32 fast step (headline) ------------------------------------------
    sim32 = Simulation(g, HLLC, "f32", p.s, p.u)
    st32 = sim32.torch_stream()
    sim32.run(W).sync()
    sim32.set_state(p.s, p.u)
    sim32.set_time()
    ms32 = _timed(sim32, K, st32)
    s32, u32 = sim32.state()
    t32 = sim32.time()
    kms32 = _kernel_ms(g, "f32", HLLC, st32)

    # -- plain FP64 step on the same workload --------------------------------
    sim64 = Simulation(g, HLLC, "f64", p.s, p.u)
    st64 = sim64.torch_stream()
    sim64.run(W).sync()
    sim64.set_state(p.s, p.u)
    sim64.set_time()
    ms64 = _timed(sim64, K, st64)
    s64, u64 = sim64.state()
    t64 = sim64.time()
    kms64 = _kernel_ms(g, "f64", HLLC, st64, reps=20)

    value = cells * K / (ms32 * 1e-3) / 1e9
    v64 = cells * K / (ms64 * 1e-3) / 1e9
    alg_bytes = 32.0 * cells
    achieved = alg_bytes / (kms32 * 1e-3) / 1e9
    traffic = _traffic("step_f32_hllc_2048x1024")

    extra = {
        "fp64_gcells": round(v64, 3),
        "fp32_vs_fp64": round(value / v64, 3),
        "l1_rho_vs_fp64": _l1(s32[..., 0], s64[..., 0]),
        "l1_p_vs_fp64": _l1(s32[..., 1], s64[..., 1]),
        "sim_time": [t32["TT"], t64["TT"]],
        "kernel_us": {"step_f32_hllc": round(kms32 * 1e3, 2), "step_f64_hllc": round(kms64 * 1e3, 2)},
        "fp64_roofline_frac": round(64.0 * cells / (kms64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }
    sim32.close()
    sim64.close()

    # -- exact solver: FP32 (+FP64 islands) vs plain FP64, fewer steps -------
    if not a.no_exact:
        Ke = max(10, min(K, 200))
        pe = problems.wind(N_GRID, M_GRID, method=EXACT, steps=Ke)
        se32 = Simulation(g, EXACT, "f32", pe.s, pe.u)
        mse32 = _timed(se32, Ke, se32.torch_stream())
        e32 = se32.state()
        stats = se32.stats()
        se64 = Simulation(g, EXACT, "f64", pe.s, pe.u)
        mse64 = _timed(se64, Ke, se64.torch_stream())
        e64 = se64.state()
        extra["exact"] = {
            "steps": Ke,
            "fp32_gcells": round(cells * Ke / (mse32 * 1e-3) / 1e9, 3),
            "fp64_gcells": round(cells * Ke / (mse64 * 1e-3) / 1e9, 3),
            "speedup": round(mse64 / mse32, 3),
            "l1_rho_vs_fp64": _l1(e32[0][..., 0], e64[0][..., 0]),
            "newton_iters_per_face": round(stats["newton_iters"] / max(1, stats["newton_faces"]), 3),
            "fp64_island_faces": stats["fp64_faces"],
        }
        se32.close()
        se64.close()

    cpu = None
    if not a.no_cpu:
        cpu = _cpu_baseline(g, p.s, p.u, HLLC, a.cpu_seconds)

    line = {
        "metric": "Gcell-updates/s @2048x1024 vs plain FP64 on GPU; L1 err vs FP64; % HBM peak",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms32 / K, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "wind/counter-flow 2048x1024, HLLC+minmod, CFL 0.3, FP32 step",
                   "grid": [N_GRID, M_GRID], "method": "hllc", "cfl": 0.3, "gamma": 5 / 3},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
