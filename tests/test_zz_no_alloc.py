"""The entry point that allocates nothing: ast_conservation_sums_noalloc
keeps its partial sums in the library's own per-GPU scratch array instead
of a caller workspace.  Through the C-ABI it must give exactly what
ast_conservation_sums gives (the same kernels, the same fixed summation
order: bit for bit), on two calls in a row, the second with a smaller grid,
for FP64 and FP32 state planes."""
import ctypes

import numpy as np
import pytest

from asteriy_2d_amd import grid as G
from asteriy_2d_amd import problems

pytestmark = pytest.mark.gpu


def _sums(fn, grid, S, U, fp32, dev):
    import torch
    from asteriy_2d_amd import native
    L = native.load()
    if fn == "ast_conservation_sums":
        out = torch.full((int(L.ast_conservation_work_doubles()),), float("nan"),
                         dtype=torch.float64, device=dev)
    else:
        out = torch.full((2,), float("nan"), dtype=torch.float64, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    native.call(fn, native.grid_ref(grid), ctypes.c_void_p(S.data_ptr()),
                ctypes.c_void_p(U.data_ptr()), 1 if fp32 else 0, ctypes.c_void_p(out.data_ptr()),
                stream)
    torch.cuda.synchronize()
    return out[:2].cpu().numpy()


@pytest.mark.parametrize("fp32", [False, True])
def test_noalloc_sums_equal_workspace_sums(dev, fp32):
    import torch
    dt = torch.float32 if fp32 else torch.float64
    results = []
    for N, M, geom in ((512, 256, G.AXISYMMETRIC), (96, 40, G.PLANAR)):  # second call smaller
        print(f"noalloc sums fp32={fp32} {N}x{M} geometry={geom}", flush=True)
        grid = G.Grid.from_extent(N, M, -1.0, 1.0, 0.8, geometry=geom, bc=(G.BC_OUTFLOW,) * 4)
        s, u = problems.random_state(grid, seed=N)
        S = torch.from_numpy(s).to(dev, dt)
        U = torch.from_numpy(u).to(dev, dt)
        a = _sums("ast_conservation_sums_noalloc", grid, S, U, fp32, dev)
        b = _sums("ast_conservation_sums", grid, S, U, fp32, dev)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), (a, b)
        # and the sums are the integrals: against numpy in FP64
        w = grid.dx * grid.dy * (grid.y(np.arange(M))[:, None] if geom == G.AXISYMMETRIC else 1.0)
        sd, ud = S.double().cpu().numpy(), U.double().cpu().numpy()
        mass = float((sd[..., 0] * w).sum())
        energy = float(((sd[..., 1] / (grid.gamma - 1) + 0.5 * sd[..., 0] *
                         (ud ** 2).sum(-1)) * w).sum())
        assert abs(a[0] - mass) <= 1e-12 * abs(mass) and abs(a[1] - energy) <= 1e-12 * abs(energy)
        results.append(a)
    assert not np.array_equal(results[0], results[1])


def test_noalloc_sums_reject_bad_arguments(dev):
    import torch
    from asteriy_2d_amd import native
    grid = G.Grid.from_extent(16, 8, 0.0, 1.0, 1.0, bc=(G.BC_OUTFLOW,) * 4)
    S = torch.ones(grid.cells * 2, dtype=torch.float64, device=dev)
    out = torch.zeros(2, dtype=torch.float64, device=dev)
    with pytest.raises(native.AsteriyError):
        native.call("ast_conservation_sums_noalloc", native.grid_ref(grid),
                    ctypes.c_void_p(S.data_ptr()), None, 0, ctypes.c_void_p(out.data_ptr()), None)
    with pytest.raises(native.AsteriyError):
        native.call("ast_conservation_sums_noalloc", native.grid_ref(grid.replace(N=1)),
                    ctypes.c_void_p(S.data_ptr()), ctypes.c_void_p(S.data_ptr()), 0,
                    ctypes.c_void_p(out.data_ptr()), None)
