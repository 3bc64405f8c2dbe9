"""The Monte-Carlo helpers of include/asteriy_mc.hpp (Gas_Dinamic.cu:1468-
1602) compiled for the host against the oracle's restatement (the same
library math, so to the last bit but for the rates' libm calls: 1e-15),
and the C-ABI's argument checks.  No GPU needed."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from asteriy_2d_amd import grid as G

from cases import mc_particles, mc_rate_pairs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("mchdr") / "mc_header_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", f"-I{ROCM}/include",
                    os.path.join(ROOT, "tests", "mc_header_check.cpp"), "-o", str(exe), "-lm"],
                   check=True)
    return exe


def _run(exe, tmp_path, mode, data, *args):
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    np.ascontiguousarray(data, dtype=np.float64).tofile(src)
    subprocess.run([str(exe), mode, str(src), str(dst), *[repr(float(a)) for a in args]],
                   check=True)
    return np.fromfile(dst)


def test_mc_rates_header_matches_oracle(ora, checker, tmp_path):
    P = mc_rate_pairs(20000, seed=81)
    out = _run(checker, tmp_path, "rates", P).reshape(3, -1)
    for k in (1, 2, 3):
        ref = np.array([ora.velosity(k, u, c) for u, c in P.T])
        assert np.all(np.abs(out[k - 1] - ref) <= 1e-15 * np.abs(ref)), k


def test_mc_cell_exit_and_belong_header_match_oracle(ora, checker, tmp_path):
    g = G.Grid.from_extent(64, 32, -3.0, 3.0, 1.5)
    P = mc_particles(g, 5000, seed=82)
    out = _run(checker, tmp_path, "exit", P, g.dx, g.dy).reshape(3, -1)
    for i in range(P.shape[1]):
        hit, mode, t = ora.peresechenie(g, *P[:, i])
        assert (int(out[2, i]), int(out[1, i])) == (int(hit), mode), i
        assert out[0, i] == t, i
    rng = np.random.default_rng(83)
    xy = np.stack([rng.uniform(g.x_min, g.x_min + (g.N - 1) * g.dx, 4000),
                   rng.uniform(0.0, g.M * g.dy * 0.999, 4000)])
    nm = _run(checker, tmp_path, "belong", xy, g.x_min, g.dx, g.y_min, g.dy).reshape(2, -1)
    for i in range(xy.shape[1]):
        assert tuple(int(v) for v in nm[:, i]) == ora.belong_point(g, *xy[:, i]), i


def test_mc_entry_points_check_arguments():
    from asteriy_2d_amd import native
    L = native.load()
    g = G.Grid.from_extent(8, 4, 0.0, 1.0, 1.0)
    buf = (ctypes.c_double * 16)()
    ibuf = (ctypes.c_int * 16)()
    assert L.ast_mc_rates_f64(4, None, buf, buf, None) == 1
    assert L.ast_mc_rates_f64(0, None, None, None, None) == 0
    assert L.ast_mc_cell_exit_f64(None, 1, buf, buf, ibuf, ibuf, None) == 1
    assert L.ast_mc_cell_exit_f64(native.grid_ref(g.replace(dx=0.0)), 1, buf, buf, ibuf, ibuf,
                                  None) == 1
    assert L.ast_mc_cell_exit_f64(native.grid_ref(g), 1, buf, buf, None, ibuf, None) == 1
    assert L.ast_mc_belong_point_f64(native.grid_ref(g), 1, None, ibuf, None) == 1
    assert L.ast_mc_belong_point_f64(native.grid_ref(g), 0, None, None, None) == 0
