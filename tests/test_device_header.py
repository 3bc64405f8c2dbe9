"""The per-face device operators of include/asteriy_device.hpp (the
reference's __device__ HLL / HLLC_2d_Korolkov / HLLC_Aleksashov /
Godunov_Solver_Alexashov, kernel.cu:1399, Gas_Dinamic.cu:1814, kernel.cu:767,
kernel.cu:128, for a caller's own kernels): compiled for the host and
checked against the oracle to 1e-12 of the natural scales on the flux
microbenchmark's faces under four normals; compiled into a caller's HIP
kernel for gfx950.  No GPU needed."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from asteriy_2d_amd import facegen
from asteriy_2d_amd import grid as G

from cases import flux_mismatch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("devhdr") / "device_header_check"
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
           f"-I{ROCM}/include", os.path.join(ROOT, "tests", "device_header_check.cpp"),
           "-o", str(exe), "-lm"]
    subprocess.run(cmd, check=True)
    return exe


@pytest.mark.parametrize("method", [G.HLL, G.HLLC, G.EXACT, G.HLLC_A])
def test_device_header_matches_oracle(ora, checker, tmp_path, method):
    F = np.ascontiguousarray(facegen.faces(4096, seed=61))
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    F.tofile(src)
    for n1, n2 in ((1.0, 0.0), (0.0, -1.0), (-0.6, 0.8), (0.0, 1.0)):
        subprocess.run([str(checker), str(src), str(dst), str(method), repr(n1), repr(n2),
                        "0.25", repr(5.0 / 3.0), "0.05"], check=True)
        out = np.fromfile(dst).reshape(5, -1)
        ref = ora.flux_batch(method, F, n1, n2, 0.25)
        assert np.array_equal(np.isnan(out), np.isnan(ref))
        bad, worst = flux_mismatch(out, ref, F, 5.0 / 3.0)
        assert bad == 0, (n1, n2, worst)


@pytest.mark.skipif(not shutil.which(os.path.join(ROCM, "bin", "hipcc")), reason="no hipcc")
def test_device_header_compiles_into_a_caller_kernel(tmp_path):
    obj = tmp_path / "caller.o"
    cmd = [os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=gfx950", "-O3", "-std=c++17",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-c", os.path.join(ROOT, "tests", "device_header_kernel.hip"), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert obj.stat().st_size > 0
