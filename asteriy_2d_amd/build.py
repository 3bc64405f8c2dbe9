"""Builds libasteriy.so: the HIP kernels for gfx950 + the C-ABI glue.

Every source is compiled by hipcc into an object in-tree (asteriy_2d_amd/_build)
and linked into asteriy_2d_amd/libasteriy.so.  The FP64 parity sources are
compiled without FMA contraction (bit parity with the reference arithmetic);
the FP32 fast path contracts a*b+c to FMA only within one source expression
(-ffp-contract=on: the arithmetic is fixed by the source, so the single-grid
step and the batched launch, which inline the same strip code into different
kernels, round identically) and uses the hardware FP32 divide/sqrt.
Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libasteriy.so")
EXE = os.path.join(HERE, "asteriy_run")
ARCH = os.environ.get("AST_OFFLOAD_ARCH", "gfx950")

COMMON = [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
          "-Wno-unused-variable", "-Wno-unused-function", "-Wno-unused-but-set-variable",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
PARITY = ["-ffp-contract=off"]
FAST = ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt"]

# (source, flag set)
SOURCES = [
    ("api.cpp", PARITY),
    ("sim.cpp", PARITY),
    ("misc.hip", PARITY),
    ("flux_f64.hip", PARITY),
    ("flux_mhd_f64.hip", PARITY),
    ("mc.hip", PARITY),
    ("step_f64.hip", PARITY),
    ("flux_f32.hip", FAST),
    ("step_f32.hip", FAST),
    ("step_f32_hll.hip", FAST),
    ("step_f32_exact.hip", FAST),
    ("step_f32_learned.hip", FAST),
    ("step_f32_detect.hip", FAST),
    ("mlp.hip", FAST),
    ("batch.hip", FAST),
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP extension cannot be built")


def _compile(src: str, flags) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *COMMON, *flags, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compiles every source (always) and links libasteriy.so."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [(s, f) for s, f in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda sf: _compile(*sf), srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the host driver (the reference's main() over the C-ABI)
    drv = os.path.join(CSRC, "driver.cpp")
    if os.path.exists(drv):
        cmd = [_hipcc(), f"--offload-arch={ARCH}", "-O2", "-std=c++17", drv, "-o", EXE + ".tmp",
               f"-L{HERE}", "-lasteriy", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(EXE + ".tmp", EXE)
    if verbose:
        print(f"built {LIB} from {len(objs)} sources and {EXE}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
