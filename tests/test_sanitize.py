"""The oracle and the library's host-side C++ under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only): oracle/sancheck.c drives every oracle
entry point over edge cases (make -C oracle sancheck); tests/host_sancheck.cpp
drives the grid validation and error reporting of api.cpp and the FP64 face
solvers of flux64.hpp compiled for the host.  Any sanitizer report aborts the
driver, so a zero exit status with its closing line is the pass."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer", "-O1", "-g"]


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1"
    return env


def _run(cmd, **kw):
    return subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)


@pytest.fixture(scope="module")
def sanitizers_available(tmp_path_factory):
    d = tmp_path_factory.mktemp("probe")
    src = d / "p.c"
    src.write_text("int main(void){return 0;}\n")
    r = _run(["gcc", *SAN, str(src), "-o", str(d / "p")])
    if r.returncode != 0:
        pytest.skip("gcc without AddressSanitizer/UBSan runtime: " + r.stderr[-200:])


def test_oracle_under_sanitizers(sanitizers_available):
    r = _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sancheck"])
    assert r.returncode == 0, r.stderr[-2000:]
    r = _run([os.path.join(ROOT, "oracle", "_san", "sancheck")], env=_env())
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-3000:])
    assert "sancheck ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_host_cpp_under_sanitizers(sanitizers_available, tmp_path):
    rocm = "/opt/rocm"
    if not os.path.exists(os.path.join(rocm, "include", "hip", "hip_runtime.h")):
        pytest.skip("no ROCm headers")
    exe = tmp_path / "host_sancheck"
    cmd = ["g++", "-std=c++17", *SAN, "-D__HIP_PLATFORM_AMD__", f"-I{rocm}/include",
           os.path.join(ROOT, "tests", "host_sancheck.cpp"),
           os.path.join(ROOT, "asteriy_2d_amd", "csrc", "api.cpp"),
           f"-L{rocm}/lib", "-lamdhip64", f"-Wl,-rpath,{rocm}/lib", "-o", str(exe)]
    r = _run(cmd)
    assert r.returncode == 0, r.stderr[-2000:]
    r = _run([str(exe)], env=_env())
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-3000:])
    assert "host sancheck ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_sanitizer_catches_a_planted_overflow(sanitizers_available, tmp_path):
    """The harness itself: an out-of-bounds read planted in a copy of a tiny
    C program aborts with a report (so a clean run above means something)."""
    src = tmp_path / "bad.c"
    src.write_text("#include <stdlib.h>\nint main(void){int* a = malloc(4 * sizeof(int));"
                   " volatile int x = a[4]; free(a); return x;}\n")
    r = _run(["gcc", *SAN, str(src), "-o", str(tmp_path / "bad")])
    assert r.returncode == 0
    r = _run([str(tmp_path / "bad")], env=_env())
    assert r.returncode != 0
    assert "AddressSanitizer" in r.stderr or "runtime error" in r.stderr
