"""CPU checks of the C-ABI library: it loads without a GPU, exports every
function include/asteriy.h declares, validates grids without a device, and
its struct layouts match the Python mirrors.  No compute calls."""
import ctypes
import math
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "asteriy.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(ast_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    from asteriy_2d_amd import native
    if not os.path.exists(native.LIB_PATH):
        from asteriy_2d_amd import build
        build.build()
    return ctypes.CDLL(native.LIB_PATH)


def test_header_declares_the_step_boundary():
    names = declared_functions()
    for n in ("ast_step_f64", "ast_step_f32", "ast_funk_time", "ast_flux_f64", "ast_flux_f32",
              "ast_sim_create", "ast_sim_run", "ast_batch_create", "ast_batch_run",
              "ast_mlp_forward", "ast_mlp_loss_grad"):
        assert n in names, n


def test_library_exports_every_declared_function(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared in include/asteriy.h but not exported: {missing}"


def test_python_binding_covers_exports():
    from asteriy_2d_amd import learned, native, sim  # noqa: F401  (register signatures)
    missing = [n for n in declared_functions() if n not in native.SIGNATURES
               and not n.startswith("ast_batch_")]
    assert not missing, missing


def test_flux_mhd_argument_checks(lib):
    """ast_flux_mhd_f64 refuses a bad method, gamma, a normal that picks no
    frame, null and overlapping buffers, before any device work."""
    import ctypes

    from asteriy_2d_amd import grid as G
    from asteriy_2d_amd import native
    L = native.load()
    buf = (ctypes.c_double * 64)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    q = ctypes.c_void_p(p.value + 8)
    far = ctypes.c_void_p(p.value + 4096)
    f = L.ast_flux_mhd_f64
    assert f(G.HLLC, 1, p, 1.0, 0.0, 0.0, 1.0, 5 / 3, 0.05, far, None) == 1
    assert f(G.HLLD, 1, p, 1.0, 0.0, 0.0, 1.0, 1.0, 0.05, far, None) == 1
    assert f(G.HLLD, 1, p, 0.05, -0.1, 0.02, 1.0, 5 / 3, 0.05, far, None) == 1
    assert b"normal" in L.ast_last_error()
    assert f(G.HLLD, 1, None, 1.0, 0.0, 0.0, 1.0, 5 / 3, 0.05, far, None) == 1
    assert f(G.HLLD, 2, p, 1.0, 0.0, 0.0, 1.0, 5 / 3, 0.05, q, None) == 1
    assert b"overlap" in L.ast_last_error()
    assert f(G.HLLC_MHD, 0, None, -1.0, 0.0, 0.0, 1.0, 5 / 3, 0.05, None, None) == 0


def test_no_device_needed_for_validation(lib):
    from asteriy_2d_amd import grid as G
    from asteriy_2d_amd import native
    L = native.load()
    assert L.ast_version().startswith(b"asteriy")
    assert L.ast_device_count() >= 0
    g = G.Grid.from_extent(16, 8, 0.0, 1.0, 1.0)
    assert L.ast_grid_check(native.grid_ref(g)) == 0
    bad = [g.replace(N=1), g.replace(gamma=1.0), g.replace(krit=0.0),
           g.replace(geometry=7), g.replace(bc=(9, 0, 0, 0)),
           g.replace(bc=(G.BC_DIRICHLET, 0, 0, 0), bc_state=[[0.0] * 4] * 4),
           g.replace(geometry=G.AXISYMMETRIC, y_min=0.0),
           g.replace(bc=(G.BC_DIRICHLET_PULSED, 0, 0, 0), bc_pulse=(1.0, 6.0, 0.5, 3.0)),
           g.replace(bc=(G.BC_DIRICHLET_PULSED, 0, 0, 0), bc_pulse=(0.01, math.inf, 0.5, 3.0)),
           g.replace(bc=(5, 0, 0, 0))]
    for b in bad:
        rc = L.ast_grid_check(native.grid_ref(b))
        assert rc != 0
        assert len(native.last_error()) > 0


def test_grid_struct_layout():
    from asteriy_2d_amd.grid import CGrid, CStats
    # int32 N, M, geometry, bc[4] -> 28 bytes, padded to 32; 4+2+16+4+2+3+4 doubles
    assert ctypes.sizeof(CGrid) == 32 + 8 * (6 + 16 + 4 + 2 + 3 + 4)
    assert CGrid.x_min.offset == 32
    assert CGrid.bc_pulse.offset == 32 + 8 * (6 + 16 + 4 + 2 + 3)
    assert ctypes.sizeof(CStats) == 10 * 8


def test_step_rejects_bad_arguments_without_gpu(lib):
    from asteriy_2d_amd import grid as G
    from asteriy_2d_amd import native
    L = native.load()
    g = G.Grid.from_extent(16, 8, 0.0, 1.0, 1.0)
    # null planes: rejected before any device work
    rc = L.ast_step_f64(native.grid_ref(g), None, None, None, None, None, None, None, 0, None, None)
    assert rc == 1 and b"null" in L.ast_last_error()
    rc = L.ast_flux_f64(7, 10, ctypes.c_void_p(8), 1.0, 0.0, 1.0, 1.4, 0.5, ctypes.c_void_p(8),
                        None, None)
    assert rc == 1


def test_pulsed_side_needs_the_time_without_gpu(lib):
    """A pulsed Dirichlet side: the FP32 and FP64 steps refuse a null TT, and
    the batch (no per-launch time) refuses the side, before any device
    work."""
    from asteriy_2d_amd import grid as G
    from asteriy_2d_amd import native
    L = native.load()
    g = G.Grid.from_extent(64, 32, 0.0, 1.0, 1.0, bc=(G.BC_DIRICHLET_PULSED, 0, 0, 0),
                           bc_state=[[1.0, 1.0, 0.5, 0.0], [0.0] * 4, [0.0] * 4, [0.0] * 4],
                           bc_pulse=(0.01, 6.283185307179586, 0.5, 3.0))
    buf = (ctypes.c_double * 16)()
    a = [ctypes.c_void_p(ctypes.addressof(buf) + 8 * k) for k in range(6)]
    rc = L.ast_step_f32(native.grid_ref(g), a[0], a[1], a[2], a[3], a[4], a[5], None, 1, None,
                        None, None)
    assert rc == 1 and b"TT" in L.ast_last_error()
    rc = L.ast_step_f64(native.grid_ref(g), a[0], a[1], a[2], a[3], a[4], a[5], None, 1, None,
                        None)
    assert rc == 1 and b"TT" in L.ast_last_error()
    cg = g.to_c()
    grids = (type(cg) * 1)(cg)
    h = ctypes.c_void_p()
    rc = L.ast_batch_create(1, grids, 1, ctypes.byref(h))
    assert rc == 1 and b"pulsed" in L.ast_last_error()


def test_driver_runs_without_gpu_as_usage_only():
    import subprocess
    from asteriy_2d_amd import build
    if not os.path.exists(build.EXE):
        build.build()
    r = subprocess.run([build.EXE, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--problem" in r.stdout
    r = subprocess.run([build.EXE, "--problem", "nope"], capture_output=True, text=True)
    assert r.returncode == 2
