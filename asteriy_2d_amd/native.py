"""ctypes binding of libasteriy.so (the C-ABI declared in include/asteriy.h).

The library is loaded from this package directory; there is no fallback: if
it is missing or a call fails, an exception is raised.  Build it with
`python -m asteriy_2d_amd.build` (or `__graft_entry__.build()`).
"""
from __future__ import annotations

import ctypes
import os

from .grid import CGrid, CStats

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libasteriy.so")

_c_int, _c_size, _c_double, _c_float, _vp = (ctypes.c_int, ctypes.c_size_t, ctypes.c_double,
                                             ctypes.c_float, ctypes.c_void_p)
_gp = ctypes.POINTER(CGrid)


class CSources(ctypes.Structure):
    """ast_sources_t: device pointers of the neutral-H planes and their
    coefficients."""
    _fields_ = [("nn2", ctypes.c_void_p), ("nn3", ctypes.c_void_p),
                ("kmom", ctypes.c_double), ("ken", ctypes.c_double)]

# name -> (restype, argtypes); every symbol declared in include/asteriy.h
SIGNATURES = {
    "ast_last_error": (ctypes.c_char_p, []),
    "ast_version": (ctypes.c_char_p, []),
    "ast_device_count": (_c_int, []),
    "ast_grid_check": (_c_int, [_gp]),
    "ast_flux_f64": (_c_int, [_c_int, _c_size, _vp, _c_double, _c_double, _c_double, _c_double,
                              _c_double, _vp, _vp, _vp]),
    "ast_flux_f32": (_c_int, [_c_int, _c_size, _vp, _c_float, _c_float, _c_float, _c_float,
                              _c_float, _vp, _vp, _vp]),
    "ast_flux_q_f64": (_c_int, [_c_int, _c_size, _vp, _c_double, _c_double, _c_double,
                                _c_double, _c_double, _vp, _vp]),
    "ast_flux_mhd_f64": (_c_int, [_c_int, _c_size, _vp, _c_double, _c_double, _c_double,
                                  _c_double, _c_double, _c_double, _vp, _vp]),
    "ast_mc_rates_f64": (_c_int, [_c_size, _vp, _vp, _vp, _vp]),
    "ast_mc_cell_exit_f64": (_c_int, [_gp, _c_size, _vp, _vp, _vp, _vp, _vp]),
    "ast_mc_belong_point_f64": (_c_int, [_gp, _c_size, _vp, _vp, _vp]),
    "ast_step_f64": (_c_int, [_gp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _vp]),
    "ast_step_f64_src": (_c_int, [_gp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int,
                                  ctypes.POINTER(CSources), _vp, _vp]),
    "ast_step_f32": (_c_int, [_gp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _vp]),
    "ast_step_f32_src": (_c_int, [_gp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_int,
                                  ctypes.POINTER(CSources), _vp, _vp]),
    "ast_funk_time": (_c_int, [_vp, _vp, _vp, _vp, _vp]),
    "ast_conservation_work_doubles": (_c_size, []),
    "ast_conservation_sums": (_c_int, [_gp, _vp, _vp, _c_int, _vp, _vp]),
    "ast_conservation_sums_noalloc": (_c_int, [_gp, _vp, _vp, _c_int, _vp, _vp]),
    "ast_convert_f64_to_f32": (_c_int, [_c_size, _vp, _vp, _vp]),
    "ast_convert_f32_to_f64": (_c_int, [_c_size, _vp, _vp, _vp]),
    "ast_flux_f32_info": (_c_int, [_c_int, _c_size, _vp, _c_float, _c_float, _c_float, _c_float,
                                   _c_float, _vp, _vp, _vp, _vp, _vp]),
    "ast_set_tile_rows": (_c_int, [_c_int]),
    "ast_get_tile_rows": (_c_int, []),
}

_lib = None


class AsteriyError(RuntimeError):
    """A C-ABI call returned a non-zero AST_* status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def load(path: str = LIB_PATH):
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: the HIP extension must be built before use "
            "(python -m asteriy_2d_amd.build); there is no CPU fallback")
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        if not hasattr(L, name):
            continue
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    _lib = L
    return L


def exported_symbols(path: str = LIB_PATH):
    L = ctypes.CDLL(path)
    return {n for n in SIGNATURES if hasattr(L, n)}


def call(name: str, *args) -> None:
    L = load()
    rc = getattr(L, name)(*args)
    if rc != 0:
        raise AsteriyError(name, rc, L.ast_last_error().decode(errors="replace"))


def last_error() -> str:
    return load().ast_last_error().decode(errors="replace")


def device_count() -> int:
    return int(load().ast_device_count())


def grid_ref(grid) -> ctypes.pointer:
    cg = grid.to_c() if not isinstance(grid, CGrid) else grid
    return ctypes.pointer(cg)


def stats_from_tensor(t) -> dict:
    """Decodes an ast_stats_t held in a torch int64 tensor of 10 entries."""
    from .grid import STATS_FIELDS
    vals = t.detach().cpu().tolist()
    return {k: int(v) & 0xFFFFFFFFFFFFFFFF for k, v in zip(STATS_FIELDS, vals)}


__all__ = ["load", "call", "AsteriyError", "device_count", "grid_ref", "CStats",
           "exported_symbols", "SIGNATURES", "LIB_PATH"]
