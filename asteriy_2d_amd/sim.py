"""Python handle on the native simulation runtime (csrc/sim.cpp).

`Simulation` owns the device planes and the device time scalars and advances
the reference's protocol -- step, then funk_time (kernel.cu:3868-3906) -- from
a captured HIP graph.  State crosses the boundary in the reference layout:
(M, N, 2) float64 arrays s = {rho, p}, u = {vx, vy}.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import native
from .grid import CStats, STATS_FIELDS

_vp, _c_int, _c_double = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_dp = ctypes.POINTER(ctypes.c_double)

SIM_SIGNATURES = {
    "ast_sim_create": (_c_int, [ctypes.c_void_p, _c_int, _c_int, ctypes.POINTER(_vp)]),
    "ast_sim_destroy": (_c_int, [_vp]),
    "ast_sim_set_state": (_c_int, [_vp, _dp, _dp]),
    "ast_sim_get_state": (_c_int, [_vp, _dp, _dp]),
    "ast_sim_device_state": (_c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "ast_sim_set_time": (_c_int, [_vp, _c_double, _c_double, _c_double, _c_int]),
    "ast_sim_get_time": (_c_int, [_vp, _dp, _dp, _dp, ctypes.POINTER(_c_int)]),
    "ast_sim_get_stats": (_c_int, [_vp, ctypes.POINTER(CStats), _c_int]),
    "ast_sim_collect_stats": (_c_int, [_vp, _c_int]),
    "ast_sim_set_mlp": (_c_int, [_vp, _vp]),
    "ast_sim_set_sources": (_c_int, [_vp, _vp, _vp, _c_double, _c_double]),
    "ast_sim_learned_modes": (_c_int, [_vp, ctypes.POINTER(_c_int)]),
    "ast_sim_set_learned_mode": (_c_int, [_vp, _c_int]),
    "ast_sim_set_tile_rows": (_c_int, [_vp, _c_int]),
    "ast_sim_stream": (_vp, [_vp]),
    "ast_sim_autotune": (_c_int, [_vp, ctypes.POINTER(_c_int), _c_int, _c_int,
                                  ctypes.POINTER(_c_int), ctypes.POINTER(ctypes.c_float)]),
    "ast_sim_run": (_c_int, [_vp, _c_int, _c_int]),
    "ast_sim_sync": (_c_int, [_vp]),
    "ast_sim_write_b31": (_c_int, [_vp, ctypes.c_char_p]),
    "ast_sim_read_b31": (_c_int, [_vp, ctypes.c_char_p]),
    "ast_sim_write_tecplot": (_c_int, [_vp, ctypes.c_char_p, _c_int]),
    "ast_sim_write_axis": (_c_int, [_vp, ctypes.c_char_p]),
    "ast_sim_save": (_c_int, [_vp, ctypes.c_char_p]),
    "ast_sim_load": (_c_int, [_vp, ctypes.c_char_p]),
}
native.SIGNATURES.update(SIM_SIGNATURES)


def _lib():
    L = native.load()
    for name, (res, args) in SIM_SIGNATURES.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


class Simulation:
    """precision: "f64" (the bit-faithful reference step) or "f32" (fast path)."""

    def __init__(self, grid, method: int, precision: str = "f64", s=None, u=None):
        self.L = _lib()
        self.grid = grid
        self.method = method
        self.precision = precision
        self._cg = grid.to_c()
        h = _vp()
        native.call("ast_sim_create", ctypes.byref(self._cg), 1 if precision == "f32" else 0,
                    method, ctypes.byref(h))
        self.h = h
        if s is not None:
            self.set_state(s, u)

    def close(self):
        if getattr(self, "h", None):
            self.L.ast_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, s, u):
        s = np.ascontiguousarray(s, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        n = self.grid.cells * 2
        if s.size != n or u.size != n:
            raise ValueError("state must hold N*M pairs")
        native.call("ast_sim_set_state", self.h, s.ctypes.data_as(_dp), u.ctypes.data_as(_dp))

    def state(self):
        s = np.empty((self.grid.M, self.grid.N, 2))
        u = np.empty_like(s)
        native.call("ast_sim_get_state", self.h, s.ctypes.data_as(_dp), u.ctypes.data_as(_dp))
        return s, u

    def device_state(self):
        a, b = _vp(), _vp()
        native.call("ast_sim_device_state", self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def set_time(self, T=1e7, T_do=1e-9, TT=0.0, i=0):
        native.call("ast_sim_set_time", self.h, T, T_do, TT, i)

    def time(self):
        T, Td, TT, i = _c_double(), _c_double(), _c_double(), _c_int()
        native.call("ast_sim_get_time", self.h, ctypes.byref(T), ctypes.byref(Td),
                    ctypes.byref(TT), ctypes.byref(i))
        return {"T": T.value, "T_do": Td.value, "TT": TT.value, "i": i.value}

    def collect_stats(self, on: bool = True):
        """Count the exact solver's work (slower kernel variants) from now on."""
        native.call("ast_sim_collect_stats", self.h, 1 if on else 0)

    def stats(self, reset=False):
        st = CStats()
        native.call("ast_sim_get_stats", self.h, ctypes.byref(st), 1 if reset else 0)
        return {k: getattr(st, k) for k in STATS_FIELDS}

    def set_mlp(self, mlp):
        native.call("ast_sim_set_mlp", self.h, _vp(mlp.buffer.data_ptr()) if mlp else None)

    def set_sources(self, nn2=None, nn3=None, kmom: float = 3.0 / 0.419231, ken: float = 0.0):
        """Neutral-H source planes of add_MK (kernel.cu:2495-2498) for every
        later step: nn2 (M, N, 3), nn3 (M, N) host float64 arrays (either may
        be None); the defaults are the reference's n_H/Kn and 0.0*n_H/Kn
        (Header.h:84-85).  An FP32 simulation with planes runs the exact
        solver without the learned start."""
        a2 = None if nn2 is None else np.ascontiguousarray(nn2, dtype=np.float64)
        a3 = None if nn3 is None else np.ascontiguousarray(nn3, dtype=np.float64)
        if a2 is not None and a2.shape != (self.grid.M, self.grid.N, 3):
            raise ValueError(f"nn2: need shape {(self.grid.M, self.grid.N, 3)}, got {a2.shape}")
        if a3 is not None and a3.shape != (self.grid.M, self.grid.N):
            raise ValueError(f"nn3: need shape {(self.grid.M, self.grid.N)}, got {a3.shape}")
        native.call("ast_sim_set_sources", self.h,
                    None if a2 is None else a2.ctypes.data_as(_vp),
                    None if a3 is None else a3.ctypes.data_as(_vp), float(kmom), float(ken))

    def learned_modes(self):
        """The learned modes (0 plain, 1 learned start) of the next two runs."""
        m = (_c_int * 2)()
        native.call("ast_sim_learned_modes", self.h, m)
        return m[0], m[1]

    def set_learned_mode(self, mode: int):
        """-1 automatic (default), 0 plain solver, 1 learned start, every run."""
        native.call("ast_sim_set_learned_mode", self.h, int(mode))

    def set_tile_rows(self, krows: int):
        native.call("ast_sim_set_tile_rows", self.h, int(krows))

    def autotune(self, candidates=None, reps: int = 10):
        """Picks the fastest FP32 strip tiling for this grid (csrc/sim.cpp);
        returns (tiling, microseconds per step): rows per wave strip, or -S
        for S strips of (nearly) equal height per 60-column band."""
        arr, n = None, 0
        if candidates:
            n = len(candidates)
            arr = (_c_int * n)(*candidates)
        r, t = _c_int(), ctypes.c_float()
        native.call("ast_sim_autotune", self.h, arr, n, int(reps), ctypes.byref(r),
                    ctypes.byref(t))
        return r.value, t.value

    @property
    def stream_ptr(self) -> int:
        return int(self.L.ast_sim_stream(self.h) or 0)

    def torch_stream(self):
        import torch
        return torch.cuda.ExternalStream(self.stream_ptr)

    def run(self, steps: int, graph: bool = True):
        native.call("ast_sim_run", self.h, int(steps), 1 if graph else 0)
        return self

    def sync(self):
        native.call("ast_sim_sync", self.h)

    def write_b31(self, path: str):
        native.call("ast_sim_write_b31", self.h, path.encode())

    def read_b31(self, path: str):
        native.call("ast_sim_read_b31", self.h, path.encode())

    def write_tecplot(self, path: str, stride: int = 2):
        native.call("ast_sim_write_tecplot", self.h, path.encode(), int(stride))

    def write_axis(self, path: str):
        native.call("ast_sim_write_axis", self.h, path.encode())

    def save(self, path: str):
        native.call("ast_sim_save", self.h, path.encode())

    def load(self, path: str):
        native.call("ast_sim_load", self.h, path.encode())
