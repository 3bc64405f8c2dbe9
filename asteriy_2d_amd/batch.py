"""Batched parameter sweep: many FP32 instances in one persistent launch
per step (csrc/batch.hip), captured in a HIP graph.

`sweep()` builds the stellar-wind / counter-flow sweep: `count` instances on
grids drawn uniformly from {256x128, 512x256, 1024x512, 2048x1024}, with the
counter-flow Mach number uniform in [1.5, 5] and the wind Mach number uniform
in [5, 20] (problems.WindParams), all HLLC + minmod.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import native, problems
from .grid import CGrid

_vp, _c_int, _dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
BATCH_SIGNATURES = {
    "ast_batch_create": (_c_int, [_c_int, ctypes.POINTER(CGrid), _c_int, ctypes.POINTER(_vp)]),
    "ast_batch_destroy": (_c_int, [_vp]),
    "ast_batch_set_state": (_c_int, [_vp, _c_int, _dp, _dp]),
    "ast_batch_get_state": (_c_int, [_vp, _c_int, _dp, _dp]),
    "ast_batch_get_time": (_c_int, [_vp, _c_int, _dp, _dp, _dp, ctypes.POINTER(_c_int)]),
    "ast_batch_set_time": (_c_int, [_vp, _c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, _c_int]),
    "ast_batch_autotune": (_c_int, [_vp, _c_int]),
    "ast_batch_run": (_c_int, [_vp, _c_int, _c_int]),
    "ast_batch_sync": (_c_int, [_vp]),
    "ast_batch_stream": (_vp, [_vp]),
    "ast_batch_cells": (ctypes.c_longlong, [_vp]),
}
native.SIGNATURES.update(BATCH_SIGNATURES)

SHAPES = [(256, 128), (512, 256), (1024, 512), (2048, 1024)]


def _lib():
    L = native.load()
    for name, (res, args) in BATCH_SIGNATURES.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


class Batch:
    def __init__(self, grids, method: int, states=None):
        self.L = _lib()
        self.grids = list(grids)
        n = len(self.grids)
        self._cg = (CGrid * n)(*[g.to_c() for g in self.grids])
        h = _vp()
        native.call("ast_batch_create", n, self._cg, int(method), ctypes.byref(h))
        self.h = h
        if states is not None:
            for k, (s, u) in enumerate(states):
                self.set_state(k, s, u)

    def close(self):
        if getattr(self, "h", None):
            self.L.ast_batch_destroy(self.h)
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

    def set_state(self, k, s, u):
        s = np.ascontiguousarray(s, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        native.call("ast_batch_set_state", self.h, int(k), s.ctypes.data_as(_dp),
                    u.ctypes.data_as(_dp))

    def state(self, k):
        g = self.grids[k]
        s = np.empty((g.M, g.N, 2))
        u = np.empty_like(s)
        native.call("ast_batch_get_state", self.h, int(k), s.ctypes.data_as(_dp),
                    u.ctypes.data_as(_dp))
        return s, u

    def time(self, k):
        T, Td, TT, i = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), _c_int()
        native.call("ast_batch_get_time", self.h, int(k), ctypes.byref(T), ctypes.byref(Td),
                    ctypes.byref(TT), ctypes.byref(i))
        return {"T": T.value, "T_do": Td.value, "TT": TT.value, "i": i.value}

    def set_time(self, k, T=1e7, T_do=1e-9, TT=0.0, i=0):
        native.call("ast_batch_set_time", self.h, int(k), T, T_do, TT, int(i))

    def autotune(self, reps: int = 5):
        native.call("ast_batch_autotune", self.h, int(reps))

    def run(self, steps: int, graph: bool = True):
        native.call("ast_batch_run", self.h, int(steps), 1 if graph else 0)
        return self

    def sync(self):
        native.call("ast_batch_sync", self.h)

    @property
    def cells(self) -> int:
        return int(self.L.ast_batch_cells(self.h))

    def torch_stream(self):
        import torch
        return torch.cuda.ExternalStream(int(self.L.ast_batch_stream(self.h)))


def sweep(count: int = 64, seed: int = 5, method: int = 1, shapes=SHAPES):
    """The instances of the sweep: list of (grid, s, u, WindParams)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        N, M = shapes[int(rng.integers(len(shapes)))]
        wp = problems.WindParams(mach_inf=float(rng.uniform(1.5, 5.0)),
                                 mach_w=float(rng.uniform(5.0, 20.0)))
        g = problems.wind_grid(N, M, wp)
        s, u = problems.wind_state(g, wp)
        out.append((g, s, u, wp))
    return out
