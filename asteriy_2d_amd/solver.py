"""Host-side mirror of the reference's operator interface for the step.

The reference exposes (kernel.cu, Gas_Dinamic.cu):
  * per-face operators  HLL / HLLC_2d_Korolkov / Godunov_Solver_Alexashov /
    HLLC_Aleksashov
    (Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad) -> dt      kernel.cu:1399, 128, 767;
                                                     Gas_Dinamic.cu:1814
  * the step kernel     Kernel_TVD(s, u, s2, u2, T, T_do, TT, method)
                                                     kernel.cu:2513
  * the time advance    funk_time(T, T_do, TT, i)    kernel.cu:1525
  * the neutral-H source terms of add_MK (nn2, nn3)   kernel.cu:2293, 2495-2498
This module offers the same names and argument meaning over torch device
tensors, each call going through the C-ABI of libasteriy.so (HIP kernels for
gfx950).  Tensors: s, u are (M, N, 2) float64 (FP64 path) or float32 (FP32
path) on the GPU, the reference's double2 planes; T, T_do, TT are 1-element
float64 device tensors; i is a 1-element int32 device tensor.
"""
from __future__ import annotations

import ctypes

from . import native
from .grid import EXACT, HLL as M_HLL, HLLC as M_HLLC, HLLC_A as M_HLLC_A
from .grid import HLLCQ as M_HLLCQ, HLLCQ_A as M_HLLCQ_A, HLLC_MHD as M_HLLC_MHD, HLLD as M_HLLD

_STATS_WORDS = 10


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _check_state(t, dtype, cells, name):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise TypeError(f"{name} must be a CUDA (HIP) tensor")
    if t.dtype != dtype or not t.is_contiguous() or t.numel() != 2 * cells:
        raise ValueError(f"{name}: need a contiguous {dtype} tensor of {cells} pairs, got "
                         f"{t.dtype} {tuple(t.shape)}")


def new_stats(device="cuda"):
    """Zeroed ast_stats_t in device memory (10 x uint64 held as int64)."""
    import torch
    return torch.zeros(_STATS_WORDS, dtype=torch.int64, device=device)


# the reference's neutral-H coupling constants n_H and Kn (Header.h:84-85):
# add_MK's momentum coefficient is n_H/Kn, its energy coefficient 0.0*n_H/Kn
N_H, KN = 3.0, 0.419231


class NeutralSources:
    """The neutral-H moment planes of add_MK (kernel.cu:2293, 2495-2498) on
    the device: nn2 (M, N, 3) float64 (the reference's double3 plane; x and y
    are read), nn3 (M, N) float64 or None; the momentum of a cell gains
    T_do*(rho*kmom)*nn2.{x,y}, its energy T_do*ken*nn3."""

    def __init__(self, nn2=None, nn3=None, kmom: float = N_H / KN, ken: float = 0.0 * N_H / KN):
        import torch
        for t, nm, last in ((nn2, "nn2", 3), (nn3, "nn3", None)):
            if t is None:
                continue
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
                    and t.is_contiguous()):
                raise TypeError(f"{nm} must be a contiguous float64 CUDA (HIP) tensor")
            if last is not None and t.shape[-1] != last:
                raise ValueError(f"{nm}: last dimension must be {last}")
        self.nn2, self.nn3, self.kmom, self.ken = nn2, nn3, float(kmom), float(ken)

    def c(self, cells: int):
        for t, nm, per in ((self.nn2, "nn2", 3), (self.nn3, "nn3", 1)):
            if t is not None and t.numel() != per * cells:
                raise ValueError(f"{nm}: need {per * cells} values, got {t.numel()}")
        return native.CSources(_ptr(self.nn2), _ptr(self.nn3), self.kmom, self.ken)


def Kernel_TVD(grid, s, u, s2, u2, T, T_do, TT, method: int, stats=None, mlp=None,
               stream=None, sources: NeutralSources | None = None):
    """One step of the reference's Kernel_TVD (kernel.cu:2513) on the GPU.

    FP64 tensors run the face-centric FP64 step (ast_step_f64, within 1e-12
    of the reference's step), with `sources` the neutral-H terms of add_MK
    (ast_step_f64_src); FP32 tensors run the fast path (ast_step_f32),
    optionally seeding the exact solver with a learned guess (`mlp`, see
    mlp.py), or with `sources` (ast_step_f32_src, no learned guess)."""
    import torch
    cells = grid.N * grid.M
    dt = s.dtype
    if dt not in (torch.float64, torch.float32):
        raise TypeError("state must be float64 or float32")
    for t, nm in ((s, "s"), (u, "u"), (s2, "s2"), (u2, "u2")):
        _check_state(t, dt, cells, nm)
    gref = native.grid_ref(grid)
    if dt == torch.float64:
        if mlp is not None:
            raise ValueError("the learned guess is served by the FP32 step only")
        if sources is not None:
            cs = sources.c(cells)
            native.call("ast_step_f64_src", gref, _ptr(s), _ptr(u), _ptr(s2), _ptr(u2), _ptr(T),
                        _ptr(T_do), _ptr(TT), int(method), ctypes.byref(cs), _ptr(stats),
                        _stream(stream))
        else:
            native.call("ast_step_f64", gref, _ptr(s), _ptr(u), _ptr(s2), _ptr(u2), _ptr(T),
                        _ptr(T_do), _ptr(TT), int(method), _ptr(stats), _stream(stream))
    elif sources is not None:
        if mlp is not None:
            raise ValueError("the FP32 step with source planes runs without the learned guess")
        cs = sources.c(cells)
        native.call("ast_step_f32_src", gref, _ptr(s), _ptr(u), _ptr(s2), _ptr(u2), _ptr(T),
                    _ptr(T_do), _ptr(TT), int(method), ctypes.byref(cs), _ptr(stats),
                    _stream(stream))
    else:
        native.call("ast_step_f32", gref, _ptr(s), _ptr(u), _ptr(s2), _ptr(u2), _ptr(T),
                    _ptr(T_do), _ptr(TT), int(method),
                    _ptr(mlp.buffer) if mlp is not None else None, _ptr(stats),
                    _stream(stream))


def funk_time(T, T_do, TT, i=None, stream=None):
    """kernel.cu:1525-1536: T_do <- T; TT += T_do; T <- 1e7; i += 1."""
    native.call("ast_funk_time", _ptr(T), _ptr(T_do), _ptr(TT), _ptr(i), _stream(stream))


def flux_batch(method: int, faces, n1=1.0, n2=0.0, rad=1.0, gamma=5.0 / 3.0, krit=0.05,
               out=None, stats=None, stream=None):
    """The per-face operators over n faces: faces is an (8, n) device tensor
    (rhoL, pL, vxL, vyL, rhoR, pR, vxR, vyR); returns (5, n): mass flux,
    energy flux, x- and y-momentum flux, dt."""
    import torch
    if faces.dim() != 2 or faces.shape[0] != 8 or not faces.is_contiguous():
        raise ValueError("faces must be a contiguous (8, n) tensor")
    n = faces.shape[1]
    if out is None:
        out = torch.empty((5, n), dtype=faces.dtype, device=faces.device)
    if faces.dtype == torch.float64:
        native.call("ast_flux_f64", int(method), n, _ptr(faces), float(n1), float(n2),
                    float(rad), float(gamma), float(krit), _ptr(out), _ptr(stats),
                    _stream(stream))
    elif faces.dtype == torch.float32:
        native.call("ast_flux_f32", int(method), n, _ptr(faces), float(n1), float(n2),
                    float(rad), float(gamma), float(krit), _ptr(out), _ptr(stats),
                    _stream(stream))
    else:
        raise TypeError("faces must be float64 or float32")
    return out


def HLL(faces, n1, n2, rad, **kw):
    """kernel.cu:1399 over a batch of faces."""
    return flux_batch(M_HLL, faces, n1, n2, rad, **kw)


def HLLC_2d_Korolkov(faces, n1, n2, rad, **kw):
    """Gas_Dinamic.cu:1814 (nul_potok=false) over a batch of faces."""
    return flux_batch(M_HLLC, faces, n1, n2, rad, **kw)


def Godunov_Solver_Alexashov(faces, n1, n2, rad, **kw):
    """kernel.cu:128 over a batch of faces."""
    return flux_batch(EXACT, faces, n1, n2, rad, **kw)


def HLLC_Aleksashov(faces, n1, n2, rad, **kw):
    """kernel.cu:767 (zero magnetic field, as add_MK calls it) over a batch of
    faces."""
    return flux_batch(M_HLLC_A, faces, n1, n2, rad, **kw)


def flux_q_batch(method: int, faces, n1=1.0, n2=0.0, rad=1.0, gamma=5.0 / 3.0, krit=0.05,
                 out=None, stream=None):
    """The passive-scalar operators over n faces: faces is a (10, n) float64
    device tensor (the 8 planes of flux_batch, then Q_L, Q_R); returns (6, n):
    the 5 planes of flux_batch, then the scalar's flux PQ.  method: HLLC (the
    live solver's scalar), HLLCQ or HLLCQ_A."""
    import torch
    if (faces.dim() != 2 or faces.shape[0] != 10 or not faces.is_contiguous()
            or faces.dtype != torch.float64):
        raise ValueError("faces must be a contiguous (10, n) float64 tensor")
    n = faces.shape[1]
    if out is None:
        out = torch.empty((6, n), dtype=faces.dtype, device=faces.device)
    native.call("ast_flux_q_f64", int(method), n, _ptr(faces), float(n1), float(n2), float(rad),
                float(gamma), float(krit), _ptr(out), _stream(stream))
    return out


def HLLCQ_Korolkov_2D(faces, n1, n2, rad, **kw):
    """Gas_Dinamic.cu:249 over a batch of faces with a passive scalar."""
    return flux_q_batch(M_HLLCQ, faces, n1, n2, rad, **kw)


def HLLCQ_Aleksashov(faces, n1, n2, rad, **kw):
    """Gas_Dinamic.cu:355 (zero magnetic field) over a batch of faces with a
    passive scalar."""
    return flux_q_batch(M_HLLCQ_A, faces, n1, n2, rad, **kw)


def flux_mhd_batch(method: int, faces, n1=1.0, n2=0.0, n3=0.0, rad=1.0, gamma=5.0 / 3.0,
                   krit=0.05, out=None, stream=None):
    """The MHD operators over n faces: faces is a (18, n) float64 device
    tensor (rho, Q, p, v1, v2, v3, Bx, By, Bz of the left state, then of the
    right); returns (10, n): mass, momentum x/y/z, field x/y/z, energy
    fluxes, the scalar's flux PQ and dt.  method: HLLC_MHD or HLLD."""
    import torch
    if (faces.dim() != 2 or faces.shape[0] != 18 or not faces.is_contiguous()
            or faces.dtype != torch.float64):
        raise ValueError("faces must be a contiguous (18, n) float64 tensor")
    n = faces.shape[1]
    if out is None:
        out = torch.empty((10, n), dtype=faces.dtype, device=faces.device)
    native.call("ast_flux_mhd_f64", int(method), n, _ptr(faces), float(n1), float(n2),
                float(n3), float(rad), float(gamma), float(krit), _ptr(out), _stream(stream))
    return out


def HLLDQ_Korolkov(faces, n1, n2, n3, rad, metod, **kw):
    """Gas_Dinamic.cu:874 over a batch of faces: metod 2 is HLLC for MHD,
    metod 3 is HLLD (the reference's numbering)."""
    if metod not in (2, 3):
        raise ValueError("metod must be 2 (HLLC for MHD) or 3 (HLLD)")
    return flux_mhd_batch(M_HLLD if metod == 3 else M_HLLC_MHD, faces, n1, n2, n3, rad, **kw)


def mc_rates(u, cp, out=None, stream=None):
    """Velosity_1/2/3 (Gas_Dinamic.cu:1566-1602) of float64 device tensors
    u, cp (n each): returns (3, n)."""
    import torch
    if u.shape != cp.shape or u.dim() != 1 or u.dtype != torch.float64 or cp.dtype != torch.float64:
        raise ValueError("u and cp must be float64 vectors of one length")
    u, cp = u.contiguous(), cp.contiguous()
    n = u.shape[0]
    if out is None:
        out = torch.empty((3, n), dtype=torch.float64, device=u.device)
    native.call("ast_mc_rates_f64", n, _ptr(u), _ptr(cp), _ptr(out), _stream(stream))
    return out


def mc_cell_exit(grid, particles, stream=None):
    """Peresechenie (Gas_Dinamic.cu:1492) of the particles, a (8, n) float64
    device tensor (x0, y0, x, y, z, Vx, Vy, Vz): returns (t, mode, hit)."""
    import torch
    if particles.dim() != 2 or particles.shape[0] != 8 or particles.dtype != torch.float64:
        raise ValueError("particles must be an (8, n) float64 tensor")
    p = particles.contiguous()
    n = p.shape[1]
    t = torch.empty(n, dtype=torch.float64, device=p.device)
    mode = torch.empty(n, dtype=torch.int32, device=p.device)
    hit = torch.empty(n, dtype=torch.int32, device=p.device)
    native.call("ast_mc_cell_exit_f64", native.grid_ref(grid), n, _ptr(p), _ptr(t), _ptr(mode),
                _ptr(hit), _stream(stream))
    return t, mode, hit


def mc_belong_point(grid, xy, stream=None):
    """Belong_point (Gas_Dinamic.cu:1559) of the points xy, a (2, n) float64
    device tensor (x, r): returns the (2, n) int32 cell indices (n, m)."""
    import torch
    if xy.dim() != 2 or xy.shape[0] != 2 or xy.dtype != torch.float64:
        raise ValueError("xy must be a (2, n) float64 tensor")
    xy = xy.contiguous()
    nm = torch.empty((2, xy.shape[1]), dtype=torch.int32, device=xy.device)
    native.call("ast_mc_belong_point_f64", native.grid_ref(grid), xy.shape[1], _ptr(xy), _ptr(nm),
                _stream(stream))
    return nm


def conservation_sums(grid, s, u, stream=None):
    """Compensated FP64 (mass, energy) integrals of the state (device)."""
    import torch
    work = torch.empty(int(native.load().ast_conservation_work_doubles()), dtype=torch.float64,
                       device=s.device)
    native.call("ast_conservation_sums", native.grid_ref(grid), _ptr(s), _ptr(u),
                1 if s.dtype == torch.float32 else 0, _ptr(work), _stream(stream))
    return work[:2]
