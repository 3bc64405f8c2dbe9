"""ctypes front-end of the CPU ORACLE (test infrastructure only).

Loads oracle/liboracle.so (built by oracle/Makefile).  Only tests/, the smoke
check and the bench's CPU-baseline leg import this module; the product never
does.  Arrays are numpy float64; state planes are (M, N, 2) views of the
reference's double2 planes s = {rho, p}, u = {vx, vy}.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class CPhys(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("ga", "g1", "g2", "gp", "gm", "krit")]


class CInfo(ctypes.Structure):
    _fields_ = [("pattern", ctypes.c_int), ("iters", ctypes.c_int),
                ("capped", ctypes.c_int), ("swapped", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in ("p_star", "u_star", "enI", "pI", "rI",
                                        "enII", "pII", "rII")]


class CMhdFan(ctypes.Structure):
    _fields_ = [("S", ctypes.c_double * 5), ("U", (ctypes.c_double * 9) * 6),
                ("F", (ctypes.c_double * 9) * 2), ("branch", ctypes.c_int)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_HOST_BUILDS = (("x86-64-v4", "liboracle_v4.so", ("avx512f", "avx512bw", "avx512dq", "avx512vl")),
                ("x86-64-v3", "liboracle_v3.so", ("avx2", "fma", "bmi2", "movbe")),
                ("x86-64-v2", "liboracle.so", ()))


def select_host_build() -> str:
    """Loads the oracle build for the widest instruction set this host's
    /proc/cpuinfo lists (for the timed CPU baseline; every build gives the
    same bits) and returns its -march level."""
    global _lib, LIB_PATH
    try:
        flags = set()
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                flags = set(line.split(":", 1)[1].split())
                break
    except OSError:
        flags = set()
    for march, name, need in _HOST_BUILDS:
        path = os.path.join(HERE, name)
        if all(f in flags for f in need) and os.path.exists(path):
            if path != LIB_PATH:
                LIB_PATH = path
                _lib = None
            lib()
            return march
    return "x86-64-v2"


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.ora_phys_init.argtypes = [ctypes.POINTER(CPhys), ctypes.c_double, ctypes.c_double]
        L.ora_minmod.restype = ctypes.c_double
        L.ora_minmod.argtypes = [ctypes.c_double, ctypes.c_double]
        L.ora_linear.restype = ctypes.c_double
        L.ora_linear.argtypes = [ctypes.c_double] * 7
        L.ora_linear2.argtypes = [ctypes.c_double] * 8 + [_dp, _dp]
        for fn in ("ora_hll", "ora_hllc", "ora_hllc_a", "ora_godunov"):
            f = getattr(L, fn)
            f.restype = ctypes.c_double
            f.argtypes = [ctypes.POINTER(CPhys), _dp, _dp, _dp, _dp, ctypes.c_double,
                          ctypes.c_double, _dp, _dp, ctypes.c_double]
        L.ora_godunov_ex.restype = ctypes.c_double
        L.ora_godunov_ex.argtypes = [ctypes.POINTER(CPhys), _dp, _dp, _dp, _dp,
                                     ctypes.c_double, ctypes.c_double, _dp, _dp,
                                     ctypes.c_double, _dp, ctypes.POINTER(CInfo)]
        L.ora_flux_batch.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_long, _dp, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, _dp, _ip, ctypes.c_int]
        L.ora_godunov_guess_batch.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_long,
                                              _dp, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, _dp, _dp, _ip, _dp]
        for f in (L.ora_hllc_q, L.ora_hllcq, L.ora_hllcq_a):
            f.restype = ctypes.c_double
            f.argtypes = [ctypes.POINTER(CPhys), _dp, _dp, _dp, _dp, ctypes.c_double,
                          ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp,
                          ctypes.c_double]
        L.ora_flux_q_batch.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_long, _dp, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, _dp, ctypes.c_int]
        L.ora_hlldq.restype = ctypes.c_double
        L.ora_hlldq.argtypes = [ctypes.POINTER(CPhys), _dp, _dp, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                _dp, _dp, ctypes.POINTER(CMhdFan)]
        L.ora_flux_mhd_batch.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_long, _dp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int]
        L.ora_minplus.restype = ctypes.c_double
        L.ora_minplus.argtypes = [ctypes.c_double, ctypes.c_double]
        L.ora_peresich.argtypes = [ctypes.c_double] * 5 + [_dp, _dp]
        L.ora_peresechenie.restype = ctypes.c_int
        L.ora_peresechenie.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 8 + [_ip, _dp]
        L.ora_belong_point.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _ip,
                                       _ip]
        for fn in ("ora_velosity_1", "ora_velosity_2", "ora_velosity_3"):
            getattr(L, fn).restype = ctypes.c_double
            getattr(L, fn).argtypes = [ctypes.c_double, ctypes.c_double]
        L.ora_step.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp,
                               ctypes.c_int, ctypes.c_int]
        L.ora_step_src.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        L.ora_step_rows.restype = ctypes.c_double
        L.ora_step_rows.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_funk_time.argtypes = [_dp, _dp, _dp, _ip]
        L.ora_max_threads.restype = ctypes.c_int
        L.toro_star.restype = ctypes.c_int
        L.toro_star.argtypes = [ctypes.c_double] * 7 + [_dp, _dp, _ip]
        L.toro_sample.argtypes = [ctypes.c_double] * 10 + [_dp, _dp, _dp]
        L.toro_sample_vacuum.restype = ctypes.c_int
        L.toro_sample_vacuum.argtypes = [ctypes.c_double] * 8 + [_dp, _dp, _dp, _dp]
        L.toro_sod_profile.argtypes = [ctypes.c_double] * 9 + [ctypes.c_long, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def phys(gamma: float, krit: float) -> CPhys:
    ph = CPhys()
    lib().ora_phys_init(ctypes.byref(ph), gamma, krit)
    return ph


def minmod(a, b):
    return lib().ora_minmod(a, b)


def linear(x1, t1, x2, t2, x3, t3, y):
    return lib().ora_linear(x1, t1, x2, t2, x3, t3, y)


def linear2(x1, t1, x2, t2, x3, t3, y1, y2):
    A, B = ctypes.c_double(), ctypes.c_double()
    lib().ora_linear2(x1, t1, x2, t2, x3, t3, y1, y2, ctypes.byref(A), ctypes.byref(B))
    return A.value, B.value


_FLUX = {0: "ora_hll", 1: "ora_hllc", 2: "ora_godunov", 3: "ora_hllc_a"}


def flux(method: int, Ls, Lu, Rs, Ru, n1, n2, rad, gamma=5.0 / 3.0, krit=0.05):
    """One face: returns (Ps, Pu, dt) as the reference operator does."""
    ph = phys(gamma, krit)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (Ls, Lu, Rs, Ru)]
    Ps, Pu = np.zeros(2), np.zeros(2)
    t = getattr(lib(), _FLUX[method])(ctypes.byref(ph), *[_p(v) for v in a], n1, n2,
                                       _p(Ps), _p(Pu), rad)
    return Ps, Pu, t


# passive-scalar operators: AST_HLLC (the live solver's scalar), AST_HLLCQ,
# AST_HLLCQ_A
_FLUX_Q = {1: "ora_hllc_q", 4: "ora_hllcq", 5: "ora_hllcq_a"}


def flux_q(method: int, Ls, Lu, Rs, Ru, QL, QR, n1, n2, rad, gamma=5.0 / 3.0, krit=0.05):
    """One face with a passive scalar: returns (Ps, Pu, PQ, dt)."""
    ph = phys(gamma, krit)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (Ls, Lu, Rs, Ru)]
    Ps, Pu, PQ = np.zeros(2), np.zeros(2), np.zeros(1)
    t = getattr(lib(), _FLUX_Q[method])(ctypes.byref(ph), *[_p(v) for v in a], QL, QR, n1, n2,
                                         _p(Ps), _p(Pu), _p(PQ), rad)
    return Ps, Pu, float(PQ[0]), t


def flux_q_batch(method: int, faces: np.ndarray, n1=1.0, n2=0.0, rad=1.0, gamma=5.0 / 3.0,
                 krit=0.05, nthreads: int = 0):
    """faces: (10, n) SoA (the 8 planes of flux_batch, Q_L, Q_R) -> out (6, n)
    (the 5 planes of flux_batch, PQ)."""
    faces = np.ascontiguousarray(faces, dtype=np.float64)
    assert faces.shape[0] == 10
    n = faces.shape[1]
    out = np.empty((6, n))
    lib().ora_flux_q_batch(method, gamma, krit, n, _p(faces), n1, n2, rad, _p(out), nthreads)
    return out


def flux_mhd(method: int, L, R, n=(1.0, 0.0, 0.0), rad=1.0, gamma=5.0 / 3.0, krit=0.05,
             fan: bool = False):
    """One MHD face (HLLDQ_Korolkov): L, R = (rho, Q, p, v1, v2, v3, Bx, By, Bz).
    Returns (P (8,), PQ, dt) or, with fan, (P, PQ, dt, fan) where fan is a
    dict: S (SL, SLL, SM, SRR, SR), U (6, 9) = UL, U*L, U**L, U**R, U*R, UR
    and F (2, 9) = FL, FR in the frame of the normal, branch (0..5)."""
    ph = phys(gamma, krit)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (L, R)]
    assert a[0].shape == (9,) and a[1].shape == (9,)
    P, PQ = np.zeros(8), np.zeros(1)
    f = CMhdFan()
    t = lib().ora_hlldq(ctypes.byref(ph), _p(a[0]), _p(a[1]), float(n[0]), float(n[1]),
                        float(n[2]), rad, int(method), _p(P), _p(PQ), ctypes.byref(f))
    if not fan:
        return P, float(PQ[0]), t
    d = {"S": np.array(f.S[:]), "U": np.array([f.U[k][:] for k in range(6)]),
         "F": np.array([f.F[k][:] for k in range(2)]), "branch": int(f.branch)}
    return P, float(PQ[0]), t, d


def flux_mhd_batch(method: int, faces: np.ndarray, n=(1.0, 0.0, 0.0), rad=1.0,
                   gamma=5.0 / 3.0, krit=0.05, nthreads: int = 0):
    """faces: (18, n) SoA (rho, Q, p, v1, v2, v3, Bx, By, Bz left, then right)
    -> out (10, n): the 8 fluxes, PQ, dt."""
    faces = np.ascontiguousarray(faces, dtype=np.float64)
    assert faces.shape[0] == 18
    m = faces.shape[1]
    out = np.empty((10, m))
    lib().ora_flux_mhd_batch(int(method), gamma, krit, m, _p(faces), float(n[0]), float(n[1]),
                             float(n[2]), rad, _p(out), nthreads)
    return out


def minplus(x, y):
    return lib().ora_minplus(x, y)


def peresich(y, z, Vy, Vz, R):
    """Gas_Dinamic.cu:1543: the times at which the ray (y, z) + t (Vy, Vz)
    meets the circle of radius R, or (-1, -1)."""
    t1, t2 = ctypes.c_double(), ctypes.c_double()
    lib().ora_peresich(y, z, Vy, Vz, R, ctypes.byref(t1), ctypes.byref(t2))
    return t1.value, t2.value


def peresechenie(grid, x0, y0, x, y, z, Vx, Vy, Vz):
    """Gas_Dinamic.cu:1492: (hit, mode, t) of the ray leaving the cell with
    centre (x0, y0) of the grid's dx, dy."""
    cg = grid.to_c()
    mode, t = ctypes.c_int(), ctypes.c_double()
    hit = lib().ora_peresechenie(ctypes.byref(cg), x0, y0, x, y, z, Vx, Vy, Vz,
                                 ctypes.byref(mode), ctypes.byref(t))
    return bool(hit), mode.value, t.value


def belong_point(grid, x, y):
    cg = grid.to_c()
    n, m = ctypes.c_int(), ctypes.c_int()
    lib().ora_belong_point(ctypes.byref(cg), x, y, ctypes.byref(n), ctypes.byref(m))
    return n.value, m.value


def velosity(k: int, u: float, cp: float) -> float:
    """Velosity_k (Gas_Dinamic.cu:1566, 1578, 1591)."""
    return getattr(lib(), f"ora_velosity_{k}")(u, cp)


def godunov_info(Ls, Lu, Rs, Ru, n1=1.0, n2=0.0, rad=1.0, gamma=5.0 / 3.0, krit=0.05,
                 guess=None):
    ph = phys(gamma, krit)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (Ls, Lu, Rs, Ru)]
    Ps, Pu = np.zeros(2), np.zeros(2)
    info = CInfo()
    g = None if guess is None else ctypes.byref(ctypes.c_double(guess))
    t = lib().ora_godunov_ex(ctypes.byref(ph), *[_p(v) for v in a], n1, n2, _p(Ps), _p(Pu),
                             rad, g, ctypes.byref(info))
    return Ps, Pu, t, info


def flux_batch(method: int, faces: np.ndarray, n1=1.0, n2=0.0, rad=1.0, gamma=5.0 / 3.0,
               krit=0.05, nthreads: int = 0, want_iters: bool = False):
    """faces: (8, n) SoA -> out (5, n) [, iters (n, 2) = (iterations, pattern)]."""
    faces = np.ascontiguousarray(faces, dtype=np.float64)
    n = faces.shape[1]
    out = np.empty((5, n))
    it = np.zeros((n, 2), dtype=np.int32) if want_iters else None
    lib().ora_flux_batch(method, gamma, krit, n, _p(faces), n1, n2, rad, _p(out),
                         it.ctypes.data_as(_ip) if it is not None else None, nthreads)
    return (out, it) if want_iters else out


def godunov_guess_batch(faces: np.ndarray, guess: np.ndarray | None, n1=1.0, n2=0.0,
                        rad=1.0, gamma=5.0 / 3.0, krit=0.05):
    """Exact solver with per-face Newton starting points (NaN = PVRS)."""
    faces = np.ascontiguousarray(faces, dtype=np.float64)
    n = faces.shape[1]
    out = np.empty((5, n))
    it = np.zeros((n, 2), dtype=np.int32)
    ps = np.empty(n)
    g = None
    if guess is not None:
        g = np.ascontiguousarray(guess, dtype=np.float64)
    lib().ora_godunov_guess_batch(gamma, krit, n, _p(faces), n1, n2, rad,
                                  _p(g) if g is not None else None, _p(out),
                                  it.ctypes.data_as(_ip), _p(ps))
    return out, it, ps


class CSrc(ctypes.Structure):
    _fields_ = [("nn2", ctypes.c_void_p), ("nn3", ctypes.c_void_p),
                ("kmom", ctypes.c_double), ("ken", ctypes.c_double)]


class Sources:
    """Neutral-H source planes of add_MK (kernel.cu:2495-2498): nn2 (M,N,3)
    and nn3 (M,N) float64 (either may be None), momentum coefficient kmom and
    energy coefficient ken (the reference: n_H/Kn and 0.0*n_H/Kn,
    Header.h:84-85)."""

    def __init__(self, nn2=None, nn3=None, kmom=3.0 / 0.419231, ken=0.0):
        self.nn2 = None if nn2 is None else np.ascontiguousarray(nn2, dtype=np.float64)
        self.nn3 = None if nn3 is None else np.ascontiguousarray(nn3, dtype=np.float64)
        self.kmom, self.ken = float(kmom), float(ken)

    def to_c(self):
        return CSrc(None if self.nn2 is None else self.nn2.ctypes.data,
                    None if self.nn3 is None else self.nn3.ctypes.data, self.kmom, self.ken)


def step(grid, s: np.ndarray, u: np.ndarray, T: float, T_do: float, method: int,
         nthreads: int = 0, src: "Sources | None" = None, TT: float = 0.0):
    """One Kernel_TVD step on the CPU.  s,u: (M,N,2) float64; src: neutral-H
    source planes (add_MK's terms) or None; TT: the simulation time (read by
    pulsed Dirichlet sides).  Returns (s2, u2, T_new)."""
    cg = grid.to_c()
    s = np.ascontiguousarray(s, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    s2, u2 = np.empty_like(s), np.empty_like(u)
    Tv = np.array([T], dtype=np.float64)
    Td = np.array([T_do], dtype=np.float64)
    cs = None
    if src is not None:
        if src.nn2 is not None:
            assert src.nn2.shape == (grid.M, grid.N, 3)
        if src.nn3 is not None:
            assert src.nn3.shape == (grid.M, grid.N)
        cs = src.to_c()
    lib().ora_step_src(ctypes.byref(cg), _p(s), _p(u), _p(s2), _p(u2), _p(Tv), _p(Td),
                       float(TT), method, ctypes.byref(cs) if cs is not None else None,
                       nthreads)
    return s2, u2, float(Tv[0])


def step_rows(grid, s, u, T_do, method, m0, m1, nthreads=0):
    cg = grid.to_c()
    s2, u2 = np.empty_like(s), np.empty_like(u)
    Td = np.array([T_do], dtype=np.float64)
    t = lib().ora_step_rows(ctypes.byref(cg), _p(s), _p(u), _p(s2), _p(u2), _p(Td), method,
                            m0, m1, nthreads)
    return s2, u2, t


def funk_time(T, T_do, TT, i):
    a = [np.array([v], dtype=np.float64) for v in (T, T_do, TT)]
    iv = np.array([i], dtype=np.int32)
    lib().ora_funk_time(*[_p(v) for v in a], iv.ctypes.data_as(_ip))
    return float(a[0][0]), float(a[1][0]), float(a[2][0]), int(iv[0])


class Sim:
    """The reference's host protocol on the CPU (kernel.cu:3868-3906): step,
    then funk_time; initial T = 1e7, T_do = 1e-9, TT = 0 (kernel.cu:3321-3324)."""

    def __init__(self, grid, s, u, method, nthreads=0, src=None):
        self.grid, self.method, self.nthreads, self.src = grid, method, nthreads, src
        self.s = np.ascontiguousarray(s, dtype=np.float64)
        self.u = np.ascontiguousarray(u, dtype=np.float64)
        self.T, self.T_do, self.TT, self.i = 1e7, 1e-9, 0.0, 0

    def advance(self, steps=1):
        for _ in range(steps):
            self.s, self.u, self.T = step(self.grid, self.s, self.u, self.T, self.T_do,
                                          self.method, self.nthreads, self.src, self.TT)
            self.T, self.T_do, self.TT, self.i = funk_time(self.T, self.T_do, self.TT, self.i)
        return self


def toro_star(gamma, dl, ul, pl, dr, ur, pr):
    pm, um, it = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    rc = lib().toro_star(gamma, dl, ul, pl, dr, ur, pr, ctypes.byref(pm), ctypes.byref(um),
                         ctypes.byref(it))
    if rc != 0:
        return None
    return pm.value, um.value


def toro_sample(gamma, dl, ul, pl, dr, ur, pr, S):
    st = toro_star(gamma, dl, ul, pl, dr, ur, pr)
    d, u, p = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().toro_sample(gamma, dl, ul, pl, dr, ur, pr, st[0], st[1], S, ctypes.byref(d),
                      ctypes.byref(u), ctypes.byref(p))
    return d.value, u.value, p.value


def toro_sample_vacuum(gamma, dl, ul, pl, dr, ur, pr, S):
    """Toro's vacuum-generation solution (sect. 4.6) at x/t = S: (d, u, p,
    (left head, left vacuum front, right vacuum front, right head)), or None
    when the data do not generate vacuum."""
    d, u, p = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    sp = np.empty(4)
    rc = lib().toro_sample_vacuum(gamma, dl, ul, pl, dr, ur, pr, S, ctypes.byref(d),
                                  ctypes.byref(u), ctypes.byref(p), _p(sp))
    if rc != 0:
        return None
    return d.value, u.value, p.value, tuple(sp)


def toro_profile(gamma, left, right, x0, t, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    d, u, p = np.empty_like(x), np.empty_like(x), np.empty_like(x)
    lib().toro_sod_profile(gamma, left[0], left[1], left[2], right[0], right[1], right[2],
                           x0, t, x.size, _p(x), _p(d), _p(u), _p(p))
    return d, u, p
