"""The learned starting point of the exact solver's newton iteration.

`LearnedGuess` holds an ast_mlp_t (device block: FP32 masters + BF16 MFMA
operands) and drives the HIP kernels of csrc/mlp.hip: the solver's own face
log (features and FP64 targets), forward (FP32 VALU or BF16 MFMA), loss and
gradient, Adam.  `train()` is the GPU training loop on logged faces of the
Riemann-flux microbenchmark distribution.  The trained parameters ship as
data/mlp_guess.npy (2369 float32).

Command line:  python -m asteriy_2d_amd.learned --train --faces 10000000
"""
from __future__ import annotations

import argparse
import ctypes
import os
import time

import numpy as np

from . import native

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PARAMS = os.path.join(HERE, "data", "mlp_guess.npy")
NPARAM = 2369

_vp, _c_int, _c_size = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
MLP_SIGNATURES = {
    "ast_mlp_bytes": (_c_size, []),
    "ast_mlp_work_floats": (_c_size, []),
    "ast_mlp_features": (_c_int, [_c_size, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  _vp, _vp, _vp, _vp]),
    "ast_mlp_forward": (_c_int, [_vp, _c_size, _vp, _vp, _c_int, _vp]),
    "ast_mlp_loss_grad": (_c_int, [_vp, _c_size, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ast_mlp_adam": (_c_int, [_vp, _vp, _vp, _c_int, ctypes.c_float, _vp]),
    "ast_mlp_refresh": (_c_int, [_vp, _vp]),
    "ast_mlp_set_params": (_c_int, [_vp, _vp, _vp]),
    "ast_mlp_get_params": (_c_int, [_vp, _vp, _vp]),
    "ast_flux_f64_guess": (_c_int, [_c_size, _vp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, _vp, _vp,
                                    _vp, _vp, _vp]),
}
native.SIGNATURES.update(MLP_SIGNATURES)


def _lib():
    L = native.load()
    for name, (res, args) in MLP_SIGNATURES.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


def _p(t):
    return _vp(t.data_ptr()) if t is not None else None


def _st(stream=None):
    import torch
    return _vp((stream or torch.cuda.current_stream()).cuda_stream)


def glorot_init(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases, in the kernels' parameter order
    (W1[32][6] b1 W2[32][32] b2 W3[32][32] b3 w4[32] b4)."""
    rng = np.random.default_rng(seed)

    def glorot(fan_out, fan_in):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        return rng.uniform(-lim, lim, (fan_out, fan_in))

    parts = [glorot(32, 6), np.zeros(32), glorot(32, 32), np.zeros(32), glorot(32, 32),
             np.zeros(32), glorot(1, 32).ravel(), np.zeros(1)]
    p = np.concatenate([x.ravel() for x in parts]).astype(np.float32)
    assert p.size == NPARAM
    return p


class LearnedGuess:
    """An ast_mlp_t in device memory."""

    def __init__(self, params: np.ndarray | None = None, device="cuda:0", seed: int = 0):
        import torch
        self.L = _lib()
        self.device = torch.device(device)
        self.buffer = torch.zeros(int(self.L.ast_mlp_bytes()), dtype=torch.uint8,
                                  device=self.device)
        self.set_params(glorot_init(seed) if params is None else params)

    @classmethod
    def load(cls, path: str = DEFAULT_PARAMS, device="cuda:0"):
        return cls(np.load(path).astype(np.float32), device)

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != NPARAM:
            raise ValueError(f"need {NPARAM} parameters")
        native.call("ast_mlp_set_params", _p(self.buffer), _vp(p.ctypes.data), _st())

    def params(self) -> np.ndarray:
        p = np.empty(NPARAM, dtype=np.float32)
        native.call("ast_mlp_get_params", _p(self.buffer), _vp(p.ctypes.data), _st())
        return p

    def save(self, path: str = DEFAULT_PARAMS):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.save(path, self.params())

    def forward(self, feats, precision: int = 1):
        """feats: (6, n) float32 device tensor -> (n,) predictions of
        log(p_star / p_ref).  precision 0 FP32 VALU, 1 BF16 MFMA."""
        import torch
        n = feats.shape[1]
        out = torch.empty(n, dtype=torch.float32, device=feats.device)
        native.call("ast_mlp_forward", _p(self.buffer), n, _p(feats.contiguous()), _p(out),
                    int(precision), _st())
        return out

    def loss_grad(self, feats, target, mask=None, work=None):
        import torch
        n = feats.shape[1]
        grad = torch.empty(NPARAM, dtype=torch.float32, device=feats.device)
        loss = torch.empty(1, dtype=torch.float32, device=feats.device)
        if work is None:
            work = torch.empty(int(self.L.ast_mlp_work_floats()), dtype=torch.float32,
                               device=feats.device)
        native.call("ast_mlp_loss_grad", _p(self.buffer), n, _p(feats), _p(target), _p(mask),
                    _p(grad), _p(loss), _p(work), _st())
        return loss, grad

    def adam(self, grad, state, step: int, lr: float):
        native.call("ast_mlp_adam", _p(self.buffer), _p(grad), _p(state), int(step), float(lr),
                    _st())


def log_faces(faces, gamma: float = 5.0 / 3.0, n1: float = 1.0, n2: float = 0.0):
    """The solver's face log: features (6, n), targets log(p*/p_ref) (n,) from
    the FP64 exact solve, two-shock mask (n,) -- all device tensors."""
    import torch
    _lib()
    n = faces.shape[1]
    feats = torch.empty((6, n), dtype=torch.float32, device=faces.device)
    target = torch.empty(n, dtype=torch.float32, device=faces.device)
    mask = torch.empty(n, dtype=torch.uint8, device=faces.device)
    native.call("ast_mlp_features", n, _p(faces.contiguous()), float(n1), float(n2), float(gamma),
                _p(feats), _p(target), _p(mask), _st())
    return feats, target, mask


def guesses(model: LearnedGuess, faces, gamma: float = 5.0 / 3.0, precision: int = 1):
    """p* guesses (device float32) for faces (8, n) with normal (1, 0):
    exp(prediction) * sqrt(pL pR)."""
    import torch
    feats, _, _ = log_faces(faces, gamma)
    pred = model.forward(feats, precision)
    pref = torch.sqrt(faces[1] * faces[5])
    return (torch.exp(pred.double()) * pref).float()


def solve_with_guess(faces, guess=None, gamma: float = 5.0 / 3.0, rad: float = 1.0,
                     krit: float = 0.05):
    """FP64 exact solver over faces (8, n), optional per-face p* guesses:
    returns (out (5, n), info (n, 2) = iterations, flags)."""
    import torch
    n = faces.shape[1]
    out = torch.empty((5, n), dtype=torch.float64, device=faces.device)
    info = torch.empty((n, 2), dtype=torch.int32, device=faces.device)
    native.call("ast_flux_f64_guess", n, _p(faces.contiguous()), 1.0, 0.0, float(rad),
                float(gamma), float(krit), _p(guess), _p(out), None, _p(info), _st())
    return out, info


def train(n_faces: int = 10_000_000, steps: int = 3000, batch: int = 1 << 16, lr: float = 3e-3,
          seed: int = 1, gamma: float = 5.0 / 3.0, log=print):
    """Logs two-shock faces of the microbenchmark distribution until n_faces
    are collected, then trains with Adam on minibatches (all on the GPU)."""
    import torch

    from . import facegen
    dev = torch.device("cuda:0")
    feats_l, targ_l = [], []
    got, s = 0, seed
    t0 = time.time()
    while got < n_faces:
        F = torch.from_numpy(facegen.faces(1 << 22, gamma=gamma, seed=s)).to(dev)
        f, t, m = log_faces(F, gamma)
        sel = m.bool()
        feats_l.append(f[:, sel])
        targ_l.append(t[sel])
        got += int(sel.sum())
        s += 1
    feats = torch.cat(feats_l, 1)[:, :n_faces].contiguous()
    targ = torch.cat(targ_l)[:n_faces].contiguous()
    log(f"logged {feats.shape[1]} two-shock faces in {time.time() - t0:.1f} s")
    n = feats.shape[1]
    n_val = min(200_000, n // 20)
    val_f, val_t = feats[:, :n_val].contiguous(), targ[:n_val].contiguous()
    tr_f, tr_t = feats[:, n_val:], targ[n_val:]
    ntr = tr_f.shape[1]
    model = LearnedGuess(seed=seed)
    state = torch.zeros(2 * NPARAM, dtype=torch.float32, device=dev)
    work = torch.empty(int(model.L.ast_mlp_work_floats()), dtype=torch.float32, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    t0 = time.time()
    for k in range(1, steps + 1):
        idx = torch.randint(0, ntr, (batch,), device=dev, generator=g)
        bf = tr_f[:, idx].contiguous()
        bt = tr_t[idx].contiguous()
        lr_k = lr * (0.5 * (1 + np.cos(np.pi * k / steps)) * 0.95 + 0.05)
        loss, grad = model.loss_grad(bf, bt, None, work)
        model.adam(grad, state, k, lr_k)
        if k % max(1, steps // 10) == 0 or k == 1:
            vl, _ = model.loss_grad(val_f, val_t, None, work)
            log(f"step {k}: train loss {loss.item():.5f}  val loss {vl.item():.5f}  "
                f"({time.time() - t0:.1f} s)")
    return model, (val_f, val_t)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", action="store_true")
    ap.add_argument("--faces", type=int, default=10_000_000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default=DEFAULT_PARAMS)
    a = ap.parse_args(argv)
    if a.train:
        model, _ = train(a.faces, a.steps)
        model.save(a.out)
        print(f"saved {a.out}")


if __name__ == "__main__":
    main()
