"""Force matching over the C ABI (mtmd_fm_*): loss and adjoint gradient with
respect to LJ (eps, sigma), harmonic bond (k, r0) and harmonic angle
(k, theta0[deg]) parameters, a torch.autograd.Function around it, and an Adam
fit. Frames and target forces live in device memory inside the library;
torch only carries the handful of parameters.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .engine import MtmdError, track
from .system import System

_dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


class ForceMatch:
    def __init__(self, system: System, cutoff: float, frames: np.ndarray, target: np.ndarray | None = None):
        self.lib = abi.load_library()
        frames = np.ascontiguousarray(frames, dtype=np.float64).reshape(-1, system.n_atoms * 3)
        self.n_frames = frames.shape[0]
        self.n_atoms = system.n_atoms
        self._csys = system.to_c()
        tgt = None if target is None else np.ascontiguousarray(target, dtype=np.float64).reshape(frames.shape)
        h = C.c_void_p()
        rc = self.lib.mtmd_fm_create(C.byref(self._csys), float(cutoff), self.n_frames, _p(frames),
                                     None if tgt is None else _p(tgt), C.byref(h))
        if rc:
            raise MtmdError(rc, self.lib.mtmd_fm_last_error(None).decode())
        self.h = h
        track(self)
        self.n_params = self.lib.mtmd_fm_num_params(self.h)

    def _chk(self, rc):
        if rc:
            raise MtmdError(rc, self.lib.mtmd_fm_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mtmd_fm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def target_from_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._chk(self.lib.mtmd_fm_target_from_params(self.h, _p(p)))

    def set_target(self, target):
        t = np.ascontiguousarray(target, dtype=np.float64).reshape(self.n_frames, -1)
        self._chk(self.lib.mtmd_fm_set_target(self.h, _p(t)))

    def forces(self, params, frame: int, fp64: bool = True) -> np.ndarray:
        p = np.ascontiguousarray(params, dtype=np.float64)
        F = np.zeros((self.n_atoms, 3))
        self._chk(self.lib.mtmd_fm_forces(self.h, _p(p), frame, int(fp64), _p(F)))
        return F

    def loss_grad(self, params, fp64: bool = False):
        p = np.ascontiguousarray(params, dtype=np.float64)
        loss = C.c_double(0.0)
        g = np.zeros(self.n_params)
        self._chk(self.lib.mtmd_fm_loss_grad(self.h, _p(p), int(fp64), C.byref(loss), _p(g)))
        return loss.value, g


def params_of(system: System) -> np.ndarray:
    """The parameter vector of a system's own force fields, in the
    mtmd_fm layout: (eps, sigma) per LJ entry, (k, r0) per bond type,
    (k, theta0) per angle type."""
    out = []
    for _, _, _, prm in system.nonbonded:
        out += [prm[0], prm[1]]
    for _, prm in system.bond_ff:
        out += [prm[0], prm[1]]
    for _, prm in system.angle_ff:
        out += [prm[0], prm[1]]
    return np.array(out, dtype=np.float64)


def loss_function(fm: ForceMatch, fp64: bool = False):
    """A torch.autograd.Function whose forward is the force-matching loss of
    `fm` and whose backward is the adjoint gradient (computed with the loss)."""
    import torch

    class _Loss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, params):
            loss, grad = fm.loss_grad(params.detach().cpu().double().numpy(), fp64)
            ctx.save_for_backward(torch.as_tensor(grad, dtype=params.dtype))
            return params.new_tensor(loss)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out * g

    return _Loss.apply


def fit(fm: ForceMatch, start: np.ndarray, iters: int = 500, lr: float = 1e-2, fp64: bool = False):
    """Adam on relative parameters theta = start * s (s starts at 1), through
    the autograd Function. Returns (theta, loss history)."""
    import torch

    f = loss_function(fm, fp64)
    base = torch.as_tensor(np.asarray(start, dtype=np.float64))
    s = torch.ones_like(base, requires_grad=True)
    opt = torch.optim.Adam([s], lr=lr)
    hist = []
    for _ in range(iters):
        opt.zero_grad()
        loss = f(base * s)
        loss.backward()
        opt.step()
        hist.append(float(loss.detach()))
    return (base * s).detach().numpy(), hist
