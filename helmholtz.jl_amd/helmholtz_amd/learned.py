"""Learned coarse-grid correction: the U-Net, its training pairs and its use
inside the SL-MG preconditioner (libhelmholtz_amd, no counterpart in
Helmholtz.jl).  Training pairs come from the plain SL-MG itself: the solver
records the residuals that reach a coarse level during FGMRES iterations and
the targets are that level's shifted-operator solves."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .solver import _ensure_handle, _validate


class UNet:
    def __init__(self, seed: int = 0):
        h = C.c_void_p()
        L.check(L.lib().hh_unet_create(int(seed), C.byref(h)))
        self.ptr = h
        self.nparams = int(L.lib().hh_unet_num_params())

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p:
            try:
                L.lib().hh_unet_destroy(p)
            except Exception:
                pass
            self.ptr = None

    def get_params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        L.check(L.lib().hh_unet_get_params(self.ptr, out.ctypes.data))
        return out

    def set_params(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        assert t.size == self.nparams
        L.check(L.lib().hh_unet_set_params(self.ptr, t.ctypes.data))

    def forward(self, feats, nx: int, ny: int) -> np.ndarray:
        f = np.ascontiguousarray(feats, dtype=np.float32)
        n = f.shape[0]
        y = np.empty((n, nx * ny, 2), dtype=np.float32)
        L.check(L.lib().hh_unet_forward(self.ptr, n, nx, ny, f.ctypes.data, y.ctypes.data))
        return y

    def loss_grad(self, feats, target, nx: int, ny: int):
        f = np.ascontiguousarray(feats, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = C.c_float()
        g = np.empty(self.nparams, dtype=np.float32)
        L.check(L.lib().hh_unet_loss_grad(self.ptr, f.shape[0], nx, ny, f.ctypes.data, t.ctypes.data, C.byref(loss),
                                          g.ctypes.data))
        return float(loss.value), g

    def adam_step(self, grad=None, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        gp = None if grad is None else np.ascontiguousarray(grad, dtype=np.float32).ctypes.data
        L.check(L.lib().hh_unet_adam_step(self.ptr, gp, lr, beta1, beta2, eps))

    def train(self, data: "Dataset", epochs: int, batch: int = 16, lr: float = 1e-3, seed: int = 0):
        losses = np.zeros(epochs, dtype=np.float32)
        L.check(L.lib().hh_unet_train(self.ptr, data.ptr, epochs, batch, lr, int(seed),
                                      losses.ctypes.data_as(C.POINTER(C.c_float))))
        return losses

    def eval(self, data: "Dataset", first: int = 0, count: int | None = None) -> float:
        n = data.size() - first if count is None else count
        err = C.c_float()
        L.check(L.lib().hh_unet_eval(self.ptr, data.ptr, first, n, C.byref(err)))
        return float(err.value)


class Dataset:
    def __init__(self, nx: int, ny: int, capacity: int):
        h = C.c_void_p()
        L.check(L.lib().hh_dataset_create(nx, ny, capacity, C.byref(h)))
        self.ptr, self.nx, self.ny = h, nx, ny

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p:
            try:
                L.lib().hh_dataset_destroy(p)
            except Exception:
                pass
            self.ptr = None

    def size(self) -> int:
        n = C.c_int64()
        L.check(L.lib().hh_dataset_size(self.ptr, C.byref(n)))
        return int(n.value)

    def add_from_solver(self, solver, sweeps: int = 12):
        L.check(L.lib().hh_dataset_add_from_solver(self.ptr, solver._handle, sweeps))


def capture_begin(solver, level: int, max_samples: int, iter_lo: int = 0, iter_hi: int = 10):
    """Record the level-`level` right-hand sides of the solver's next solves."""
    _validate(solver)
    _ensure_handle(solver)
    L.check(L.lib().hh_solver_capture_begin(solver._handle, level, max_samples, iter_lo, iter_hi))


def capture_count(solver) -> int:
    n = C.c_int64()
    L.check(L.lib().hh_solver_capture_count(solver._handle, C.byref(n)))
    return int(n.value)


def attach(solver, net: UNet | None, level: int = 2):
    """Use `net` as the coarse-grid correction at `level` (None detaches)."""
    _validate(solver)
    _ensure_handle(solver)
    solver._net_ref = net
    L.check(L.lib().hh_solver_attach_unet(solver._handle, None if net is None else net.ptr, level))


def policy(solver, always: bool):
    """How the attached network is used: the two-restart trial (always=False,
    the default) or every restart cycle with no trial (always=True, the
    ablation: only non-finite results switch to the plain cycle)."""
    _validate(solver)
    _ensure_handle(solver)
    L.check(L.lib().hh_solver_learned_policy(solver._handle, L.HH_LEARN_ALWAYS if always else L.HH_LEARN_TRIAL))


def workspace(solver):
    """(activations, staging) device addresses of the solver's own network
    workspace (0 where the network has not run on it yet): inspection hook."""
    a, b = C.c_void_p(), C.c_void_p()
    L.check(L.lib().hh_solver_learned_workspace(solver._handle, C.byref(a), C.byref(b)))
    return int(a.value or 0), int(b.value or 0)
