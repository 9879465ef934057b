"""Learned warm start: network handle, training data and training loop.

The network (include/o3d_nn.h) maps two planes, the residual of the previous
pressure r = rhs - L p_prev (times dx^2) and the previous increment
e_prev = p_prev - p_prev2, each divided by its RMS, to the increment
e = p_new - p_prev divided by RMS(e_prev); its output is a learned correction
added to the second plane, so the untrained network starts from the linear
extrapolation e = e_prev.  All three planes are block averages over
WARM_COARSEN^3 cells: the network runs on the coarse grid and the solver
interpolates its correction back.  Training pairs come from the solver's own
runs (o3d_solver_capture): every step of a run yields (r, e_prev, e) with
p_new the SOR solution to eps, reduced to the network's input and target by
the same library call the solver uses (o3d_warm_features).

Device memory is plumbing through torch tensors; all compute is libo3d.so.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import capi
from .config import Config


def _lib():
    lib = capi.load()
    for name in ("o3d_nn_create", "o3d_nn_destroy", "o3d_nn_resize", "o3d_nn_num_params",
                 "o3d_nn_get_params", "o3d_nn_set_params", "o3d_nn_get_grads",
                 "o3d_nn_zero_grad", "o3d_nn_forward", "o3d_nn_backward", "o3d_nn_mse",
                 "o3d_nn_adam_step", "o3d_nn_last_timing"):
        getattr(lib, name).restype = C.c_int
    return lib


NN_EXPORTS = ["o3d_nn_create", "o3d_nn_destroy", "o3d_nn_resize", "o3d_nn_num_params",
              "o3d_nn_get_params", "o3d_nn_set_params", "o3d_nn_get_grads", "o3d_nn_zero_grad",
              "o3d_nn_forward", "o3d_nn_backward", "o3d_nn_mse", "o3d_nn_adam_step",
              "o3d_nn_last_timing"]


class Network:
    """The U-Net-lite of libo3d.so for one grid size (re-sizable)."""

    def __init__(self, nx: int, ny: int, nz: int, seed: int = 0):
        self.lib = _lib()
        self.h = C.c_void_p(0)
        capi.check(self.lib.o3d_nn_create(nx, ny, nz, C.c_ulonglong(seed), C.byref(self.h)))
        n = C.c_longlong(0)
        capi.check(self.lib.o3d_nn_num_params(self.h, C.byref(n)))
        self.nparam = n.value
        self.shape = (nz, ny, nx)
        self.t = 0

    def close(self):
        if self.h:
            self.lib.o3d_nn_destroy(self.h)
            self.h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def resize(self, nx, ny, nz):
        capi.check(self.lib.o3d_nn_resize(self.h, nx, ny, nz))
        self.shape = (nz, ny, nx)

    # parameters ------------------------------------------------------------
    def params(self) -> np.ndarray:
        out = np.empty(self.nparam, dtype=np.float32)
        capi.check(self.lib.o3d_nn_get_params(self.h, capi.ptr(out)))
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        capi.check(self.lib.o3d_nn_set_params(self.h, capi.ptr(p)))

    def grads(self) -> np.ndarray:
        out = np.empty(self.nparam, dtype=np.float32)
        capi.check(self.lib.o3d_nn_get_grads(self.h, capi.ptr(out)))
        return out

    def save(self, path: str):
        np.save(path, self.params(), allow_pickle=False)

    def load(self, path: str):
        self.set_params(np.load(path, allow_pickle=False))

    # compute (torch cuda tensors) -----------------------------------------
    def forward(self, inp, out):
        capi.check(self.lib.o3d_nn_forward(self.h, capi.ptr(inp), capi.ptr(out)))

    def backward(self, dout):
        """Back-propagates dout through the last forward pass (whose input
        tensor must still hold its values) and accumulates the gradients."""
        capi.check(self.lib.o3d_nn_backward(self.h, capi.ptr(dout)))

    def zero_grad(self):
        capi.check(self.lib.o3d_nn_zero_grad(self.h))

    def mse(self, out, target, dout=None, scale: float = 1.0) -> float:
        loss = C.c_double(0.0)
        capi.check(self.lib.o3d_nn_mse(self.h, capi.ptr(out), capi.ptr(target), C.c_float(scale),
                                       capi.ptr(dout), C.byref(loss)))
        return loss.value

    def adam(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.t += 1
        capi.check(self.lib.o3d_nn_adam_step(self.h, C.c_float(lr), C.c_float(b1),
                                             C.c_float(b2), C.c_float(eps), self.t))

    def last_timing(self) -> Tuple[float, float]:
        ms, fl = C.c_double(0), C.c_double(0)
        capi.check(self.lib.o3d_nn_last_timing(self.h, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value

    def attach(self, solver):
        capi.check(capi.load().o3d_solver_set_warm_model(solver.handle, self.h))


def load_or_none(path: str, shape=(256, 256, 256)) -> Optional[Network]:
    """The trained network for a solver grid `shape` (nx, ny, nz), or None
    when `path` does not exist.  The network itself runs on the grid
    coarsened by capi.WARM_COARSEN."""
    if not os.path.exists(path):
        return None
    F = capi.WARM_COARSEN
    net = Network(*[s // F for s in shape])
    net.load(path)
    return net


# ------------------------------------------------------------- training data
def normalize(r, eprev, e, dx: float):
    """Network input and target as the solver forms them (o3d_warm_features):
    returns (inp [2, nc], target [nc], scale) on the coarse grid, or None
    when e_prev vanishes.  r, eprev, e: float32 device planes of shape
    (nz, ny, nx) or flat with `dims` set."""
    import torch
    nz, ny, nx = r.shape
    F = capi.WARM_COARSEN
    nc = (nx // F) * (ny // F) * (nz // F)
    inp = torch.empty(2, nc, dtype=torch.float32, device=r.device)
    tgt = torch.empty(nc, dtype=torch.float32, device=r.device)
    scale = C.c_double(0.0)
    rc = capi.check(capi.load().o3d_warm_features(
        capi.ptr(r), capi.ptr(eprev), capi.ptr(e), nx, ny, nz, C.c_double(dx), capi.ptr(inp),
        capi.ptr(tgt), C.byref(scale)))
    if rc == 1:
        return None
    return inp, tgt, scale.value


def capture_run(cfg: Config, steps: int, precision: int, init=None, skip: int = 2,
                keep_every: int = 1, store_bf16: bool = False, reasons=None):
    """Run the solver and keep (inp, target) samples of every kept step.
    init: initial fields, or a callable that initialises the solver (the
    x-Dirichlet jet also sets its inflow).  reasons: a list that receives
    each step's stop reason."""
    import torch
    from .solver import Solver, taylor_green
    s = Solver(cfg, precision)
    if init is None:
        s.upload(*taylor_green(cfg))
    elif callable(init):
        init(s)
    else:
        s.upload(*init)
    r = torch.empty((cfg.nz, cfg.ny, cfg.nx), dtype=torch.float32, device="cuda")
    ep = torch.empty_like(r)
    e = torch.empty_like(r)
    capi.check(capi.load().o3d_solver_capture(s.handle, capi.ptr(r), capi.ptr(ep), capi.ptr(e)))
    dx = cfg.spacing()[0]
    samples = []
    iters = []
    last_kmax = -10
    for it in range(1, steps + 1):
        info = s.step(it)
        if info["diverged"] or info["has_nan"]:
            break   # the run left the regime the warm start is meant for
        iters.append(info["sor_exit_iter"])
        if reasons is not None:
            reasons.append(info["sor_reason"])
        if info["sor_reason"] == 2:
            # A solve stopped by kmax leaves an unconverged pressure: neither
            # its increment nor the next one (whose input it is) is a sample.
            last_kmax = it
            continue
        if it > skip and it > last_kmax + 1 and (it - skip) % keep_every == 0:
            z = normalize(r, ep, e, dx)
            if z is not None:
                inp, tgt, _ = z
                if bool(torch.isfinite(inp).all()) and bool(torch.isfinite(tgt).all()):
                    samples.append((inp, tgt))
    capi.check(capi.load().o3d_solver_capture(s.handle, None, None, None))
    s.close()
    return samples, iters


def relative_errors(net: Network, samples) -> Dict[str, float]:
    """Mean relative L2 error of the predicted increment, against the
    previous-pressure guess (e = 0) and the linear extrapolation
    (e = e_prev)."""
    import torch
    out = torch.empty_like(samples[0][1])
    errs_nn, errs_ex = [], []
    for inp, tgt in samples:
        net.forward(inp, out)
        nt = torch.linalg.norm(tgt).item()
        errs_nn.append(torch.linalg.norm(out - tgt).item() / nt)
        errs_ex.append(torch.linalg.norm(inp[1] - tgt).item() / nt)
    return {"nn": float(np.mean(errs_nn)), "extrapolation": float(np.mean(errs_ex)),
            "previous": 1.0}


def train(net: Network, samples, epochs: int = 3, lr: float = 3e-4, batch: int = 4,
          seed: int = 0, log=None, log_every: int = 0) -> List[float]:
    """Adam on the mean squared error, mini-batches of `batch` samples
    (gradients accumulate across the batch).  With `log`, reports each epoch
    and, when log_every > 0, every log_every mini-batches."""
    import torch
    rng = np.random.default_rng(seed)
    out = torch.empty_like(samples[0][1])
    dout = torch.empty_like(out)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot = 0.0
        t0 = time.perf_counter()
        for b0 in range(0, len(order), batch):
            net.zero_grad()
            idx = order[b0:b0 + batch]
            for i in idx:
                inp, tgt = samples[i]
                net.forward(inp, out)
                tot += net.mse(out, tgt, dout, scale=1.0 / len(idx))
                net.backward(dout)
            net.adam(lr)
            nb = b0 // batch + 1
            if log and log_every > 0 and nb % log_every == 0:
                log(f"epoch {ep}: {b0 + len(idx)}/{len(order)} samples, "
                    f"mse {tot / (b0 + len(idx)):.4e} ({time.perf_counter() - t0:.1f} s)")
        hist.append(tot / len(order))
        if log:
            log(f"epoch {ep}: mse {hist[-1]:.4e} ({time.perf_counter() - t0:.1f} s)")
    return hist
