"""Checkpoint / restart of the device-resident solver.

A checkpoint is the reference's own fields.bin (src/IOfunctions.f90:360-404,
written by fields_io.write_fields, so the reference driver can restart from
it with typesim = 0) plus a sidecar `<path>.history.npz` holding what the
reference does not save: the Adams-Bashforth right-hand sides of the last
three steps, the history depth, omega, the step number and the time
(o3d_solver_history).  With the previous pressure as warm start (the
reference's behaviour) restoring both continues a run bit for bit.  The
extrapolated and learned warm starts are not saved: they re-prime after a
restart (the first step starts from the previous pressure), so such a run
continues exactly as every restart from the same checkpoint does, and
differs from the uninterrupted run only by where its solves start (all end
on the same stop rule).  A restart from fields.bin alone starts again with
an Euler step, as the reference's restart does.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Tuple

import numpy as np

from . import capi
from .fields_io import read_fields, write_fields


def _history_buffer(solver) -> np.ndarray:
    dtype = np.float64 if solver.precision == 0 else np.float32
    return np.empty((9,) + (solver.cfg.nz, solver.cfg.ny, solver.cfg.nx), dtype=dtype)


def save(solver, path: str, itime: int, time: float) -> None:
    """Write fields.bin at `path` and the history sidecar after step itime."""
    ux, uy, uz, pp = solver.download()
    write_fields(path, solver.cfg, ux, uy, uz, pp, time=time)
    hist = _history_buffer(solver)
    state = np.zeros(2)
    capi.check(capi.load().o3d_solver_history(solver.handle, 0, capi.ptr(hist),
                                              state.ctypes.data_as(C.POINTER(C.c_double))))
    np.savez(path + ".history.npz", hist=hist, depth=np.int64(state[0]), omega=state[1],
             itime=np.int64(itime), time=np.float64(time),
             precision=np.int64(solver.precision))


def load(solver, path: str) -> Tuple[int, float]:
    """Restore a checkpoint into `solver` (same grid and precision form).
    Returns (itime, time) of the step it was written after; without a
    sidecar, (0, time of fields.bin) and an empty history (Euler restart)."""
    time, f = read_fields(path, solver.cfg)
    solver.upload(f["ux"], f["uy"], f["uz"], f["pp"])
    side = path + ".history.npz"
    if not os.path.exists(side):
        return 0, time
    with np.load(side, allow_pickle=False) as z:
        if int(z["precision"]) != solver.precision:
            raise ValueError("checkpoint precision form differs from the solver's")
        hist = np.ascontiguousarray(z["hist"])
        if hist.shape != _history_buffer(solver).shape or hist.dtype != _history_buffer(
                solver).dtype:
            raise ValueError(f"history of shape {hist.shape} / {hist.dtype} does not fit")
        state = np.array([float(z["depth"]), float(z["omega"])])
        capi.check(capi.load().o3d_solver_history(solver.handle, 1, capi.ptr(hist),
                                                  state.ctypes.data_as(C.POINTER(C.c_double))))
        return int(z["itime"]), float(z["time"])
