"""fields.bin checkpoints in the reference's on-disk format.

save_fields (src/IOfunctions.f90:360-404) writes an unformatted stream file:
float64 time, int32 nx, ny, nz, float64 x(nx), y(ny), z(nz), then ux, uy, uz,
pp, phi as float64 column-major (nx, ny, nz) arrays; read_fields (:406-470)
reads it back (a dimension mismatch is reported and the file is not read).
Arrays here have shape (nz, ny, nx), which is the same memory order.
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from .config import Config


def write_fields(path: str, cfg: Config, ux, uy, uz, pp, phi: Optional[np.ndarray] = None,
                 time: float = 0.0) -> None:
    dx, dy, dz = cfg.spacing()
    x = cfg.x0 + np.arange(cfg.nx) * dx
    y = cfg.y0 + np.arange(cfg.ny) * dy
    z = cfg.z0 + np.arange(cfg.nz) * dz
    shape = (cfg.nz, cfg.ny, cfg.nx)
    if phi is None:
        phi = np.zeros(shape)
    with open(path, "wb") as fh:
        fh.write(np.float64(time).tobytes())
        fh.write(np.array([cfg.nx, cfg.ny, cfg.nz], dtype=np.int32).tobytes())
        for a in (x, y, z):
            fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        for a in (ux, uy, uz, pp, phi):
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.shape != shape:
                raise ValueError(f"field shape {a.shape} != {shape}")
            fh.write(a.tobytes())


def read_fields(path: str, cfg: Optional[Config] = None) -> Tuple[float, dict]:
    """Returns (time, dict with x, y, z, ux, uy, uz, pp, phi).  With cfg
    given, a dimension mismatch raises (the reference only prints)."""
    with open(path, "rb") as fh:
        data = fh.read()
    off = 0
    time = float(np.frombuffer(data, np.float64, 1, off)[0])
    off += 8
    nx, ny, nz = (int(v) for v in np.frombuffer(data, np.int32, 3, off))
    off += 12
    if cfg is not None and (nx, ny, nz) != (cfg.nx, cfg.ny, cfg.nz):
        raise ValueError(f"fields.bin has {nx}x{ny}x{nz}, parameters.o3d has "
                         f"{cfg.nx}x{cfg.ny}x{cfg.nz}")
    out = {}
    for name, n in (("x", nx), ("y", ny), ("z", nz)):
        out[name] = np.frombuffer(data, np.float64, n, off).copy()
        off += 8 * n
    n = nx * ny * nz
    for name in ("ux", "uy", "uz", "pp", "phi"):
        out[name] = np.frombuffer(data, np.float64, n, off).reshape(nz, ny, nx).copy()
        off += 8 * n
    return time, out
