"""Host side of the time loop: the reference's driver sequence on the device.

``Solver`` owns one device-resident solver of libo3d.so built from a
``Config`` (the reference's parameters.o3d) and runs whole time steps of
src/osinco3d_main.f90:97-128 on the GPU.  The initial fields are the
reference's (Taylor-Green: src/initial_conditions.f90:157-162) or come from a
fields.bin checkpoint (typesim = 0).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Tuple

import numpy as np

from . import capi
from .config import Config

FP64 = 0
MIXED = 1
WARM_PREVIOUS = 0
WARM_LEARNED = 1
WARM_EXTRAPOLATE = 2

SOR_REASONS = {0: "eps", 1: "stagnation", 2: "kmax"}


def grid(cfg: Config):
    """Node coordinates x(i) = x0 + (i-1) dx (src/initialization.f90:211-219)."""
    dx, dy, dz = cfg.spacing()
    x = cfg.x0 + np.arange(cfg.nx, dtype=np.float64) * dx
    y = cfg.y0 + np.arange(cfg.ny, dtype=np.float64) * dy
    z = cfg.z0 + np.arange(cfg.nz, dtype=np.float64) * dz
    return x, y, z


def taylor_green(cfg: Config) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 and the
    pressure 1/16 (cos 2x + cos 2y)(cos 2z + 2)
    (src/initial_conditions.f90:157-162).  Arrays of shape (nz, ny, nx)."""
    x, y, z = grid(cfg)
    amp = cfg.ratio * cfg.u0 / cfg.l0
    sx, cx = np.sin(x), np.cos(x)
    sy, cy = np.sin(y), np.cos(y)
    cz = np.cos(z)
    ux = ((amp * sx)[None, None, :] * cy[None, :, None]) * cz[:, None, None]
    uy = ((-amp * cx)[None, None, :] * sy[None, :, None]) * cz[:, None, None]
    uz = np.zeros_like(ux)
    pp = 0.0625 * (np.cos(2 * x)[None, None, :] + np.cos(2 * y)[None, :, None]) * \
        (np.cos(2 * z)[:, None, None] + 2.0)
    return (np.ascontiguousarray(ux), np.ascontiguousarray(uy), uz,
            np.ascontiguousarray(pp))


def splitmix64_uniform(n: int, seed: int) -> np.ndarray:
    """n uniform doubles in [0, 1) from the SplitMix64 sequence of `seed`
    (state seed + k * 0x9E3779B97F4A7C15, k = 1..n; top 53 bits)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(1, n + 1, dtype=np.uint64) *
                               np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def mixing_layer(cfg: Config, seed: int = 1234, amplitude: float = 0.05):
    """Temporal mixing layer: u = 0.5 U0 tanh(2 y / l0) plus uniform noise of
    amplitude `amplitude` U0 on all three components (SplitMix64, `seed`),
    p = 0.  Written to fields.bin for the reference (typesim = 0)."""
    x, y, z = grid(cfg)
    sh = (cfg.nz, cfg.ny, cfg.nx)
    n = cfg.nx * cfg.ny * cfg.nz
    noise = splitmix64_uniform(3 * n, seed).reshape(3, *sh)
    a = amplitude * cfg.u0
    ux = 0.5 * cfg.u0 * np.tanh(2.0 * y / cfg.l0)[None, :, None] + a * (2.0 * noise[0] - 1.0)
    uy = a * (2.0 * noise[1] - 1.0)
    uz = a * (2.0 * noise[2] - 1.0)
    return (np.ascontiguousarray(ux), np.ascontiguousarray(uy), np.ascontiguousarray(uz),
            np.zeros(sh))


def jet_profile(cfg: Config, y: np.ndarray) -> np.ndarray:
    """The planar-jet profile of src/initial_conditions.f90:176-242:
    u = (u1 + u2)/2 + (u1 - u2)/2 tanh(t3 (1 - 2|y|/l0)), u1 = u0,
    u2 = ratio u0, theta0 = l0/30, t3 = l0/(4 theta0)."""
    u1 = cfg.u0
    u2 = u1 * cfg.ratio
    theta0 = 1.0 / 30.0 * cfg.l0
    t3 = cfg.l0 / (4.0 * theta0)
    return 0.5 * (u1 + u2) + 0.5 * (u1 - u2) * np.tanh(t3 * (1.0 - 2.0 * np.abs(y) / cfg.l0))


def planar_jet(cfg: Config):
    """Initial field of the x-Dirichlet jet: the inflow profile in the whole
    box (x-independent, as the reference's initialize_planar_jet), v = w = 0,
    p = 0."""
    x, y, z = grid(cfg)
    sh = (cfg.nz, cfg.ny, cfg.nx)
    ux = np.broadcast_to(jet_profile(cfg, y)[None, :, None], sh)
    return (np.ascontiguousarray(ux), np.zeros(sh), np.zeros(sh), np.zeros(sh))


def jet_inflow(cfg: Config, amplitude: float = 0.05, seed: int = 1234):
    """Inflow of the jet: (base3, weight, amp, period, seed, uc) for
    o3d_solver_set_inflow.  base3 = the profile (u) and zeros (v, w); the
    noise weight is the normalised shear 1 - tanh^2 of the profile's
    argument (1 in the shear layers, 0 in the core and the coflow), as the
    reference weights its initial noise by the shear profile
    (src/initial_conditions.f90:515-545); amp = amplitude u0 (5 %); period
    l0 / u0; uc = the trapezoid-weighted mean inflow velocity (the outflow's
    convective velocity)."""
    _, y, _ = grid(cfg)
    P = (cfg.nz, cfg.ny)
    u = jet_profile(cfg, y)
    t3 = cfg.l0 / (4.0 * (cfg.l0 / 30.0))
    w = 1.0 - np.tanh(t3 * (1.0 - 2.0 * np.abs(y) / cfg.l0)) ** 2
    base = np.zeros((3,) + P)
    base[0] = u[None, :]
    weight = np.ascontiguousarray(np.broadcast_to(w[None, :], P))
    wy = np.ones(cfg.ny)
    if cfg.bc()[1] != 0:
        wy[0] = wy[-1] = 0.5
    uc = float((wy * u).sum() / wy.sum())
    return base, weight, amplitude * cfg.u0, cfg.l0 / cfg.u0, seed, uc


def make_params(cfg: Config, precision: int = FP64, warm: int = WARM_PREVIOUS,
                inner_max: int = 0, coarse_cycles: int = 0,
                smooth_sweeps: int = 0, coarse_cycles_learned: int = 0,
                adaptive: int = 0, rhs_as_given: int = 0,
                refine_cycles: int = 0) -> capi.O3DParams:
    p = capi.O3DParams()
    p.nx, p.ny, p.nz = cfg.nx, cfg.ny, cfg.nz
    p.dx, p.dy, p.dz = cfg.spacing()
    p.bc = (C.c_int * 3)(*cfg.bc())
    p.itscheme = cfg.itscheme
    dt = cfg.time_step_overrides()
    p.dt = dt
    p.re = cfg.re
    c2 = Config(**{**cfg.__dict__, "dt": dt})
    adt, bdt, cdt = c2.ab_coefficients()
    p.adt = (C.c_double * 3)(*adt)
    p.bdt = (C.c_double * 3)(*bdt)
    p.cdt = (C.c_double * 3)(*cdt)
    p.iles = cfg.iles
    p.cs = cfg.cs
    p.delta = cfg.delta()
    p.omega = cfg.omega
    p.eps = cfg.eps
    p.kmax = cfg.kmax
    p.idyn = cfg.idyn
    p.precision = precision
    p.warm_start = warm
    p.inner_max = inner_max
    p.coarse_cycles = coarse_cycles
    p.smooth_sweeps = smooth_sweeps
    p.coarse_cycles_learned = coarse_cycles_learned
    p.adaptive_extrapolation = adaptive
    p.rhs_as_given = rhs_as_given
    p.refine_cycles = refine_cycles
    return p


class Solver:
    """Device-resident time stepping of one configuration."""

    def __init__(self, cfg: Config, precision: int = FP64, warm: int = WARM_PREVIOUS,
                 inner_max: int = 0, coarse_cycles: int = 0, smooth_sweeps: int = 0,
                 coarse_cycles_learned: int = 0, adaptive: int = 0, rhs_as_given: int = 0,
                 refine_cycles: int = 0):
        """rhs_as_given: in the mixed form on a grid with a free-slip axis,
        solve the divergence's rhs unprojected, as the FP64 form does (0:
        remove its weighted mean, so every solve can end on eps)."""
        self.cfg = cfg
        self.precision = precision
        self.lib = capi.load()
        self.params = make_params(cfg, precision, warm, inner_max, coarse_cycles,
                                  smooth_sweeps, coarse_cycles_learned, adaptive, rhs_as_given,
                                  refine_cycles)
        self.handle = C.c_void_p(0)
        capi.check(self.lib.o3d_solver_create(C.byref(self.params), C.byref(self.handle)))
        self.dt = self.params.dt

    def close(self):
        if self.handle:
            self.lib.o3d_solver_destroy(self.handle)
            self.handle = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, ux=None, uy=None, uz=None, pp=None):
        capi.check(self.lib.o3d_solver_upload(self.handle, capi.ptr(ux), capi.ptr(uy),
                                              capi.ptr(uz), capi.ptr(pp)))

    def download(self):
        shape = (self.cfg.nz, self.cfg.ny, self.cfg.nx)
        out = [np.empty(shape) for _ in range(4)]
        capi.check(self.lib.o3d_solver_download(self.handle, *[capi.ptr(a) for a in out]))
        return out

    def init_taylor_green(self):
        self.upload(*taylor_green(self.cfg))

    def set_inflow(self, base3, weight, amp, period, seed, uc):
        """x-Dirichlet inflow (o3d_solver_set_inflow)."""
        b = np.ascontiguousarray(base3, dtype=np.float64)
        w = np.ascontiguousarray(weight, dtype=np.float64)
        capi.check(self.lib.o3d_solver_set_inflow(self.handle, capi.ptr(b), capi.ptr(w),
                                                  C.c_double(amp), C.c_double(period),
                                                  C.c_ulonglong(seed), C.c_double(uc)))

    def init_jet(self, amplitude: float = 0.05, seed: int = 1234):
        """The planar jet's initial field and inflow."""
        self.upload(*planar_jet(self.cfg))
        self.set_inflow(*jet_inflow(self.cfg, amplitude, seed))

    def step(self, itime: int) -> dict:
        info = capi.O3DStepInfo()
        rc = capi.check(self.lib.o3d_solver_step(self.handle, int(itime), C.byref(info)))
        d = info.to_dict()
        d["rc"] = rc
        return d

    def statistics(self, t: float) -> np.ndarray:
        out = np.zeros(17)
        capi.check(self.lib.o3d_solver_statistics(self.handle, C.c_double(t),
                                                  out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def field_ptr(self, which: int) -> int:
        return int(self.lib.o3d_solver_field(self.handle, int(which)) or 0)

    def field(self, which: int) -> np.ndarray:
        """Host copy of solver field `which` (o3d_solver_field numbering):
        float64 for the pressure and the FP64 form, float32 otherwise."""
        dt = np.float64 if which == 6 or self.precision == FP64 else np.float32
        out = np.empty((self.cfg.nz, self.cfg.ny, self.cfg.nx), dt)
        capi.check(self.lib.o3d_copy(out.ctypes.data_as(C.c_void_p),
                                     C.c_void_p(self.field_ptr(which)), C.c_size_t(out.nbytes)))
        return out

    def pressure_stats(self) -> np.ndarray:
        """function_stats of the solver's pressure on the device (min, max,
        mean, then the 1-based position of the first maximum)."""
        out = np.zeros(6)
        capi.check(self.lib.o3d_function_stats(C.c_void_p(self.field_ptr(6)), self.cfg.nx,
                                               self.cfg.ny, self.cfg.nz,
                                               out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def sor_period(self, iterations: int = 200) -> float:
        """Mean time (us) of one FP32 red-black iteration over `iterations`
        back-to-back iterations (o3d_solver_sor_period; mixed form)."""
        us = C.c_double(0)
        capi.check(self.lib.o3d_solver_sor_period(self.handle, int(iterations), C.byref(us)))
        return us.value

    def sor_timing(self, launches: bool = False):
        """(ms, sweeps) of the last step's complete SOR batches per precision
        [FP64, FP32]; with launches=True also the kernel launch counts."""
        ms = (C.c_double * 2)()
        sw = (C.c_int * 2)()
        ln = (C.c_int * 2)()
        capi.check(self.lib.o3d_solver_sor_timing(self.handle, ms, sw, ln))
        if launches:
            return list(ms), list(sw), list(ln)
        return list(ms), list(sw)
