"""ctypes binding of libo3d.so (the C ABI declared in include/o3d.h).

Arrays may be numpy arrays (host memory, staged through device buffers by the
library) or torch tensors on the GPU (used in place).  Fields use the
reference's memory order: a numpy/torch array of shape (nz, ny, nx),
C-contiguous, is the Fortran (nx, ny, nz) column-major array.

The library must be built (``__graft_entry__.build()``); nothing here falls
back to a CPU implementation: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libo3d.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class O3DError(RuntimeError):
    """A runtime or argument error reported by libo3d (status < 0)."""


class O3DParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("bc", C.c_int * 3),
        ("itscheme", C.c_int),
        ("dt", C.c_double), ("re", C.c_double),
        ("adt", C.c_double * 3), ("bdt", C.c_double * 3), ("cdt", C.c_double * 3),
        ("iles", C.c_int),
        ("cs", C.c_double), ("delta", C.c_double),
        ("omega", C.c_double), ("eps", C.c_double),
        ("kmax", C.c_int), ("idyn", C.c_int),
        ("precision", C.c_int),
        ("warm_start", C.c_int),
        ("inner_max", C.c_int),
        ("coarse_cycles", C.c_int),
        ("smooth_sweeps", C.c_int),
        ("coarse_cycles_learned", C.c_int),
        ("adaptive_extrapolation", C.c_int),
        ("rhs_as_given", C.c_int),
        ("refine_cycles", C.c_int),
    ]


class O3DStepInfo(C.Structure):
    _fields_ = [
        ("sor_exit_iter", C.c_int), ("sor_reason", C.c_int),
        ("sweeps_f64", C.c_int), ("sweeps_f32", C.c_int),
        ("refinements", C.c_int), ("warm_used", C.c_int),
        ("diverged", C.c_int), ("has_nan", C.c_int),
        ("sor_dmax", C.c_double), ("omega", C.c_double),
        ("poisson_ms", C.c_double), ("step_ms", C.c_double),
        ("warm_dmax", C.c_double * 2),
        ("divu_pred", C.c_double * 6), ("divu", C.c_double * 6),
        ("vel_min", C.c_double * 3), ("vel_max", C.c_double * 3), ("vel_amax", C.c_double * 3),
        ("cfl", C.c_double * 3),
        ("nu_t", C.c_double * 6),
        ("extrap_beta", C.c_double),
        ("flow_rate", C.c_double * 2),
        ("extrap_beta_ls", C.c_double),
    ]

    def to_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


# Exported symbols of include/o3d.h (checked by the CPU test suite).
EXPORTS = [
    "o3d_init", "o3d_set_stream", "o3d_last_error", "o3d_sync", "o3d_version",
    "o3d_stream_probe", "o3d_register", "o3d_unregister", "o3d_mark_host_dirty",
    "o3d_sync_host", "o3d_transfer_stats",
    "o3d_der", "o3d_der_f32", "o3d_divergence", "o3d_nu_t", "o3d_predict_velocity",
    "o3d_poisson", "o3d_poisson_f32", "o3d_correct_pression", "o3d_correct_velocity",
    "o3d_function_stats",
    "o3d_statistics", "o3d_statistics_noalloc", "o3d_transeq", "o3d_rotational", "o3d_q_criterion", "o3d_solver_create",
    "o3d_solver_destroy", "o3d_solver_upload",
    "o3d_solver_download", "o3d_solver_step", "o3d_solver_statistics", "o3d_solver_field",
    "o3d_solver_sor_timing", "o3d_solver_sor_period", "o3d_solver_set_warm_model",
    "o3d_solver_capture",
    "o3d_warm_features", "o3d_solver_history", "o3d_multigrid", "o3d_poisson_multigrid",
    "o3d_solver_set_inflow", "o3d_copy",
]
NN_EXPORTS = ["o3d_nn_create", "o3d_nn_destroy", "o3d_nn_resize", "o3d_nn_num_params",
              "o3d_nn_get_params", "o3d_nn_set_params", "o3d_nn_get_grads", "o3d_nn_zero_grad",
              "o3d_nn_forward", "o3d_nn_backward", "o3d_nn_mse", "o3d_nn_adam_step",
              "o3d_nn_last_timing"]

_lib = None


def load(path: str = LIB_PATH):
    """Load libo3d.so (once).  Raises when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise O3DError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(path)
    lib.o3d_last_error.restype = C.c_char_p
    lib.o3d_version.restype = C.c_char_p
    lib.o3d_solver_field.restype = C.c_void_p
    lib.o3d_solver_field.argtypes = [_vp, C.c_int]
    for name in ("o3d_solver_destroy", "o3d_solver_upload", "o3d_solver_download",
                 "o3d_solver_step", "o3d_solver_statistics", "o3d_solver_sor_timing",
                 "o3d_solver_set_warm_model", "o3d_solver_capture", "o3d_warm_features",
                 "o3d_solver_history", "o3d_solver_set_inflow"):
        getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def last_error() -> str:
    return load().o3d_last_error().decode()


def check(rc: int) -> int:
    if rc < 0:
        raise O3DError(last_error())
    return rc


def ptr(a) -> C.c_void_p:
    """Raw data pointer of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return C.c_void_p(0)
    if isinstance(a, np.ndarray):
        assert a.flags.c_contiguous
        return C.c_void_p(a.ctypes.data)
    # torch tensor
    assert a.is_contiguous()
    return C.c_void_p(a.data_ptr())


def _dims(a):
    s = tuple(a.shape)[-3:]
    return int(s[2]), int(s[1]), int(s[0])


def _bc(bc: Sequence[int]):
    return (C.c_int * 3)(*[int(b) for b in bc])


def _d(v: float):
    return C.c_double(float(v))


WARM_COARSEN = 4   # O3D_WARM_COARSEN


def set_stream(stream_handle: int) -> None:
    check(load().o3d_set_stream(C.c_void_p(stream_handle)))


# ------------------------------------------------------------ operators
def der(which: int, df, f, h: float) -> None:
    lib = load()
    nx, ny, nz = _dims(f)
    fn = lib.o3d_der_f32 if str(getattr(f, "dtype", "")).endswith("float32") else lib.o3d_der
    check(fn(int(which), ptr(df), ptr(f), nx, ny, nz, _d(h)))


def divergence(divf, fx, fy, fz, d, bc, odd=1, divisor=1.0, want_stats=True):
    nx, ny, nz = _dims(fx)
    st = np.zeros(6)
    check(load().o3d_divergence(ptr(divf), ptr(fx), ptr(fy), ptr(fz), *map(_d, d), nx, ny, nz,
                                _bc(bc), int(odd), _d(divisor),
                                st.ctypes.data_as(_dp) if want_stats else None))
    return st


def nu_t(out, ux, uy, uz, d, cs, delta, bc):
    nx, ny, nz = _dims(ux)
    st = np.zeros(6)
    check(load().o3d_nu_t(ptr(out), ptr(ux), ptr(uy), ptr(uz), *map(_d, d), _d(cs), _d(delta),
                          nx, ny, nz, _bc(bc), st.ctypes.data_as(_dp)))
    return st


def predict_velocity(up, u, F, re, adt, bdt, cdt, itime, itscheme, d, iles, cs, delta, nut, bc):
    """up, u: 3 arrays each; F: 3 history arrays of shape (3, nz, ny, nx)."""
    nx, ny, nz = _dims(u[0])
    arr = lambda v: (C.c_double * 3)(*[float(x) for x in v])
    rc = load().o3d_predict_velocity(ptr(up[0]), ptr(up[1]), ptr(up[2]), ptr(u[0]), ptr(u[1]),
                                     ptr(u[2]), ptr(F[0]), ptr(F[1]), ptr(F[2]), _d(re),
                                     arr(adt), arr(bdt), arr(cdt), int(itime), int(itscheme),
                                     *map(_d, d), nx, ny, nz, int(iles), _d(cs), _d(delta),
                                     ptr(nut), _bc(bc))
    return check(rc)


def poisson(pp, rhs, d, omega, eps, kmax, idyn, solver):
    """Returns (exit_iter, dmax, reason, omega)."""
    nx, ny, nz = _dims(pp)
    om = C.c_double(omega)
    it = C.c_int(0)
    dm = C.c_double(0)
    rs = C.c_int(0)
    code = {"0000": 0, "0011": 1, "111111": 2}.get(solver, solver)
    check(load().o3d_poisson(ptr(pp), ptr(rhs), *map(_d, d), nx, ny, nz, C.byref(om), _d(eps),
                             int(kmax), int(idyn), int(code), C.byref(it), C.byref(dm),
                             C.byref(rs)))
    return it.value, dm.value, rs.value, om.value


def poisson_multigrid(pp, rhs, d, solver, npre=5, npost=4, tol=1e-8, max_cycles=100,
                      nlevels=100, with_reason=False):
    """Multigrid solve of the SOR solvers' Poisson problem (pp in place,
    from its current value) until max |rhs - L pp| < tol, stagnation or
    max_cycles cycles; the interface of solve_poisson_multigrid
    (src/poisson_multigrid.f90:10-18).  Returns (cycles, max |r|), or
    (cycles, max |r|, reason) with with_reason (0 tol, 1 stagnation,
    2 max_cycles)."""
    nx, ny, nz = _dims(pp)
    cy = C.c_int(0)
    rm = C.c_double(0)
    rs = C.c_int(-1)
    code = {"0000": 0, "0011": 1, "111111": 2}.get(solver, solver)
    check(load().o3d_poisson_multigrid(ptr(pp), ptr(rhs), *map(_d, d), nx, ny, nz, int(code),
                                       int(nlevels), int(npre), int(npost), _d(tol),
                                       int(max_cycles), C.byref(cy), C.byref(rm), C.byref(rs)))
    if with_reason:
        return cy.value, rm.value, rs.value
    return cy.value, rm.value


def multigrid(u, f, h, per, cycles):
    """V(2,2) multigrid cycles on the cell-centred 7-point Poisson problem
    L u = f from u = 0 (u written in place); h = (hx, hy, hz), per = three
    flags (periodic axis, else homogeneous Neumann)."""
    nx, ny, nz = _dims(u)
    check(load().o3d_multigrid(ptr(u), ptr(f), nx, ny, nz, *map(_d, h),
                               (C.c_int * 3)(*[int(bool(q)) for q in per]), int(cycles)))


def poisson_f32(pp, rhs, d, omega, eps, kmax, idyn, solver, variant=0):
    """FP32 red-black SOR; returns (exit_iter, dmax, reason, omega).
    variant: 0 fused iteration per launch, 1 two iterations per launch,
    2 colour-split half sweeps (identical iterates)."""
    nx, ny, nz = _dims(pp)
    om = C.c_double(omega)
    it = C.c_int(0)
    dm = C.c_double(0)
    rs = C.c_int(0)
    code = {"0000": 0, "0011": 1, "111111": 2}.get(solver, solver)
    check(load().o3d_poisson_f32(ptr(pp), ptr(rhs), *map(_d, d), nx, ny, nz, C.byref(om),
                                 _d(eps), int(kmax), int(idyn), int(code), int(variant),
                                 C.byref(it), C.byref(dm), C.byref(rs)))
    return it.value, dm.value, rs.value, om.value


def correct_pression(pp, up, d, dt, omega, eps, kmax, idyn, bc):
    nx, ny, nz = _dims(pp)
    om = C.c_double(omega)
    it = C.c_int(0)
    dm = C.c_double(0)
    rs = C.c_int(0)
    check(load().o3d_correct_pression(ptr(pp), ptr(up[0]), ptr(up[1]), ptr(up[2]), *map(_d, d),
                                      nx, ny, nz, _d(dt), C.byref(om), _d(eps), int(kmax),
                                      int(idyn), _bc(bc), C.byref(it), C.byref(dm),
                                      C.byref(rs)))
    return it.value, dm.value, rs.value, om.value


def correct_velocity(u, up, pp, dt, d, bc):
    """Returns (diverged, vstats[3][4])."""
    nx, ny, nz = _dims(pp)
    vs = np.zeros(12)
    rc = check(load().o3d_correct_velocity(ptr(u[0]), ptr(u[1]), ptr(u[2]), ptr(up[0]),
                                           ptr(up[1]), ptr(up[2]), ptr(pp), _d(dt), *map(_d, d),
                                           nx, ny, nz, _bc(bc), vs.ctypes.data_as(_dp)))
    return rc, vs.reshape(3, 4)


def function_stats(f):
    nx, ny, nz = _dims(f)
    st = np.zeros(6)
    check(load().o3d_function_stats(ptr(f), nx, ny, nz, st.ctypes.data_as(_dp)))
    return st


def transeq(phi, u, src, fphi, re, sc, adt, bdt, cdt, itime, itscheme, d, iles, nu_t, bc):
    """Scalar transport (src/integration.f90:332-468) in place on phi and
    fphi (shape (3, nz, ny, nx), slot 0 = newest)."""
    nx, ny, nz = _dims(phi)
    arr = lambda v: (C.c_double * 3)(*[float(x) for x in v])
    return check(load().o3d_transeq(ptr(phi), ptr(u[0]), ptr(u[1]), ptr(u[2]), ptr(src),
                                    ptr(fphi), _d(re), _d(sc), arr(adt), arr(bdt), arr(cdt),
                                    int(itime), int(itscheme), *map(_d, d), nx, ny, nz,
                                    int(iles), ptr(nu_t), _bc(bc)))


def rotational(ux, uy, uz, d, bc):
    """curl u (src/differential_operators.f90:40-76): numpy arrays in,
    (rotx, roty, rotz) out."""
    nx, ny, nz = _dims(ux)
    out = [np.empty_like(ux) for _ in range(3)]
    check(load().o3d_rotational(*[ptr(o) for o in out], ptr(ux), ptr(uy), ptr(uz), *map(_d, d),
                                nx, ny, nz, _bc(bc)))
    return out


def q_criterion(ux, uy, uz, d, bc):
    """Q criterion (src/differential_operators.f90:78-108)."""
    nx, ny, nz = _dims(ux)
    q = np.empty_like(ux)
    check(load().o3d_q_criterion(ptr(q), ptr(ux), ptr(uy), ptr(uz), *map(_d, d), nx, ny, nz,
                                 _bc(bc)))
    return q


def stream_probe(bytes_per_buffer=1 << 30, reps=20):
    """Hand-written streaming triad (2 reads + 1 write per element) over three
    device buffers: (GB/s, ms per launch)."""
    g = C.c_double(0)
    m = C.c_double(0)
    check(load().o3d_stream_probe(C.c_size_t(int(bytes_per_buffer)), int(reps), C.byref(g),
                                  C.byref(m)))
    return g.value, m.value


def statistics(ux, uy, uz, d, re, t, bc, noalloc=False):
    """statistics_calc (src/utils.f90:243-375): the 17 values of a stats.dat
    row.  noalloc=True calls o3d_statistics_noalloc (device tensors only;
    partial sums in the library's fixed device array)."""
    nx, ny, nz = _dims(ux)
    out = np.zeros(17)
    fn = load().o3d_statistics_noalloc if noalloc else load().o3d_statistics
    check(fn(ptr(ux), ptr(uy), ptr(uz), nx, ny, nz, *map(_d, d), _d(re), _d(t), _bc(bc),
             out.ctypes.data_as(_dp)))
    return out
