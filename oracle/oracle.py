"""ctypes bindings for the oracle libraries (test infrastructure only).

Fields are numpy float64 arrays of shape (nz, ny, nx), C-contiguous, so that
memory order is the reference's column-major (nx, ny, nz) with x fastest.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libo3dref.so")
REF_RUN = os.path.join(HERE, "_ref", "o3d_ref_run")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _dims(a: np.ndarray) -> Tuple[int, int, int]:
    nz, ny, nx = a.shape[-3:]
    return nx, ny, nz


def _bc(bc: Sequence[int]):
    return (C.c_int * 3)(*[int(b) for b in bc])


SOLVER_CODE = {"0000": 0, "0011": 1, "111111": 2}


class Oracle:
    """The C restatement (o3d_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle oracle`")
        self.lib = C.CDLL(path)

    def der1(self, f, axis, kind, h):
        df = np.empty_like(f)
        self.lib.orc_der1(_p(df), _p(f), *_dims(f), int(axis), int(kind), C.c_double(h))
        return df

    def der2(self, f, axis, kind, h):
        df = np.empty_like(f)
        self.lib.orc_der2(_p(df), _p(f), *_dims(f), int(axis), int(kind), C.c_double(h))
        return df

    def divergence(self, fx, fy, fz, d, bc, odd=1):
        out = np.empty_like(fx)
        self.lib.orc_divergence(_p(out), _p(fx), _p(fy), _p(fz), *[C.c_double(v) for v in d],
                                *_dims(fx), _bc(bc), int(odd))
        return out

    def transeq(self, phi, u, src, fphi, re, sc, adt, bdt, cdt, itime, itscheme, d, iles,
                nu_t, bc):
        """In place on phi and fphi (3, nz, ny, nx)."""
        arr = lambda v: (C.c_double * 3)(*[float(x) for x in v])
        return self.lib.orc_transeq(_p(phi), _p(u[0]), _p(u[1]), _p(u[2]), _p(src), _p(fphi),
                                    C.c_double(re), C.c_double(sc), arr(adt), arr(bdt),
                                    arr(cdt), int(itime), int(itscheme),
                                    *[C.c_double(v) for v in d], *_dims(phi), int(iles),
                                    _p(nu_t), _bc(bc))

    def rotational(self, ux, uy, uz, d, bc):
        out = [np.empty_like(ux) for _ in range(3)]
        self.lib.orc_rotational(*[_p(o) for o in out], _p(ux), _p(uy), _p(uz),
                                *[C.c_double(v) for v in d], *_dims(ux), _bc(bc))
        return out

    def q_criterion(self, ux, uy, uz, d, bc):
        out = np.empty_like(ux)
        self.lib.orc_q_criterion(_p(out), _p(ux), _p(uy), _p(uz), *[C.c_double(v) for v in d],
                                 *_dims(ux), _bc(bc))
        return out

    def nu_t(self, ux, uy, uz, d, cs, delta, bc):
        out = np.empty_like(ux)
        self.lib.orc_nu_t(_p(out), _p(ux), _p(uy), _p(uz), *[C.c_double(v) for v in d],
                          C.c_double(cs), C.c_double(delta), *_dims(ux), _bc(bc))
        return out

    def predict(self, ux, uy, uz, fux, fuy, fuz, re, adt, bdt, cdt, itime, itscheme, d,
                iles, cs, delta, nu_t, bc):
        """fu?: shape (3, nz, ny, nx), slot 0 = newest; updated in place."""
        up = [np.empty_like(ux) for _ in range(3)]
        arr = lambda v: (C.c_double * 3)(*v)
        rc = self.lib.orc_predict(_p(up[0]), _p(up[1]), _p(up[2]), _p(ux), _p(uy), _p(uz),
                                  _p(fux), _p(fuy), _p(fuz), C.c_double(re), arr(adt),
                                  arr(bdt), arr(cdt), int(itime), int(itscheme),
                                  *[C.c_double(v) for v in d], *_dims(ux), int(iles),
                                  C.c_double(cs), C.c_double(delta), _p(nu_t), _bc(bc))
        if rc != 0:
            raise ValueError(f"itscheme: {itscheme} unrecognized")
        return up

    def sor(self, pp, rhs, d, solver, omega, eps, kmax, idyn=0, red_black=False,
            log_len=0):
        """In-place SOR on pp.  Returns (exit_iter, dmax, reason, omega, log)."""
        om = C.c_double(omega)
        dm = C.c_double(0.0)
        reason = C.c_int(0)
        code = SOLVER_CODE[solver] if isinstance(solver, str) else int(solver)
        if red_black:
            log = np.zeros(max(log_len, 1))
            it = self.lib.orc_sor_rb(_p(pp), _p(rhs), *[C.c_double(v) for v in d], *_dims(pp),
                                     code, C.byref(om), C.c_double(eps), int(kmax), int(idyn),
                                     C.byref(dm), C.byref(reason), _p(log), int(log_len))
            return it, dm.value, reason.value, om.value, log[:log_len]
        it = self.lib.orc_sor_lex(_p(pp), _p(rhs), *[C.c_double(v) for v in d], *_dims(pp), code,
                                  C.byref(om), C.c_double(eps), int(kmax), int(idyn),
                                  C.byref(dm), C.byref(reason))
        return it, dm.value, reason.value, om.value, None

    def residual(self, pp, rhs, d, solver):
        r = np.empty_like(pp)
        code = SOLVER_CODE[solver] if isinstance(solver, str) else int(solver)
        self.lib.orc_residual(_p(r), _p(pp), _p(rhs), *[C.c_double(v) for v in d], *_dims(pp),
                              code)
        return r

    def correct_velocity(self, ux, uy, uz, uxp, uyp, uzp, pp, dt, d, bc):
        rc = self.lib.orc_correct_velocity(_p(ux), _p(uy), _p(uz), _p(uxp), _p(uyp), _p(uzp),
                                           _p(pp), C.c_double(dt),
                                           *[C.c_double(v) for v in d], *_dims(ux), _bc(bc))
        return rc

    def function_stats(self, f):
        st = np.zeros(6)
        self.lib.orc_function_stats(_p(f), *_dims(f), _p(st))
        return st

    def statistics(self, ux, uy, uz, d, re, t, bc):
        out = np.zeros(17)
        self.lib.orc_statistics(_p(ux), _p(uy), _p(uz), *_dims(ux), *[C.c_double(v) for v in d],
                                C.c_double(re), C.c_double(t), _bc(bc), _p(out))
        return out

    # ---- x-Dirichlet inflow / outflow (no reference counterpart)
    def splitmix64(self, seed, counter):
        self.lib.orc_splitmix64.restype = C.c_ulonglong
        return self.lib.orc_splitmix64(C.c_ulonglong(seed), C.c_longlong(counter))

    def noise_uniform(self, seed, counter):
        self.lib.orc_noise_uniform.restype = C.c_double
        return self.lib.orc_noise_uniform(C.c_ulonglong(seed), C.c_longlong(counter))

    def jet_edges(self, up, u, base, w, amp, cph, sph, seed, cu, bc):
        """In place on up = [ux*, uy*, uz*]; base (3, nz, ny), w (nz, ny).
        Returns (Q_in, Q_out before the shift, W)."""
        flux = np.zeros(3)
        self.lib.orc_jet_edges(_p(up[0]), _p(up[1]), _p(up[2]), _p(u[0]), _p(u[1]), _p(u[2]),
                               _p(np.ascontiguousarray(base)), _p(np.ascontiguousarray(w)),
                               C.c_double(amp), C.c_double(cph), C.c_double(sph),
                               C.c_ulonglong(seed), C.c_double(cu), *_dims(u[0]), _bc(bc),
                               _p(flux))
        return tuple(flux)

    def weighted_mean(self, f, mirror):
        self.lib.orc_weighted_mean.restype = C.c_double
        return self.lib.orc_weighted_mean(_p(f), *_dims(f), _bc(mirror))


class Reference:
    """The reference's own Fortran routines (oracle/_ref/libo3dref.so)."""

    # which = axis*6 + (order-1)*3 + variant (variant 0 _00, 1 p_11, 2 i_11)
    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref`")
        self.lib = C.CDLL(path)

    def set_bc(self, bc, sim2d=0):
        self.lib.ref_set_bc(int(bc[0]), int(bc[1]), int(bc[2]), int(sim2d))

    def der(self, f, axis, order, kind, h):
        df = np.empty_like(f)
        which = axis * 6 + (order - 1) * 3 + kind
        self.lib.ref_der(int(which), *_dims(f), _p(f), _p(df), C.c_double(h))
        return df

    def divergence(self, fx, fy, fz, d, odd=1):
        out = np.empty_like(fx)
        self.lib.ref_divergence(_p(out), _p(fx), _p(fy), _p(fz), *[C.c_double(v) for v in d],
                                *_dims(fx), int(odd))
        return out

    def transeq(self, phi, u, src, fphi, re, sc, adt, bdt, cdt, itime, itscheme, d, iles,
                nu_t):
        """The reference's transeq, in place on phi and fphi."""
        arr = lambda v: (C.c_double * 3)(*[float(x) for x in v])
        self.lib.ref_transeq(_p(phi), _p(u[0]), _p(u[1]), _p(u[2]), _p(src), _p(fphi),
                             C.c_double(re), C.c_double(sc), arr(adt), arr(bdt), arr(cdt),
                             int(itime), int(itscheme), *[C.c_double(v) for v in d],
                             *_dims(phi), int(iles), _p(nu_t))

    def rotational(self, ux, uy, uz, d):
        out = [np.empty_like(ux) for _ in range(3)]
        self.lib.ref_rotational(*[_p(o) for o in out], _p(ux), _p(uy), _p(uz),
                                *[C.c_double(v) for v in d], *_dims(ux))
        return out

    def q_criterion(self, ux, uy, uz, d):
        out = np.empty_like(ux)
        self.lib.ref_q_criterion(_p(out), _p(ux), _p(uy), _p(uz), *[C.c_double(v) for v in d],
                                 *_dims(ux))
        return out

    def nu_t(self, ux, uy, uz, d, cs, delta):
        out = np.empty_like(ux)
        self.lib.ref_nu_t(_p(out), _p(ux), _p(uy), _p(uz), *[C.c_double(v) for v in d],
                          C.c_double(cs), C.c_double(delta), *_dims(ux))
        return out

    def ab_coefficients(self, dt):
        a, b, c = (C.c_double * 3)(), (C.c_double * 3)(), (C.c_double * 3)()
        self.lib.ref_ab_coefficients(C.c_double(dt), a, b, c)
        return list(a), list(b), list(c)

    def predict(self, ux, uy, uz, fux, fuy, fuz, re, adt, bdt, cdt, itime, itscheme, d,
                iles, cs, delta, nu_t):
        """fu? of shape (3, nz, ny, nx) in memory = Fortran (nx,ny,nz,3)."""
        up = [np.empty_like(ux) for _ in range(3)]
        arr = lambda v: (C.c_double * 3)(*v)
        self.lib.ref_predict(_p(up[0]), _p(up[1]), _p(up[2]), _p(ux), _p(uy), _p(uz),
                             _p(fux), _p(fuy), _p(fuz), C.c_double(re), arr(adt), arr(bdt),
                             arr(cdt), int(itime), int(itscheme), *[C.c_double(v) for v in d],
                             *_dims(ux), int(iles), C.c_double(cs), C.c_double(delta), _p(nu_t))
        return up

    def poisson(self, pp, rhs, d, omega, eps, kmax, idyn=0):
        om = C.c_double(omega)
        self.lib.ref_poisson(_p(pp), _p(rhs), *[C.c_double(v) for v in d], *_dims(pp),
                             C.byref(om), C.c_double(eps), int(kmax), int(idyn))
        return om.value

    def correct_pression(self, pp, uxp, uyp, uzp, d, dt, omega, eps, kmax, idyn=0):
        om = C.c_double(omega)
        self.lib.ref_correct_pression(_p(pp), _p(uxp), _p(uyp), _p(uzp),
                                      *[C.c_double(v) for v in d], *_dims(pp), C.c_double(dt),
                                      C.byref(om), C.c_double(eps), int(kmax), int(idyn))
        return om.value

    def correct_velocity(self, ux, uy, uz, uxp, uyp, uzp, pp, dt, d):
        self.lib.ref_correct_velocity(_p(ux), _p(uy), _p(uz), _p(uxp), _p(uyp), _p(uzp), _p(pp),
                                      C.c_double(dt), *[C.c_double(v) for v in d], *_dims(ux))

    def function_stats(self, f):
        st = np.zeros(6)
        self.lib.ref_function_stats(_p(f), *_dims(f), _p(st))
        return st


def load_oracle() -> Oracle:
    return Oracle()


def load_reference() -> Optional[Reference]:
    return Reference() if os.path.exists(REF_SO) else None


def ref_run_path() -> Optional[str]:
    return REF_RUN if os.path.exists(REF_RUN) else None


def read_ref_run(path: str):
    """Parse o3d_ref_run output: list of (itime, time, ux, uy, uz, pp)."""
    out = []
    with open(path, "rb") as fh:
        data = fh.read()
    off = 0
    while off < len(data):
        itime, nx, ny, nz = np.frombuffer(data, np.int32, 4, off)
        off += 16
        t = float(np.frombuffer(data, np.float64, 1, off)[0])
        off += 8
        n = int(nx) * int(ny) * int(nz)
        arr = np.frombuffer(data, np.float64, 4 * n, off).reshape(4, int(nz), int(ny), int(nx))
        off += 32 * n
        out.append((int(itime), t, arr[0].copy(), arr[1].copy(), arr[2].copy(), arr[3].copy()))
    return out
