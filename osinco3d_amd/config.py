"""parameters.o3d handling: the reference's namelist configuration.

The reference reads nine Fortran namelist groups, in this fixed order, from
``parameters.o3d`` (src/initialization.f90:114-151).  Only a FlowParam read
error stops the reference; the other groups keep their previous values when
absent.  This module writes and reads that file so that the reference CPU path
and this package run the same configuration, and derives the grid constants the
way ``variables()`` does (src/initialization.f90:153-224).
"""
from __future__ import annotations

import dataclasses
import math
import re
from typing import Dict, List, Tuple

# Group order and members, as declared at src/initialization.f90:117-125.
NAMELIST_GROUPS: List[Tuple[str, Tuple[str, ...]]] = [
    ("FlowParam", ("u0", "l0", "re", "typesim", "ratio")),
    ("Domain", ("nx", "ny", "nz", "xlx", "yly", "zlz", "x0", "y0", "z0")),
    ("BoundaryConditions", ("nbcx1", "nbcxn", "nbcy1", "nbcyn", "nbcz1", "nbczn", "sim2d")),
    ("AdvanceTime", ("itscheme", "dt", "itstart", "itstop", "irestart")),
    ("PoissonEq", ("multigrid", "omega", "idyn", "eps", "kmax")),
    ("Scalar", ("nscr", "sc")),
    ("VisuParameters", ("nfre", "nsve", "initstat", "xpro", "ypro", "zpro")),
    ("InitInflow", ("ici", "init_noise_x", "init_noise_y", "init_noise_z")),
    ("LES", ("iles", "cs")),
]

PERIODIC = 0   # src/initialization.f90:77
FREE_SLIP = 1  # src/initialization.f90:78
DIRICHLET = 2  # x only; listed by the reference's README, refused by its schemes()


@dataclasses.dataclass
class Config:
    """All namelist inputs of the reference (names as in the namelists)."""
    # FlowParam
    u0: float = 1.0
    l0: float = 1.0
    re: float = 1600.0
    typesim: int = 2
    ratio: float = 1.0
    # Domain
    nx: int = 64
    ny: int = 64
    nz: int = 64
    xlx: float = 2 * math.pi
    yly: float = 2 * math.pi
    zlz: float = 2 * math.pi
    x0: float = 0.0
    y0: float = 0.0
    z0: float = 0.0
    # BoundaryConditions
    nbcx1: int = PERIODIC
    nbcxn: int = PERIODIC
    nbcy1: int = PERIODIC
    nbcyn: int = PERIODIC
    nbcz1: int = PERIODIC
    nbczn: int = PERIODIC
    sim2d: int = 0
    # AdvanceTime
    itscheme: int = 2
    dt: float = 0.01
    itstart: int = 1
    itstop: int = 100
    irestart: int = 0
    # PoissonEq
    multigrid: int = 0
    omega: float = 1.5
    idyn: int = 0
    eps: float = 1e-8
    kmax: int = 5000
    # Scalar
    nscr: int = 0
    sc: float = 1.0
    # VisuParameters
    nfre: int = 1000000
    nsve: int = 1000000
    initstat: float = 0.0
    xpro: float = 0.0
    ypro: float = 0.0
    zpro: float = 0.0
    # InitInflow
    ici: int = 0
    init_noise_x: float = 0.0
    init_noise_y: float = 0.0
    init_noise_z: float = 0.0
    # LES
    iles: int = 0
    cs: float = 0.17

    # ---- derived constants, as variables() computes them -----------------
    def spacing(self) -> Tuple[float, float, float]:
        """dx = xlx/(nx-1) etc.: the grid includes both end nodes
        (src/initialization.f90:182-184)."""
        return (self.xlx / float(self.nx - 1), self.yly / float(self.ny - 1),
                self.zlz / float(self.nz - 1))

    def delta(self) -> float:
        """LES filter width (dx*dy*dz)**(1/3) (src/initialization.f90:193)."""
        dx, dy, dz = self.spacing()
        return (dx * dy * dz) ** (1.0 / 3.0)

    def ab_coefficients(self) -> Tuple[List[float], List[float], List[float]]:
        """adt, bdt, cdt for Euler, AB2, AB3 (src/initialization.f90:194-202)."""
        dt = self.dt
        adt = [dt, 3.0 * dt / 2.0, 23.0 * dt / 12.0]
        bdt = [0.0, -1.0 * dt / 2.0, -16.0 * dt / 12.0]
        cdt = [0.0, 0.0, 5.0 * dt / 12.0]
        return adt, bdt, cdt

    def bc(self) -> Tuple[int, int, int]:
        """Per-axis boundary kind; both faces of an axis must agree, as
        schemes() requires (src/initialization.f90:226-282)."""
        out = []
        for a, b, name in ((self.nbcx1, self.nbcxn, "x"), (self.nbcy1, self.nbcyn, "y"),
                           (self.nbcz1, self.nbczn, "z")):
            ok = (PERIODIC, FREE_SLIP, DIRICHLET) if name == "x" else (PERIODIC, FREE_SLIP)
            if a != b or a not in ok:
                raise ValueError(f"Unrecognized boundary types for {name}-axis: {a}, {b}")
            out.append(a)
        return tuple(out)

    def poisson_kind(self) -> str:
        """Which SOR solver schemes() binds: it looks at x and y only
        (src/initialization.f90:283-301).  '' means no solver is bound."""
        bx, by, bz = self.bc()
        if bx == DIRICHLET:
            # The x-Dirichlet kinds (no reference counterpart): 3 + (y
            # mirrored) + 2 (z mirrored), pressure mirrored in x.
            return 3 + (by != PERIODIC) + 2 * (bz != PERIODIC)
        if bx == PERIODIC and by == PERIODIC:
            return "0000"
        if bx == PERIODIC and by == FREE_SLIP:
            return "0011"
        if bx == FREE_SLIP and by == FREE_SLIP:
            return "111111"
        return ""

    def time_step_overrides(self) -> float:
        """Return dt after the reference's silent viscous override
        (src/initialization.f90:186-209): if cnu > cfl then dt = 0.1 dmin^2 re."""
        dmin = min(self.spacing())
        cfl = self.dt * self.u0 / dmin
        cnu = 1.0 / self.re * (self.dt / (dmin * dmin))
        if cnu > cfl:
            return 0.1 * dmin * dmin * self.re
        return self.dt

    # ---- file I/O ---------------------------------------------------------
    def to_namelist(self) -> str:
        lines = ["! parameters.o3d written by osinco3d_amd.config"]
        for group, names in NAMELIST_GROUPS:
            lines.append(f"&{group}")
            for n in names:
                v = getattr(self, n)
                if isinstance(v, bool):
                    v = int(v)
                if isinstance(v, int):
                    lines.append(f"{n} = {v}")
                else:
                    lines.append(f"{n} = " + repr(float(v)).replace("e", "d"))
            lines.append("/End")
            lines.append("")
        return "\n".join(lines)

    def write(self, path: str) -> None:
        with open(path, "w") as fh:
            fh.write(self.to_namelist())

    @classmethod
    def read(cls, path: str) -> "Config":
        with open(path) as fh:
            return cls.from_namelist(fh.read())

    @classmethod
    def from_namelist(cls, text: str) -> "Config":
        """Parse the subset of namelist syntax the reference's files use:
        ``&Group``, ``key = value`` lines, ``!`` comments, ``/`` terminator.
        Unknown keys inside a known group are an error for the reference's
        reader (the whole group read fails and keeps old values); here they
        are ignored for the same effect on known keys."""
        cfg = cls()
        fields = {f.name: f for f in dataclasses.fields(cls)}
        group = None
        known = dict(NAMELIST_GROUPS)
        for raw in text.splitlines():
            line = raw.split("!", 1)[0].strip()
            if not line:
                continue
            if line.startswith("&"):
                group = line[1:].strip()
                continue
            if line.startswith("/"):
                group = None
                continue
            if group is None or group not in known:
                continue
            for key, val in re.findall(r"(\w+)\s*=\s*([^,\s]+)", line):
                key = key.lower()
                if key not in known[group]:
                    continue
                typ = fields[key].type
                sval = val.lower().replace("d", "e")
                if typ in ("int", int):
                    setattr(cfg, key, int(float(sval)))
                else:
                    setattr(cfg, key, float(sval))
        return cfg


def exact_periodic_length(n: int, period: float = 2 * math.pi) -> float:
    """Domain length that makes a node-inclusive periodic grid exactly
    periodic: the reference wraps index n+1 to 1 (src/derivation.f90:34-56),
    so the period is n*dx and xlx must be period*(n-1)/n."""
    return period * (n - 1) / n


def sor_omega(n: int) -> float:
    """omega = 2/(1+sin(pi/n)), the usual SOR optimum for an n-point axis."""
    return 2.0 / (1.0 + math.sin(math.pi / n))


def tgv_config(n: int, steps: int, itscheme: int, cfl: float = 0.25,
               re: float = 1600.0, eps: float = 1e-8, kmax: int = 5000) -> Config:
    """Taylor-Green vortex in an exactly periodic [0, 2pi)^3 box."""
    L = exact_periodic_length(n)
    cfg = Config(u0=1.0, l0=1.0, re=re, typesim=2, ratio=1.0,
                 nx=n, ny=n, nz=n, xlx=L, yly=L, zlz=L,
                 itscheme=itscheme, itstart=1, itstop=steps,
                 omega=sor_omega(n), idyn=0, eps=eps, kmax=kmax,
                 nfre=steps + 1000000, nsve=steps + 1000000, initstat=0.0,
                 iles=0)
    dx = cfg.spacing()[0]
    cfg.dt = cfl * dx / cfg.u0
    return cfg


def mixing_layer_config(nx: int = 384, ny: int = 256, nz: int = 128, steps: int = 500,
                        cfl: float = 0.25, re: float = 3000.0, cs: float = 0.17,
                        eps: float = 1e-8, kmax: int = 5000,
                        lengths: Tuple[float, float, float] = (60.0, 40.0, 20.0)) -> Config:
    """Temporal mixing layer, 60x40x20 (delta = 1) unless `lengths`, periodic
    x and z, free-slip y (y centred on 0), Smagorinsky LES; the initial field
    comes from fields.bin (typesim = 0)."""
    cfg = Config(u0=1.0, l0=1.0, re=re, typesim=0, ratio=2.0,
                 nx=nx, ny=ny, nz=nz, xlx=lengths[0], yly=lengths[1], zlz=lengths[2],
                 x0=0.0, y0=-0.5 * lengths[1], z0=0.0,
                 nbcy1=FREE_SLIP, nbcyn=FREE_SLIP,
                 itscheme=2, itstart=1, itstop=steps,
                 omega=sor_omega(max(nx, ny, nz)), idyn=0, eps=eps, kmax=kmax,
                 nfre=steps + 1000000, nsve=steps + 1000000, iles=1, cs=cs)
    cfg.dt = cfl * min(cfg.spacing()) / cfg.u0
    return cfg


def jet_config(nx: int = 512, ny: int = 384, nz: int = 256, steps: int = 500,
               cfl: float = 0.25, re: float = 3000.0, cs: float = 0.17, eps: float = 1e-8,
               kmax: int = 5000) -> Config:
    """Planar jet with x-Dirichlet inflow / outflow: jet width l0 = 1 and
    centre velocity u0 = 1 on a 16 x 12 x 8 box (y in [-6, 6], z exactly
    periodic, free-slip y), coflow u0 / 10 (the reference's planar-jet
    profile with ratio = 0.1, typesim = 3), Smagorinsky LES, AB3."""
    zl = exact_periodic_length(nz, 8.0)
    cfg = Config(u0=1.0, l0=1.0, re=re, typesim=3, ratio=0.1,
                 nx=nx, ny=ny, nz=nz, xlx=16.0, yly=12.0, zlz=zl,
                 x0=0.0, y0=-6.0, z0=0.0,
                 nbcx1=DIRICHLET, nbcxn=DIRICHLET, nbcy1=FREE_SLIP, nbcyn=FREE_SLIP,
                 itscheme=3, itstart=1, itstop=steps,
                 omega=sor_omega(max(nx, ny, nz)), idyn=0, eps=eps, kmax=kmax,
                 nfre=steps + 1000000, nsve=steps + 1000000, iles=1, cs=cs)
    cfg.dt = cfl * min(cfg.spacing()) / cfg.u0
    return cfg
