#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden/ from the reference's own code.

Run in a container that holds the reference tree (it needs oracle/_ref/,
built by `make -C oracle ref` from the reference sources):

    python tests/golden/gen_golden.py

Every fixture is inputs + the reference's outputs, saved with numpy
(allow_pickle=False):
  derivatives.npz   18 stencil routines of src/derivation.f90 on a seeded
                    random field
  operators.npz     divergence, calculate_nu_t, predict_velocity (Euler,
                    AB2, AB3, LES on/off), correct_velocity, function_stats,
                    for periodic, y-free-slip and all-free-slip boundaries
  sor.npz           the three lexicographic SOR solvers: iterates after a
                    fixed number of sweeps and the converged fields
  run_*.npz         whole reference time loops (o3d_ref_run) on small grids:
                    fields after selected steps and the printed SOR exit
                    iterations
  c1_tgv64.npz      config 1 (TGV Re=1600 64^3, AB2, 100 steps): fields at
                    4096 seeded sample points after steps 10, 50 and 100,
                    plus statistics_calc-style sums of the full fields
"""
from __future__ import annotations

import os
import re
import resource
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import load_reference, ref_run_path  # noqa: E402
from oracle.oracle import read_ref_run  # noqa: E402
from osinco3d_amd.config import Config, tgv_config  # noqa: E402
from osinco3d_amd.fields_io import write_fields  # noqa: E402


def rnd(shape, seed, scale=1.0):
    return np.ascontiguousarray(scale * np.random.default_rng(seed).standard_normal(shape))


def gen_derivatives(ref):
    f = rnd((9, 11, 13), 101)
    out = {"f": f, "h": np.array(0.173)}
    for axis in range(3):
        for order in (1, 2):
            for kind in range(3):
                ref.set_bc([1 if kind else 0] * 3)
                out[f"d_{axis}_{order}_{kind}"] = ref.der(f, axis, order, kind, 0.173)
    np.savez_compressed(os.path.join(HERE, "derivatives.npz"), **out)


def gen_operators(ref):
    sh = (8, 8, 8)
    d = (0.11, 0.13, 0.17)
    out = {"d": np.array(d)}
    u = [rnd(sh, 200 + c) for c in range(3)]
    for c in range(3):
        out[f"u{c}"] = u[c]
    p = rnd(sh, 210)
    out["p"] = p
    dt = 0.01
    a, b, c3 = ref.ab_coefficients(dt)
    out["adt"], out["bdt"], out["cdt"] = np.array(a), np.array(b), np.array(c3)
    for bc in ((0, 0, 0), (0, 1, 0), (1, 1, 1)):
        tag = "".join(map(str, bc))
        ref.set_bc(bc)
        out[f"div1_{tag}"] = ref.divergence(*u, d, 1)
        out[f"div0_{tag}"] = ref.divergence(*u, d, 0)
        out[f"nut_{tag}"] = ref.nu_t(*u, d, 0.17, 0.2)
        for its in (1, 2, 3):
            for iles in (0, 1):
                F = [rnd((3,) + sh, 220 + q) for q in range(3)]
                nut = np.zeros(sh)
                for itime in (1, 2, 3):
                    up = ref.predict(*u, *F, 1600.0, a, b, c3, itime, its, d, iles, 0.17, 0.2,
                                     nut)
                    key = f"pred_{tag}_{its}_{iles}_{itime}"
                    for q in range(3):
                        out[f"{key}_up{q}"] = up[q]
                        if itime == 3:
                            out[f"{key}_F{q}"] = F[q].copy()
                    out[f"{key}_nut"] = nut.copy()
        v = [np.empty(sh) for _ in range(3)]
        ref.correct_velocity(*v, *u, p, 0.02, d)
        for q in range(3):
            out[f"corr_{tag}_u{q}"] = v[q]
    out["fstats"] = ref.function_stats(u[0])
    np.savez_compressed(os.path.join(HERE, "operators.npz"), **out)


def gen_sor(ref):
    out = {}
    d = (0.3, 0.35, 0.4)
    for solver, sh, bc in (("0000", (10, 12, 14), (0, 0, 0)), ("0011", (10, 9, 14), (0, 1, 0)),
                           ("111111", (9, 11, 13), (1, 1, 1))):
        ref.set_bc(bc)
        rhs = rnd(sh, 300)
        rhs -= rhs.mean()
        p0 = rnd(sh, 301)
        out[f"{solver}_rhs"] = rhs
        out[f"{solver}_p0"] = p0
        for kmax in (1, 7):
            p = p0.copy()
            ref.poisson(p, rhs, d, 1.6, 1e-30, kmax, 0)
            out[f"{solver}_k{kmax}"] = p
        p = p0.copy()
        ref.poisson(p, rhs, d, 1.6, 1e-9, 5000, 0)
        out[f"{solver}_conv"] = p
    out["d"] = np.array(d)
    np.savez_compressed(os.path.join(HERE, "sor.npz"), **out)


def run_reference(cfg: Config, dump_every: int, fields=None):
    exe = ref_run_path()
    with tempfile.TemporaryDirectory() as d:
        cfg.write(os.path.join(d, "parameters.o3d"))
        if fields is not None:
            write_fields(os.path.join(d, "fields.bin"), cfg, *fields)

        def unlimited():
            resource.setrlimit(resource.RLIMIT_STACK,
                               (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
        r = subprocess.run([exe, "out.bin", str(dump_every)], cwd=d, capture_output=True,
                           text=True, preexec_fn=unlimited)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        recs = read_ref_run(os.path.join(d, "out.bin"))
    iters = [int(m.group(1)) for m in re.finditer(r"Exit iteration\s*:\s*(\d+)", r.stdout)]
    return recs, iters


def save_run(name, cfg, recs, iters):
    out = {"iters": np.array(iters), "params": np.frombuffer(cfg.to_namelist().encode(), np.uint8)}
    for itime, t, ux, uy, uz, pp in recs:
        out[f"t{itime}"] = np.array(t)
        out[f"ux{itime}"], out[f"uy{itime}"], out[f"uz{itime}"], out[f"pp{itime}"] = ux, uy, uz, pp
    np.savez_compressed(os.path.join(HERE, name), **out)


def gen_runs():
    cfg = tgv_config(16, 20, 2)
    recs, iters = run_reference(cfg, 10)
    save_run("run_tgv16_ab2.npz", cfg, recs, iters)

    cfg = tgv_config(16, 12, 3)
    cfg.nbcx1 = cfg.nbcxn = cfg.nbcy1 = cfg.nbcyn = cfg.nbcz1 = cfg.nbczn = 1
    cfg.xlx = cfg.yly = cfg.zlz = np.pi
    cfg.iles = 1
    cfg.dt = 0.01
    cfg.omega = 1.7
    recs, iters = run_reference(cfg, 4)
    save_run("run_tgv16fs_ab3_les.npz", cfg, recs, iters)

    cfg = Config(nx=24, ny=16, nz=12, xlx=6.0, yly=4.0, zlz=3.0, y0=-2.0, nbcy1=1, nbcyn=1,
                 itscheme=2, dt=0.02, itstop=10, re=3000.0, iles=1, cs=0.17, omega=1.8,
                 eps=1e-8, kmax=3000, typesim=0)
    sh = (cfg.nz, cfg.ny, cfg.nx)
    rng = np.random.default_rng(1234)
    y = cfg.y0 + np.arange(cfg.ny) * cfg.spacing()[1]
    ux = 0.5 * np.tanh(2 * y)[None, :, None] + 0.05 * (2 * rng.random(sh) - 1)
    uy = 0.05 * (2 * rng.random(sh) - 1)
    uz = 0.05 * (2 * rng.random(sh) - 1)
    recs, iters = run_reference(cfg, 5, fields=(ux, uy, uz, np.zeros(sh)))
    save_run("run_mixing24_ab2_les.npz", cfg, recs, iters)


def gen_c1():
    cfg = tgv_config(64, 100, 2)
    recs, iters = run_reference(cfg, 10)
    rng = np.random.default_rng(64)
    idx = rng.choice(64 ** 3, size=4096, replace=False)
    out = {"idx": idx, "iters": np.array(iters),
           "params": np.frombuffer(cfg.to_namelist().encode(), np.uint8)}
    for itime, t, ux, uy, uz, pp in recs:
        if itime not in (0, 10, 50, 100):
            continue
        pm = pp - pp.mean()
        out[f"t{itime}"] = np.array(t)
        for name, f in (("ux", ux), ("uy", uy), ("uz", uz), ("pm", pm)):
            out[f"{name}{itime}"] = f.reshape(-1)[idx]
        out[f"ek{itime}"] = np.array(0.5 * np.mean(ux * ux + uy * uy + uz * uz))
        out[f"umax{itime}"] = np.array([ux.max(), uy.max(), uz.max()])
    np.savez_compressed(os.path.join(HERE, "c1_tgv64.npz"), **out)


def main():
    ref = load_reference()
    if ref is None or ref_run_path() is None:
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` with the reference tree")
    gen_derivatives(ref)
    gen_operators(ref)
    gen_sor(ref)
    gen_runs()
    gen_c1()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
