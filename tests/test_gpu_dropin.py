"""The Fortran drop-in on the GPU: the reference's own initialization,
functions, IOfunctions and utils modules (built from its sources, with
schemes() binding its procedure pointers to this package's derivation and
poisson modules) and the reference time loop, with predict_velocity,
correct_pression and correct_velocity served by libo3d.so.  Compared with
the reference's own runs (golden fixtures)."""
from __future__ import annotations

import os
import re
import resource
import subprocess
import tempfile

import numpy as np
import pytest

from oracle.oracle import read_ref_run
from osinco3d_amd.config import Config
from osinco3d_amd.fields_io import write_fields

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "o3d_dropin_run")


def _run(cfg, dump_every, fields=None, stats=False, timeout=120):
    with tempfile.TemporaryDirectory() as d:
        cfg.write(os.path.join(d, "parameters.o3d"))
        if fields is not None:
            write_fields(os.path.join(d, "fields.bin"), cfg, *fields)
        if stats:
            os.makedirs(os.path.join(d, "outputs"))

        def unlimited():
            resource.setrlimit(resource.RLIMIT_STACK,
                               (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
        r = subprocess.run([EXE, "out.bin", str(dump_every)], cwd=d, capture_output=True,
                           text=True, preexec_fn=unlimited, timeout=timeout)
        assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
        recs = read_ref_run(os.path.join(d, "out.bin"))
        rows = None
        if stats and os.path.exists(os.path.join(d, "outputs", "stats.dat")):
            rows = np.loadtxt(os.path.join(d, "outputs", "stats.dat"), comments="#", ndmin=2)
    iters = [int(m.group(1)) for m in re.finditer(r"Exit iteration\s*:\s*(\d+)", r.stdout)]
    if stats:
        return recs, iters, r.stdout, rows
    return recs, iters, r.stdout


@pytest.mark.parametrize("name,dump", [("run_tgv16_ab2.npz", 10), ("run_tgv16fs_ab3_les.npz", 4),
                                       ("run_mixing24_ab2_les.npz", 5)])
def test_dropin_matches_reference_run(name, dump):
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/o3d_dropin_run not built (needs the reference tree)")
    g = np.load(os.path.join(ROOT, "tests", "golden", name), allow_pickle=False)
    cfg = Config.from_namelist(bytes(g["params"]).decode())
    fields = None
    if cfg.typesim == 0:
        fields = (g["ux0"], g["uy0"], g["uz0"], g["pp0"])
    print(f"[gpu] drop-in {name}", flush=True)
    recs, iters, out = _run(cfg, dump, fields)
    assert "Predict velocity" in out and "Exit iteration" in out
    stagnated = "No converged" in out
    tol_u, tol_p = (2e-3, 2e-2) if stagnated else (2e-6, 2e-5)
    assert len(iters) == cfg.itstop
    for itime, t, ux, uy, uz, pp in recs:
        if itime == 0:
            continue
        for a, k in ((ux, "ux"), (uy, "uy"), (uz, "uz")):
            assert np.abs(a - g[f"{k}{itime}"]).max() <= tol_u, (itime, k)
        pr = g[f"pp{itime}"]
        assert np.abs((pp - pp.mean()) - (pr - pr.mean())).max() <= tol_p


def test_dropin_multigrid_switch():
    """multigrid = 1 in parameters.o3d: correct_pression runs the multigrid
    solver (o3d_poisson_multigrid) to max |r| < eps; the run follows the
    reference's SOR run of the same case within the eps-tied tolerance."""
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/o3d_dropin_run not built (needs the reference tree)")
    g = np.load(os.path.join(ROOT, "tests", "golden", "run_tgv16_ab2.npz"), allow_pickle=False)
    cfg = Config.from_namelist(bytes(g["params"]).decode())
    cfg.multigrid = 1
    print("[gpu] drop-in run_tgv16_ab2.npz with multigrid = 1", flush=True)
    recs, iters, out = _run(cfg, 10)
    cyc = [int(m.group(1)) for m in re.finditer(r"Multigrid cycles\s*:\s*(\d+)", out)]
    assert len(cyc) == cfg.itstop and not iters
    print(f"[gpu] multigrid cycles per step: {cyc}", flush=True)
    for itime, t, ux, uy, uz, pp in recs:
        if itime == 0:
            continue
        for a, k in ((ux, "ux"), (uy, "uy"), (uz, "uz")):
            assert np.abs(a - g[f"{k}{itime}"]).max() <= 2e-6, (itime, k)
        pr = g[f"pp{itime}"]
        assert np.abs((pp - pp.mean()) - (pr - pr.mean())).max() <= 2e-5


def test_dropin_multigrid_incompatible_rhs_converges():
    """multigrid = 1 on the mixing-layer run (free-slip y, whose rhs the
    mirror cannot satisfy, so the reference's SOR ends every solve on its
    stagnation rule): the multigrid solve works on the compatible part of
    the rhs and ends every step on eps (or stagnation) far below kmax, and
    the run follows the reference's SOR run within the stagnation-tied
    tolerance."""
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/o3d_dropin_run not built (needs the reference tree)")
    g = np.load(os.path.join(ROOT, "tests", "golden", "run_mixing24_ab2_les.npz"),
                allow_pickle=False)
    cfg = Config.from_namelist(bytes(g["params"]).decode())
    cfg.multigrid = 1
    assert cfg.kmax >= 1000
    print("[gpu] drop-in run_mixing24_ab2_les.npz with multigrid = 1", flush=True)
    recs, iters, out = _run(cfg, 5, (g["ux0"], g["uy0"], g["uz0"], g["pp0"]))
    m = [(int(a), int(b)) for a, b in
         re.findall(r"Multigrid cycles\s*:\s*(\d+),\s*\S+,\s*(\d+)", out)]
    print(f"[gpu] multigrid (cycles, reason) per step: {m}", flush=True)
    assert len(m) == cfg.itstop
    assert all(why in (0, 1) and cyc < 200 for cyc, why in m)
    for itime, t, ux, uy, uz, pp in recs:
        if itime == 0:
            continue
        for a, k in ((ux, "ux"), (uy, "uy"), (uz, "uz")):
            err = np.abs(a - g[f"{k}{itime}"]).max()
            print(f"[gpu] step {itime} {k}: max |diff| {err:.2e}", flush=True)
            assert err <= 5e-3, (itime, k)


def test_dropin_config1_tgv64_100_steps_device_resident():
    """Config 1 (TGV Re 1600, 64^3, AB2, 100 steps) through the unchanged
    reference modules and the drop-in: the fields after steps 10, 50 and 100
    follow the reference's run (tolerance tied to eps), statistics_calc
    (every 25 steps, through the drop-in derivative pointers) gives the
    reference's kinetic energy, and with the run's fields registered in the
    device registry no field crosses to the device after the first step:
    each step copies back only what host code reads (velocity, pressure and
    the two diagnostic divergences)."""
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/o3d_dropin_run not built (needs the reference tree)")
    g = np.load(os.path.join(ROOT, "tests", "golden", "c1_tgv64.npz"), allow_pickle=False)
    cfg = Config.from_namelist(bytes(g["params"]).decode())
    idx = g["idx"]
    print("[gpu] drop-in config 1 (64^3, 100 steps)", flush=True)
    recs, iters, out, rows = _run(cfg, 10, stats=True, timeout=300)
    assert len(iters) == 100
    by = {itime: (ux, uy, uz, pp) for itime, t, ux, uy, uz, pp in recs}
    for itime in (10, 50, 100):
        ux, uy, uz, pp = by[itime]
        for a, k in ((ux, "ux"), (uy, "uy"), (uz, "uz")):
            err = np.abs(a.reshape(-1)[idx] - g[f"{k}{itime}"]).max()
            assert err <= 1e-6, (itime, k, err)
        pm = pp - pp.mean()
        assert np.abs(pm.reshape(-1)[idx] - g[f"pm{itime}"]).max() <= 1e-5
    # statistics_calc rows at steps 25, 50, 75, 100: t, E_k, ...
    assert rows is not None and rows.shape[0] >= 4
    t100, ek100 = rows[-1, 0], rows[-1, 1]
    assert abs(t100 - 100 * cfg.dt) <= 1e-9
    assert abs(ek100 - float(g["ek100"])) <= 1e-8 * abs(float(g["ek100"]))
    sec = [float(m.group(2)) for m in re.finditer(r"O3D_STEP_SECONDS\s+(\d+)\s+(\S+)", out)]
    tr = [(int(m.group(1)), float(m.group(2)), float(m.group(3))) for m in
          re.finditer(r"O3D_STEP_TRANSFER\s+(\d+)\s+(\S+)\s+(\S+)", out)]
    field = 8.0 * cfg.nx * cfg.ny * cfg.nz
    steady = [(h, d) for it, h, d in tr if it > 1 and it % 25 != 0]
    print(f"[gpu] drop-in 64^3: {len(sec) / sum(sec):.1f} steps/s (host-timed loop), "
          f"steady-state transfer per step h2d {steady[0][0] / field:.1f} fields, d2h "
          f"{steady[0][1] / field:.1f} fields", flush=True)
    assert all(h == 0.0 for h, d in steady)
    assert all(d == 6 * field for h, d in steady)
