"""Whole time steps on the GPU against the oracle.

The FP64 solver uses red-black SOR, so it is compared bit for bit with the
oracle's composition of the same steps in red-black order, and against the
reference's lexicographic order within a tolerance tied to eps (the two
orderings converge to the same field up to the constant null space).
"""
from __future__ import annotations

import numpy as np
import pytest

from osinco3d_amd.config import Config, tgv_config
from osinco3d_amd.solver import FP64, MIXED, Solver, taylor_green
from tests.helpers import OracleStepper, rand_fields

pytestmark = pytest.mark.gpu


def _cfg_free_slip_les():
    cfg = tgv_config(16, 6, 3)
    cfg.nbcx1 = cfg.nbcxn = cfg.nbcy1 = cfg.nbcyn = cfg.nbcz1 = cfg.nbczn = 1
    cfg.xlx = cfg.yly = cfg.zlz = np.pi
    cfg.iles = 1
    cfg.dt = 0.01
    cfg.omega = 1.7
    return cfg


def _cfg_mixing():
    cfg = Config(nx=24, ny=16, nz=12, xlx=6.0, yly=4.0, zlz=3.0, y0=-2.0, nbcy1=1, nbcyn=1,
                 itscheme=2, dt=0.02, re=3000.0, iles=1, cs=0.17, omega=1.8, eps=1e-8,
                 kmax=3000, typesim=0)
    return cfg


def _init(cfg, seed=0):
    if cfg.typesim == 2:
        return list(taylor_green(cfg))
    sh = (cfg.nz, cfg.ny, cfg.nx)
    u = rand_fields(sh, 3, seed=seed, scale=0.05)
    y = cfg.y0 + np.arange(cfg.ny) * cfg.spacing()[1]
    u[0] += 0.5 * np.tanh(2 * y)[None, :, None]
    return [*u, np.zeros(sh)]


@pytest.mark.parametrize("make", [lambda: tgv_config(16, 6, 2), lambda: tgv_config(16, 6, 3),
                                  _cfg_free_slip_les, _cfg_mixing])
def test_fp64_steps_bit_exact(orc, make):
    cfg = make()
    print(f"[gpu] solver bit-exact {cfg.nx}x{cfg.ny}x{cfg.nz} bc={cfg.bc()} "
          f"scheme={cfg.itscheme} les={cfg.iles}", flush=True)
    ux, uy, uz, pp = _init(cfg)
    s = Solver(cfg, FP64)
    s.upload(ux, uy, uz, pp)
    os_ = OracleStepper(orc, cfg, red_black=True)
    u = [ux.copy(), uy.copy(), uz.copy()]
    p = pp.copy()
    for itime in range(1, 6):
        info = s.step(itime)
        u, p, up, oi = os_.step(itime, *u, p)
        g = s.download()
        for c in range(3):
            assert np.array_equal(g[c], u[c]), (itime, c, np.abs(g[c] - u[c]).max())
        assert np.array_equal(g[3], p)
        assert info["sor_exit_iter"] == oi["exit_iter"]
        assert info["sor_reason"] == oi["reason"]
        assert info["diverged"] == 0
        div = orc.divergence(*u, cfg.spacing(), list(cfg.bc()), 1)
        st = orc.function_stats(div)
        assert info["divu"][0] == st[0] and info["divu"][1] == st[1]
        assert list(info["divu"][3:]) == list(st[3:])
        for c in range(3):
            assert info["vel_max"][c] == u[c].max() and info["vel_min"][c] == u[c].min()
    s.close()


def test_fp64_close_to_reference_ordering(orc):
    """Reference (lexicographic) vs GPU (red-black) after 5 TGV steps."""
    cfg = tgv_config(16, 5, 2, eps=1e-12, kmax=20000)
    ux, uy, uz, pp = taylor_green(cfg)
    s = Solver(cfg, FP64)
    s.upload(ux, uy, uz, pp)
    lex = OracleStepper(orc, cfg, red_black=False)
    u = [ux.copy(), uy.copy(), uz.copy()]
    p = pp.copy()
    for itime in range(1, 6):
        s.step(itime)
        u, p, _, _ = lex.step(itime, *u, p)
    g = s.download()
    for c in range(3):
        assert np.abs(g[c] - u[c]).max() <= 1e-8
    assert np.abs((g[3] - g[3].mean()) - (p - p.mean())).max() <= 1e-8


def test_mixed_precision_tracks_fp64():
    cfg = tgv_config(32, 10, 3)
    ux, uy, uz, pp = taylor_green(cfg)
    a = Solver(cfg, FP64)
    b = Solver(cfg, MIXED)
    a.upload(ux, uy, uz, pp)
    b.upload(ux, uy, uz, pp)
    for itime in range(1, 11):
        ia = a.step(itime)
        ib = b.step(itime)
        assert ib["sor_reason"] == 0 and ib["sor_dmax"] < cfg.eps
        assert ib["sweeps_f32"] > 0 and ib["sweeps_f64"] >= 1
    sa = a.statistics(10 * a.dt)
    sb = b.statistics(10 * b.dt)
    assert abs(sb[1] - sa[1]) <= 1e-6 * abs(sa[1])     # kinetic energy
    assert abs(sb[4] - sa[4]) <= 1e-5 * abs(sa[4])     # enstrophy
    ga, gb = a.download(), b.download()
    for c in range(3):
        assert np.abs(ga[c] - gb[c]).max() <= 1e-5
    a.close()
    b.close()


def _gold(name):
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", name), allow_pickle=False)


@pytest.mark.parametrize("name", ["run_tgv16_ab2.npz", "run_tgv16fs_ab3_les.npz",
                                  "run_mixing24_ab2_les.npz"])
def test_fp64_matches_reference_runs(name):
    """GPU (red-black SOR) against whole reference runs (lexicographic SOR):
    the orderings change the iterates, not the converged fields, so the
    fields agree to a tolerance tied to eps when SOR converges.  With a
    free-slip y axis the reference's 2nd-order Neumann Laplacian is not
    compatible with the 6th-order divergence, SOR stops on stagnation with
    dmax ~ 1e-2 (both orderings), and the fields agree only to that level."""
    g = _gold(name)
    cfg = Config.from_namelist(bytes(g["params"]).decode())
    print(f"[gpu] {name}", flush=True)
    s = Solver(cfg, FP64)
    s.upload(g["ux0"], g["uy0"], g["uz0"], g["pp0"])
    stagnated = False
    for itime in range(1, cfg.itstop + 1):
        info = s.step(itime)
        stagnated |= info["sor_reason"] == 1
        if f"ux{itime}" in g:
            out = s.download()
            tol_u, tol_p = (2e-3, 2e-2) if stagnated else (2e-6, 2e-5)
            for c, k in enumerate(("ux", "uy", "uz")):
                assert np.abs(out[c] - g[f"{k}{itime}"]).max() <= tol_u, (itime, k)
            pr = g[f"pp{itime}"]
            assert np.abs((out[3] - out[3].mean()) - (pr - pr.mean())).max() <= tol_p
    assert stagnated == ("mixing" in name)
    s.close()


def test_config1_tgv64_matches_reference():
    """Config 1: TGV Re=1600 64^3 AB2, 100 steps; fields at 4096 sampled
    nodes after steps 10, 50, 100 against the reference CPU run, tolerance
    tied to eps = 1e-8.  Red-black ordering reaches dmax < eps in fewer
    sweeps than the reference's lexicographic ordering (about 0.65x here);
    the test asserts it never needs more."""
    g = _gold("c1_tgv64.npz")
    cfg = Config.from_namelist(bytes(g["params"]).decode())
    idx = g["idx"]
    s = Solver(cfg, FP64)
    s.init_taylor_green()
    iters = []
    for itime in range(1, 101):
        info = s.step(itime)
        iters.append(info["sor_exit_iter"])
        assert info["sor_reason"] == 0
        if itime in (10, 50, 100):
            out = s.download()
            pm = out[3] - out[3].mean()
            for c, k in enumerate(("ux", "uy", "uz")):
                err = np.abs(out[c].reshape(-1)[idx] - g[f"{k}{itime}"]).max()
                print(f"[gpu] c1 step {itime} {k} max err {err:.2e}", flush=True)
                assert err <= 1e-6
            assert np.abs(pm.reshape(-1)[idx] - g[f"pm{itime}"]).max() <= 1e-5
            ek = 0.5 * np.mean(out[0] ** 2 + out[1] ** 2 + out[2] ** 2)
            assert abs(ek - float(g[f"ek{itime}"])) <= 1e-8 * abs(float(g[f"ek{itime}"]))
    ref_iters = np.array(g["iters"], dtype=float)
    print(f"[gpu] c1 iterations: gpu mean {np.mean(iters):.1f}, reference mean "
          f"{ref_iters.mean():.1f}", flush=True)
    assert np.mean(iters) <= ref_iters.mean()
    assert np.all(np.array(iters) <= 1.1 * ref_iters)
    s.close()


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_checkpoint_restart_continues_bit_for_bit(tmp_path, precision):
    """fields.bin + the AB-history sidecar restore a run exactly: 4 steps,
    checkpoint, a fresh solver restored from it, 4 more steps == 8 steps in
    one go (AB3 uses the two previous right-hand sides after the restart)."""
    from osinco3d_amd import checkpoint
    cfg = tgv_config(16, 8, 3)
    a = Solver(cfg, precision)
    a.init_taylor_green()
    for it in range(1, 9):
        a.step(it)
    ref = a.download()
    a.close()
    b = Solver(cfg, precision)
    b.init_taylor_green()
    for it in range(1, 5):
        b.step(it)
    path = str(tmp_path / "fields.bin")
    checkpoint.save(b, path, 4, 4 * b.dt)
    b.close()
    c = Solver(cfg, precision)
    itime, t = checkpoint.load(c, path)
    assert itime == 4 and abs(t - 4 * c.dt) < 1e-15
    for it in range(itime + 1, 9):
        c.step(it)
    got = c.download()
    c.close()
    for x, y in zip(got, ref):
        assert np.array_equal(x, y)


def test_checkpoint_restart_with_warm_start_is_reproducible(tmp_path):
    """The extrapolated warm start re-primes on restart: a fresh solver and
    one that has already stepped (its increment history stale), both
    restored from the same checkpoint, continue bit for bit alike, and stay
    close to the uninterrupted run (same stop rule, other starting points)."""
    from osinco3d_amd import checkpoint
    from osinco3d_amd.solver import WARM_EXTRAPOLATE
    cfg = tgv_config(16, 8, 3)
    a = Solver(cfg, MIXED, WARM_EXTRAPOLATE)
    a.init_taylor_green()
    for it in range(1, 5):
        a.step(it)
    path = str(tmp_path / "fields.bin")
    checkpoint.save(a, path, 4, 4 * a.dt)
    for it in range(5, 9):
        a.step(it)
    ref = a.download()
    outs = []
    for stepped in (False, True):
        b = Solver(cfg, MIXED, WARM_EXTRAPOLATE)
        b.init_taylor_green()
        if stepped:
            for it in range(1, 7):
                b.step(it)
        itime, _ = checkpoint.load(b, path)
        for it in range(itime + 1, 9):
            b.step(it)
        outs.append(b.download())
        b.close()
    a.close()
    for x, y in zip(*outs):
        assert np.array_equal(x, y)
    for x, y in zip(outs[0][:3], ref[:3]):
        assert np.abs(x - y).max() <= 1e-6 * np.abs(y).max()
