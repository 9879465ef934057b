"""What the warm starts gain, on the configuration they are meant for:
Taylor-Green Re 1600 at 256^3, mixed-precision form, SOR to eps = 1e-8.

The method without them starts each solve from the previous step's pressure
(the reference's behaviour).  The extrapolated guess p_prev + e_prev is
measured on its own, and the learned form (the network's coarse correction
of that extrapolation, the previous pressure as fallback) beside it, so the
network's own share is visible: the network may not cost sweeps against the
extrapolation it corrects, and on held-out solves of a run it was not
trained on its prediction of the increment must beat the extrapolation's.
The coarse-grid correction (multigrid on the 2^3 block averages of the
guess's residual) with Gauss-Seidel sweeps opening the FP32 solve is
measured after each.  No test here claims that the network beats the
coarse-grid correction: it does not (measured figures in the design notes).
The solve's stop rule is the same in every form, so the fields agree to a
tolerance tied to eps.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from osinco3d_amd.config import tgv_config
from osinco3d_amd.solver import MIXED, WARM_EXTRAPOLATE, WARM_LEARNED, WARM_PREVIOUS, Solver

pytestmark = pytest.mark.gpu

N, STEPS, SKIP = 256, 16, 3
COARSE = (2, 8)
WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "osinco3d_amd", "warmstart_weights.npy")


def run(warm, net=None, coarse=(0, 0)):
    cfg = tgv_config(N, STEPS, 3)
    s = Solver(cfg, MIXED, warm, coarse_cycles=coarse[0], smooth_sweeps=coarse[1])
    if net is not None:
        net.attach(s)
    s.init_taylor_green()
    iters, used = [], []
    for it in range(1, STEPS + 1):
        info = s.step(it)
        assert info["sor_reason"] in (0, 1) and not info["diverged"]
        if it > SKIP:
            iters.append(info["sweeps_f32"] + info["sweeps_f64"])
            used.append(info["warm_used"])
    fields = s.download()
    s.close()
    return np.array(iters), np.mean(used), fields


def test_warm_starts_reduce_sor_sweeps():
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained warm-start weights in the package")
    from osinco3d_amd import warmstart
    net = warmstart.load_or_none(WEIGHTS, (N, N, N))
    try:
        it_prev, _, f_prev = run(WARM_PREVIOUS)
        it_ex, used_ex, _ = run(WARM_EXTRAPOLATE)
        it_nn, used, f_nn = run(WARM_LEARNED, net)
        it_mg, used_mg, f_mg = run(WARM_EXTRAPOLATE, coarse=COARSE)
        it_nm, used_nm, f_nm = run(WARM_LEARNED, net, COARSE)
    finally:
        net.close()
    print(f"[warm] SOR sweeps/step at {N}^3: previous {it_prev.mean():.1f}, "
          f"extrapolation {it_ex.mean():.1f}, learned {it_nn.mean():.1f} (accepted on "
          f"{used:.2f} of the steps; network's own gain "
          f"{100 * (1 - it_nn.mean() / it_ex.mean()):.1f} %), extrapolation + coarse-grid "
          f"correction {it_mg.mean():.1f}, learned + coarse-grid correction "
          f"{it_nm.mean():.1f}", flush=True)
    assert used_ex == 1.0 and used == 1.0 and used_mg == 1.0 and used_nm == 1.0
    assert it_ex.mean() <= 0.8 * it_prev.mean()      # the extrapolation's gain
    assert it_nn.mean() <= it_ex.mean()              # the network may not lose sweeps
    assert it_mg.mean() <= 0.5 * it_ex.mean()        # the coarse-grid correction's gain
    assert it_nm.mean() <= 0.5 * it_ex.mean()
    for f in (f_nn, f_mg, f_nm):
        for c in range(3):
            assert np.abs(f[c] - f_prev[c]).max() <= 1e-6


def test_network_beats_extrapolation_on_unseen_solves():
    """The network's own target: the increment e = p_new - p_prev of the
    solves (block averages on its grid, divided by RMS(e_prev)).  On solves
    of a run it was not trained on (Taylor-Green Re 1000 at 128^3), its
    relative L2 error is below that of the linear extrapolation e = e_prev
    it corrects."""
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained warm-start weights in the package")
    from osinco3d_amd import warmstart
    cfg = tgv_config(128, 60, 3, re=1000.0)
    samples, iters = warmstart.capture_run(cfg, 60, MIXED, skip=10)
    net = warmstart.load_or_none(WEIGHTS, (128, 128, 128))
    try:
        err = warmstart.relative_errors(net, samples)
    finally:
        net.close()
    print(f"[warm] unseen TGV Re 1000 128^3, {len(samples)} solves: relative error of the "
          f"increment: network {err['nn']:.4f}, extrapolation {err['extrapolation']:.4f}, "
          f"previous pressure {err.get('previous', float('nan')):.4f}", flush=True)
    assert err["nn"] < err["extrapolation"]


def test_coarse_grid_correction_on_incompatible_mixing_layer():
    """The mixing layer's rhs is not compatible with the free-slip mirror:
    the FP64 form's solves end on the stagnation rule.  The mixed form
    removes the rhs's weighted mean (the part the mirrored operator cannot
    reach), so with the coarse-grid correction every solve ends on eps; with
    rhs_as_given it solves the incompatible rhs as FP64 does and every solve
    ends on the stagnation rule, well before kmax.  Both track FP64 (the
    projection changes the velocity by far less than the eps-tied
    tolerance: about 2e-9 on a 48x32x16 oracle run)."""
    from osinco3d_amd.config import mixing_layer_config
    from osinco3d_amd.solver import FP64, mixing_layer
    cfg = mixing_layer_config(96, 64, 32, steps=8)
    init = mixing_layer(cfg)
    runs = {}
    for name, prec, warm, coarse, as_given in (
            ("fp64", FP64, WARM_PREVIOUS, (0, 0), 0),
            ("mg", MIXED, WARM_EXTRAPOLATE, COARSE, 0),
            ("mg, rhs as given", MIXED, WARM_EXTRAPOLATE, COARSE, 1)):
        s = Solver(cfg, prec, warm, coarse_cycles=coarse[0], smooth_sweeps=coarse[1],
                   rhs_as_given=as_given)
        s.upload(*init)
        info = [s.step(it) for it in range(1, 9)]
        st = s.statistics(8 * s.dt)
        s.close()
        runs[name] = (info, st)
        print(f"[mix] {name}: sweeps/step {np.mean([i['sor_exit_iter'] for i in info]):.1f}, "
              f"reasons {sorted(set(i['sor_reason'] for i in info))}, E_k {st[1]:.12e}",
              flush=True)
    info_64, st_64 = runs["fp64"]
    assert all(i["sor_reason"] == 1 for i in info_64)
    info_p, st_p = runs["mg"]
    assert all(i["sor_reason"] == 0 for i in info_p)
    info_g, st_g = runs["mg, rhs as given"]
    assert all(i["sor_reason"] == 1 for i in info_g)
    assert max(i["sor_exit_iter"] for i in info_g) < cfg.kmax // 2
    for st in (st_p, st_g):
        assert abs(st[1] - st_64[1]) <= 1e-8 * abs(st_64[1])


def _lap7(f, d):
    """The solver's 7-point periodic Laplacian in FP64, arrays (nz, ny, nx)."""
    ox, oy, oz = (1.0 / h ** 2 for h in d)
    return (ox * (np.roll(f, 1, 2) + np.roll(f, -1, 2)) + oy * (np.roll(f, 1, 1) + np.roll(f, -1, 1))
            + oz * (np.roll(f, 1, 0) + np.roll(f, -1, 0)) - 2.0 * (ox + oy + oz) * f)


def test_adaptive_extrapolation_weight_is_least_squares():
    """The adaptive guess p_prev + e_prev + beta (e_prev - e_prev2): the
    weight of step k + 1 is the one minimising the L2 residual of the guess
    at step k, || r_lin - beta L d || with d = e_prev - e_prev2, i.e.
    beta* = (r_q + beta_k L d) . L d / || L d ||^2 from the residual r_q of
    the guess step k used.  Recomputed here in numpy from the captured r_q
    and increments, then clamped to [0, 1.5] as the solver does."""
    import torch

    from osinco3d_amd import capi
    n, steps = 64, 14
    cfg = tgv_config(n, steps, 3)
    d = cfg.spacing()
    s = Solver(cfg, MIXED, WARM_EXTRAPOLATE, adaptive=1)
    s.init_taylor_green()
    r = torch.empty((n, n, n), dtype=torch.float32, device="cuda")
    ep = torch.empty_like(r)
    e = torch.empty_like(r)
    capi.check(capi.load().o3d_solver_capture(s.handle, capi.ptr(r), capi.ptr(ep), capi.ptr(e)))
    infos, rs, eps_ = [], [], []
    for it in range(1, steps + 1):
        infos.append(s.step(it))
        rs.append(r.cpu().numpy().astype(np.float64))
        eps_.append(ep.cpu().numpy())
    capi.check(capi.load().o3d_solver_capture(s.handle, None, None, None))
    s.close()
    checked = 0
    for k in range(3, steps - 1):     # 0-based step index; two increments exist from k = 2
        dd = (eps_[k] - eps_[k - 1]).astype(np.float64)   # FP32 difference, as the kernel forms it
        ld = _lap7(dd, d)
        beta_k = infos[k]["extrap_beta"]
        want = float(((rs[k] + beta_k * ld) * ld).sum() / (ld * ld).sum())
        got = infos[k + 1]["extrap_beta_ls"]
        print(f"[adapt] step {k + 2}: beta used {infos[k + 1]['extrap_beta']:.4f}, least squares "
              f"{got:.6f} (numpy {want:.6f})", flush=True)
        assert abs(got - want) <= 1e-6 * max(1.0, abs(want)), (k, got, want)
        assert infos[k + 1]["extrap_beta"] == min(1.5, max(0.0, got))
        checked += 1
    assert checked >= 8


def _residual_order(q, rhs32, d):
    """rhs - L q in FP64 with the residual kernels' operation order
    (ox (xm + xp) + oy (ym + yp) + oz (zm + zp) + A c), rounded to FP32."""
    ox, oy, oz = (1.0 / h ** 2 for h in d)
    A = -(2.0 * ox + 2.0 * oy + 2.0 * oz)
    lap = (ox * (np.roll(q, 1, 2) + np.roll(q, -1, 2)) + oy * (np.roll(q, 1, 1) + np.roll(q, -1, 1))
           + oz * (np.roll(q, 1, 0) + np.roll(q, -1, 0)) + A * q)
    return (rhs32.astype(np.float64) - lap).astype(np.float32)


def _residual_bc(q, rhs32, d, per):
    """rhs - L q with the SOR neighbour rules (periodic wrap, or the
    free-slip mirror q(-1) = q(1), q(n) = q(n - 2)) and the residual
    kernels' operation order, rounded to FP32; q is (nz, ny, nx) and per
    holds the periodic flags of x, y, z."""
    def nb(axis, s):
        n = q.shape[axis]
        idx = np.arange(n) + s
        if per[2 - axis]:
            idx = idx % n
        else:
            idx = np.where(idx < 0, -idx, np.where(idx > n - 1, 2 * (n - 1) - idx, idx))
        return np.take(q, idx, axis=axis)
    ox, oy, oz = (1.0 / h ** 2 for h in d)
    A = -(2.0 * ox + 2.0 * oy + 2.0 * oz)
    lap = (ox * (nb(2, -1) + nb(2, 1)) + oy * (nb(1, -1) + nb(1, 1))
           + oz * (nb(0, -1) + nb(0, 1)) + A * q)
    return (rhs32.astype(np.float64) - lap).astype(np.float32)


@pytest.mark.parametrize("kind", ["periodic", "free-slip y"])
def test_residual_kernel_is_bit_exact(kind):
    """The LDS-tiled residual r = rhs - L p (the first step of a warm form,
    before any increment exists) against numpy, bit for bit, from a random
    pressure on a 72 x (38 or 26) x 22 grid: partial tiles in x (64 + 8) and
    y (rows of 4) and a partial last z chunk (8 + 8 + 6 planes), so every
    halo duty, the periodic wrap and the free-slip mirror are exercised."""
    import torch

    from osinco3d_amd import capi
    from osinco3d_amd.config import mixing_layer_config
    from osinco3d_amd.solver import taylor_green
    if kind == "periodic":
        cfg = tgv_config(32, 2, 3)
        cfg.nx, cfg.ny, cfg.nz = 72, 38, 22
        per = (1, 1, 1)
    else:
        cfg = mixing_layer_config(72, 26, 22, steps=2)
        per = (1, 0, 1)
    shape = (cfg.nz, cfg.ny, cfg.nx)
    print(f"[gpu] residual kernel bit-exact {kind} {shape}", flush=True)
    d = cfg.spacing()
    u = taylor_green(cfg)[:3]
    p0 = np.random.default_rng(17).standard_normal(shape)
    s = Solver(cfg, MIXED, WARM_EXTRAPOLATE)
    s.upload(*u, p0)
    r = torch.empty(shape, dtype=torch.float32, device="cuda")
    ep = torch.empty_like(r)
    e = torch.empty_like(r)
    capi.check(capi.load().o3d_solver_capture(s.handle, capi.ptr(r), capi.ptr(ep), capi.ptr(e)))
    s.step(1)
    rhs = s.field(7)
    want = _residual_bc(p0, rhs, d, per)
    got = r.cpu().numpy()
    capi.check(capi.load().o3d_solver_capture(s.handle, None, None, None))
    s.close()
    assert np.array_equal(got, want), np.abs(got - want).max()


@pytest.mark.parametrize("adaptive", [0, 1])
def test_guess_residual_is_bit_exact(adaptive):
    """The guess the extrapolated warm start forms and its FP32 residual,
    which open every solve of the warm forms: q = p_prev + e_prev (+ beta
    (e_prev - e_prev2), the FP32 difference) in FP64 and r = rhs - L q with
    the kernels' operation order, recomputed in numpy from the solver's own
    fields and compared bit for bit; and each step's increment
    e = p_new - p_prev, with p_new after the periodic mean shift (its mean
    at round-off)."""
    import torch

    from osinco3d_amd import capi
    n, steps = 32, 8
    cfg = tgv_config(n, steps, 3)
    d = cfg.spacing()
    s = Solver(cfg, MIXED, WARM_EXTRAPOLATE, adaptive=adaptive)
    s.init_taylor_green()
    r = torch.empty((n, n, n), dtype=torch.float32, device="cuda")
    ep = torch.empty_like(r)
    e = torch.empty_like(r)
    capi.check(capi.load().o3d_solver_capture(s.handle, capi.ptr(r), capi.ptr(ep), capi.ptr(e)))
    eps_prev = None
    checked = 0
    for it in range(1, steps + 1):
        p_prev = s.field(6)
        info = s.step(it)
        rhs = s.field(7)
        e1 = ep.cpu().numpy()
        # The step's increment, formed in the velocity correction pass from
        # the pressure after its mean shift: e = p_new - p_prev in FP32.
        p_new = s.field(6)
        assert abs(p_new.mean()) <= 1e-12 * np.abs(p_new).max(), it
        assert np.array_equal(e.cpu().numpy(), (p_new - p_prev).astype(np.float32)), it
        if it >= 2:
            q = p_prev + e1.astype(np.float64)
            if adaptive and it >= 3 and info["extrap_beta"] != 0.0:
                dd = (e1 - eps_prev).astype(np.float64)
                q = (p_prev + e1.astype(np.float64)) + info["extrap_beta"] * dd
            want = _residual_order(q, rhs, d)
            got = r.cpu().numpy()
            assert np.array_equal(got, want), (it, np.abs(got - want).max())
            checked += 1
        eps_prev = e1
    capi.check(capi.load().o3d_solver_capture(s.handle, None, None, None))
    s.close()
    assert checked == steps - 1


def test_fp64_form_with_extrapolated_guess():
    """The FP64 form accepts the extrapolated warm start too (the guess and
    residual kernels in their FP64-rhs instances, the increment formed in
    the FP64 correction pass): the solves end on the reference's rules, the
    fields track the FP64 form started from the previous pressure to the
    eps-tied tolerance with fewer sweeps, and each step's increment is
    p_new - p_prev in FP32, bit for bit."""
    import torch

    from osinco3d_amd import capi
    from osinco3d_amd.solver import FP64
    n, steps = 32, 8
    cfg = tgv_config(n, steps, 3)
    runs = {}
    for name, warm in (("previous", WARM_PREVIOUS), ("extrapolation", WARM_EXTRAPOLATE)):
        s = Solver(cfg, FP64, warm)
        s.init_taylor_green()
        cap = None
        if warm == WARM_EXTRAPOLATE:
            cap = [torch.empty((n, n, n), dtype=torch.float32, device="cuda") for _ in range(3)]
            capi.check(capi.load().o3d_solver_capture(s.handle, *[capi.ptr(t) for t in cap]))
        sweeps = []
        for it in range(1, steps + 1):
            p_prev = s.field(6)
            info = s.step(it)
            assert info["sor_reason"] in (0, 1) and not info["diverged"], (name, it)
            sweeps.append(info["sor_exit_iter"])
            if cap is not None:
                e = cap[2].cpu().numpy()
                assert np.array_equal(e, (s.field(6) - p_prev).astype(np.float32)), it
        if cap is not None:
            capi.check(capi.load().o3d_solver_capture(s.handle, None, None, None))
        runs[name] = (np.array(sweeps), s.download())
        s.close()
    (sp, fp), (se, fe) = runs["previous"], runs["extrapolation"]
    print(f"[warm] FP64 form at {n}^3: sweeps/step previous {sp.mean():.1f}, extrapolation "
          f"{se.mean():.1f}", flush=True)
    assert se[2:].mean() < sp[2:].mean()
    for c in range(3):
        assert np.abs(fe[c] - fp[c]).max() <= 1e-6
