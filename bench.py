#!/usr/bin/env python3
"""Benchmark of the osinco3d time-step path on one MI355X.

Workload: Taylor-Green vortex, Re = 1600, 256^3 exactly periodic box, AB3,
CFL 0.25, SOR omega = 2/(1+sin(pi/256)), eps = 1e-8, kmax = 5000.  A step is
one full time step of the reference driver (predict_velocity,
correct_pression with the Poisson solve to eps, correct_velocity and the
driver's per-step diagnostics), with all fields resident in HBM.

Each requested form is timed over the same steps:
  fp64   plain FP64 (the drop-in; bit-identical operators to the reference)
  mixed  FP32 stencils / AB update / SOR sweeps, FP64-refined Poisson
  extrap mixed + warm start by linear extrapolation p_prev + (p_prev - p_prev2)
  mg     the adaptive extrapolation p_prev + e_prev + beta (e_prev - e_prev2)
         (beta the least-squares weight of the previous step), then the
         coarse-grid correction of the guess (multigrid V-cycles on the 2^3
         block averages of its residual) and Gauss-Seidel sweeps opening the
         FP32 correction solve
  nn0    mixed + learned warm start (the trained network of
         osinco3d_amd/warmstart_weights.npy; skipped when it is missing)
  nn     the learned warm start with every stage of mg: adaptive extrapolation,
         the network's estimate as the start of the coarse-grid correction,
         Gauss-Seidel opening (the headline)
  mgl    mg with the linear extrapolation (beta = 0)
  nnl    nn with the linear extrapolation
  nn1    nn with one V-cycle (instead of coarse_cycles) in the first coarse-grid
         correction, which starts from the network's coarse estimate
The JSON line's value is the steps/s of the headline form (the last of
--forms); every form's figures are included under "forms".

late_window: the forms fp64, mg and nn run the first 975 steps of the
1000-step run untimed by events (wall-clocked as a block), then steps
976-1000 are timed: steps/s of that late window, the 1000-step mean steps/s,
sweeps per step, and E_k / enstrophy at step 1000 against the FP64 form.
The mixing workload runs the same series over 500 steps (--late-total).

roofline: the dominant work of the headline form is the red-black SOR
iteration (in the mixed forms the FP32 colour-split red and black launches,
sor_split_kernel<0> and <1>).  Its algorithmic bytes per iteration are those
of a red-black sweep that reads the whole iterate once per colour: (2 reads
of p, 1 read of rhs, 1 write of p) = 4 values per cell, 16 B/cell in FP32 and
32 B/cell in FP64, times the cell count; the floor of one read of p and rhs
and one write of p (12 / 24 B/cell) is reported beside it.  Its average
duration is measured with HIP events on the launch stream over complete
batches of iterations inside the timed steps (the start event follows each
batch's first iteration), and, for the FP32 iteration, right after the
timed steps over 200 back-to-back iterations of the solver's own correction
problem (o3d_solver_sor_period, one event interval of 199 iterations): the
in-step intervals hold only 7 to 10 iterations each, and the two events'
own processing adds about 3 us per iteration to them.  The roofline uses
the back-to-back figure and reports the in-step one beside it; `traffic` is the measured HBM
bytes of one iteration (both launches) from the committed rocprofv3 --pmc
summary.  The ceiling beside the nominal 8 TB/s is a hand-written streaming
triad over three 1 GiB buffers (o3d_stream_probe).

cpu_baseline: the reference Fortran (oracle/_ref/o3d_ref_run, built from the
reference sources at -O3, single-threaded) on one host core: one full step
of the same 256^3 Taylor-Green case (kmax = 5000, the SOR to eps = 1e-8),
wall-clocked, with its phases timed by the harness.  The former model (a
kmax-capped step extrapolated by the lexicographic/red-black sweep ratio
of config 1) is kept beside it as a cross-check.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import resource
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TGV 256³ steps/s + Poisson ms & iters to eps=1e-8; plain FP64 vs mixed+NN"
DEFAULT_MODEL = os.path.join(ROOT, "osinco3d_amd", "warmstart_weights.npy")
HBM_PEAK_GBS = 8000.0
MFMA_BF16_TFS = 2500.0


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--forms", default="fp64,mixed,extrap,mg,nn0,nn")
    ap.add_argument("--coarse-cycles", type=int, default=2,
                    help="V-cycles of the coarse-grid correction (forms mg, nn)")
    ap.add_argument("--smooth-sweeps", type=int, default=8,
                    help="Gauss-Seidel sweeps opening the FP32 solve (forms mg, nn)")
    ap.add_argument("--workload", choices=("tgv", "mixing", "jet", "poisson"), default="tgv",
                    help="tgv: the headline Taylor-Green case; mixing: the 384x256x128 LES "
                         "temporal mixing layer (free-slip y, AB2), no CPU baseline; "
                         "jet: the 512x384x256 planar jet with x-Dirichlet inflow / outflow "
                         "(LES, AB3; forms fp64, mixed, mg unless --forms), no CPU baseline; "
                         "poisson: one n^3 pressure solve from zero, FP64 SOR against the "
                         "multigrid solver (multigrid = 1)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-kmax", type=int, default=5000,
                    help="kmax of the reference's measured 256^3 step (5000: a full step)")
    ap.add_argument("--late-forms", default="fp64,mg,nn",
                    help="forms run through the late-window series ('' to skip)")
    ap.add_argument("--late-total", type=int, default=None,
                    help="steps of the late-window series (default: 1000 for tgv, 500 for "
                         "mixing, none for jet)")
    ap.add_argument("--late-window", type=int, default=25)
    ap.add_argument("--model", default=DEFAULT_MODEL)
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sor_split_kernel_tgv256_pmc.md")


def measured_traffic(which):
    """HBM bytes per SOR launch from the committed rocprofv3 --pmc summary
    (lines "TRAFFIC f32 <bytes>" / "TRAFFIC f64 <bytes>": FETCH_SIZE +
    WRITE_SIZE per dispatch, scaled by the same counters' reading for a copy
    of known size), or None when there is no such summary."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                m = re.match(r"^TRAFFIC (f32|f64) ([0-9.eE+]+)", line.strip())
                if m and m.group(1) == which:
                    return float(m.group(2))
    except OSError:
        pass
    return None


def pct(v, q):
    v = sorted(v)
    if not v:
        return None
    k = min(len(v) - 1, max(0, int(round(q / 100.0 * (len(v) - 1)))))
    return v[k]


LEARNED_FORMS = ("nn", "nn0", "nn1", "nnl")
COARSE_FORMS = ("mg", "mgl", "nn", "nn1", "nnl")
WARM_FORMS = LEARNED_FORMS + ("extrap", "mg", "mgl")


def make_solver(form, cfg, model_path, coarse):
    """The solver of a bench form and its network (None for the forms
    without one); (None, None) when a learned form has no model."""
    from osinco3d_amd import solver as S

    prec = S.FP64 if form == "fp64" else S.MIXED
    warm = S.WARM_PREVIOUS
    if form in LEARNED_FORMS:
        warm = S.WARM_LEARNED
    elif form in ("extrap", "mg", "mgl"):
        warm = S.WARM_EXTRAPOLATE
    cc, gs = coarse if form in COARSE_FORMS else (0, 0)
    net = None
    if form in LEARNED_FORMS:
        from osinco3d_amd import warmstart
        net = warmstart.load_or_none(model_path, (cfg.nx, cfg.ny, cfg.nz))
        if net is None:
            return None, None
    s = S.Solver(cfg, prec, warm, coarse_cycles=cc, smooth_sweeps=gs,
                 coarse_cycles_learned=1 if form == "nn1" else 0,
                 adaptive=1 if form in ("mg", "nn", "nn1") else 0)
    if net is not None:
        net.attach(s)
    return s, net


def run_form(form, cfg, steps, warmup, model_path, init=None, coarse=(0, 0)):
    import torch

    s, net = make_solver(form, cfg, model_path, coarse)
    if s is None:
        print(f"[bench] form {form} skipped: no model at {model_path}", file=sys.stderr,
              flush=True)
        return None
    cc, gs = coarse if form in COARSE_FORMS else (0, 0)
    if init is None:
        s.init_taylor_green()
    elif callable(init):
        init(s)
    else:
        s.upload(*init)
    for it in range(1, warmup + 1):
        s.step(it)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = []
    sor_ms = [0.0, 0.0]
    sor_sw = [0, 0]
    sor_ln = [0, 0]
    nn_ms, nn_fl, nn_n = 0.0, 0.0, 0
    for k in range(1, steps + 1):
        infos.append(s.step(warmup + k))
        ms, sw, ln = s.sor_timing(launches=True)
        for q in range(2):
            sor_ms[q] += ms[q]
            sor_sw[q] += sw[q]
            sor_ln[q] += ln[q]
        if net is not None:
            a, b = net.last_timing()
            nn_ms += a
            nn_fl += b
            nn_n += 1
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    t_end = (warmup + steps) * s.dt
    st = s.statistics(t_end)
    # The FP32 iteration back to back: 200 iterations of the solver's own
    # correction problem between two HIP events (after the timed steps; the
    # in-step batches hold only 7 to 10 iterations per event interval).
    period_us = None
    if form != "fp64" and cfg.nx % 8 == 0:
        period_us = s.sor_period(200)
    s.close()
    elapsed = t1 - t0
    n = cfg.nx * cfg.ny * cfg.nz
    iters = [i["sor_exit_iter"] for i in infos]
    pms = [i["poisson_ms"] for i in infos]
    out = {
        "steps_per_s": steps / elapsed,
        "ms_per_step": 1e3 * elapsed / steps,
        "mcell_updates_per_s": n * steps / elapsed / 1e6,
        "poisson_iters": {"mean": statistics.mean(iters), "p50": pct(iters, 50),
                          "p95": pct(iters, 95)},
        "poisson_ms": {"mean": statistics.mean(pms), "p50": pct(pms, 50), "p95": pct(pms, 95)},
        "sweeps_f64_mean": statistics.mean(i["sweeps_f64"] for i in infos),
        "sweeps_f32_mean": statistics.mean(i["sweeps_f32"] for i in infos),
        "refinements_mean": statistics.mean(i["refinements"] for i in infos),
        "sor_reasons": sorted({i["sor_reason"] for i in infos}),
        "diverged": any(i["diverged"] for i in infos),
        "kinetic_energy": st[1],
        "enstrophy": st[4],
        "t_end": t_end,
    }
    # SOR iterations: average duration from events over complete batches of
    # launches (one launch = one red-black iteration: the red and the black
    # colour-split launch, sor_split_kernel<0> and <1>).  Algorithmic bytes
    # of an iteration: a red-black sweep reading the whole iterate once per
    # colour, the rhs once and writing the iterate once, 4 values per cell
    # (32 B/cell FP64, 16 B/cell FP32); the floor of one read of p and rhs
    # and one write of p (24 / 12 B/cell) beside it.
    for q, name, bpc, floor in ((0, "sor_f64", 32, 24), (1, "sor_f32", 16, 12)):
        if sor_ln[q] > 0:
            avg_ms = sor_ms[q] / sor_ln[q]
            gbs = bpc * n / (avg_ms * 1e-3) / 1e9
            out[name] = {"in_step_us_per_iteration": 1e3 * avg_ms,
                         "avg_launch_us": 1e3 * avg_ms, "launches_timed": sor_ln[q],
                         "iterations_per_launch": sor_sw[q] / sor_ln[q],
                         "us_per_iteration": 1e3 * sor_ms[q] / sor_sw[q],
                         "algorithmic_bytes_per_launch": bpc * n, "achieved_GBps": gbs,
                         "floor_bytes_per_launch": floor * n,
                         "floor_GBps": floor * n / (avg_ms * 1e-3) / 1e9}
            if q == 1 and period_us:
                # The roofline figure: the back-to-back period, whose event
                # interval spans 199 iterations.
                avg_ms = period_us * 1e-3
                out[name].update({
                    "avg_launch_us": period_us, "timing": "back-to-back, 199 iterations",
                    "achieved_GBps": bpc * n / (avg_ms * 1e-3) / 1e9,
                    "floor_GBps": floor * n / (avg_ms * 1e-3) / 1e9})
    if cc or gs:
        out["coarse_cycles"] = cc
        out["smooth_sweeps"] = gs
    if form in WARM_FORMS:
        out["extrap_beta_mean"] = statistics.mean(i["extrap_beta"] for i in infos)
        out["warm_used_frac"] = statistics.mean(i["warm_used"] for i in infos)
        out["warm_residual_l2"] = {"previous": statistics.mean(i["warm_dmax"][0] for i in infos),
                                   "guess": statistics.mean(i["warm_dmax"][1] for i in infos)}
    if net is not None and nn_n and nn_ms > 0:
        tfs = nn_fl / (nn_ms * 1e-3) / 1e12
        out["nn_inference"] = {"ms_per_step": nn_ms / nn_n, "flops": nn_fl / nn_n,
                               "tflops": tfs, "mfma_util_bf16_dense": tfs / MFMA_BF16_TFS}
    if net is not None:
        net.close()
    return out


def stream_bandwidth(nbytes=1 << 30, reps=20):
    """HBM ceiling: the library's hand-written streaming triad (two reads and
    one write per element) over three 1 GiB device buffers, far beyond the
    256 MB Infinity Cache, timed with HIP events."""
    from osinco3d_amd import capi
    return capi.stream_probe(nbytes, reps)[0]


def run_series(form, cfg, total, window, model_path, coarse, init=None):
    """The first total - window steps of a run (wall-clocked as one block),
    then the last `window` steps timed on their own: the late-window figures
    and the whole-run mean."""
    import torch

    s, net = make_solver(form, cfg, model_path, coarse)
    if s is None:
        return None
    if init is None:
        s.init_taylor_green()
    elif callable(init):
        init(s)
    else:
        s.upload(*init)
    first = total - window
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sweeps_all = []
    for it in range(1, first + 1):
        sweeps_all.append(s.step(it)["sor_exit_iter"])
        if it % 250 == 0:
            print(f"[bench] late window {form}: step {it}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    infos = [s.step(it) for it in range(first + 1, total + 1)]
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    st = s.statistics(total * s.dt)
    s.close()
    if net is not None:
        net.close()
    sweeps_all += [i["sor_exit_iter"] for i in infos]
    beta = [i["extrap_beta"] for i in infos]
    late = [i["sor_exit_iter"] for i in infos]
    q = len(sweeps_all) // 4
    return {
        "late_steps": [first + 1, total],
        "late_steps_per_s": window / (t2 - t1),
        "late_ms_per_step": 1e3 * (t2 - t1) / window,
        "late_sweeps_mean": statistics.mean(late),
        "late_refinements_mean": statistics.mean(i["refinements"] for i in infos),
        "late_poisson_ms_mean": statistics.mean(i["poisson_ms"] for i in infos),
        "mean_steps_per_s": total / (t2 - t0),
        "mean_sweeps": statistics.mean(sweeps_all),
        "sweeps_by_quarter": [statistics.mean(sweeps_all[k * q:(k + 1) * q]) for k in range(4)],
        "sor_reasons": sorted({i["sor_reason"] for i in infos}),
        "late_extrap_beta_mean": statistics.mean(beta),
        "kinetic_energy": st[1],
        "enstrophy": st[4],
        "t_end": total * s.dt,
    }


def lex_over_red_black():
    """Sweeps the reference's lexicographic SOR needs per red-black sweep to
    reach the same dmax < eps: the reference's config-1 run (TGV 64^3, 100
    steps, committed fixture) against the FP64 path on the same 100 steps."""
    import numpy as np
    from osinco3d_amd.config import Config
    from osinco3d_amd.solver import FP64, Solver
    g = np.load(os.path.join(ROOT, "tests", "golden", "c1_tgv64.npz"), allow_pickle=False)
    cfg = Config.from_namelist(bytes(g["params"]).decode())
    s = Solver(cfg, FP64)
    s.init_taylor_green()
    rb = [s.step(it)["sor_exit_iter"] for it in range(1, cfg.itstop + 1)]
    s.close()
    return float(np.mean(g["iters"])) / float(np.mean(rb))


def cpu_baseline(n, kmax, iters_per_step, ratio):
    """Reference Fortran, one core, one full n^3 step (kmax = 5000 by
    default: the SOR runs to eps), wall-clocked; the older model (a capped
    step extrapolated with the measured lexicographic/red-black ratio) is
    returned beside it as a cross-check."""
    from oracle import ref_run_path
    from osinco3d_amd.config import tgv_config

    exe = ref_run_path()
    if exe is None:
        return None
    cfg = tgv_config(n, 1, 3)
    cfg.kmax = kmax
    with tempfile.TemporaryDirectory() as d:
        cfg.write(os.path.join(d, "parameters.o3d"))

        def unlimited_stack():
            resource.setrlimit(resource.RLIMIT_STACK,
                               (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
        cmd = [exe, "-"]
        if os.path.exists("/usr/bin/taskset"):
            cmd = ["/usr/bin/taskset", "-c", "0"] + cmd
        t0 = time.perf_counter()
        r = subprocess.run(cmd, cwd=d, capture_output=True, text=True,
                           preexec_fn=unlimited_stack, timeout=600)
        wall = time.perf_counter() - t0
    if r.returncode != 0:
        return {"error": f"reference exited with {r.returncode}"}
    ph = None
    iters = None
    t_meas = None
    reason = None
    for line in r.stdout.splitlines():
        if line.startswith("O3D_PHASE_SECONDS"):
            ph = [float(x) for x in line.split()[2:]]
        if line.startswith("O3D_STEP_SECONDS"):
            t_meas = float(line.split()[2])
        m = re.match(r"\s*Exit iteration\s*:\s*(\d+)", line)
        if m:
            iters = int(m.group(1))
        if "No converged" in line:
            reason = "kmax or stagnation"
    if ph is None or iters is None or t_meas is None:
        return {"error": "could not parse the reference output"}
    predict, pression, correct, div, diag = ph
    sweeps = min(iters, kmax)
    t_iter = (pression - div) / sweeps
    t_other = predict + div + correct + div + diag
    lex_iters = iters_per_step * ratio
    t_model = t_other + lex_iters * t_iter
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    full = iters <= kmax and reason is None
    return {
        "value": 1.0 / t_meas, "unit": "steps/s", "cores": 1, "kind": "reference",
        "sample": (f"reference Fortran (amdflang -O3) on 1 core of '{cpu_model}' "
                   f"({os.cpu_count()} visible): one {'full ' if full else ''}{n}^3 TGV step "
                   f"(step 1, kmax={kmax}: {iters} lexicographic SOR sweeps"
                   f"{' to eps=1e-8' if full else ', capped'}) measured at {t_meas:.2f} s "
                   f"({wall:.1f} s wall for the process); phases predict {predict:.2f} s, "
                   f"correct_pression {pression:.2f} s ({t_iter:.3f} s/sweep), correct "
                   f"{correct:.2f} s, diagnostics {div + diag:.2f} s"),
        "seconds_per_step": t_meas,
        "sweeps": iters,
        "full_step": full,
        "model_cross_check": {
            "seconds_per_step": t_model,
            "rule": (f"{t_other:.2f} s + {lex_iters:.1f} sweeps x {t_iter:.3f} s, where "
                     f"{lex_iters:.1f} = the FP64 GPU path's {iters_per_step:.1f} red-black "
                     f"sweeps per step (timed window) x {ratio:.3f}, the reference's "
                     f"lexicographic/red-black sweep ratio measured on config 1"),
        },
    }


def run_poisson(args):
    """One periodic n^3 pressure solve from zero by the FP64 red-black SOR
    (o3d_poisson, eps 1e-8, omega 2/(1+sin(pi/n))) and by the multigrid solver
    (o3d_poisson_multigrid, npre 5, npost 4, max |r| < 1e-6).  The exact
    solution is the Taylor-Green pressure (cos 2x + cos 2y)(cos 2z + 2)/16 on
    [0, 2pi)^3; its rhs is the 7-point Laplacian in closed form (each cosine
    of wavenumber 2 is an eigenfunction with eigenvalue -(2 - 2 cos 2h)/h^2).
    Fields are device tensors; the errors are max |p - q| after removing the
    mean (the periodic problem fixes p up to a constant).  Each solver runs
    twice and the second run is reported (the first includes one-time set-up)."""
    import math

    import torch
    from osinco3d_amd import capi
    n = args.n
    h = 2 * math.pi / n
    lam = (2 - 2 * math.cos(2 * h)) / h ** 2
    x = torch.arange(n, dtype=torch.float64, device="cuda") * h
    cx = torch.cos(2 * x)
    A = cx[None, None, :] + cx[None, :, None]
    B = cx[:, None, None] + 2.0
    q = A * B / 16
    rhs = ((-lam * A * B - lam * A * cx[:, None, None]) / 16).contiguous()
    q = q - q.mean()
    d = (h, h, h)
    out = {}
    for name in ("sor", "multigrid", "sor", "multigrid"):  # the second pair is timed
        p = torch.zeros((n, n, n), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if name == "sor":
            it, dm, rs, _ = capi.poisson(p, rhs, d, 2 / (1 + math.sin(math.pi / n)), 1e-8, 50000,
                                         0, "0000")
            info = {"sweeps": it, "dmax": dm, "reason": rs}
        else:
            cyc, rmax = capi.poisson_multigrid(p, rhs, d, "0000", npre=5, npost=4, tol=1e-6,
                                               max_cycles=200)
            info = {"cycles": cyc, "max_residual": rmax}
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        err = float((p - p.mean() - q).abs().max())
        info.update({"ms": ms, "max_error": err})
        out[name] = info
        print(f"[bench] poisson {name}: {ms:.1f} ms, {info}", file=sys.stderr, flush=True)
    line = {"metric": f"periodic {n}^3 pressure solve from zero, ms (FP64 SOR vs multigrid)",
            "value": out["multigrid"]["ms"], "unit": "ms", "n_gpus": 1, "steps": 1,
            "warmup": 0, "ms_per_step": out["multigrid"]["ms"], "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
            "data": "synthetic (Taylor-Green pressure, closed-form rhs)",
            "config": {"workload": f"Poisson solve {n}^3 periodic from zero"},
            "speedup_vs_sor": out["sor"]["ms"] / out["multigrid"]["ms"], "solvers": out}
    print(json.dumps(line), flush=True)


def main():
    args = parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}, "
              f"WORLD_SIZE {world}); the path has no multi-GPU mode.", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    from osinco3d_amd.config import mixing_layer_config, tgv_config

    if args.workload == "poisson":
        run_poisson(args)
        return
    init = None
    if args.workload == "jet":
        from osinco3d_amd.config import jet_config
        cfg = jet_config()
        init = lambda s: s.init_jet()  # noqa: E731
        args.no_cpu = True
        jet_model = os.path.join(ROOT, "osinco3d_amd", "warmstart_weights_jet.npy")
        if args.model == DEFAULT_MODEL and os.path.exists(jet_model):
            args.model = jet_model  # fine-tuned on the jet's own solves
        if args.forms == "fp64,mixed,extrap,mg,nn0,nn":
            args.forms = "fp64,mixed,mg" + (",nn" if os.path.exists(args.model) else "")
        workload = ("Planar jet, 512x384x256 (16x12x8 l0), x-Dirichlet inflow (tanh top hat, "
                    "velocity ratio 10, 5% seeded noise) and convective outflow, free-slip y, "
                    "periodic z, Smagorinsky LES cs=0.17, Re=3000, AB3, CFL 0.25, SOR eps=1e-8")
        data = "synthetic (planar-jet profile; SplitMix64 seed 1234 inflow noise)"
    elif args.workload == "mixing":
        from osinco3d_amd.solver import mixing_layer
        cfg = mixing_layer_config()
        init = mixing_layer(cfg)
        args.no_cpu = True
        workload = ("Temporal mixing layer, 384x256x128 (60x40x20), free-slip y, Smagorinsky LES "
                    "cs=0.17, Re=3000, AB2, CFL 0.25, SOR eps=1e-8 (solver 0011)")
        data = "synthetic (tanh profile + SplitMix64 seed 1234 uniform noise, 0.05 U0)"
    else:
        cfg = tgv_config(args.n, 1000, 3)
        workload = (f"Taylor-Green vortex Re=1600, {args.n}^3 periodic, AB3, CFL 0.25, "
                    f"SOR eps=1e-8")
        data = "synthetic (analytic Taylor-Green initial field)"
    copy_gbs = stream_bandwidth()
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    results = {}
    for form in forms:
        print(f"[bench] form {form}: {cfg.nx}x{cfg.ny}x{cfg.nz}, warmup {args.warmup}, "
              f"steps {args.steps}", file=sys.stderr, flush=True)
        r = run_form(form, cfg, args.steps, args.warmup, args.model, init,
                     (args.coarse_cycles, args.smooth_sweeps))
        if r is not None:
            results[form] = r
            print(f"[bench] form {form}: {r['steps_per_s']:.2f} steps/s, "
                  f"{r['poisson_iters']['mean']:.1f} iters/step", file=sys.stderr, flush=True)
    head_form = [f for f in forms if f in results][-1]
    head = results[head_form]
    ref = results.get("fp64")
    if ref is not None:
        for f, r in results.items():
            r["speedup_vs_fp64"] = r["steps_per_s"] / ref["steps_per_s"]
            r["kinetic_energy_rel_err_vs_fp64"] = abs(r["kinetic_energy"] - ref["kinetic_energy"]) / abs(ref["kinetic_energy"])
            r["enstrophy_rel_err_vs_fp64"] = abs(r["enstrophy"] - ref["enstrophy"]) / abs(ref["enstrophy"])
    which = "f32" if head_form != "fp64" else "f64"
    dom = head.get("sor_" + which)
    roofline = None
    if dom:
        kname = ("sor_split_kernel<float,0> + <float,1> (one colour-split red-black "
                 "iteration, f32)" if which == "f32" else
                 "sor_split_kernel<double,0> + <double,1> (one red-black iteration, f64)")
        roofline = {"bound": "hbm", "kernel": kname,
                    "achieved": dom["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": dom["achieved_GBps"] / HBM_PEAK_GBS,
                    "traffic": measured_traffic(which),
                    "algorithmic_bytes": dom["algorithmic_bytes_per_launch"],
                    "avg_launch_us": dom["avg_launch_us"],
                    "timing": dom.get("timing", "in-step event batches"),
                    "in_step_us_per_iteration": dom["in_step_us_per_iteration"],
                    "floor_bytes": dom["floor_bytes_per_launch"],
                    "frac_floor": dom["floor_GBps"] / HBM_PEAK_GBS,
                    "frac_of_stream_probe": dom["achieved_GBps"] / copy_gbs,
                    "stream_probe_GBps": copy_gbs}
    late = None
    late_forms = [f.strip() for f in args.late_forms.split(",") if f.strip()]
    late_total = args.late_total
    if late_total is None:
        late_total = {"tgv": 1000, "mixing": 500}.get(args.workload, 0)
    args.late_total = late_total
    if late_forms and late_total > args.late_window:
        late = {}
        for form in late_forms:
            print(f"[bench] late window {form}: {args.late_total} steps, last "
                  f"{args.late_window} timed", file=sys.stderr, flush=True)
            r = run_series(form, cfg, args.late_total, args.late_window, args.model,
                           (args.coarse_cycles, args.smooth_sweeps), init)
            if r is not None:
                late[form] = r
                print(f"[bench] late window {form}: {r['late_steps_per_s']:.2f} steps/s late, "
                      f"{r['mean_steps_per_s']:.2f} mean, {r['late_sweeps_mean']:.1f} sweeps",
                      file=sys.stderr, flush=True)
        ref64 = late.get("fp64")
        if ref64 is not None:
            for f, r in late.items():
                r["late_speedup_vs_fp64"] = r["late_steps_per_s"] / ref64["late_steps_per_s"]
                r["mean_speedup_vs_fp64"] = r["mean_steps_per_s"] / ref64["mean_steps_per_s"]
                r["kinetic_energy_rel_err_vs_fp64"] = (
                    abs(r["kinetic_energy"] - ref64["kinetic_energy"]) / abs(ref64["kinetic_energy"]))
                r["enstrophy_rel_err_vs_fp64"] = (
                    abs(r["enstrophy"] - ref64["enstrophy"]) / abs(ref64["enstrophy"]))
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline (reference Fortran, 1 core)", file=sys.stderr, flush=True)
        iters = (ref or head)["poisson_iters"]["mean"]
        cpu = cpu_baseline(args.n, args.cpu_kmax, iters, lex_over_red_black())
    metric = {"tgv": METRIC,
              "mixing": "Temporal mixing layer 384x256x128 LES steps/s + Poisson ms & iters",
              "jet": "Planar jet 512x384x256 x-Dirichlet LES steps/s + Poisson ms & iters"
              }[args.workload]
    line = {
        "metric": metric,
        "value": head["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head_form == "fp64" else "f32+f64",
        "data": data,
        "config": {"workload": workload, "grid": [cfg.nx, cfg.ny, cfg.nz],
                   "form": head_form, "parallelism": "none (1 GPU)",
                   "warm_start": {"coarse_cycles": head.get("coarse_cycles", 0),
                                  "smooth_sweeps": head.get("smooth_sweeps", 0)}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "late_window": late,
        "forms": results,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
