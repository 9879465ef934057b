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
  extrap2 mixed + quadratic extrapolation p_prev + 2 e_prev - e_prev2
  mg     extrap + the coarse-grid correction of the guess (multigrid V-cycles
         on the 2^3 block averages of its residual) and Gauss-Seidel sweeps
         opening the FP32 correction solve
  nn0    mixed + learned warm start (the trained network of
         osinco3d_amd/warmstart_weights.npy; skipped when it is missing)
  nn     nn0 + the coarse-grid correction and Gauss-Seidel sweeps of mg (the
         headline: every warm-start stage)
The JSON line's value is the steps/s of the headline form (the last of
--forms); every form's figures are included under "forms".

roofline: the dominant work of the headline form is the red-black SOR
iteration (in the mixed forms the FP32 colour-split red and black launches,
sor_split_kernel<0> and <1>).  Its algorithmic bytes per iteration are
(read p, read rhs, write p) = 3 values per cell: 24 B/cell in FP64, 12 B/cell
in FP32, times the cell count.  Its average duration is measured with HIP
events on the launch stream over complete batches of iterations inside the
timed steps; `traffic` is the measured HBM bytes of one iteration (both
launches) from the committed rocprofv3 --pmc summary.

cpu_baseline: the reference Fortran (oracle/_ref/o3d_ref_run, built from the
reference sources at -O3, single-threaded) on one host core: a bounded sample
of one 256^3 step with kmax capped, whose phases are timed; the per-step
time is extrapolated with the measured per-sweep SOR cost and the number of
lexicographic sweeps the reference needs (the FP64 form's red-black count
times the lexicographic/red-black ratio measured on config 1).
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
    ap.add_argument("--workload", choices=("tgv", "mixing", "poisson"), default="tgv",
                    help="tgv: the headline Taylor-Green case; mixing: the 384x256x128 LES "
                         "temporal mixing layer (free-slip y, AB2), no CPU baseline; "
                         "poisson: one n^3 pressure solve from zero, FP64 SOR against the "
                         "multigrid solver (multigrid = 1)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-kmax", type=int, default=10)
    ap.add_argument("--model", default=os.path.join(ROOT, "osinco3d_amd", "warmstart_weights.npy"))
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


def run_form(form, cfg, steps, warmup, model_path, init=None, coarse=(0, 0)):
    import torch
    from osinco3d_amd import solver as S

    prec = S.FP64 if form == "fp64" else S.MIXED
    warm = {"nn": S.WARM_LEARNED, "nn0": S.WARM_LEARNED, "extrap": S.WARM_EXTRAPOLATE,
            "mg": S.WARM_EXTRAPOLATE, "extrap2": S.WARM_EXTRAPOLATE2}.get(form, S.WARM_PREVIOUS)
    cc, gs = coarse if form in ("mg", "nn") else (0, 0)
    net = None
    if form in ("nn", "nn0"):
        from osinco3d_amd import warmstart
        net = warmstart.load_or_none(model_path, (cfg.nx, cfg.ny, cfg.nz))
        if net is None:
            print(f"[bench] form nn skipped: no model at {model_path}", file=sys.stderr,
                  flush=True)
            return None
    s = S.Solver(cfg, prec, warm, coarse_cycles=cc, smooth_sweeps=gs)
    if net is not None:
        net.attach(s)
    if init is None:
        s.init_taylor_green()
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
    # launches.  The algorithmic bytes of one iteration are one read of the
    # iterate and the rhs and one write of the iterate (24 B/cell FP64,
    # 12 B/cell FP32).  An FP64 iteration is one sor_kernel launch; an FP32
    # iteration of the mixed forms is the red and the black colour-split
    # launch (sor_split_kernel<0>, <1>), which together move 16 B/cell.
    for q, name, bpc in ((0, "sor_f64", 24), (1, "sor_f32", 12)):
        if sor_ln[q] > 0:
            avg_ms = sor_ms[q] / sor_ln[q]
            gbs = bpc * n / (avg_ms * 1e-3) / 1e9
            out[name] = {"avg_launch_us": 1e3 * avg_ms, "launches_timed": sor_ln[q],
                         "iterations_per_launch": sor_sw[q] / sor_ln[q],
                         "us_per_iteration": 1e3 * sor_ms[q] / sor_sw[q],
                         "algorithmic_bytes_per_launch": bpc * n, "achieved_GBps": gbs}
    if cc or gs:
        out["coarse_cycles"] = cc
        out["smooth_sweeps"] = gs
    if form in ("nn", "nn0", "extrap", "extrap2", "mg"):
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


def copy_bandwidth(nbytes=1 << 30, reps=20):
    """Achievable HBM bandwidth: a 1 GiB device-to-device copy (read + write
    counted), timed with events on the current stream."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(1.0)
    for _ in range(3):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2 * nbytes / (ms * 1e-3) / 1e9


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
    """Reference Fortran, one core, one n^3 step with kmax capped."""
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
    for line in r.stdout.splitlines():
        if line.startswith("O3D_PHASE_SECONDS"):
            ph = [float(x) for x in line.split()[2:]]
        m = re.match(r"\s*Exit iteration\s*:\s*(\d+)", line)
        if m:
            iters = int(m.group(1))
    if ph is None or iters is None:
        return {"error": "could not parse the reference output"}
    predict, pression, correct, div, diag = ph
    sweeps = min(iters, kmax)
    t_iter = (pression - div) / sweeps
    t_other = predict + div + correct + div + diag
    lex_iters = iters_per_step * ratio
    t_step = t_other + lex_iters * t_iter
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": 1.0 / t_step, "unit": "steps/s", "cores": 1, "kind": "reference",
        "sample": (f"reference Fortran (amdflang -O3) on 1 core of '{cpu_model}' "
                   f"({os.cpu_count()} visible): one {n}^3 TGV step with kmax={kmax} "
                   f"({wall:.1f} s wall); phases predict {predict:.2f} s, correct_pression "
                   f"{pression:.2f} s ({sweeps} sweeps, {t_iter:.3f} s/sweep), correct "
                   f"{correct:.2f} s, diagnostics {div + diag:.2f} s; step time = "
                   f"{t_other:.2f} s + {lex_iters:.1f} sweeps x {t_iter:.3f} s, where "
                   f"{lex_iters:.1f} = the FP64 GPU path's {iters_per_step:.1f} red-black "
                   f"sweeps x {ratio:.3f}, the reference's lexicographic/red-black sweep "
                   f"ratio measured on config 1 (TGV 64^3, 100 steps)"),
        "seconds_per_step": t_step,
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
    if args.workload == "mixing":
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
    copy_gbs = copy_bandwidth()
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
        kname = ("sor_split_kernel<0> + <1> (one colour-split red-black iteration, f32)"
                 if which == "f32" else "sor_kernel (one red-black iteration, f64)")
        roofline = {"bound": "hbm", "kernel": kname,
                    "achieved": dom["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": dom["achieved_GBps"] / HBM_PEAK_GBS,
                    "traffic": measured_traffic(which),
                    "algorithmic_bytes": dom["algorithmic_bytes_per_launch"],
                    "frac_of_measured_copy": dom["achieved_GBps"] / copy_gbs,
                    "measured_copy_GBps": copy_gbs}
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline (reference Fortran, 1 core)", file=sys.stderr, flush=True)
        iters = (ref or head)["poisson_iters"]["mean"]
        cpu = cpu_baseline(args.n, args.cpu_kmax, iters, lex_over_red_black())
    line = {
        "metric": METRIC,
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
        "forms": results,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
