"""Train the warm-start network of the Poisson solve on the solver's own runs.

    python -m osinco3d_amd.train_warmstart [--n 128] [--samples 2000]
        [--epochs 12] [--out osinco3d_amd/warmstart_weights.npy] [--eval-n 256]
    python -m osinco3d_amd.train_warmstart --config jet [--grid 512,384,256]
        [--init weights.npy] [--out osinco3d_amd/warmstart_weights_jet.npy]

Training pairs (inputs: the residual of the extrapolated pressure and e_prev;
target: the increment e; see warmstart.py) come from seeded runs at n^3
(SOR eps 1e-8, mixed precision form, CFL 0.25): Taylor-Green (periodic, AB3)
at Re 1600 over the vortex breakdown, at Re 400, and at Re 1600 with the
axes permuted so the network does not learn one orientation; and two
temporal mixing layers (20^3 delta box, free-slip y, LES, AB2, Re 3000,
the C3 initial field with SplitMix64 seeds 1234 and 4321), 2000 pairs in
all by default.  Every
sample stays resident in device memory: the network's input and target, 3
float32 planes on the grid coarsened by capi.WARM_COARSEN (384 KiB per
sample at 128^3).  Every tenth step of each run is held out to report the
relative error of the predicted increment against the linear extrapolation
e = e_prev, and the saved weights are those of the epoch with the smallest
held-out error.

Evaluation: the same solver steps with the previous-pressure, the
extrapolated and the learned warm starts, reporting SOR sweeps per step; for
the jet, a run with an inflow seed absent from the training runs.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import capi, warmstart
from .config import jet_config, mixing_layer_config, tgv_config
from .solver import (MIXED, WARM_EXTRAPOLATE, WARM_LEARNED, WARM_PREVIOUS, Solver, mixing_layer,
                     taylor_green)

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_weights.npy")


def log(msg: str):
    print(f"[train] {msg}", flush=True)


def permuted(fields):
    """Taylor-Green field with the axes permuted cyclically: v_x(x,y,z) =
    u_z(y,z,x), v_y = u_x(y,z,x), v_z = u_y(y,z,x) (still divergence free);
    arrays are (nz, ny, nx)."""
    ux, uy, uz, pp = fields
    t = lambda a: np.ascontiguousarray(np.transpose(a, (1, 2, 0)))
    return t(uz), t(ux), t(uy), t(pp)


def runs(n: int, total: int):
    """(label, config, initial fields, steps) with about `total` steps."""
    specs = [("tgv Re1600", 1600.0, False, 0.35), ("tgv Re400", 400.0, False, 0.15),
             ("tgv Re1600 permuted", 1600.0, True, 0.20)]
    out = []
    for label, re, perm, frac in specs:
        steps = max(8, int(round(total * frac)))
        cfg = tgv_config(n, steps, 3, re=re)
        init = taylor_green(cfg)
        if perm:
            init = permuted(init)
        out.append((label, cfg, init, steps))
    for seed in (1234, 4321):
        steps = max(8, int(round(total * 0.15)))
        cfg = mixing_layer_config(n, n, n, steps=steps, lengths=(20.0, 20.0, 20.0))
        out.append((f"mixing layer seed {seed}", cfg, mixing_layer(cfg, seed=seed), steps))
    return out


JET_GRID = (128, 96, 64)


def jet_runs(total: int, grid=JET_GRID):
    """The planar jet with x-Dirichlet inflow on `grid` (by default the
    512x384x256 box at a quarter of its resolution), seeds 1234 and 4321 of
    the inflow noise: the fine-tuning data of the jet's network."""
    out = []
    for seed in (1234, 4321):
        steps = max(8, total // 2)
        cfg = jet_config(*grid, steps=steps)
        out.append((f"jet {grid} seed {seed}", cfg,
                    (lambda s, seed=seed: s.init_jet(seed=seed)), steps))
    return out


def collect(n: int, total: int, config: str = "tgv+mixing", grid=JET_GRID):
    train, held = [], []
    for label, cfg, init, steps in (jet_runs(total, grid) if config == "jet" else runs(n, total)):
        t0 = time.perf_counter()
        reasons = []
        samples, iters = warmstart.capture_run(cfg, steps + 2, MIXED, init=init, reasons=reasons)
        train += [x for i, x in enumerate(samples) if i % 10 != 9]
        held += samples[9::10]
        log(f"{label}: {len(samples)} samples of {len(iters)} steps in "
            f"{time.perf_counter() - t0:.1f} s, SOR sweeps/step mean {np.mean(iters):.1f}, "
            f"stop reasons eps/stagnation/kmax {[reasons.count(k) for k in (0, 1, 2)]}")
    return train, held


def sweeps_with(cfg, init, steps, warm, net=None, skip=3, coarse=(0, 0)):
    s = Solver(cfg, MIXED, warm, coarse_cycles=coarse[0], smooth_sweeps=coarse[1])
    if net is not None:
        net.attach(s)
    if callable(init):
        init(s)
    else:
        s.upload(*init)
    its, used = [], []
    for it in range(1, steps + 1):
        info = s.step(it)
        if it > skip:
            its.append(info["sor_exit_iter"])
            used.append(info["warm_used"])
    s.close()
    return float(np.mean(its)), float(np.mean(used))


def evaluate(net, n: int, steps: int, re: float = 1600.0):
    cfg = tgv_config(n, steps, 3, re=re)
    init = taylor_green(cfg)
    res = {}
    for name, warm, nn in (("previous", WARM_PREVIOUS, None),
                           ("extrapolation", WARM_EXTRAPOLATE, None),
                           ("learned", WARM_LEARNED, net)):
        res[name] = sweeps_with(cfg, init, steps, warm, nn)
        log(f"eval {n}^3 Re{re:g}: {name}: {res[name][0]:.1f} sweeps/step "
            f"(guess accepted {res[name][1]:.2f})")
    return res


def evaluate_jet(net, grid, steps: int, seed: int = 777):
    """The jet with an inflow seed absent from the training runs: relative
    error of the predicted increment on its solves, then SOR sweeps per step
    with each warm start, alone and under the coarse-grid correction."""
    cfg = jet_config(*grid, steps=steps)
    init = lambda s: s.init_jet(seed=seed)  # noqa: E731
    samples, _ = warmstart.capture_run(cfg, steps, MIXED, init=init)
    err = warmstart.relative_errors(net, samples)
    log(f"eval jet {grid} unseen seed {seed}: {len(samples)} solves, relative error of the "
        f"increment: network {err['nn']:.4f}, extrapolation {err['extrapolation']:.4f}")
    res = {"relative_error": err}
    for name, warm, nn, coarse in (("extrapolation", WARM_EXTRAPOLATE, None, (0, 0)),
                                   ("learned", WARM_LEARNED, net, (0, 0)),
                                   ("extrapolation + coarse-grid", WARM_EXTRAPOLATE, None, (2, 8)),
                                   ("learned + coarse-grid", WARM_LEARNED, net, (2, 8))):
        res[name] = sweeps_with(cfg, init, steps, warm, nn, coarse=coarse)
        log(f"eval jet {grid} unseen seed {seed}: {name}: {res[name][0]:.1f} sweeps/step "
            f"(guess accepted {res[name][1]:.2f})")
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--lr", type=float, default=1e-4)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default="", help="start from these weights (.npy)")
    ap.add_argument("--out", default="",
                    help="default osinco3d_amd/warmstart_weights.npy (_jet.npy for --config jet)")
    ap.add_argument("--eval-n", type=int, default=256)
    ap.add_argument("--eval-steps", type=int, default=16)
    ap.add_argument("--config", choices=("tgv+mixing", "jet"), default="tgv+mixing",
                    help="training data: Taylor-Green and mixing-layer runs at n^3 (C4), or "
                         "the x-Dirichlet planar jet at 128x96x64 (fine-tuning for C5, "
                         "usually with --init)")
    ap.add_argument("--grid", default="",
                    help="jet grid NX,NY,NZ (default 128,96,64)")
    a = ap.parse_args(argv)
    jgrid = tuple(int(v) for v in a.grid.split(",")) if a.grid else JET_GRID
    if not a.out:
        a.out = DEFAULT_OUT.replace(".npy", "_jet.npy") if a.config == "jet" else DEFAULT_OUT

    import torch
    if not torch.cuda.is_available():
        sys.exit("train_warmstart needs a HIP device")
    t_start = time.perf_counter()
    train, held = collect(a.n, a.samples, a.config, jgrid)
    grid = jgrid if a.config == "jet" else (a.n, a.n, a.n)
    log(f"{len(train)} training and {len(held)} held-out samples on {grid} "
        f"({time.perf_counter() - t_start:.1f} s to capture)")
    F = capi.WARM_COARSEN
    net = warmstart.Network(grid[0] // F, grid[1] // F, grid[2] // F, seed=a.seed)
    if a.init:
        net.load(a.init)
    log(f"network: {net.nparam} parameters")
    before = warmstart.relative_errors(net, held)
    log(f"held-out relative error before: {before}")
    # Model selection on the held-out error: the saved weights are those of
    # the best epoch (the initial, extrapolation-like network included).
    best_err, best_params, best_ep = before["nn"], net.params(), -1
    t0 = time.perf_counter()
    for ep in range(a.epochs):
        lr = a.lr if ep < a.epochs - 1 else a.lr * 0.3
        hist = warmstart.train(net, train, epochs=1, lr=lr, batch=a.batch, seed=a.seed + ep,
                               log=log, log_every=50)
        ms, fl = net.last_timing()
        err = warmstart.relative_errors(net, held)
        if np.isfinite(err["nn"]) and err["nn"] < best_err:
            best_err, best_params, best_ep = err["nn"], net.params(), ep
        log(f"epoch {ep}: lr {lr:.1e} train mse {hist[-1]:.4e}, held-out {err}, "
            f"forward {ms:.2f} ms = {fl / (ms * 1e-3) / 1e12:.1f} TFLOP/s, "
            f"elapsed {time.perf_counter() - t0:.0f} s")
    net.set_params(best_params)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.save(a.out)
    log(f"saved {a.out} (epoch {best_ep}, held-out relative error {best_err:.4f}, "
        f"extrapolation {before['extrapolation']:.4f})")
    del train, held
    torch.cuda.empty_cache()
    log(f"training wall time {time.perf_counter() - t0:.1f} s")
    if a.config == "jet":
        if a.eval_steps > 0:
            evaluate_jet(net, grid, a.eval_steps)
    elif a.eval_n > 0:
        evaluate(net, a.eval_n, a.eval_steps)
    net.close()


if __name__ == "__main__":
    main()
