"""The warm-start network on the GPU against its oracle, training, and the
warm start inside the solver."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import nn_oracle as N

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def _params(seed=0):
    L, head, n = N.layout()
    rng = np.random.default_rng(seed)
    P = np.zeros(n)
    for li, (ci, co, _, wo, bo) in enumerate(L):
        w = rng.uniform(-1, 1, (co, 27, ci)) * np.sqrt(6 / (27 * ci)) * 0.8
        if li == 0:
            w[:, :, 2:] = 0.0
        P[wo:wo + co * 27 * ci] = w.reshape(-1)
        P[bo:bo + co] = rng.standard_normal(co) * 0.05
    P[head:head + 17] = rng.standard_normal(17) * 0.3
    return P.astype(np.float32)


def _net(torch, shape, P):
    from osinco3d_amd.warmstart import Network
    nz, ny, nx = shape
    net = Network(nx, ny, nz)
    net.set_params(P)
    return net


# (32, 64, 64): its 64x64x32 and 32x32x16 levels run the LDS-staged
# convolution (four- and one-row bricks, the decoder's up-sampling fused
# into it), the coarser ones the K split among a block's waves; (64, 64,
# 64), the network of the 256^3 solver: the LDS-staged kernel at 64^3 and
# 32^3, the K split at 16^3 and 8^3.
@pytest.mark.parametrize("shape", [(16, 16, 16), (8, 16, 32), (32, 64, 64), (64, 64, 64)])
def test_forward_matches_oracle(torch, shape):
    print(f"[gpu] nn forward {shape}", flush=True)
    P = _params()
    rng = np.random.default_rng(3)
    inp = rng.standard_normal((2,) + shape).astype(np.float32)
    net = _net(torch, shape, P)
    x = torch.from_numpy(inp).cuda().reshape(2, -1).contiguous()
    out = torch.empty(x.shape[1], device="cuda")
    net.forward(x, out)
    got = out.cpu().numpy().reshape(shape)
    exp = N.Net(P.astype(np.float64)).forward(inp.astype(np.float64))
    err = np.abs(got - exp).max() / np.abs(exp).max()
    print(f"[gpu] forward rel err {err:.2e}", flush=True)
    assert err < 2e-2
    assert np.corrcoef(got.ravel(), exp.ravel())[0, 1] > 0.9995


# The gradients of this randomly initialised, BF16-rounded network are very
# sensitive: a 1e-3 relative perturbation of the input changes the oracle's
# own weight gradients by up to 21 % (16^3) and 14 % (32x64x64), through the
# BF16 roundings and leaky-ReLU slopes it flips.  At 16^3 the GPU's FP32
# sums round like the oracle's almost everywhere (forward 7e-4, gradients
# within 8e-3); on the larger grid, whose fine levels run the LDS-staged
# kernel forward and backward, more roundings differ (forward 1.7e-3).
# There the oracle's own gradients, under 3e-4 and 1e-3 relative input
# noise, move by up to 16.5 % with cosines down to 0.9934 (layers 2, 4, 5),
# so the check is set at that floor, the one a wrong tap, halo or channel
# would fail by far: every layer's gradient within cos 0.99 and 20 % of
# the oracle's.
BWD_TOL = {(16, 16, 16): (5e-2, 0.999), (32, 64, 64): (0.2, 0.99)}


@pytest.mark.parametrize("shape", list(BWD_TOL))
def test_backward_matches_oracle(torch, shape):
    print(f"[gpu] nn backward {shape}", flush=True)
    P = _params(1)
    rng = np.random.default_rng(4)
    inp = rng.standard_normal((2,) + shape).astype(np.float32)
    dout = rng.standard_normal(shape).astype(np.float32)
    net = _net(torch, shape, P)
    x = torch.from_numpy(inp).cuda().reshape(2, -1).contiguous()
    out = torch.empty(x.shape[1], device="cuda")
    net.zero_grad()
    net.forward(x, out)
    net.backward(torch.from_numpy(dout).cuda().reshape(-1).contiguous())
    g = net.grads().astype(np.float64)
    on = N.Net(P.astype(np.float64))
    on.forward(inp.astype(np.float64))
    go = on.backward(dout.astype(np.float64))
    L, head, n = N.layout()
    bad = []
    for li, (ci, co, _, wo, bo) in enumerate(L):
        a, b = g[wo:wo + co * 27 * ci], go[wo:wo + co * 27 * ci]
        if li == 0:
            keep = np.zeros((co, 27, ci), bool)
            keep[:, :, :2] = True
            a, b = a[keep.ravel()], b[keep.ravel()]
        rel = np.abs(a - b).max() / np.abs(b).max()
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        bb = np.abs(g[bo:bo + co] - go[bo:bo + co]).max()
        print(f"[gpu] layer {li} dW rel {rel:.2e} cos {cos:.6f} bias err {bb:.2e} of "
              f"{np.abs(go[bo:bo + co]).max():.2e}", flush=True)
        tol_rel, tol_cos = BWD_TOL[shape]
        if not (rel < tol_rel and cos > tol_cos and
                bb <= tol_rel * np.abs(go[bo:bo + co]).max() + 1e-3):
            bad.append(li)
    assert not bad, bad
    assert np.allclose(g[head:], go[head:], rtol=2e-2, atol=1e-3 * np.abs(go[head:]).max())


def test_adam_matches_closed_form(torch):
    shape = (8, 8, 16)
    P = _params(2)
    net = _net(torch, shape, P)
    rng = np.random.default_rng(5)
    x = torch.from_numpy(rng.standard_normal((2,) + shape).astype(np.float32)).cuda().reshape(2, -1)
    tgt = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).cuda().reshape(-1)
    out = torch.empty_like(tgt)
    dout = torch.empty_like(tgt)
    net.zero_grad()
    net.forward(x.contiguous(), out)
    net.mse(out, tgt, dout)
    net.backward(dout)
    g = net.grads().astype(np.float64)
    net.adam(1e-3)
    w1 = net.params().astype(np.float64)
    exp, _, _ = N.adam(P.astype(np.float64), g, np.zeros_like(g), np.zeros_like(g), 1e-3, 0.9,
                       0.999, 1e-8, 1)
    assert np.abs(w1 - exp).max() < 1e-6


def test_training_reduces_loss_and_warm_start_runs(torch):
    """Train on a short captured TGV run at 64^3 (the network runs on the
    16^3 coarse grid), then use the network as the solver's warm start: every
    step still converges to eps, and the fields match the run without it to a
    tolerance tied to eps."""
    from osinco3d_amd import warmstart as W
    from osinco3d_amd.config import tgv_config
    from osinco3d_amd.solver import FP64, MIXED, Solver, taylor_green
    cfg = tgv_config(64, 40, 3)
    samples, iters = W.capture_run(cfg, 30, MIXED)
    assert len(samples) >= 20
    assert tuple(samples[0][0].shape) == (2, 16 ** 3)
    net = W.Network(16, 16, 16, seed=7)
    before = W.relative_errors(net, samples)
    hist = W.train(net, samples, epochs=4, lr=3e-4, batch=4)
    after = W.relative_errors(net, samples)
    print(f"[gpu] nn train mse {hist}, rel err before {before} after {after}", flush=True)
    assert hist[-1] < hist[0]
    assert after["nn"] < before["nn"]
    a = Solver(cfg, MIXED)
    b = Solver(cfg, MIXED)
    a.upload(*taylor_green(cfg))
    b.upload(*taylor_green(cfg))
    net.attach(b)
    used = 0
    for it in range(1, 21):
        ia = a.step(it)
        ib = b.step(it)
        assert ib["sor_reason"] == 0 and ia["sor_reason"] == 0
        used += ib["warm_used"]
        if it > 3:
            assert ib["warm_dmax"][0] > 0
    print(f"[gpu] warm start used on {used} of 20 steps", flush=True)
    ga, gb = a.download(), b.download()
    for c in range(3):
        assert np.abs(ga[c] - gb[c]).max() < 1e-6


def test_warm_features_are_block_averages(torch):
    """o3d_warm_features: the network's input and target are block averages
    over F^3 cells (F = WARM_COARSEN) of r dx^2 F^2 / RMS(e_prev),
    e_prev / RMS(e_prev) and e / RMS(e_prev), with the RMS over the fine grid
    in FP64."""
    from osinco3d_amd import capi
    from osinco3d_amd.warmstart import normalize
    rng = np.random.default_rng(5)
    sh = (16, 24, 32)                      # (nz, ny, nx)
    r, ep, e = [rng.standard_normal(sh).astype(np.float32) for _ in range(3)]
    dx = 0.1
    inp, tgt, scale = normalize(*[torch.from_numpy(a).cuda() for a in (r, ep, e)], dx)
    F = capi.WARM_COARSEN

    def pool(a):
        a = a.astype(np.float64)
        return a.reshape(sh[0] // F, F, sh[1] // F, F, sh[2] // F, F).mean(axis=(1, 3, 5))

    re = np.sqrt(np.mean(ep.astype(np.float64) ** 2))
    assert abs(scale - re) <= 1e-12 * re
    got = inp.cpu().numpy()
    np.testing.assert_allclose(got[0], (pool(r) * dx * dx * F * F / re).reshape(-1), rtol=1e-5,
                               atol=1e-6)
    np.testing.assert_allclose(got[1], (pool(ep) / re).reshape(-1), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(tgt.cpu().numpy(), (pool(e) / re).reshape(-1), rtol=1e-5,
                               atol=1e-6)
    z = torch.zeros(sh, dtype=torch.float32, device="cuda")
    assert normalize(z, z, z, dx) is None
