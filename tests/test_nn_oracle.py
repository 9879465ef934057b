"""Pinning the network oracle (CPU): its analytic gradients against central
finite differences of its own forward pass (rounding off, float64), and its
Adam step against the closed form of one step from zero moments."""
from __future__ import annotations

import numpy as np

from oracle import nn_oracle as N


def _params(seed=0):
    L, head, n = N.layout()
    rng = np.random.default_rng(seed)
    P = np.zeros(n)
    for ci, co, _, wo, bo in L:
        P[wo:wo + co * 27 * ci] = rng.uniform(-1, 1, co * 27 * ci) * np.sqrt(6 / (27 * ci)) * 0.8
        P[bo:bo + co] = rng.standard_normal(co) * 0.05
    P[head:head + 17] = rng.standard_normal(17) * 0.3
    return L, head, P


def test_layout_size():
    L, head, n = N.layout()
    assert n == head + 17 and len(L) == 14
    assert n == 1320913


def test_gradients_match_finite_differences():
    L, head, P = _params()
    rng = np.random.default_rng(1)
    inp = rng.standard_normal((2, 8, 8, 8))
    net = N.Net(P, rounding=False)
    out = net.forward(inp)
    dout = rng.standard_normal(out.shape)
    G = net.backward(dout)
    picks = [L[0][3] + 1, L[3][3] + 50, L[6][4] + 7, L[7][3] + 100, L[9][3] + 77,
             L[12][3] + 999, L[13][4] + 3, head + 3, head + 16]
    for idx in picks:
        e = 1e-5
        P2 = P.copy()
        P2[idx] += e
        f1 = (N.Net(P2, False).forward(inp) * dout).sum()
        P2[idx] -= 2 * e
        f2 = (N.Net(P2, False).forward(inp) * dout).sum()
        fd = (f1 - f2) / (2 * e)
        assert abs(G[idx] - fd) <= 1e-6 * max(1.0, abs(fd)), (idx, G[idx], fd)


def test_conv_adjoint_identity():
    """<conv(x), g> == <x, conv_t(g)> for random tensors (periodic)."""
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 6, 8, 16))
    g = rng.standard_normal((4, 6, 8, 32))
    W = rng.standard_normal(32 * 27 * 16)
    lhs = (N.conv(x, W, 16, 32) * g).sum()
    rhs = (x * N.conv_t(g, W, 16, 32)).sum()
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159, 65504.0])
    r = N.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0            # ties to even
    assert r[2] == 1.0 + 2 ** -7                  # rounds up
    assert abs(r[3] + 3.140625) < 1e-12


def test_adam_first_step():
    w, m, v = N.adam(np.array([1.0, -2.0]), np.array([0.5, -4.0]), np.zeros(2), np.zeros(2),
                     1e-2, 0.9, 0.999, 1e-8, 1)
    # First step from zero moments moves each weight by lr * sign(g).
    assert np.allclose(w, [1.0 - 1e-2, -2.0 + 1e-2], atol=1e-9)
