"""Test infrastructure only: numpy restatement of the warm-start network.

Same architecture, parameter layout and rounding points as the HIP
implementation (osinco3d_amd/csrc/o3d_nn.hip): BF16 activations and weights
(round to nearest even), sums in float64 here (FP32 on the GPU), periodic
padding, leaky-ReLU slope 0.1, average pooling, nearest up-sampling,
additive skips before the activation, FP32 1x1x1 head added to the second
input plane (the network predicts a correction to it).  It checks the GPU's
forward pass, its backward pass (activation and weight gradients) and Adam.

The network is the build's own (the reference has no learned component), so
this file restates the build's design, and the tests pin it independently:
finite differences of its own forward pass check its gradients.
"""
from __future__ import annotations

import numpy as np

SLOPE = 0.1
SPEC = [(16, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2),
        (64, 128, 3), (128, 128, 3), (128, 64, 2), (64, 64, 2), (64, 32, 1), (32, 32, 1),
        (32, 16, 0), (16, 16, 0)]


def layout():
    """Offsets of W[co][27][ci] and b[co] per layer, then the head (16 + 1)."""
    out = []
    off = 0
    for ci, co, lvl in SPEC:
        w = off
        off += co * 27 * ci
        b = off
        off += co
        out.append((ci, co, lvl, w, b))
    return out, off, off + 17


def bf16(a):
    a32 = np.asarray(a, dtype=np.float32)
    bits = a32.view(np.uint32).astype(np.uint64)
    bits = (bits + 0x7FFF + ((bits >> 16) & 1)) & 0xFFFF0000
    return bits.astype(np.uint32).view(np.float32).astype(np.float64)


def taps():
    for t in range(27):
        yield t, t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1


def shift(x, dx, dy, dz):
    """x[v + off] for x of shape (Z, Y, X, C), periodic."""
    return np.roll(x, shift=(-dz, -dy, -dx), axis=(0, 1, 2))


def conv(x, W, ci, co):
    """y[v][o] = sum_t sum_c x[v + off_t][c] W[o][t][c]; W flat (co, 27, ci)."""
    Wr = W.reshape(co, 27, ci)
    y = np.zeros(x.shape[:3] + (co,))
    for t, dx, dy, dz in taps():
        y += shift(x, dx, dy, dz) @ Wr[:, t, :].T
    return y


def conv_t(g, W, ci, co):
    """Adjoint of conv: gx[u][c] = sum_t sum_o g[u - off_t][o] W[o][t][c]."""
    Wr = W.reshape(co, 27, ci)
    gx = np.zeros(g.shape[:3] + (ci,))
    for t, dx, dy, dz in taps():
        gx += shift(g, -dx, -dy, -dz) @ Wr[:, t, :]
    return gx


def wgrad(dz, x, ci, co):
    g = np.zeros((co, 27, ci))
    for t, dx, dy, dz_ in taps():
        xs = shift(x, dx, dy, dz_)
        g[:, t, :] = np.einsum("zyxo,zyxc->oc", dz, xs)
    return g.reshape(-1)


def act(v):
    return np.where(v > 0, v, SLOPE * v)


def dact(y):
    return np.where(y > 0, 1.0, SLOPE)


def pool(x):
    Z, Y, X, C = x.shape
    return x.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).mean(axis=(1, 3, 5))


def up(x):
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def pool_t(g):
    return up(g) / 8.0


def up_t(g):
    Z, Y, X, C = g.shape
    return g.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).sum(axis=(1, 3, 5))


class Net:
    def __init__(self, params, rounding=True):
        self.P = np.asarray(params, dtype=np.float64)
        self.L, self.head, self.n = layout()
        self.r = bf16 if rounding else (lambda x: np.asarray(x, dtype=np.float64))

    def w(self, i):
        ci, co, _, wo, bo = self.L[i]
        return self.r(self.P[wo:wo + co * 27 * ci]), self.P[bo:bo + co]

    def layer(self, i, x, skip=None):
        ci, co, _, _, _ = self.L[i]
        W, b = self.w(i)
        v = conv(x, W, ci, co) + b
        if skip is not None:
            v = v + skip
        return self.r(act(v))

    def forward(self, inp):
        """inp: (2, Z, Y, X) float planes -> out (Z, Y, X)."""
        Z, Y, X = inp.shape[1:]
        x = np.zeros((Z, Y, X, 16))
        x[..., 0] = self.r(inp[0])
        x[..., 1] = self.r(inp[1])
        a = {"xin": x}
        a["e0a"] = self.layer(0, x)
        a["e0"] = self.layer(1, a["e0a"])
        a["p1"] = self.r(pool(a["e0"]))
        a["e1a"] = self.layer(2, a["p1"])
        a["e1"] = self.layer(3, a["e1a"])
        a["p2"] = self.r(pool(a["e1"]))
        a["e2a"] = self.layer(4, a["p2"])
        a["e2"] = self.layer(5, a["e2a"])
        a["p3"] = self.r(pool(a["e2"]))
        a["e3a"] = self.layer(6, a["p3"])
        a["e3"] = self.layer(7, a["e3a"])
        a["u2"] = up(a["e3"])
        a["d2a"] = self.layer(8, a["u2"], a["e2"])
        a["d2"] = self.layer(9, a["d2a"])
        a["u1"] = up(a["d2"])
        a["d1a"] = self.layer(10, a["u1"], a["e1"])
        a["d1"] = self.layer(11, a["d1a"])
        a["u0"] = up(a["d1"])
        a["d0a"] = self.layer(12, a["u0"], a["e0"])
        a["d0"] = self.layer(13, a["d0a"])
        wh = self.P[self.head:self.head + 16]
        bh = self.P[self.head + 16]
        self.a = a
        return a["d0"] @ wh + bh + inp[1]

    def backward(self, dout):
        """Gradient of sum(dout * out) w.r.t. the flat parameters."""
        a = self.a
        G = np.zeros(self.n)
        wh = self.P[self.head:self.head + 16]
        G[self.head:self.head + 16] = np.einsum("zyx,zyxc->c", dout, a["d0"])
        G[self.head + 16] = dout.sum()
        g = {k: np.zeros_like(v) for k, v in a.items()}
        g["d0"] += dout[..., None] * wh

        def lb(i, gy, y, x, gx_name=None, skip_name=None):
            ci, co, _, wo, bo = self.L[i]
            d32 = gy * dact(y)
            d = self.r(d32)
            if skip_name:
                g[skip_name] += d32
            G[bo:bo + co] += d32.sum(axis=(0, 1, 2))
            G[wo:wo + co * 27 * ci] += wgrad(d, x, ci, co)
            if gx_name:
                W, _ = self.w(i)
                g[gx_name] += conv_t(d, W, ci, co)

        lb(13, g["d0"], a["d0"], a["d0a"], "d0a")
        lb(12, g["d0a"], a["d0a"], a["u0"], "u0", "e0")
        g["d1"] += up_t(g["u0"])
        lb(11, g["d1"], a["d1"], a["d1a"], "d1a")
        lb(10, g["d1a"], a["d1a"], a["u1"], "u1", "e1")
        g["d2"] += up_t(g["u1"])
        lb(9, g["d2"], a["d2"], a["d2a"], "d2a")
        lb(8, g["d2a"], a["d2a"], a["u2"], "u2", "e2")
        g["e3"] += up_t(g["u2"])
        lb(7, g["e3"], a["e3"], a["e3a"], "e3a")
        lb(6, g["e3a"], a["e3a"], a["p3"], "p3")
        g["e2"] += pool_t(g["p3"])
        lb(5, g["e2"], a["e2"], a["e2a"], "e2a")
        lb(4, g["e2a"], a["e2a"], a["p2"], "p2")
        g["e1"] += pool_t(g["p2"])
        lb(3, g["e1"], a["e1"], a["e1a"], "e1a")
        lb(2, g["e1a"], a["e1a"], a["p1"], "p1")
        g["e0"] += pool_t(g["p1"])
        lb(1, g["e0"], a["e0"], a["e0a"], "e0a")
        lb(0, g["e0a"], a["e0a"], a["xin"])
        return G


def adam(w, g, m, v, lr, b1, b2, eps, t):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v
