"""Test infrastructure (oracle), not product code: a numpy restatement of the
learned preconditioner's network, used by tests/ to check the HIP forward and
backward kernels. Only tests and smoke checks import it.

The network has no counterpart in the reference (its only preconditioner is
HODLR, outside this project). Definition, matching include/lsg_learn.h:
n 3x3 convolutions, NHWC activations, weights [cout][ky][kx][cin] applied at
offsets ((ky-1) d, (kx-1) d) with zero padding, ReLU on every layer but the
last, MSE loss mean((out - target)^2). With quantize=True the forward rounds
inputs, weights and stored hidden activations to BF16 (round to nearest even)
exactly where the device kernels store BF16, and the backward treats the
rounding as the identity (the device backward reads the same BF16 values).
"""
import numpy as np


def bf16(a):
    """Round float32 values to the nearest BF16 value (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32)


class Layout:
    def __init__(self, n_layers, width, dilations):
        self.layers = []
        off = 0
        for l in range(n_layers):
            cin = 3 if l == 0 else width
            cout = 2 if l == n_layers - 1 else width
            w_off = off
            off += cout * 9 * cin
            b_off = off
            off += cout
            self.layers.append(dict(cin=cin, cout=cout, dil=dilations[l], relu=l < n_layers - 1, w_off=w_off,
                                    b_off=b_off))
        self.nparams = off

    def unpack(self, p):
        out = []
        for L in self.layers:
            w = p[L["w_off"]:L["w_off"] + L["cout"] * 9 * L["cin"]].reshape(L["cout"], 3, 3, L["cin"])
            b = p[L["b_off"]:L["b_off"] + L["cout"]]
            out.append((w, b))
        return out


def _shift(x, dy, dx):
    """s[b, y, x, :] = x[b, y + dy, x + dx, :] with zeros outside."""
    B, G, _, C = x.shape
    s = np.zeros_like(x)
    ys, ye = max(0, -dy), min(G, G - dy)
    xs, xe = max(0, -dx), min(G, G - dx)
    if ys < ye and xs < xe:
        s[:, ys:ye, xs:xe, :] = x[:, ys + dy:ye + dy, xs + dx:xe + dx, :]
    return s


def conv(x, w, b, dil):
    """y[.., co] = b[co] + sum_{ky,kx,ci} w[co, ky, kx, ci] x[.., y+(ky-1)d, x+(kx-1)d, ci]"""
    y = np.zeros(x.shape[:3] + (w.shape[0],), dtype=np.float64) + b
    for ky in range(3):
        for kx in range(3):
            y += _shift(x, (ky - 1) * dil, (kx - 1) * dil) @ w[:, ky, kx, :].T
    return y


def forward(layout, p, x, quantize=True):
    """Returns (out, acts): acts[l] is the input of layer l (as stored)."""
    q = bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))
    params = layout.unpack(np.asarray(p, dtype=np.float64))
    a = q(x).astype(np.float64)
    acts = [a]
    out = None
    for L, (w, b) in zip(layout.layers, params):
        wq = q(w.astype(np.float32)).astype(np.float64) if quantize else w
        y = conv(a, wq, b, L["dil"])
        if L["relu"]:
            y = np.maximum(y, 0.0)
            a = q(y.astype(np.float32)).astype(np.float64) if quantize else y
            acts.append(a)
        else:
            out = y
    return out, acts


def loss_grad(layout, p, x, target, quantize=True):
    """MSE loss and its gradient with respect to the parameter vector p."""
    params = layout.unpack(np.asarray(p, dtype=np.float64))
    q = bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))
    out, acts = forward(layout, p, x, quantize)
    n = out.size
    d = out - target
    loss = float(np.sum(d * d) / n)
    g = 2.0 * d / n
    grad = np.zeros(layout.nparams)
    for l in range(len(layout.layers) - 1, -1, -1):
        L = layout.layers[l]
        w, _ = params[l]
        wq = q(w.astype(np.float32)).astype(np.float64) if quantize else w
        if L["relu"]:
            g = g * (acts[l + 1] > 0)
        a = acts[l]
        gw = np.zeros_like(w)
        for ky in range(3):
            for kx in range(3):
                s = _shift(a, (ky - 1) * L["dil"], (kx - 1) * L["dil"])
                gw[:, ky, kx, :] = np.einsum("byxo,byxi->oi", g, s)
        grad[L["w_off"]:L["w_off"] + gw.size] = gw.reshape(-1)
        grad[L["b_off"]:L["b_off"] + L["cout"]] = g.sum(axis=(0, 1, 2))
        if l == 0:
            break
        gin = np.zeros(a.shape)
        for ky in range(3):
            for kx in range(3):
                gin += _shift(g, -(ky - 1) * L["dil"], -(kx - 1) * L["dil"]) @ wq[:, ky, kx, :]
        g = gin
    return loss, grad


def cheb_lagrange(nc, t):
    """Lagrange basis on the nc first-kind Chebyshev nodes -cos((k + 1/2) pi / nc) at points t."""
    tn = -np.cos((np.arange(nc) + 0.5) * np.pi / nc)
    t = np.atleast_1d(np.asarray(t, dtype=float))
    W = np.ones((t.size, nc))
    for a in range(nc):
        for b in range(nc):
            if b != a:
                W[:, a] *= (t - tn[b]) / (tn[a] - tn[b])
    return W
