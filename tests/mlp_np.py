"""FP64 CPU reference of the surrogate MLP (test infrastructure).

Same architecture, parameter layout and loss as gft_mlp_* (include/gft.h):
hidden layers z = h W^T + b, h = GELU(z) (exact, erf-based); heads
y = h W_h^T + b_h; loss = sum w (y - t)^2 / (n_out * batch * heads).

With bf16=True the forward and backward passes round to BF16 at the same
points as the GPU kernels (inputs, weights, activations, dY, dZ), with FP64
arithmetic everywhere else, so GPU and CPU differ only by the GPU's FP32
accumulation order; with bf16=False it is the plain FP64 network, whose
analytic gradient is checked against central differences.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def bf16(a):
    """Round to BF16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def layout(n_in, width, n_hidden, n_heads, n_out):
    out = []
    for l in range(n_hidden):
        out += [(width, n_in if l == 0 else width), (width,)]
    out += [(n_heads * n_out, width), (n_heads * n_out,)]
    return out


def unflatten(p, shapes):
    out, o = [], 0
    for sh in shapes:
        n = int(np.prod(sh))
        out.append(np.asarray(p[o:o + n], dtype=np.float64).reshape(sh))
        o += n
    return out


class MlpRef:
    def __init__(self, n_in=8, width=256, n_hidden=4, n_heads=4, n_out=256):
        self.n_in, self.width, self.n_hidden, self.n_heads, self.n_out = n_in, width, n_hidden, n_heads, n_out
        self.shapes = layout(n_in, width, n_hidden, n_heads, n_out)
        self.n_params = sum(int(np.prod(s)) for s in self.shapes)

    def init(self, seed=0):
        """He-style initialisation (what the surrogate driver uses), flat float32."""
        rng = np.random.default_rng(seed)
        parts = []
        for sh in self.shapes:
            if len(sh) == 2:
                parts.append(rng.standard_normal(sh) * np.sqrt(2.0 / sh[1]))
            else:
                parts.append(np.zeros(sh))
        return np.concatenate([p.ravel() for p in parts]).astype(np.float32)

    def forward(self, p, x, bf=False):
        r = (lambda a: bf16(a)) if bf else (lambda a: np.asarray(a, dtype=np.float64))
        ws = unflatten(p, self.shapes)
        h = r(x)
        zs, hs = [], [h]
        for l in range(self.n_hidden):
            W, b = ws[2 * l], ws[2 * l + 1]
            z = h @ r(W).T + b
            h = r(gelu(z))
            zs.append(z)
            hs.append(h)
        Wh, bh = ws[-2], ws[-1]
        y = h @ r(Wh).T + bh
        return y.reshape(len(x), self.n_heads, self.n_out), (zs, hs, ws)

    def loss_grad(self, p, x, t, w, bf=False):
        r = (lambda a: bf16(a)) if bf else (lambda a: np.asarray(a, dtype=np.float64))
        B = len(x)
        y, (zs, hs, ws) = self.forward(p, x, bf)
        scale = 1.0 / (self.n_out * B * self.n_heads)
        d = y - np.asarray(t, np.float64)[:, None, :]
        wt = np.asarray(w, np.float64)[:, :, None]
        loss = float((wt * d * d).sum() * scale)
        dY = r((2.0 * wt * d * scale).reshape(B, -1))
        grads = [None] * len(ws)
        grads[-2] = dY.T @ hs[-1]
        grads[-1] = dY.sum(axis=0)
        dH = dY @ r(ws[-2])
        for l in range(self.n_hidden - 1, -1, -1):
            dZ = r(dH * gelu_grad(zs[l]))
            grads[2 * l] = dZ.T @ hs[l]
            grads[2 * l + 1] = dZ.sum(axis=0)
            dH = dZ @ r(ws[2 * l])
        return loss, np.concatenate([g.ravel() for g in grads])
