"""numpy restatement of the surrogate MLP (test infrastructure): the
forward pass, the mean-squared-error loss and its backward pass, in FP64,
optionally rounding the operands of every matrix product to BF16 (round to
nearest even) where the GPU kernels do, so that the GPU's products can be
compared at FP32-accumulation accuracy."""
import math

import numpy as np
from scipy.special import erf


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unflatten(p, n_in, width, n_hidden, n_out):
    sizes = [n_in] + [width] * n_hidden + [n_out]
    Ws, bs, off = [], [], 0
    for k, n in zip(sizes[:-1], sizes[1:]):
        Ws.append(np.asarray(p[off:off + k * n], dtype=np.float64).reshape(k, n))
        off += k * n
        bs.append(np.asarray(p[off:off + n], dtype=np.float64))
        off += n
    return Ws, bs


def flatten(Ws, bs):
    return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in zip(Ws, bs)])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(Ws, bs, x, emulate=True):
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    h, zs, hs = r(x), [], []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        hs.append(h)
        z = h @ r(W) + b
        zs.append(z)
        h = r(gelu(z)) if l + 1 < len(Ws) else z
    return h, zs, hs


def loss_and_grads(Ws, bs, x, t, emulate=True):
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    y, zs, hs = forward(Ws, bs, x, emulate)
    n = y.size
    loss = float(np.mean((y - t) ** 2))
    dz = 2.0 * (y - t) / n
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        if l + 1 < len(Ws):
            dz = dh * gelu_grad(zs[l])
        gW[l] = hs[l].T @ r(dz)
        gb[l] = dz.sum(axis=0)
        if l > 0:
            dh = r(dz) @ r(Ws[l]).T
    return loss, gW, gb
