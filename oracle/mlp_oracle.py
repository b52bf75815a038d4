"""ORACLE -- TEST INFRASTRUCTURE ONLY.

CPU (float64 numpy) forward and backward of the load predictor's network, used to check the
device kernels' gradients. Same definition as the device: y = W3 relu(W2 relu(x W1 + b1) + b2) + b3,
weights rounded to BF16 (round-to-nearest-even) as the device copy is, masked MSE
loss = sum(m * (y - t)^2) / sum(m). The network has no counterpart in the reference.
Pinned in tests/test_predictor.py by central finite differences on a tiny network.
"""
import numpy as np


def bf16_round(w):
    """Round float32 values to the nearest BF16 (ties to even), returned as float64."""
    u = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, dims):
    n_in, h1, h2, n_out = dims
    o = 0
    out = []
    for shape in ((n_in, h1), (h1,), (h1, h2), (h2,), (h2, n_out), (n_out,)):
        n = int(np.prod(shape))
        out.append(np.asarray(params[o:o + n], dtype=np.float64).reshape(shape))
        o += n
    return out


def loss_and_grad(params, dims, x, t, mask=None, round_weights=True):
    W1, b1, W2, b2, W3, b3 = unpack(params, dims)
    if round_weights:
        W1, W2, W3 = (bf16_round(W.astype(np.float32)) for W in (W1, W2, W3))
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    m = np.ones_like(t) if mask is None else np.asarray(mask, dtype=np.float64)
    z1 = x @ W1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ W3 + b3
    cnt = max(m.sum(), 1.0)
    d = (y - t) * m
    loss = (d * d).sum() / cnt
    dy = 2 * d / cnt
    gW3 = h2.T @ dy
    gb3 = dy.sum(0)
    dh2 = (dy @ W3.T) * (z2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (z1 > 0)
    gW1 = x.T @ dh1
    gb1 = dh1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g, y
