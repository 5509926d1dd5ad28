"""CPU oracle of the MLP warm start (TEST INFRASTRUCTURE ONLY: only tests and
the smoke check use it; the product runs the HIP kernels).

Dense layers with GELU between them and a mean-squared-error loss, in
float64, with the quantisation points of the GPU kernels: the GEMM inputs
(network input, hidden activations, weights, output gradients and hidden
gradients) are rounded to BF16 (round to nearest even) before the products,
accumulation and everything else is exact-ish (float64 here, float32 on the
GPU).  The network has no counterpart in the reference; the oracle's
backward pass is pinned by central finite differences of its own forward
(tests/test_mlp_oracle.py).
"""
import math

import numpy as np

try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover
    _erf = np.vectorize(math.erf)


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unpack(p, dims):
    Ws, bs, o = [], [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        Ws.append(np.asarray(p[o: o + a * b], dtype=np.float64).reshape(a, b))
        o += a * b
        bs.append(np.asarray(p[o: o + b], dtype=np.float64))
        o += b
    return Ws, bs


def forward(p, dims, x, quantise=True):
    q = bf16 if quantise else (lambda v: np.asarray(v, dtype=np.float64))
    Ws, bs = unpack(p, dims)
    h = q(x)
    Zs, Hs = [], [h]
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = h @ q(W) + b
        Zs.append(z)
        if l + 1 < len(Ws):
            h = q(gelu(z))
            Hs.append(h)
    return Zs[-1], Zs, Hs


def loss_grad(p, dims, x, t, quantise=True):
    """Mean squared error and its gradient (flattened like the parameters)."""
    q = bf16 if quantise else (lambda v: np.asarray(v, dtype=np.float64))
    Ws, bs = unpack(p, dims)
    y, Zs, Hs = forward(p, dims, x, quantise)
    B, Nout = y.shape
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.sum(d * d) / (B * Nout))
    dz = 2.0 * d / (B * Nout)
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        dzq = q(dz)
        gW[l] = Hs[l].T @ dzq
        gb[l] = dz.sum(axis=0)
        if l > 0:
            dz = (dzq @ q(Ws[l]).T) * gelu_d(Zs[l - 1])
    g = np.concatenate([np.concatenate([w.ravel(), b_]) for w, b_ in zip(gW, gb)])
    return loss, g
