"""FP64 NumPy reference of the subgrid-closure MLP (11-64-64-64-9, GELU),
forward and backward by hand, with optional BF16 rounding of every MFMA
operand so that the device kernels can be compared to it tightly.

TEST INFRASTRUCTURE ONLY (the checker of compchan3dmpi_amd/csrc/cc_mlp.hip);
its gradient is pinned by central finite differences in tests/test_mlp_ref.py.
"""
from __future__ import annotations

import math

import numpy as np

NP = 10448
OFF = dict(W1=(0, 16, 64), b1=(1024, 64), W2=(1088, 64, 64), b2=(5184, 64),
           W3=(5248, 64, 64), b3=(9344, 64), W4=(9408, 64, 16), b4=(10432, 16))
_erf = np.vectorize(math.erf, otypes=[np.float64])


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    out = {}
    for k, v in OFF.items():
        o = v[0]
        n = int(np.prod(v[1:]))
        out[k] = P[o:o + n].reshape(v[1:])
    return out


def pack(d):
    P = np.zeros(NP)
    for k, v in OFF.items():
        o = v[0]
        n = int(np.prod(v[1:]))
        P[o:o + n] = np.asarray(d[k]).reshape(-1)
    return P


def bf16(x):
    """Round to the nearest BF16 (ties to even), as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def dgelu(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def forward(P, X, emulate_bf16=False):
    """X (N, 16) -> Y (N, 16) and the pre-activations [Z1, Z2, Z3]."""
    w = unpack(P)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    a = np.asarray(X, dtype=np.float64)
    Zs = []
    for L in (1, 2, 3):
        z = r(a) @ r(w[f"W{L}"]) + w[f"b{L}"]
        Zs.append(z)
        a = gelu(z)
    y = r(a) @ r(w["W4"]) + w["b4"]
    return y, Zs


def loss_grad(P, X, T, emulate_bf16=False):
    """Mean over samples and the 9 outputs of (Y - T)^2 and its gradient."""
    w = unpack(P)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    X = np.asarray(X, dtype=np.float64)
    N = X.shape[0]
    y, Zs = forward(P, X, emulate_bf16)
    e = (y - T)[:, :9]
    loss = float((e ** 2).sum() / (9 * N))
    dy = np.zeros_like(y)
    dy[:, :9] = 2.0 * e / (9 * N)
    acts = [X] + [gelu(z) for z in Zs]
    g = {}
    d = dy
    for L in (4, 3, 2, 1):
        g[f"W{L}"] = r(acts[L - 1]).T @ r(d)
        g[f"b{L}"] = d.sum(axis=0)
        if L > 1:
            d = (r(d) @ r(w[f"W{L}"]).T) * dgelu(Zs[L - 2])
    g["W1"][11:, :] = 0.0
    g["W4"][:, 9:] = 0.0
    g["b4"][9:] = 0.0
    return loss, pack(g)
