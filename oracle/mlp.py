"""CPU restatement of the warm-start MLP — test infrastructure only (the
product's network lives in diffphd_amd/csrc/phd_mlp.hip).

Per-vertex MLP 13 -> 128 -> 128 -> 128 -> 3 with exact (erf) GELU. The
quantized mode rounds to BF16 (nearest even) at the points where the MFMA
kernels do: weights, every layer input, and the output gradients that feed the
weight and input products; biases, pre-activations and accumulation stay
wide (FP64 here, FP32 on the GPU). The unquantized mode is plain FP64 and is
what the finite-difference checks pin."""
import math

import numpy as np

IN, INP, H, OUT, OUTP = 13, 16, 128, 3, 16
_SHAPES = [("W1", (INP, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
           ("W4", (H, OUTP)), ("b4", (OUTP,))]
TOTAL = sum(int(np.prod(s)) for _, s in _SHAPES)

_erf = np.vectorize(math.erf)


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    out, o = {}, 0
    for name, shp in _SHAPES:
        k = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in _SHAPES])


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def _forward(p, X, quant):
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(p)
    x0 = q(X)
    z1 = x0 @ q(P["W1"]) + P["b1"]
    h1 = q(gelu(z1))
    z2 = h1 @ q(P["W2"]) + P["b2"]
    h2 = q(gelu(z2))
    z3 = h2 @ q(P["W3"]) + P["b3"]
    h3 = q(gelu(z3))
    o = h3 @ q(P["W4"]) + P["b4"]
    return P, (x0, z1, h1, z2, h2, z3, h3), o


def forward(p, X, quant=True):
    """X: n x 16 standardized features -> n x 3."""
    return _forward(p, X, quant)[2][:, :OUT]


def loss_grad(p, X, T, quant=True):
    """Mean over n x 3 of (out - T)^2 and its gradient (flat, TOTAL)."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    P, (x0, z1, h1, z2, h2, z3, h3), o = _forward(p, X, quant)
    n = X.shape[0]
    e = o[:, :OUT] - T
    loss = float(np.sum(e * e) / (3.0 * n))
    d4 = np.zeros((n, OUTP))
    d4[:, :OUT] = 2.0 * e / (3.0 * n)
    g = {}
    g["b4"] = d4.sum(0)
    g["W4"] = h3.T @ q(d4)
    dz3 = (q(d4) @ q(P["W4"]).T) * gelu_d(z3)
    g["b3"] = dz3.sum(0)
    g["W3"] = h2.T @ q(dz3)
    dz2 = (q(dz3) @ q(P["W3"]).T) * gelu_d(z2)
    g["b2"] = dz2.sum(0)
    g["W2"] = h1.T @ q(dz2)
    dz1 = (q(dz2) @ q(P["W2"]).T) * gelu_d(z1)
    g["b1"] = dz1.sum(0)
    g["W1"] = x0.T @ q(dz1)
    return loss, pack(g)
