"""FP64 reference of the probe-predictor MLP (test infrastructure only).

p = sigmoid(W3 gelu(W2 gelu(W1 (s x) + b1) + b2) + b3), weights out-major ([out x in]),
s the input scale (1 unless given),
GELU with erf, loss = mean over the batch of sum over clusters of the pos-weighted BCE
    l = w*y*softplus(-z) + (1-y)*softplus(z).
Parameters are packed as [W1, b1, W2, b2, W3, b3] (the layout liberq's erq_mlp uses).
The analytic backward here is checked against central finite differences in tests.
"""
import math

import numpy as np

_erf = np.vectorize(math.erf)


def unpack(theta, din, hid, dout):
    o = 0
    out = []
    for shape in ((hid, din), (hid,), (hid, hid), (hid,), (dout, hid), (dout,)):
        n = int(np.prod(shape))
        out.append(theta[o:o + n].reshape(shape))
        o += n
    return out


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def softplus(z):
    return np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z)))


def forward(theta, x, din, hid, dout, xs=1.0):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(theta, np.float64), din, hid, dout)
    x = np.asarray(x, np.float64) * xs
    z1 = x @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    z3 = h2 @ W3.T + b3
    return dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3)


def loss(theta, x, y, din, hid, dout, w, xs=1.0):
    f = forward(theta, x, din, hid, dout, xs)
    z, y = f["z3"], np.asarray(y, np.float64)
    return float((w * y * softplus(-z) + (1 - y) * softplus(z)).sum() / x.shape[0])


def grad(theta, x, y, din, hid, dout, w, xs=1.0):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(theta, np.float64), din, hid, dout)
    f = forward(theta, x, din, hid, dout, xs)
    B = x.shape[0]
    p = 1.0 / (1.0 + np.exp(-f["z3"]))
    y = np.asarray(y, np.float64)
    d3 = (w * y * (p - 1) + (1 - y) * p) / B
    gW3 = d3.T @ f["h2"]
    gb3 = d3.sum(0)
    d2 = (d3 @ W3) * gelu_grad(f["z2"])
    gW2 = d2.T @ f["h1"]
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * gelu_grad(f["z1"])
    gW1 = d1.T @ f["x"]
    gb1 = d1.sum(0)
    return np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])


def round_bf16(a):
    """float -> nearest-even bfloat16, returned as float64 (the GPU's (__bf16) cast)."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def grad_bf16(theta, x, y, din, hid, dout, w, xs=1.0):
    """Loss and gradient with the GPU's precision steps: BF16 weights, inputs (scaled in
    FP32 first), activations and backward GEMM operands; FP32 pre-activations; everything
    else in FP64."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(theta, np.float64), din, hid, dout)
    W1b, W2b, W3b = round_bf16(W1), round_bf16(W2), round_bf16(W3)
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    xb = round_bf16(np.float32(xs) * np.asarray(x, np.float32))
    z1 = f32(xb @ W1b.T + b1)
    h1 = round_bf16(gelu(z1))
    z2 = f32(h1 @ W2b.T + b2)
    h2 = round_bf16(gelu(z2))
    z3 = f32(h2 @ W3b.T + b3)
    B = x.shape[0]
    y = np.asarray(y, np.float64)
    L = float((w * y * softplus(-z3) + (1 - y) * softplus(z3)).sum() / B)
    p = 1.0 / (1.0 + np.exp(-z3))
    d3 = round_bf16((w * y * (p - 1) + (1 - y) * p) / B)
    gW3, gb3 = d3.T @ h2, d3.sum(0)
    d2 = round_bf16(f32(d3 @ W3b) * gelu_grad(z2))
    gW2, gb2 = d2.T @ h1, d2.sum(0)
    d1 = round_bf16(f32(d2 @ W2b) * gelu_grad(z1))
    gW1, gb1 = d1.T @ xb, d1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return L, g, 1.0 / (1.0 + np.exp(-z3))
