"""Pins the FP64 MLP oracle: analytic gradient against central finite differences, the
loss at known points, and GELU against its closed form."""
import math

import numpy as np

from oracle import mlp_oracle as mo


def _setup(seed=0, din=32, hid=32, dout=16, B=8):
    rng = np.random.default_rng(seed)
    n = hid * din + hid + hid * hid + hid + dout * hid + dout
    theta = rng.standard_normal(n) * 0.3
    x = rng.standard_normal((B, din))
    y = (rng.random((B, dout)) < 0.2).astype(np.float64)
    return theta, x, y, din, hid, dout


def test_gradient_matches_finite_differences():
    theta, x, y, din, hid, dout = _setup()
    g = mo.grad(theta, x, y, din, hid, dout, 3.0)
    rng = np.random.default_rng(1)
    idx = rng.choice(theta.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (mo.loss(tp, x, y, din, hid, dout, 3.0) - mo.loss(tm, x, y, din, hid, dout, 3.0)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_loss_known_answers():
    din, hid, dout = 32, 32, 16
    n = hid * din + hid + hid * hid + hid + dout * hid + dout
    theta = np.zeros(n)
    x = np.ones((4, din))
    y = np.zeros((4, dout))
    y[:, 0] = 1
    # all logits 0: each cluster contributes log 2, the positive one weighted by w
    assert math.isclose(mo.loss(theta, x, y, din, hid, dout, 2.0), (dout - 1 + 2.0) * math.log(2),
                        rel_tol=1e-12)


def test_gelu_closed_form():
    z = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    ref = np.array([0.5 * t * (1 + math.erf(t / math.sqrt(2))) for t in z])
    np.testing.assert_allclose(mo.gelu(z), ref, rtol=1e-14)
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    h = 1e-6
    np.testing.assert_allclose(mo.gelu_grad(z), (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h),
                               rtol=1e-6)


def test_round_bf16_matches_torch():
    import torch
    rng = np.random.default_rng(3)
    a = (rng.standard_normal(10000) * np.exp(rng.uniform(-20, 20, 10000))).astype(np.float32)
    a[:4] = [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0]   # ties to even
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mo.round_bf16(a), ref)


def test_bf16_gradient_tracks_fp64_gradient():
    theta, x, y, din, hid, dout = _setup(seed=4)
    L, g, _ = mo.grad_bf16(theta, x, y, din, hid, dout, 3.0)
    g64 = mo.grad(theta, x, y, din, hid, dout, 3.0)
    assert abs(L - mo.loss(theta, x, y, din, hid, dout, 3.0)) < 0.02 * abs(L)
    cos = g @ g64 / np.linalg.norm(g) / np.linalg.norm(g64)
    assert cos > 0.999


def test_pos_weighted_bce_optimum_is_shifted_logit():
    """For a constant logit z and label rate p, the pos-weighted BCE
    -w p log s(z) - (1-p) log(1-s(z)) is minimised at z = ln w + logit p; the learned
    probe predictor takes ln w off its output bias on that basis."""
    din, hid, dout = 4, 4, 1
    n = hid * din + hid + hid * hid + hid + dout * hid + dout
    x = np.zeros((10, din))
    y = np.zeros((10, dout))
    y[:3] = 1.0                                    # p = 0.3
    w = 5.0
    zs = np.linspace(-3, 4, 70001)
    theta = np.zeros(n)
    best = None
    for z in zs[::100]:
        theta[-1] = z
        L = mo.loss(theta, x, y, din, hid, dout, w)
        if best is None or L < best[0]:
            best = (L, z)
    expect = math.log(w) + math.log(0.3 / 0.7)
    assert abs(best[1] - expect) < 0.011, (best, expect)


def test_input_scale_equals_scaled_first_layer():
    """MLP(s x) is the network with W1 scaled by s; its W1 gradient is s times that one's."""
    theta, x, y, din, hid, dout = _setup(seed=6)
    xs = 2.75
    t2 = theta.copy()
    t2[:hid * din] *= xs
    assert math.isclose(mo.loss(theta, x, y, din, hid, dout, 3.0, xs),
                        mo.loss(t2, x, y, din, hid, dout, 3.0), rel_tol=1e-12)
    g = mo.grad(theta, x, y, din, hid, dout, 3.0, xs)
    g2 = mo.grad(t2, x, y, din, hid, dout, 3.0)
    np.testing.assert_allclose(g[:hid * din], xs * g2[:hid * din], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(g[hid * din:], g2[hid * din:], rtol=1e-9, atol=1e-12)
