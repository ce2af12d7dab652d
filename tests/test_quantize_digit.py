"""The digit quantization of the kernels takes the IEEE float quotient u / delta of
high_acc_quantize16 (space.hpp:695-711) as float((double)u * (1.0 / (double)delta)).

The claim: the double product is within 2^-52 (relative) of the exact quotient, and a
quotient of two floats is never within 2^-49 (relative) of a float rounding boundary
unless it is exactly representable, so both round to the same float.  Checked here
against numpy's IEEE float32 division on random residual scales and on quotients placed
a few ulps either side of every kind of half-integer, where the digit flips.  The kernels
then round the quotient with roundf; that it equals the reference's rounding in double is
checked for every float in range, through a numpy form of roundf that is itself checked
against the C library's.
"""
import ctypes
import ctypes.util
import numpy as np


def _digits(tmp):
    r = tmp.astype(np.float64) + 0.5 - (tmp < 0).astype(np.float64)
    return np.trunc(r).astype(np.int64)


def _digits_one_add(tmp):
    # the kernels' form: one addition of +-0.5
    r = tmp.astype(np.float64) + np.where(tmp < 0, -0.5, 0.5)
    return np.trunc(r).astype(np.int64)


def test_single_addition_rounding():
    rng = np.random.default_rng(13)
    parts = [(rng.random(2_000_000) * 16384 - 8192).astype(np.float32),
             (np.arange(-8192, 8193) + 0.5).astype(np.float32),
             np.arange(-8192, 8193).astype(np.float32)]
    for e in range(-149, 14):  # every binade, both signs, values next to half-integers
        x = np.float32(2.0) ** e * (1 + rng.random(2000)).astype(np.float32)
        parts += [x, -x, np.nextafter(x, np.float32(0)), np.nextafter(x, np.float32(np.inf))]
    h = (np.arange(-8192, 8192) + 0.5).astype(np.float32)
    parts += [np.nextafter(h, np.float32(-np.inf)), np.nextafter(h, np.float32(np.inf))]
    t = np.concatenate(parts).astype(np.float32)
    t = t[np.abs(t) <= 8192.5]
    np.testing.assert_array_equal(_digits_one_add(t), _digits(t))


def _check(u, delta):
    ref = (u / delta).astype(np.float32)  # IEEE float32 division
    rd = 1.0 / delta.astype(np.float64)
    got = (u.astype(np.float64) * rd).astype(np.float32)
    assert ref.dtype == np.float32
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(_digits(got), _digits(ref))


def test_random_residuals():
    rng = np.random.default_rng(11)
    n = 4_000_000
    vmax = (0.02 + rng.random(n)).astype(np.float32)
    delta = (vmax / np.float32(8191.0)).astype(np.float32)
    u = ((rng.random(n) * 2 - 1).astype(np.float32) * vmax).astype(np.float32)
    _check(u, delta)


def test_near_half_integers():
    rng = np.random.default_rng(12)
    n = 400_000
    vmax = (0.02 + rng.random(n)).astype(np.float32)
    delta = (vmax / np.float32(8191.0)).astype(np.float32)
    k = rng.integers(0, 8191, n)
    u0 = ((k + 0.5) * delta.astype(np.float64)).astype(np.float32)
    for step in range(-3, 4):
        u = u0.copy()
        for _ in range(abs(step)):
            u = np.nextafter(u, np.float32(np.inf if step > 0 else -np.inf)).astype(np.float32)
        _check(u, delta)
        _check(-u, delta)


def test_extremes():
    # delta of the constant residual (|q - c| <= 1e-5) and the full digit range
    for D in (64, 128, 1536, 3072):
        value = np.float32(1.0 / np.sqrt(np.float32(D)))
        delta = np.full(3, value / np.float32(8191.0), np.float32)
        u = np.array([value, -value, 0.0], np.float32)
        _check(u, delta)
        assert list(_digits((u / delta).astype(np.float32))) == [8191, -8191, 0]


def test_roundf_equals_double_rounding_exhaustive():
    # the kernels round the quotient with roundf (half away from zero); the reference adds
    # 0.5 - [tmp < 0] in double and truncates.  Every non-negative float up to 8192.5;
    # both forms are odd functions of tmp, so the negative side mirrors it (sampled below).
    lim = int(np.float32(8192.5).view(np.uint32))
    step = 1 << 26
    for s in range(0, lim + 1, step):
        t = np.arange(s, min(lim + 1, s + step), dtype=np.uint32).view(np.float32)
        r = np.trunc(t)
        got = (r + (np.abs(t - r) >= np.float32(0.5)).astype(np.float32)).astype(np.int64)
        np.testing.assert_array_equal(got, _digits(t))
    rng = np.random.default_rng(14)
    t = (rng.random(1_000_000) * -8192.5).astype(np.float32)
    h = -(np.arange(0, 8192) + 0.5).astype(np.float32)
    t = np.concatenate([t, h, np.nextafter(h, np.float32(0)), np.nextafter(h, np.float32(-np.inf))])
    r = np.trunc(t)
    got = (r - (np.abs(t - r) >= np.float32(0.5)).astype(np.float32)).astype(np.int64)
    np.testing.assert_array_equal(got, _digits(t))


def test_numpy_roundf_is_libm_roundf():
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    libm.roundf.restype = ctypes.c_float
    libm.roundf.argtypes = [ctypes.c_float]
    rng = np.random.default_rng(15)
    h = (np.arange(-8192, 8192) + 0.5).astype(np.float32)
    t = np.concatenate([(rng.random(4000) * 16385 - 8192.5).astype(np.float32), h,
                        np.nextafter(h, np.float32(np.inf)), np.nextafter(h, np.float32(-np.inf))])
    r = np.trunc(t)
    mine = r + np.where(np.abs(t - r) >= np.float32(0.5), np.sign(t), np.float32(0))
    want = np.array([libm.roundf(float(v)) for v in t], np.float32)
    np.testing.assert_array_equal(mine, want)
