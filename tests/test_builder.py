"""Index builder (liberq_builder.so) and the HA file format, on the CPU.

The builder writes the layout the search reads; the checks here are the ones the
mathematics fixes: sign bits, code ranges, the xipnorm closed form, optimality of the
EX-bit rescale search against brute force, and byte-exact file round trips.
"""
import numpy as np
import pytest

from oracle import oracle as orc


def _one_cluster(n, D, ex, seed=0):
    import ctypes
    from erq import _lib
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((n, D)).astype(np.float32)
    onorm = rng.uniform(0.5, 2.0, n).astype(np.float32)
    nb = (n + 31) // 32
    blocks = np.zeros(nb * (4 * D + 128), np.uint8)
    longc = np.zeros(n * D * ex // 8, np.uint8)
    xip = np.zeros(n, np.float32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = _lib.builder().erqb_quantize_cluster(p(r), p(onorm), n, D, ex, p(blocks), p(longc),
                                              p(xip))
    assert rc == 0
    return r, onorm, blocks, longc, xip


def _magnitude_codes(r, longc, D, ex):
    """o_bar per dim: e if the sign bit is 1, else R-1-e (Quantizer.hpp:479-485)."""
    R = 1 << ex
    out = []
    for v in range(r.shape[0]):
        e = orc.ex_decode(longc[v * D * ex // 8:(v + 1) * D * ex // 8], D, ex).astype(np.int64)
        out.append(np.where(r[v] > 0, e, R - 1 - e))
    return np.array(out)


@pytest.mark.parametrize("ex", [2, 3, 4, 6, 7, 8])
def test_block_codes_are_sign_bits_and_factors(ex):
    D, n = 128, 45
    r, onorm, blocks, longc, xip = _one_cluster(n, D, ex)
    bb = 4 * D + 128
    for b in range((n + 31) // 32):
        blk = blocks[b * bb:(b + 1) * bb]
        vs = range(32 * b, min(n, 32 * b + 32))
        bits = np.zeros((32, D), np.uint8)
        for i, v in enumerate(vs):
            bits[i] = r[v] > 0
        np.testing.assert_array_equal(blk[:4 * D], orc.pack_block_codes(bits, D))
        fac = blk[4 * D:].view(np.float32)
        np.testing.assert_array_equal(fac[:len(vs)], onorm[list(vs)])
        assert np.all(fac[len(vs):] == 0)


@pytest.mark.parametrize("ex", [2, 3, 4, 6, 7, 8])
def test_xipnorm_closed_form(ex):
    D, n = 192, 20
    r, onorm, blocks, longc, xip = _one_cluster(n, D, ex, seed=1)
    ob = _magnitude_codes(r, longc, D, ex)
    assert ob.min() >= 0 and ob.max() <= (1 << ex) - 1
    absu = np.abs(r.astype(np.float64)) / np.linalg.norm(r.astype(np.float64), axis=1,
                                                          keepdims=True)
    numer = ((ob + 0.5) * absu).sum(1)
    np.testing.assert_allclose(xip, 2.0 * onorm / numer, rtol=2e-6)


@pytest.mark.parametrize("ex", [2, 3, 4])
def test_rescale_search_is_optimal(ex):
    """The codes maximise <o_bar + 1/2, |u|> / ||o_bar + 1/2|| over every rescale factor
    (brute force over a dense grid of t plus every breakpoint)."""
    D, n = 64, 6
    r, onorm, blocks, longc, xip = _one_cluster(n, D, ex, seed=2)
    ob = _magnitude_codes(r, longc, D, ex)
    top = (1 << ex) - 1
    for v in range(n):
        o = np.abs(r[v].astype(np.float64))
        o /= np.linalg.norm(o)
        def score(c):
            return ((c + 0.5) * o).sum() / np.sqrt(((c + 0.5) ** 2).sum())
        got = score(ob[v])
        tmax = (top + 10) / o.max()
        ts = np.concatenate([np.linspace(1e-6, tmax, 4000),
                             ((np.arange(top + 1)[:, None] + 1) / o[None, :]).ravel() + 1e-9])
        best = max(score(np.minimum(np.floor(t * o + 1e-5), top)) for t in ts if t < tmax)
        assert got >= best - 1e-9


def test_index_file_round_trip(small_data, tmp_path):
    from erq import index
    x, q, c, assign = small_data
    idx = index.build_index(x, c, assign, bits=5, seed=3)
    p = tmp_path / "x.index"
    idx.save(str(p))
    raw = p.read_bytes()
    D = idx.D
    hdr = np.frombuffer(raw[:32], "<u8")
    assert hdr.tolist() == [idx.N, 128, idx.K, 4]
    expect = (32 + 8 * idx.K + 4 * D * D + 4 * idx.K * D + idx.short_data.size +
              idx.long_code.size + 4 * idx.N + 4 * idx.N)
    assert len(raw) == expect
    idx2 = index.IVFIndex.load(str(p))
    for f in ("cluster_sizes", "P", "centroids", "short_data", "long_code", "ex_factor", "ids"):
        np.testing.assert_array_equal(getattr(idx, f), getattr(idx2, f))
    p2 = tmp_path / "y.index"
    idx2.save(str(p2))
    assert p2.read_bytes() == raw


def test_rotation_is_orthogonal():
    from erq import index
    P = index.random_rotation(128, 3).astype(np.float64)
    np.testing.assert_allclose(P @ P.T, np.eye(128), atol=1e-5)


def test_ids_grouped_by_cluster(small_data):
    from erq import index
    x, q, c, assign = small_data
    idx = index.build_index(x, c, assign, bits=4, seed=3)
    vs = idx.vec_start()
    for k in range(idx.K):
        ids = idx.ids[vs[k]:vs[k + 1]]
        assert np.all(assign[ids] == k)
        assert np.all(np.diff(ids.astype(np.int64)) > 0)   # IVF.hpp:121-131 push order


def test_zero_residual_vector_is_quantized():
    """A vector equal to its centroid (singleton cluster) has a zero residual; the
    builder must not divide by the zero maximum (regression: it used to spin)."""
    import ctypes
    from erq import _lib
    D, n, ex = 128, 3, 3
    r = np.random.default_rng(4).standard_normal((n, D)).astype(np.float32)
    r[1] = 0.0
    onorm = np.array([1.0, 0.0, 2.0], np.float32)
    blocks = np.zeros(4 * D + 128, np.uint8)
    longc = np.zeros(n * D * ex // 8, np.uint8)
    xip = np.zeros(n, np.float32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert _lib.builder().erqb_quantize_cluster(p(r), p(onorm), n, D, ex, p(blocks), p(longc),
                                                p(xip)) == 0
    e = orc.ex_decode(longc[D * ex // 8:2 * D * ex // 8], D, ex)
    assert np.all(e == (1 << ex) - 1)        # magnitude 0, sign bit 0 -> flipped code
    assert xip[1] == 0.0
    r[2, 5] = np.nan
    assert _lib.builder().erqb_quantize_cluster(p(r), p(onorm), n, D, ex, p(blocks), p(longc),
                                                p(xip)) == -2


def _fast_quantize_heap(o, ex):
    """DataQuantizer::fast_quantize (Quantizer.hpp:138-219) as written: a min-heap of
    (t, dimension) events, in double precision like the C++."""
    import heapq
    import math
    eps, n_enum, top = 1e-5, 10, (1 << ex) - 1
    o = [float(v) for v in o]
    D = len(o)
    max_o = max(o)
    t_start = float(top // 3) / max_o
    t_end = float(top + n_enum) / max_o
    cur = [int(t_start * v + eps) for v in o]
    sqr_den = D * 0.25
    num = 0.0
    for c, v in zip(cur, o):
        sqr_den += c * c + c
        num += (c + 0.5) * v
    heap = [((c + 1) / v if v > 0 else math.inf, i) for i, (c, v) in enumerate(zip(cur, o))]
    heapq.heapify(heap)
    best_ip, best_t = 0.0, 0.0
    while heap:
        t, i = heapq.heappop(heap)
        cur[i] += 1
        v = cur[i]
        sqr_den += 2 * v
        num += o[i]
        ip = num / math.sqrt(sqr_den)
        if ip > best_ip:
            best_ip, best_t = ip, t
        if v < top and o[i] > 0:
            tn = (v + 1) / o[i]
            if tn < t_end:
                heapq.heappush(heap, (tn, i))
    codes = [min(int(best_t * v + eps), top) for v in o]
    numer = 0.0
    for c, v in zip(codes, o):
        numer += (c + 0.5) * v
    ipn = np.float32(1.0 / numer)
    return np.array(codes, np.uint8), ipn if np.isfinite(ipn) else np.float32(1.0)


@pytest.mark.parametrize("ex", [2, 3, 4, 6, 7, 8])
def test_rescale_search_follows_the_heap_order(ex):
    """The builder lists the events and orders them instead of keeping a heap; codes and
    factor equal the heap form bit for bit, with zero magnitudes and many equal breakpoints
    (magnitudes on a coarse grid, so (v+1)/o ties across dimensions)."""
    import ctypes
    from erq import _lib
    rng = np.random.default_rng(40 + ex)
    D = 128
    cases = [np.abs(rng.standard_normal(D)).astype(np.float32) for _ in range(3)]
    grid = (rng.integers(0, 9, D) / 8.0).astype(np.float32)
    grid[rng.integers(0, D, 5)] = 0.0
    cases.append(grid)
    cases.append(np.full(D, 0.25, np.float32))
    for o in cases:
        o = o / np.float32(np.linalg.norm(o.astype(np.float64)))
        o = o.astype(np.float32)
        code = np.zeros(D, np.uint8)
        ipn = _lib.builder().erqb_quantize_magnitudes(
            o.ctypes.data_as(ctypes.c_void_p), D, ex, code.ctypes.data_as(ctypes.c_void_p))
        want_code, want_ipn = _fast_quantize_heap(o, ex)
        np.testing.assert_array_equal(code, want_code)
        assert np.float32(ipn) == want_ipn
