"""The AVX-512 CPU baseline (oracle/erq_cpu512.cpp) equals the scalar oracle bit for bit:
ids and distances of whole searches at every EX, D = 64 .. 3072 (the u16 wraparound of the
crafted index included), ragged clusters and k beyond the probed vectors; the single-thread
and the OpenMP runs agree; the rotation matches a float64 GEMM within the FP32 bound; the
ratio matches tools.hpp:16-51's definition on a hand-computed case."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.skipif(orc.cpu512() is None, reason="no AVX-512 on this CPU")


def _idx(n, dim, ncl, bits, seed=1, nq=12, sizes=None):
    from erq import data, index
    x, cen = data.mixture_isotropic(n, dim, ncl, 0.3, seed=seed)
    q, _ = data.mixture_isotropic(nq, dim, ncl, 0.3, seed=seed + 1, centres=cen)
    c, assign = data.kmeans(x, ncl, 4, seed=seed + 3)
    if sizes is not None:
        assign = np.repeat(np.arange(len(sizes)), sizes)[:n].astype(np.int64)
    idx = index.build_index(x, c, assign, bits=bits, seed=seed + 2)
    qp = np.zeros((nq, idx.D), np.float32)
    qp[:, :dim] = q
    return idx, (qp.astype(np.float64) @ idx.P.astype(np.float64)).astype(np.float32)


def _same(idx, qr, k, nprobe):
    oi = orc.OracleIndex(idx)
    ids, d, _ = orc.cpu512_search(oi, qr, k, nprobe)
    for i in range(qr.shape[0]):
        oid, od = oi.search(qr[i], k, nprobe)
        np.testing.assert_array_equal(ids[i], oid)
        np.testing.assert_array_equal(d[i].view(np.uint32), od.view(np.uint32))
    ids2, d2, _ = orc.cpu512_search(oi, qr, k, nprobe, threads=4)
    np.testing.assert_array_equal(ids2, ids)
    np.testing.assert_array_equal(d2.view(np.uint32), d.view(np.uint32))


@pytest.mark.parametrize("bits", [3, 4, 5, 7, 8, 9])
def test_every_ex_d128(bits):
    idx, qr = _idx(5000, 128, 20, bits)
    _same(idx, qr, 50, 6)


def test_d1536_and_d1000_padding():
    idx, qr = _idx(3000, 1536, 12, 4, nq=6)
    _same(idx, qr, 100, 5)
    idx, qr = _idx(3000, 1000, 12, 8, nq=6)
    assert idx.D == 1024
    _same(idx, qr, 20, 4)


def test_ragged_and_large_k():
    sizes = [0, 1, 31, 32, 33, 95, 64, 744]
    idx, qr = _idx(sum(sizes), 64, len(sizes), 4, sizes=sizes)
    for k, npb in ((1, 8), (300, 8), (1024, 8)):
        _same(idx, qr, k, npb)


def test_wrap_d3072_crafted():
    from crafted import wrap_index
    idx, q, _ = wrap_index()
    _same(idx, q, 20, idx.K)


def test_rotation_within_fp32_bound():
    rng = np.random.default_rng(4)
    D = 256
    P = np.linalg.qr(rng.standard_normal((D, D)))[0].astype(np.float32)
    q = rng.standard_normal((9, 200)).astype(np.float32)
    out, us = orc.cpu512_rotate(q, P)
    qp = np.zeros((9, D))
    qp[:, :200] = q
    ref = qp @ P.astype(np.float64)
    bound = D * 2.0 ** -23 * (np.abs(qp) @ np.abs(P.astype(np.float64)))
    assert np.all(np.abs(out - ref) <= bound) and us > 0


def test_ratio_definition():
    # the sorted true / returned distances paired largest with largest; pairs whose true
    # distance is below 1e-5 do not count; the mean times k
    gt = np.array([1.0, 4.0, 0.0, 9.0], np.float32)
    ann = np.array([16.0, 1.0, 4.0, 9.0], np.float32)
    # sorted: gt 0,1,4,9  ann 1,4,9,16 -> pairs (9,16) (4,9) (1,4), (0,1) skipped
    want = (np.sqrt(16 / 9) + np.sqrt(9 / 4) + np.sqrt(4 / 1)) / 3 * 4
    assert abs(orc.ratio(gt, ann) - want) < 1e-12
    assert orc.ratio(np.zeros(3, np.float32), np.ones(3, np.float32)) == 3.0


def test_distance_ratio_matches_definition():
    import torch
    from erq import data
    rng = np.random.default_rng(9)
    g = rng.uniform(0, 4, (7, 10)).astype(np.float32)
    g[0, :3] = 0.0                      # pairs with a zero true distance do not count
    g[1] = 0.0                          # no pair counts: 1
    a = (g + rng.uniform(0, 1, g.shape)).astype(np.float32)
    want = np.mean([orc.ratio(g[i], a[i]) / 10 for i in range(7)])
    got = data.distance_ratio(torch.from_numpy(g), torch.from_numpy(a))
    assert abs(got - want) < 1e-12
