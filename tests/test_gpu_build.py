"""GPU index builder (erq_build.hip through the C-ABI) against the host builder (builder.cpp,
itself pinned byte for byte against the reference's heap-ordered fast_quantize in
test_builder.py): the same long codes, xipnorm and short blocks from the same residuals, and
an end-to-end build whose residuals are exact (identity rotation) equal to the host build."""
import ctypes

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _torch():
    import torch
    return torch


def _host_quantize(r, onorm, sizes, D, ex):
    from erq import _lib
    N = r.shape[0]
    nb = int(((sizes + 31) // 32).sum())
    short = np.zeros(nb * (4 * D + 128), np.uint8)
    longc = np.zeros(N * D * ex // 8, np.uint8)
    xip = np.zeros(N, np.float32)
    vp = ctypes.c_void_p
    rc = _lib.builder().erqb_quantize_all(
        r.ctypes.data_as(vp), onorm.ctypes.data_as(vp), sizes.ctypes.data_as(vp),
        ctypes.c_size_t(sizes.size), ctypes.c_size_t(D), ctypes.c_int(ex), short.ctypes.data_as(vp),
        longc.ctypes.data_as(vp), xip.ctypes.data_as(vp), ctypes.c_int(16))
    assert rc == 0
    return short, longc, xip


def _residuals(n, D, seed):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((n, D)).astype(np.float32)
    r[0] = 0.0                                     # zero residual: codes 0, factor 1
    r[1] = np.where(np.arange(D) % 2, 1.0, -1.0)   # all |u| equal: one t for every dim
    r[2, :D // 2] = 0.0                            # exact zeros: breakpoints at t = inf
    r[3] = np.where(np.arange(D) < D // 3, 2.0, -0.5)   # two magnitudes: large tie groups
    r[4] *= 1e-20                                  # tiny magnitudes
    r[5] = rng.integers(-3, 4, D).astype(np.float32)   # few distinct values
    onorm = (np.abs(rng.standard_normal(n)) + 0.1).astype(np.float32)
    return r, onorm


@pytest.mark.parametrize("D,ex,n", [(64, 2, 700), (128, 3, 2000), (1024, 4, 600),
                                    (1536, 3, 3000), (1536, 8, 300), (768, 6, 500),
                                    (512, 7, 500), (3072, 8, 64)])
def test_quantize_and_pack_match_host_builder(D, ex, n):
    from erq import gpu
    torch = _torch()
    print(f"build D={D} EX={ex} n={n}", flush=True)
    r, onorm = _residuals(n, D, seed=D + ex)
    rng = np.random.default_rng(ex)
    cut = np.sort(rng.choice(np.arange(1, n), 4, replace=False))
    sizes = np.diff(np.concatenate([[0], cut, [n]])).astype(np.uint64)  # 5 clusters
    want_s, want_l, want_x = _host_quantize(r, onorm, sizes, D, ex)
    rd, od = torch.from_numpy(r).cuda(), torch.from_numpy(onorm).cuda()
    lc, xip = gpu.build_quantize_dev(rd, od, ex)
    short = gpu.build_pack_dev(rd, od, sizes)
    lc = lc.cpu().numpy().reshape(-1)
    bad = np.nonzero((lc.reshape(n, -1) != want_l.reshape(n, -1)).any(1))[0]
    assert bad.size == 0, f"long codes differ for vectors {bad[:10]}"
    np.testing.assert_array_equal(xip.cpu().numpy().view(np.uint32), want_x.view(np.uint32))
    np.testing.assert_array_equal(short.cpu().numpy(), want_s)


def test_quantize_rejects_non_finite():
    from erq import gpu
    from erq._lib import ErqError
    torch = _torch()
    r = np.ones((3, 64), np.float32)
    r[1, 5] = np.nan
    with pytest.raises(ErqError):
        gpu.build_quantize_dev(torch.from_numpy(r).cuda(), torch.ones(3, device="cuda"), 3)


@pytest.mark.parametrize("dim,bits", [(128, 4), (960, 9)])
def test_build_index_identity_rotation_matches_host(dim, bits):
    """With P = I the GEMMs are exact, so the GPU residuals are the host's x - c and the
    whole index must match the host builder byte for byte."""
    from erq import data, gpu
    from erq.index import padded_dim
    from oracle import oracle as O
    torch = _torch()
    x, cen = data.mixture_isotropic(6000, dim, 24, 0.5, seed=dim)
    c, assign = data.kmeans(x, 24, 4, seed=5)
    D = padded_dim(dim)
    P = np.eye(D, dtype=np.float32)
    idx = gpu.build_index_dev(torch.from_numpy(x).cuda(), torch.from_numpy(c.astype(np.float32)).cuda(),
                              torch.from_numpy(assign.astype(np.int32)).cuda(), bits, P)
    order = np.argsort(assign, kind="stable")
    np.testing.assert_array_equal(idx.ids, order.astype(np.uint32))
    np.testing.assert_array_equal(idx.cluster_sizes, np.bincount(assign, minlength=24))
    cpad = np.zeros((24, D), np.float32)
    cpad[:, :dim] = c
    np.testing.assert_array_equal(idx.centroids, cpad)
    xs = np.zeros((6000, D), np.float32)
    xs[:, :dim] = x[order]
    r = xs - cpad[assign[order]]
    # ||o - c|| as Quantizer.hpp:399-401: f32(sqrt((double)L2Sqr)), L2Sqr16 for dim % 16 == 0
    on = np.array([np.float32(np.sqrt(np.float64(O.l2sqr16(x[i], c[assign[i]]))))
                   for i in order], np.float32)
    want_s, want_l, want_x = _host_quantize(r, on, idx.cluster_sizes, D, bits - 1)
    np.testing.assert_array_equal(idx.long_code, want_l)
    np.testing.assert_array_equal(idx.ex_factor.view(np.uint32), want_x.view(np.uint32))
    np.testing.assert_array_equal(idx.short_data, want_s)


def test_build_index_random_rotation_searches_like_host():
    """A random rotation: residuals differ from the numpy path by GEMM rounding only; codes
    agree almost everywhere and the GPU-built index searches to the same recall."""
    from erq import data, gpu, index
    from oracle import oracle as O
    torch = _torch()
    x, cen = data.mixture_isotropic(20000, 128, 64, 0.3, seed=1)
    q, _ = data.mixture_isotropic(64, 128, 64, 0.3, seed=2, centres=cen)
    c, assign = data.kmeans(x, 64, 5, seed=4)
    host = index.build_index(x, c, assign, 4, seed=3)
    dev = gpu.build_index_dev(torch.from_numpy(x).cuda(), torch.from_numpy(c.astype(np.float32)).cuda(),
                              torch.from_numpy(assign.astype(np.int32)).cuda(), 4, host.P)
    np.testing.assert_array_equal(dev.ids, host.ids)
    np.testing.assert_allclose(dev.centroids, host.centroids, atol=1e-5)
    diff = (dev.long_code.reshape(20000, -1) != host.long_code.reshape(20000, -1)).any(1).mean()
    assert diff < 0.05
    gt, _ = data.exact_knn_np(x, q, 100)
    qr = q @ host.P
    rec = {}
    for name, idx in (("host", host), ("dev", dev)):
        ids = O.OracleIndex(idx).search_batch(qr, 100, np.full(64, 16, np.uint32))[0]
        rec[name] = data.recall_at_k(ids, gt, 100)
    assert abs(rec["dev"] - rec["host"]) < 0.01, rec


def _host_quantize_nonha(r, x2, sizes, D, ex):
    from erq import _lib
    N = r.shape[0]
    nb = int(((sizes + 31) // 32).sum())
    short = np.zeros(nb * (4 * D + 512), np.uint8)
    longc = np.zeros(N * D * ex // 8, np.uint8)
    xip = np.zeros(N, np.float32)
    vp = ctypes.c_void_p
    rc = _lib.builder().erqb_quantize_all_nonha(
        r.ctypes.data_as(vp), x2.ctypes.data_as(vp), sizes.ctypes.data_as(vp),
        ctypes.c_size_t(sizes.size), ctypes.c_size_t(D), ctypes.c_int(ex), short.ctypes.data_as(vp),
        longc.ctypes.data_as(vp), xip.ctypes.data_as(vp), ctypes.c_int(16))
    assert rc == 0
    return short, longc, xip


@pytest.mark.parametrize("dim,bits", [(128, 4), (960, 9), (1536, 3)])
def test_build_index_nonha_identity_rotation_matches_host(dim, bits):
    """The non-HA format (four factors per vector) built on the GPU equals the host builder's
    byte for byte when the residuals are exact (P = I): codes, ||o - c||^2, fac_ip, sumxb,
    fac_err and xipnorm; and the index searches exactly like the oracle's non-HA searcher."""
    from erq import data, gpu
    from erq.index import padded_dim
    from oracle import oracle as O
    torch = _torch()
    n = 4000
    x, cen = data.mixture_isotropic(n, dim, 16, 0.5, seed=dim + 1)
    c, assign = data.kmeans(x, 16, 4, seed=6)
    D = padded_dim(dim)
    P = np.eye(D, dtype=np.float32)
    idx = gpu.build_index_dev(torch.from_numpy(x).cuda(), torch.from_numpy(c.astype(np.float32)).cuda(),
                              torch.from_numpy(assign.astype(np.int32)).cuda(), bits, P,
                              high_acc=False)
    assert idx.factors == 4
    order = np.argsort(assign, kind="stable")
    np.testing.assert_array_equal(idx.ids, order.astype(np.uint32))
    cpad = np.zeros((16, D), np.float32)
    cpad[:, :dim] = c
    xs = np.zeros((n, D), np.float32)
    xs[:, :dim] = x[order]
    r = xs - cpad[assign[order]]
    x2 = np.array([O.l2sqr16(x[i], c[assign[i]]) for i in order], np.float32)
    want_s, want_l, want_x = _host_quantize_nonha(r, x2, idx.cluster_sizes, D, bits - 1)
    np.testing.assert_array_equal(idx.long_code, want_l)
    np.testing.assert_array_equal(idx.ex_factor.view(np.uint32), want_x.view(np.uint32))
    np.testing.assert_array_equal(idx.short_data, want_s)
    # searched on the GPU, equal to the oracle's non-HA searcher
    from erq.gpu import IVF
    q, _ = data.mixture_isotropic(16, dim, 16, 0.5, seed=dim + 2, centres=cen)
    qp = np.zeros((16, D), np.float32)
    qp[:, :dim] = q
    g = IVF.from_index(idx)
    ids, d = g.search_dev(torch.from_numpy(qp).cuda(), 20, 4)
    torch.cuda.synchronize()
    oi = O.OracleIndex(idx)
    for i in range(16):
        oid, od = oi.search(qp[i], 20, 4)
        np.testing.assert_array_equal(ids[i].cpu().numpy().view(np.uint32), oid)
        np.testing.assert_array_equal(d[i].cpu().numpy().view(np.uint32), od.view(np.uint32))
    g.close()
