"""GPU parity: the HIP search path (through the C-ABI) against the CPU oracle.

Bar: top-k ids and FP32 distances bit-identical to the oracle on the same index and
rotated queries (integer FastScan sums, code unpacking, pool order and the literal
float formulas all match); probe lists identical; rotation within the FP32 GEMM bound.
"""
import numpy as np
import pytest

from conftest import build_cached, gpu_available, rotate_np

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _torch():
    import torch
    return torch


def _build(key, n, dim, ncl, bits, seed=1, nq=48):
    from erq import data, index

    def make():
        x, cen = data.mixture_isotropic(n, dim, ncl, 0.3, seed=seed)
        q, _ = data.mixture_isotropic(nq, dim, ncl, 0.3, seed=seed + 1, centres=cen)
        c, assign = data.kmeans(x, ncl, 5, seed=seed + 3)
        idx = index.build_index(x, c, assign, bits=bits, seed=seed + 2)
        return idx, q
    return build_cached(key, make)


def _oracle_run(idx, qr, k, nprobe):
    from oracle import oracle as orc
    oi = orc.OracleIndex(idx)
    ids = np.empty((qr.shape[0], k), np.uint32)
    ds = np.empty((qr.shape[0], k), np.float32)
    pids = np.empty((qr.shape[0], nprobe), np.uint32)
    pds = np.empty((qr.shape[0], nprobe), np.float32)
    for i in range(qr.shape[0]):
        pids[i], pds[i] = oi.centroid_topn(qr[i], nprobe)
        ids[i], ds[i] = oi.search_probes(qr[i], k, pids[i], pds[i])
    return ids, ds, pids, pds


def _gpu_run(idx, qr, k, nprobe, mode=0):
    from erq.gpu import IVF
    torch = _torch()
    g = IVF.from_index(idx)
    g.set_plain_kernel(mode)
    q = torch.from_numpy(qr).cuda()
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    ids, d = g.search_dev(q, k, nprobe, stats=stats)
    pid, pd = g.probe_dev(q, nprobe)
    torch.cuda.synchronize()
    out = (ids.cpu().numpy().view(np.uint32), d.cpu().numpy(), pid.cpu().numpy().view(np.uint32),
           pd.cpu().numpy(), stats.cpu().numpy())
    g.close()
    return out


def _assert_parity(idx, q, k, nprobe, label):
    """Both integer kernels of the per-query scan (LUT, int8 MFMA) against the oracle."""
    qr = rotate_np(idx, q)
    oid, od, opid, opd = _oracle_run(idx, qr, k, nprobe)
    for mode, name in ((1, "lut"), (2, "mfma")):
        gid, gd, gpid, gpd, st = _gpu_run(idx, qr, k, nprobe, mode)
        lab = f"{label} [{name}]"
        print(f"[{lab}] k={k} nprobe={nprobe} stats={st.tolist()}", flush=True)
        np.testing.assert_array_equal(gpid, opid, err_msg=f"{lab}: probe ids")
        np.testing.assert_array_equal(gpd.view(np.uint32), opd.view(np.uint32),
                                      err_msg=f"{lab}: probe distances")
        np.testing.assert_array_equal(gid, oid, err_msg=f"{lab}: result ids")
        np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32),
                                      err_msg=f"{lab}: result distances")
        assert st[3] == q.shape[0] * nprobe


@pytest.mark.parametrize("bits", [3, 4, 5, 7, 8, 9])
def test_scan_parity_all_ex_layouts(bits):
    idx, q = _build(("d128", bits), 20000, 128, 64, bits)
    _assert_parity(idx, q, 100, 8, f"D128 B{bits}")


@pytest.mark.parametrize("k,nprobe", [(1, 1), (10, 4), (100, 16), (200, 64), (64, 3)])
def test_scan_parity_k_nprobe(k, nprobe):
    idx, q = _build(("d128", 4), 20000, 128, 64, 4)
    _assert_parity(idx, q, k, nprobe, "D128 B4")


def test_scan_parity_d1024_b8():
    idx, q = _build(("d960", 8), 6000, 960, 32, 8, nq=24)
    _assert_parity(idx, q, 100, 8, "D960->1024 B8")


@pytest.mark.parametrize("bits", [4, 9])
def test_scan_parity_d1536(bits):
    idx, q = _build(("d1536", bits), 6000, 1536, 32, bits, nq=24)
    _assert_parity(idx, q, 100, 6, f"D1536 B{bits}")


def test_scan_parity_d3072_wrap_path():
    idx, q = _build(("d3072", 4), 3000, 3072, 16, 4, nq=16)
    _assert_parity(idx, q, 50, 4, "D3072 B4")


def test_scan_with_given_probe_lists_uneven():
    """erq_scan_dev with per-query probe lists of different lengths (CSR offsets)."""
    from erq.gpu import IVF
    from oracle import oracle as orc
    torch = _torch()
    idx, q = _build(("d128", 4), 20000, 128, 64, 4)
    qr = rotate_np(idx, q)
    oi = orc.OracleIndex(idx)
    rng = np.random.default_rng(7)
    npq = rng.integers(1, 20, size=qr.shape[0])
    off = np.concatenate([[0], np.cumsum(npq)]).astype(np.uint32)
    pid = np.empty(off[-1], np.uint32)
    pd = np.empty(off[-1], np.float32)
    oid = np.empty((qr.shape[0], 50), np.uint32)
    od = np.empty((qr.shape[0], 50), np.float32)
    for i in range(qr.shape[0]):
        pid[off[i]:off[i + 1]], pd[off[i]:off[i + 1]] = oi.centroid_topn(qr[i], int(npq[i]))
        oid[i], od[i] = oi.search_probes(qr[i], 50, pid[off[i]:off[i + 1]],
                                         pd[off[i]:off[i + 1]])
    g = IVF.from_index(idx)
    t = lambda a: torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).cuda()
    ids, d = g.scan_dev(torch.from_numpy(qr).cuda(), 50, t(off), t(pid), t(pd))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), oid)
    np.testing.assert_array_equal(d.cpu().numpy().view(np.uint32), od.view(np.uint32))
    g.close()


def test_tiny_and_ragged_clusters():
    """Clusters of 0, 1, 31, 32, 33 and 95 vectors (tail masks, empty clusters)."""
    from erq import index
    rng = np.random.default_rng(3)
    sizes = [0, 1, 31, 32, 33, 95, 7, 64]
    dim = 64
    cen = rng.standard_normal((len(sizes), dim)).astype(np.float32) * 3
    xs, lab = [], []
    for k, s in enumerate(sizes):
        xs.append(cen[k] + 0.5 * rng.standard_normal((s, dim)).astype(np.float32))
        lab += [k] * s
    x = np.concatenate(xs)
    perm = rng.permutation(x.shape[0])
    x, lab = x[perm], np.asarray(lab)[perm]
    idx = index.build_index(x, cen, lab, bits=4, seed=5)
    q = cen[rng.integers(0, len(sizes), 20)] + rng.standard_normal((20, dim)).astype(np.float32)
    _assert_parity(idx, q, 40, len(sizes), "ragged")
    _assert_parity(idx, q, 300, len(sizes), "ragged k>N")


def test_duplicate_vectors_tie_order():
    """Every vector 24 times over: equal codes give equal distances, so the pool's order
    among ties (a new entry before equal older ones) decides which ids are kept."""
    from erq import data, index
    base, cen = data.mixture_isotropic(300, 128, 6, 0.3, seed=21)
    x = np.repeat(base, 24, axis=0)
    perm = np.random.default_rng(22).permutation(x.shape[0])
    x = x[perm]
    c, assign = data.kmeans(x, 6, 4, seed=23)
    idx = index.build_index(x, c, assign, bits=4, seed=24)
    q, _ = data.mixture_isotropic(24, 128, 6, 0.3, seed=25, centres=cen)
    _assert_parity(idx, q, 100, 3, "duplicates")


def test_host_pointer_search_matches_batch():
    """erq_search (IVF::search drop-in, host pointers) equals the device batch path."""
    from erq.gpu import IVF
    idx, q = _build(("d128", 4), 20000, 128, 64, 4)
    qr = rotate_np(idx, q)
    g = IVF.from_index(idx)
    bids, bd = g.search_batch(qr, 100, 8)
    for i in range(4):
        ids, d = g.search(qr[i], 100, 8, with_dists=True)
        np.testing.assert_array_equal(ids, bids[i])
        np.testing.assert_array_equal(d.view(np.uint32), bd[i].view(np.uint32))
    g.close()


def test_rotation_gemm_bound():
    from erq.gpu import IVF
    torch = _torch()
    idx, q = _build(("d960", 8), 6000, 960, 32, 8, nq=24)
    g = IVF.from_index(idx)
    rng = np.random.default_rng(11)
    qq = rng.standard_normal((300, 960)).astype(np.float32)
    out = g.rotate_dev(torch.from_numpy(qq).cuda())
    torch.cuda.synchronize()
    out = out.cpu().numpy()
    qp = np.zeros((300, idx.D))
    qp[:, :960] = qq
    ref = qp @ idx.P.astype(np.float64)
    bound = idx.D * 2.0 ** -23 * (np.abs(qp) @ np.abs(idx.P.astype(np.float64)))
    assert np.all(np.abs(out - ref) <= bound + 1e-30)
    host = g.rotate(qq[:3])
    np.testing.assert_allclose(host, ref[:3], rtol=0, atol=float(bound[:3].max()))
    g.close()


def test_centroid_distance_gemm():
    """The MFMA contraction ||q||^2 + ||c||^2 - 2 q.c against FP64."""
    from erq.gpu import IVF
    torch = _torch()
    idx, q = _build(("d1536", 4), 6000, 1536, 32, 4, nq=24)
    qr = rotate_np(idx, q)
    g = IVF.from_index(idx)
    a = g.centroid_dist_dev(torch.from_numpy(qr).cuda())
    torch.cuda.synchronize()
    a = a.cpu().numpy()
    C = idx.centroids.astype(np.float64)
    Q = qr.astype(np.float64)
    ref = (Q * Q).sum(1)[:, None] + (C * C).sum(1)[None, :] - 2 * Q @ C.T
    scale = (Q * Q).sum(1)[:, None] + (C * C).sum(1)[None, :]
    assert np.all(np.abs(a - ref) <= (4 * idx.D + 64) * 2.0 ** -24 * scale)
    g.close()


def test_query_order_does_not_change_results():
    """ERQ_OPT_PLAIN_ORDER (queries taken by scan cost, heaviest first) only reorders
    work: ids, distances and counters equal those of the order as given."""
    from erq.gpu import IVF
    torch = _torch()
    idx, q = _build(("d128", 4), 20000, 128, 64, 4)
    qr = torch.from_numpy(rotate_np(idx, q)).cuda()
    g = IVF.from_index(idx)
    out = []
    for order in (0, 1):
        g.set_option(5, order)
        st = torch.zeros(8, dtype=torch.int64, device="cuda")
        ids, d = g.search_dev(qr, 100, 12, stats=st)
        torch.cuda.synchronize()
        out.append((ids.cpu(), d.cpu(), st.cpu()))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
    assert torch.equal(out[0][2], out[1][2])
    g.close()


def test_host_batch_with_per_query_nprobe():
    """erq_search_batch_nprobe (host pointers, each query its own nprobe): the exact replay
    equals IVF::search per query; the batched flag equals the oracle's round schedule."""
    from erq.gpu import IVF
    from oracle import oracle as orc
    idx, q = _build(("d128", 4), 20000, 128, 64, 4)
    qr = rotate_np(idx, q)
    rng = np.random.default_rng(17)
    npq = rng.integers(1, 40, qr.shape[0]).astype(np.uint32)
    oi = orc.OracleIndex(idx)
    g = IVF.from_index(idx)
    print("[per-query nprobe] exact replay", flush=True)
    ids, d = g.search_batch_nprobe(qr, 50, npq)
    for i in range(qr.shape[0]):
        oid, od = oi.search(qr[i], 50, int(npq[i]))
        np.testing.assert_array_equal(ids[i], oid, err_msg=f"query {i}")
        np.testing.assert_array_equal(d[i].view(np.uint32), od.view(np.uint32))
    print("[per-query nprobe] batched rounds", flush=True)
    bids, bd = g.search_batch_nprobe(qr, 50, npq, batched=True)
    for i in range(qr.shape[0]):
        pid, pd = oi.centroid_topn(qr[i], int(npq[i]))
        oid, od = oi.search_rounds(qr[i], 50, pid, pd)
        np.testing.assert_array_equal(bids[i], oid, err_msg=f"query {i}")
        np.testing.assert_array_equal(bd[i].view(np.uint32), od.view(np.uint32))
    # repeated calls reuse the handle's device buffers: same answers
    ids2, _ = g.search_batch_nprobe(qr, 50, npq)
    np.testing.assert_array_equal(ids2, ids)
    g.close()
