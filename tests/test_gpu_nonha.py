"""GPU parity of the non-HA searcher (Searcher.hpp:14-236, the scan IVF::search runs when the
reference is built without HIGH_ACC_FAST_SCAN): the HIP kernel of erq_nonha.hip through the
C-ABI against the oracle's orc_nonha_search_probes, on the same index, rotated queries and
probe lists.

Both integer kernels (the u16 LUT, and the int8 MFMA over the bit-plane codes up to D = 2048)
are run where they apply.
Bar: top-k ids and FP32 distances bit-identical (the FastScan integers of accumulate_robust,
the unsigned 14-bit digits, the four-factor lower bound, the per-vector running threshold and
the pool order), and the same block / vector / survivor counts.  Shapes: every EX at D = 128,
D = 1000 -> 1024 with EX = 7, D = 1536, D = 3072 (natural data, and a crafted index whose
half-sums wrap mod 2^16), ragged clusters, k = 1 .. 512, a file written by IVFIndex.save
read back by erq_index_load, and the batched entry points refusing the format."""
import numpy as np
import pytest

from conftest import build_cached, gpu_available, rotate_np

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _build(key, n, dim, ncl, bits, seed=1, nq=24, sizes=None):
    from erq import data, index

    def make():
        x, cen = data.mixture_isotropic(n, dim, ncl, 0.3, seed=seed)
        q, _ = data.mixture_isotropic(nq, dim, ncl, 0.3, seed=seed + 1, centres=cen)
        c, assign = data.kmeans(x, ncl, 5, seed=seed + 3)
        if sizes is not None:  # force the cluster sizes: vectors reassigned in order
            assign = np.repeat(np.arange(len(sizes)), sizes)[:n].astype(np.int64)
        idx = index.build_index(x, c, assign, bits=bits, seed=seed + 2, high_acc=False)
        return idx, q
    return build_cached(("nonha",) + key, make)


def _oracle(idx, qr, k, nprobe):
    from oracle import oracle as orc
    oi = orc.OracleIndex(idx)
    nq = qr.shape[0]
    ids = np.empty((nq, k), np.uint32)
    ds = np.empty((nq, k), np.float32)
    tot = np.zeros(3, np.int64)
    for i in range(nq):
        ids[i], ds[i], st = oi.search(qr[i], k, nprobe, stats=True)
        tot += np.asarray(st, np.int64)
    return ids, ds, tot


def _gpu(idx, qr, k, nprobe, mode=0):
    import torch
    from erq.gpu import IVF
    g = IVF.from_index(idx)
    assert not g.high_acc()
    g.set_plain_kernel(mode)   # 0 auto (int8 MFMA up to D = 2048), 1 LUT
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    ids, d = g.search_dev(torch.from_numpy(qr).cuda(), k, nprobe, stats=stats)
    torch.cuda.synchronize()
    out = ids.cpu().numpy().view(np.uint32), d.cpu().numpy(), stats.cpu().numpy()
    g.close()
    return out


def _check(idx, qr, k, nprobe, tag, modes=(0,)):
    oid, od, ost = _oracle(idx, qr, k, nprobe)
    for mode in modes:
        print(f"[nonha {tag} D={idx.D} EX={idx.EX} k={k} nprobe={nprobe} kernel {mode}]",
              flush=True)
        gid, gd, gst = _gpu(idx, qr, k, nprobe, mode)
        np.testing.assert_array_equal(gid, oid, err_msg=tag)
        np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32), err_msg=tag)
        assert tuple(gst[:3]) == tuple(ost), (tag, mode, gst.tolist(), ost.tolist())


@pytest.mark.parametrize("bits", [3, 4, 5, 7, 8, 9])
def test_nonha_every_ex_d128(bits):
    idx, q = _build(("d128", bits), 6000, 128, 24, bits)
    qr = rotate_np(idx, q)
    for k, npb in ((1, 3), (10, 6), (100, 12)):
        _check(idx, qr, k, npb, f"d128 b{bits}", modes=(0, 1))


def test_nonha_d1000_ex7():
    idx, q = _build(("d1000",), 4000, 1000, 16, 8)
    assert idx.D == 1024  # zero-padded dims
    _check(idx, rotate_np(idx, q), 50, 8, "d1000", modes=(0, 1))


def test_nonha_d1536_ex3():
    idx, q = _build(("d1536",), 4000, 1536, 16, 4)
    _check(idx, rotate_np(idx, q), 100, 10, "d1536", modes=(0, 1))


def test_nonha_d3072_ex8_natural():
    idx, q = _build(("d3072",), 2000, 3072, 8, 9, nq=8)
    _check(idx, rotate_np(idx, q), 10, 4, "d3072")


def test_nonha_ragged_clusters_and_large_k():
    sizes = [0, 1, 31, 32, 33, 95, 64, 744]
    idx, q = _build(("ragged",), sum(sizes), 64, len(sizes), 4, sizes=sizes)
    assert list(idx.cluster_sizes) == sizes
    qr = rotate_np(idx, q)
    for k, npb in ((1, 8), (7, 3), (300, 8), (1024, 8)):
        _check(idx, qr, k, npb, "ragged", modes=(0, 1))


@pytest.mark.parametrize("k", [20, 512])
def test_nonha_wrap_d3072_crafted(k):
    """Half-sums that wrap mod 2^16 (low and high byte planes, per codebook parity): at k =
    512 >= N every vector enters the pool, so every integer is checked through its distance."""
    from crafted import wrap_index_nonha
    from oracle import oracle as orc
    idx, q, bits = wrap_index_nonha()
    # some vectors' reference integers differ from the exact inner product with the digits
    r = orc.minus_sum16(q[0], idx.centroids[0])[1]
    vl, vr = min(0.0, float(r.min())), max(0.0, float(r.max()))
    width = float(np.float32(np.float32(vr) - np.float32(vl)) / np.float32(16383))
    dig = orc.scalar_quantize16(r, vl, width)
    lut = orc.pack_lut_u16(dig)
    codes = orc.pack_block_codes(bits[:32], idx.D)
    ints = orc.nonha_block_ints(codes, lut, idx.D).astype(np.int64)
    exact = bits[:32].astype(np.int64) @ dig.astype(np.int64)
    assert np.any(ints != exact) and np.any(ints == exact)
    _check(idx, q, k, idx.K, "wrap")


def test_nonha_file_load_and_batched_refusal(tmp_path):
    import torch
    from erq import _lib
    from erq.gpu import IVF
    idx, q = _build(("d128", 5), 6000, 128, 24, 5)
    path = tmp_path / "nonha.index"
    idx.save(path)
    g = IVF.load(path)
    assert not g.high_acc()
    qr = rotate_np(idx, q)
    oid, od, _ = _oracle(idx, qr, 40, 5)
    for i in range(4):  # the host-pointer drop-in, one query at a time
        ids, d = g.search(qr[i], 40, 5, with_dists=True)
        np.testing.assert_array_equal(ids, oid[i])
        np.testing.assert_array_equal(d.view(np.uint32), od[i].view(np.uint32))
    qd = torch.from_numpy(qr).cuda()
    with pytest.raises(RuntimeError, match="EUNSUPPORTED"):
        g.search_batched_dev(qd, 40, 5)
    torch.cuda.synchronize()
    assert "non-HA" in _lib.lib().erq_last_error().decode()
    g.close()
