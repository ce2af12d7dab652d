"""The largest workload's shape at reduced N: D = 3072 unit-norm power-law mixture vectors,
8192 clusters (some empty), B = 9 (EX = 8), k = 10 and 100.  The per-query replay (LUT and
int8-MFMA kernels) and the batched rounds against the oracle, bit for bit."""
import numpy as np
import pytest

from conftest import build_cached, gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _c5():
    from erq import data, index

    def make():
        rng = np.random.default_rng(5)
        N, dim, K = 24576, 3072, 8192
        x, mix = data.mixture_powerlaw(N + 12, dim, 512, 0.5, seed=51, alpha=1.0)
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        q, x = x[N:], x[:N]
        assign = rng.integers(0, K, N)
        cen = np.zeros((K, dim), np.float32)
        np.add.at(cen, assign, x)
        cen /= np.maximum(np.bincount(assign, minlength=K), 1)[:, None]
        idx = index.build_index(x, cen, assign, bits=9, seed=52)
        return idx, idx_rotate(idx, q)
    return build_cached("c5", make)


def idx_rotate(idx, q):
    from conftest import rotate_np
    return rotate_np(idx, q)


@pytest.mark.parametrize("k", [10, 100])
def test_c5_shape_per_query(k):
    import torch
    from erq.gpu import IVF
    from oracle import oracle as orc
    idx, qr = _c5()
    nprobe = 128
    oi = orc.OracleIndex(idx)
    oid = np.empty((qr.shape[0], k), np.uint32)
    od = np.empty((qr.shape[0], k), np.float32)
    for i in range(qr.shape[0]):
        oid[i], od[i] = oi.search(qr[i], k, nprobe)
    g = IVF.from_index(idx)
    for mode, name in ((2, "mfma"), (1, "lut")):
        g.set_plain_kernel(mode)
        print(f"[C5 shape D3072 K8192 B9 k={k} nprobe={nprobe} {name}]", flush=True)
        ids, d = g.search_dev(torch.from_numpy(qr).cuda(), k, nprobe)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), oid, err_msg=name)
        np.testing.assert_array_equal(d.cpu().numpy().view(np.uint32), od.view(np.uint32),
                                      err_msg=name)
    g.close()


@pytest.mark.parametrize("k", [10, 100])
def test_c5_shape_batched(k):
    import torch
    from erq.gpu import IVF
    from oracle import oracle as orc
    idx, qr = _c5()
    nprobe = 200
    oi = orc.OracleIndex(idx)
    oid = np.empty((qr.shape[0], k), np.uint32)
    od = np.empty((qr.shape[0], k), np.float32)
    tot = np.zeros(3, np.int64)
    for i in range(qr.shape[0]):
        pid, pd = oi.centroid_topn(qr[i], nprobe)
        oid[i], od[i], st = oi.search_rounds(qr[i], k, pid, pd, stats=True)
        tot += np.asarray(st, np.int64)
    g = IVF.from_index(idx)
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    print(f"[C5 shape batched k={k} nprobe={nprobe}]", flush=True)
    ids, d = g.search_batched_dev(torch.from_numpy(qr).cuda(), k, nprobe, stats=stats)
    torch.cuda.synchronize()
    st = stats.cpu().numpy()
    np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), oid)
    np.testing.assert_array_equal(d.cpu().numpy().view(np.uint32), od.view(np.uint32))
    assert (st[0], st[2], st[4]) == tuple(int(v) for v in tot), (st.tolist(), tot.tolist())
    g.close()
