"""GPU parity where the reference's u16 FastScan lanes wrap (D = 3072, FastScan.hpp:264-290),
on the hand-built index of tests/crafted.py (wrapped and unwrapped vectors mixed, wrapped
vectors in every result): the per-query LUT kernel (which reduces each half mod 2^16), the
per-query int8-MFMA kernel and the batched rounds (both of which correct the exact integer
of every vector the lower bound passes) against the oracle, bit for bit, with the same
survivor counts."""
import numpy as np
import pytest

from conftest import gpu_available
from crafted import wrap_index

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


@pytest.mark.parametrize("k", [10, 20])
def test_wrap_per_query_lut_and_mfma_match_oracle(k):
    import torch
    from erq.gpu import IVF
    from oracle import oracle as orc
    idx, q, _ = wrap_index()
    oi = orc.OracleIndex(idx)
    oid = np.empty((q.shape[0], k), np.uint32)
    od = np.empty((q.shape[0], k), np.float32)
    osurv = 0
    for i in range(q.shape[0]):
        oid[i], od[i], st = oi.search(q[i], k, idx.K, stats=True)
        osurv += st[2]
    g = IVF.from_index(idx)
    qd = torch.from_numpy(q).cuda()
    for mode, name in ((1, "lut"), (2, "mfma")):
        g.set_plain_kernel(mode)
        stats = torch.zeros(8, dtype=torch.int64, device="cuda")
        print(f"[wrap D3072 k={k} {name}]", flush=True)
        ids, d = g.search_dev(qd, k, idx.K, stats=stats)
        torch.cuda.synchronize()
        st = stats.cpu().numpy()
        np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), oid, err_msg=name)
        np.testing.assert_array_equal(d.cpu().numpy().view(np.uint32), od.view(np.uint32),
                                      err_msg=name)
        assert st[2] == osurv, (name, st.tolist(), osurv)
    g.close()


def test_wrap_batched_rounds_match_oracle():
    import torch
    from erq.gpu import IVF
    from oracle import oracle as orc
    idx, q, _ = wrap_index()
    oi = orc.OracleIndex(idx)
    k = 20
    oid = np.empty((q.shape[0], k), np.uint32)
    od = np.empty((q.shape[0], k), np.float32)
    tot = np.zeros(3, np.int64)
    for i in range(q.shape[0]):
        pid, pd = oi.centroid_topn(q[i], idx.K)
        oid[i], od[i], st = oi.search_rounds(q[i], k, pid, pd, stats=True)
        tot += np.asarray(st, np.int64)
    g = IVF.from_index(idx)
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    print("[wrap D3072 batched]", flush=True)
    ids, d = g.search_batched_dev(torch.from_numpy(q).cuda(), k, idx.K, stats=stats)
    torch.cuda.synchronize()
    st = stats.cpu().numpy()
    np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), oid)
    np.testing.assert_array_equal(d.cpu().numpy().view(np.uint32), od.view(np.uint32))
    assert (st[0], st[2], st[4]) == tuple(int(v) for v in tot), (st.tolist(), tot.tolist())
    g.close()
