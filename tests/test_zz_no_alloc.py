"""erq_search_dev_noalloc (the search whose intermediates live in the library's fixed
per-device scratch array) against erq_search_dev, through the C-ABI: bit-identical ids,
distances and counters on two calls in a row, the second with fewer queries (the scratch
keeps the first call's contents), and no device memory taken by the calls."""
import ctypes

import numpy as np
import pytest

from conftest import build_cached, gpu_available, rotate_np

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _index():
    from erq import data, index

    def make():
        x, cen = data.mixture_isotropic(20000, 128, 64, 0.3, seed=61)
        q, _ = data.mixture_isotropic(300, 128, 64, 0.3, seed=62, centres=cen)
        c, assign = data.kmeans(x, 64, 5, seed=63)
        idx = index.build_index(x, c, assign, bits=4, seed=64)
        return idx, rotate_np(idx, q)
    return build_cached("noalloc", make)


def _run(fn, g, qd, k, nprobe, torch):
    from erq._lib import check
    from erq.gpu import _ptr, _stream
    nq = qd.shape[0]
    ids = torch.empty((nq, k), dtype=torch.int32, device="cuda")
    d = torch.empty((nq, k), dtype=torch.float32, device="cuda")
    st = torch.zeros(8, dtype=torch.int64, device="cuda")
    check(fn(g._h, _ptr(qd), nq, k, nprobe, _ptr(ids), _ptr(d), _ptr(st), _stream()))
    torch.cuda.synchronize()
    return ids.cpu().numpy(), d.cpu().numpy(), st.cpu().numpy()


@pytest.mark.parametrize("plain_kernel", [0, 1])
def test_noalloc_search_equals_search_dev(plain_kernel):
    import torch
    from erq import _lib
    from erq.gpu import IVF
    L = _lib.lib()
    idx, qr = _index()
    g = IVF.from_index(idx)
    g.set_plain_kernel(plain_kernel)
    k, nprobe = 100, 12
    q1 = torch.from_numpy(qr).cuda()
    q2 = torch.from_numpy(qr[:37].copy()).cuda()
    for step, qd in enumerate((q1, q2)):
        print(f"[noalloc] call {step}: nq={qd.shape[0]} k={k} nprobe={nprobe} "
              f"kernel={plain_kernel}", flush=True)
        ref = _run(L.erq_search_dev, g, qd, k, nprobe, torch)
        free0 = torch.cuda.mem_get_info()[0]
        got = _run(L.erq_search_dev_noalloc, g, qd, k, nprobe, torch)
        free1 = torch.cuda.mem_get_info()[0]
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1].view(np.uint32), ref[1].view(np.uint32))
        np.testing.assert_array_equal(got[2], ref[2])
        assert free1 >= free0, (free0, free1)
    g.close()


def test_noalloc_rejects_beyond_capacity():
    import torch
    from erq import _lib
    from erq.gpu import IVF, _ptr, _stream
    L = _lib.lib()
    idx, qr = _index()
    g = IVF.from_index(idx)
    cap = L.erq_search_noalloc_capacity()
    per_q = L.erq_search_noalloc_bytes(g._h, 1000, 64) / 1000.0
    nq = int(cap / per_q) + 1000
    assert L.erq_search_noalloc_bytes(g._h, nq, 64) > cap
    qd = torch.zeros((nq, idx.D), dtype=torch.float32, device="cuda")
    ids = torch.empty((nq, 10), dtype=torch.int32, device="cuda")
    rc = L.erq_search_dev_noalloc(g._h, _ptr(qd), nq, 10, 64, _ptr(ids), None, None, _stream())
    assert rc == 6 and b"scratch" in L.erq_last_error()
    g.close()
