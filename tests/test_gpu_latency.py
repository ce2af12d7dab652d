"""The low-latency scan (ERQ_OPT_LATENCY_NQ: every probed vector estimated in parallel, one
wave per query replays the probes) against the oracle's IVF::search and against the per-query
scan's statistics, bit for bit."""
import numpy as np
import pytest

from conftest import gpu_available, rotate_np
from test_gpu_parity import _build, _oracle_run

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

OPT_LATENCY_NQ = 8


def _torch():
    import torch
    return torch


def _run(idx, qr, k, nprobe, latency_nq):
    from erq.gpu import IVF
    torch = _torch()
    g = IVF.from_index(idx)
    g.set_option(OPT_LATENCY_NQ, latency_nq)
    q = torch.from_numpy(qr).cuda()
    st = torch.zeros(8, dtype=torch.int64, device="cuda")
    ids, d = g.search_dev(q, k, nprobe, stats=st)
    torch.cuda.synchronize()
    out = ids.cpu().numpy().view(np.uint32), d.cpu().numpy(), st.cpu().numpy()
    g.close()
    return out


def _check(idx, q, k, nprobe, label):
    qr = rotate_np(idx, q)
    oid, od, _, _ = _oracle_run(idx, qr, k, nprobe)
    lid, ld, lst = _run(idx, qr, k, nprobe, 4096)   # every batch size: the latency scan
    sid, sd, sst = _run(idx, qr, k, nprobe, 0)      # the per-query scan
    print(f"[{label}] k={k} nprobe={nprobe} stats latency={lst.tolist()} plain={sst.tolist()}",
          flush=True)
    np.testing.assert_array_equal(lid, oid, err_msg=f"{label}: ids")
    np.testing.assert_array_equal(ld.view(np.uint32), od.view(np.uint32), err_msg=f"{label}: dists")
    np.testing.assert_array_equal(lst[:4], sst[:4], err_msg=f"{label}: blocks/vectors/survivors/probes")


@pytest.mark.parametrize("bits", [3, 4, 5, 7, 8, 9])
def test_latency_all_ex_layouts(bits):
    idx, q = _build(("d128", bits), 20000, 128, 64, bits)
    _check(idx, q, 100, 8, f"D128 B{bits}")


@pytest.mark.parametrize("k,nprobe", [(1, 1), (10, 4), (100, 16), (200, 64), (64, 3)])
def test_latency_k_nprobe(k, nprobe):
    idx, q = _build(("d128", 4), 20000, 128, 64, 4)
    _check(idx, q, k, nprobe, "D128 B4")


def test_latency_d1024_b8():
    idx, q = _build(("d960", 8), 6000, 960, 32, 8, nq=24)
    _check(idx, q, 100, 8, "D960->1024 B8")


@pytest.mark.parametrize("bits", [4, 9])
def test_latency_d1536(bits):
    idx, q = _build(("d1536", bits), 6000, 1536, 32, bits, nq=24)
    _check(idx, q, 100, 6, f"D1536 B{bits}")


def test_latency_host_single_query_and_default():
    """erq_search (one query, host pointers) takes the latency scan by default (16)."""
    from erq.gpu import IVF
    idx, q = _build(("d128", 4), 20000, 128, 64, 4)
    qr = rotate_np(idx, q)
    oid, od, _, _ = _oracle_run(idx, qr[:5], 50, 12)
    g = IVF.from_index(idx)
    assert g.option(OPT_LATENCY_NQ) == 16
    for i in range(5):
        ids, d = g.search(qr[i], 50, 12, with_dists=True)
        np.testing.assert_array_equal(np.asarray(ids).view(np.uint32), oid[i])
        np.testing.assert_array_equal(np.asarray(d, np.float32).view(np.uint32),
                                      od[i].view(np.uint32))
    g.close()


def test_latency_wide_d_falls_back():
    """Above D = 2048 the small batch takes the per-query scan (same results)."""
    idx, q = _build(("d3072", 4), 2000, 3072, 16, 4, nq=4)
    qr = rotate_np(idx, q)
    oid, od, _, _ = _oracle_run(idx, qr, 50, 4)
    lid, ld, _ = _run(idx, qr, 50, 4, 4096)
    np.testing.assert_array_equal(lid, oid)
    np.testing.assert_array_equal(ld.view(np.uint32), od.view(np.uint32))
