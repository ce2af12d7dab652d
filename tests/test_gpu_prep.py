"""GPU dataset preparation (erq_knn.hip through the C-ABI) against the oracle: the ground truth
of python/compute_gt.py and the k-means of python/ivf.py, bit for bit."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _torch():
    import torch
    return torch


def _mix(n, d, ncl, sigma, seed, centres=None):
    from erq import data
    x, cen = data.mixture_isotropic(n, d, ncl, sigma, seed=seed, centres=centres)
    return x.astype(np.float32), cen


@pytest.mark.parametrize("n,dim,nq,k", [(20000, 100, 64, 100),   # padded rows (100 -> 112)
                                       (6000, 1536, 16, 1000),   # wide rows, k = 1000
                                       (20000, 256, 8, 4096),    # > 4096 candidates: keys in HBM
                                       (300, 64, 5, 300)])       # k = n: the whole row
def test_knn_matches_oracle(n, dim, nq, k):
    from erq import gpu
    from oracle import oracle as O
    torch = _torch()
    print(f"knn n={n} dim={dim} nq={nq} k={k}", flush=True)
    x, cen = _mix(n, dim, 32, 0.5, seed=n + dim)
    q, _ = _mix(nq, dim, 32, 0.5, seed=n + dim + 1, centres=cen)
    ids, d = gpu.knn_dev(torch.from_numpy(x).cuda(), torch.from_numpy(q).cuda(), k)
    want_i, want_d = O.knn(x, q, k)
    np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), want_i)
    np.testing.assert_array_equal(d.cpu().numpy(), want_d)


def test_knn_duplicates_tie_by_id():
    from erq import gpu
    from oracle import oracle as O
    torch = _torch()
    rng = np.random.default_rng(5)
    x = rng.standard_normal((5000, 32)).astype(np.float32)
    x[1000:3000] = x[7]              # 2000 copies of row 7 plus row 7 itself
    q = (x[7] + 0.01).astype(np.float32)[None, :].repeat(3, 0)
    ids, d = gpu.knn_dev(torch.from_numpy(x).cuda(), torch.from_numpy(q).cuda(), 1500)
    want_i, want_d = O.knn(x, q, 1500)
    np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), want_i)
    np.testing.assert_array_equal(d.cpu().numpy(), want_d)
    assert want_i[0, 0] == 7 and want_i[0, 1] == 1000


def test_kmeans_assign_and_update_match_oracle():
    from erq import gpu
    from oracle import oracle as O
    torch = _torch()
    x, cen = _mix(30000, 100, 64, 0.4, seed=3)
    c = (cen[:64] + 0.05).astype(np.float32)
    c[63] = 1e3                           # a centroid nobody is nearest to
    xd = torch.from_numpy(x).cuda()
    a, d = gpu.kmeans_assign_dev(xd, torch.from_numpy(c).cuda())
    wa, wd = O.kmeans_assign(x, c)
    np.testing.assert_array_equal(a.cpu().numpy().view(np.uint32), wa)
    np.testing.assert_array_equal(d.cpu().numpy(), wd)
    cd = torch.from_numpy(c).cuda()
    cnt = gpu.kmeans_update_dev(xd, a, cd)
    wc, wcnt = O.kmeans_update(x, wa, c)
    assert wcnt[63] == 0
    np.testing.assert_array_equal(cnt.cpu().numpy().view(np.uint32), wcnt)
    np.testing.assert_array_equal(cd.cpu().numpy(), wc)


@pytest.mark.parametrize("n,dim,K,iters", [(20000, 64, 64, 5), (12000, 200, 40, 3)])
def test_kmeans_matches_oracle(n, dim, K, iters):
    from erq import gpu
    from oracle import oracle as O
    torch = _torch()
    print(f"kmeans n={n} dim={dim} K={K} iters={iters}", flush=True)
    x, _ = _mix(n, dim, K, 0.5, seed=dim)
    init = gpu.kmeans_init_rows(n, K, 4)
    c, a, d = gpu.kmeans_dev(torch.from_numpy(x).cuda(), K, iters, init)
    wc, wa, wd = O.kmeans(x, K, iters, init)
    np.testing.assert_array_equal(c.cpu().numpy(), wc)
    np.testing.assert_array_equal(a.cpu().numpy().view(np.uint32), wa)
    np.testing.assert_array_equal(d.cpu().numpy(), wd)


def test_dataset_files(tmp_path):
    """python/ivf.py + python/compute_gt.py + src/create_index.cpp over a dataset directory."""
    from erq import IVFIndex, prep, vecio
    from oracle import oracle as O
    x, cen = _mix(3000, 40, 16, 0.4, seed=7)
    q, _ = _mix(20, 40, 16, 0.4, seed=8, centres=cen)
    vecio.write_fvecs(tmp_path / "toy_base.fvecs", x)
    vecio.write_fvecs(tmp_path / "toy_query.fvecs", q)
    prep.gt_files(str(tmp_path), "toy", k=100)
    gt = vecio.read_ivecs(tmp_path / "toy_groundtruth.ivecs")
    np.testing.assert_array_equal(gt.view(np.uint32), O.knn(x, q, 100)[0])
    prep.ivf_files(str(tmp_path), "toy", 16, iters=8, seed=3)
    c = vecio.read_fvecs(tmp_path / "toy_centroid_16.fvecs")
    cid = vecio.read_ivecs(tmp_path / "toy_cluster_id_16.ivecs")
    dist = vecio.read_fvecs(tmp_path / "toy_dist_to_centroid_16.fvecs")
    from erq import gpu
    wc, wa, wd = O.kmeans(x, 16, 8, gpu.kmeans_init_rows(3000, 16, 3))
    np.testing.assert_array_equal(c, wc)
    np.testing.assert_array_equal(cid[:, 0].view(np.uint32), wa)
    np.testing.assert_array_equal(dist[:, 0], np.sqrt(wd))
    prep.create_index_files(str(tmp_path), "toy", 16, 5, seed=2)
    back = IVFIndex.load(tmp_path / "ivf_exhaf5.index")
    assert back.N == 3000 and back.K == 16
