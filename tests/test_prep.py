"""Dataset preparation (python/ivf.py, python/compute_gt.py, src/create_index.cpp) over the
dataset directory's files, on the CPU."""
import numpy as np

from erq import IVFIndex, data, prep, vecio


def _dataset(tmp_path, n=3000, nq=20, d=40, seed=7):
    x, cen = data.mixture_isotropic(n, d, 16, 0.4, seed=seed)
    q, _ = data.mixture_isotropic(nq, d, 16, 0.4, seed=seed + 1, centres=cen)
    vecio.write_fvecs(tmp_path / "toy_base.fvecs", x)
    vecio.write_fvecs(tmp_path / "toy_query.fvecs", q)
    return x.astype(np.float32), q.astype(np.float32)


def test_compute_gt_matches_fp64(tmp_path):
    x, q = _dataset(tmp_path)
    prep.gt_files(str(tmp_path), "toy", k=100, device="cpu")
    gt = vecio.read_ivecs(tmp_path / "toy_groundtruth.ivecs")
    assert gt.shape == (20, 100)
    d64 = ((q[:, None, :].astype(np.float64) - x[None, :, :]) ** 2).sum(-1)
    want = np.sort(d64, axis=1)[:, :100]
    got = np.take_along_axis(d64, gt.astype(np.int64), 1)
    # ranks agree up to FP32 near-ties: the distance at each rank is the true one
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-4)
    assert all(len(set(r)) == 100 for r in gt.tolist())


def test_ivf_files_and_index(tmp_path):
    x, _ = _dataset(tmp_path)
    prep.ivf_files(str(tmp_path), "toy", 16, iters=8, seed=3, device="cpu")
    c = vecio.read_fvecs(tmp_path / "toy_centroid_16.fvecs")
    cid = vecio.read_ivecs(tmp_path / "toy_cluster_id_16.ivecs")
    dist = vecio.read_fvecs(tmp_path / "toy_dist_to_centroid_16.fvecs")
    assert c.shape == (16, 40) and cid.shape == (3000, 1) and dist.shape == (3000, 1)
    d64 = ((x[:, None, :].astype(np.float64) - c[None, :, :].astype(np.float64)) ** 2).sum(-1)
    own = d64[np.arange(3000), cid[:, 0]]
    assert np.all(own <= d64.min(1) + 1e-4)                 # nearest centroid
    np.testing.assert_allclose(dist[:, 0], np.sqrt(own), rtol=1e-4, atol=1e-5)
    idx = prep.create_index_files(str(tmp_path), "toy", 16, 5, seed=2)
    back = IVFIndex.load(tmp_path / "ivf_exhaf5.index")
    assert back.EX == 4 and back.N == 3000 and back.DIM == 40 and back.D == 64
    np.testing.assert_array_equal(back.cluster_sizes,
                                  np.bincount(cid[:, 0], minlength=16).astype(np.uint64))
    np.testing.assert_array_equal(back.long_code, idx.long_code)
