"""Dataset preparation (python/ivf.py, python/compute_gt.py, src/create_index.cpp) on the CPU:
the oracle's k-means and ground truth pinned against scipy / scikit-learn and against
order-separating known answers, and the dataset files written from the oracle's clusters."""
import numpy as np
import pytest
from scipy.spatial.distance import cdist

from erq import IVFIndex, data, prep, vecio
from oracle import oracle as O


def _mix(n, d, seed, centres=None):
    x, cen = data.mixture_isotropic(n, d, 16, 0.4, seed=seed, centres=centres)
    return x.astype(np.float32), cen


def test_knn_oracle_matches_scipy():
    x, cen = _mix(3000, 40, 7)
    q, _ = _mix(20, 40, 8, cen)
    ids, d = O.knn(x, q, 100)
    d64 = cdist(q.astype(np.float64), x.astype(np.float64), "sqeuclidean")
    want = np.argsort(d64, axis=1, kind="stable")[:, :100]
    np.testing.assert_array_equal(ids, want)
    np.testing.assert_allclose(d, np.take_along_axis(d64, want, 1), rtol=1e-6)


def test_knn_oracle_ties_by_id():
    x = np.zeros((50, 16), np.float32)
    x[10:] = 1.0                      # 40 equal rows at distance 16 from the origin
    x[:10] = 2.0                      # 10 equal rows at distance 64
    ids, d = O.knn(x, np.zeros((1, 16), np.float32), 45)
    np.testing.assert_array_equal(ids[0], list(range(10, 50)) + list(range(0, 5)))
    np.testing.assert_array_equal(d[0], [16.0] * 40 + [64.0] * 5)


def test_l2sqr16_f64_chain_and_butterfly_order():
    # lane chains: lane 0 holds 2^54 and lane 1 three ones -> 2^54 + 3 rounds to 2^54 + 4;
    # a sequential (or pairwise) sum over the dims loses every 1 -> 2^54
    y = np.zeros(48, np.float32)
    x = np.zeros(48, np.float32)
    x[0] = 2.0 ** 27
    x[1] = x[17] = x[33] = 1.0
    assert O.l2sqr16_f64(x, y) == 2.0 ** 54 + 4
    seq = 0.0
    for v in x.astype(np.float64):
        seq = seq + v * v
    assert seq == 2.0 ** 54
    # butterfly: lanes 8, 9, 10 hold one each; xor-8 first adds lane 8 to 2^54 (lost), the
    # rest meet 2^54 one at a time -> 2^54; pairwise over lanes would give (1 + 1) + 1 = 3
    x = np.zeros(16, np.float32)
    x[0] = 2.0 ** 27
    x[8] = x[9] = x[10] = 1.0
    assert O.l2sqr16_f64(x, np.zeros(16, np.float32)) == 2.0 ** 54


def test_kmeans_assign_oracle_matches_scipy_and_ties():
    x, cen = _mix(2000, 40, 3)
    c = cen[:12].astype(np.float32)
    a, d = O.kmeans_assign(x, c)
    d64 = cdist(x.astype(np.float64), c.astype(np.float64), "sqeuclidean")
    np.testing.assert_array_equal(a, d64.argmin(1))
    np.testing.assert_allclose(d, d64.min(1), rtol=1e-5)
    # a duplicated centroid: the lower id wins
    c2 = np.concatenate([c, c[3:4]])
    a2, _ = O.kmeans_assign(x, c2)
    assert not np.any(a2 == 12)
    np.testing.assert_array_equal(a2, a)


def test_kmeans_update_sums_in_id_order():
    # members 0, 1, 2 of cluster 0 hold 2^53, 1, -2^53 in dim 0: in id order 2^53 + 1 rounds
    # to 2^53 and the mean is 0; in reverse order the sum is 1
    x = np.zeros((4, 16), np.float32)
    x[0, 0], x[1, 0], x[2, 0] = 2.0 ** 53, 1.0, -(2.0 ** 53)
    x[3] = 5.0
    c = np.full((3, 16), 7.0, np.float32)
    new, cnt = O.kmeans_update(x, np.array([0, 0, 0, 1], np.uint32), c)
    np.testing.assert_array_equal(cnt, [3, 1, 0])
    assert new[0, 0] == 0.0
    np.testing.assert_array_equal(new[1], np.full(16, 5.0, np.float32))
    np.testing.assert_array_equal(new[2], c[2])          # empty: unchanged
    rev = 0.0
    for v in (-(2.0 ** 53), 1.0, 2.0 ** 53):
        rev = rev + v
    assert rev == 1.0


def test_kmeans_oracle_matches_sklearn_lloyd():
    from sklearn.cluster import KMeans
    x, _ = _mix(3000, 24, 11)
    init = np.sort(np.random.default_rng(4).choice(3000, 16, replace=False))
    c, a, _ = O.kmeans(x, 16, 6, init)
    km = KMeans(16, init=x[init].astype(np.float64), n_init=1, max_iter=6, tol=0.0,
                algorithm="lloyd").fit(x.astype(np.float64))
    np.testing.assert_array_equal(a, km.labels_)
    np.testing.assert_allclose(c, km.cluster_centers_, rtol=1e-5, atol=1e-5)


def test_dataset_files_from_oracle_clusters(tmp_path):
    """src/create_index.cpp's inputs in the layout python/ivf.py writes, built into an index."""
    x, _ = _mix(3000, 40, 7)
    init = np.sort(np.random.default_rng(3).choice(3000, 16, replace=False))
    c, cid, d2 = O.kmeans(x, 16, 8, init)
    vecio.write_fvecs(tmp_path / "toy_base.fvecs", x)
    vecio.write_fvecs(tmp_path / "toy_centroid_16.fvecs", c)
    vecio.write_ivecs(tmp_path / "toy_cluster_id_16.ivecs", cid.astype(np.int32)[:, None])
    vecio.write_fvecs(tmp_path / "toy_dist_to_centroid_16.fvecs", np.sqrt(d2)[:, None])
    idx = prep.create_index_files(str(tmp_path), "toy", 16, 5, seed=2)
    back = IVFIndex.load(tmp_path / "ivf_exhaf5.index")
    assert back.EX == 4 and back.N == 3000 and back.DIM == 40 and back.D == 64
    np.testing.assert_array_equal(back.cluster_sizes,
                                  np.bincount(cid, minlength=16).astype(np.uint64))
    np.testing.assert_array_equal(back.long_code, idx.long_code)


def test_prep_needs_the_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="GPU"):
        prep.compute_gt(np.zeros((4, 8), np.float32), np.zeros((1, 8), np.float32), 2)
