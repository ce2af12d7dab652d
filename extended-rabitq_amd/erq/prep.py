"""Dataset preparation of the reference's python/ scripts, on liberq.so's HIP kernels.

ivf_cluster      python/ivf.py:8-35 (faiss IVF{K},Flat training + nearest-centroid search):
                 Lloyd k-means on the GPU (erq_kmeans_dev: L2Sqr16-nearest assignment, ties
                 by id; cluster means summed in double in row order), then each vector's
                 nearest centroid and the distance to it.  Clusters differ from faiss's
                 (another trainer, another seed); the files have the same meaning and format.
compute_gt       python/compute_gt.py:6-18: the ids of each query's k nearest base vectors
                 by L2, nearest first (erq_knn_dev: distances in double, ties by id).
The *_files variants read and write the dataset directory as the scripts and
src/create_index.cpp / src/test_search.cpp name its files:
  {name}_base.fvecs, {name}_query.fvecs, {name}_groundtruth.ivecs,
  {name}_centroid_{K}.fvecs, {name}_cluster_id_{K}.ivecs, {name}_dist_to_centroid_{K}.fvecs.
Both need the GPU: there is no CPU fallback.  torch holds the device buffers.
"""
import os

import numpy as np

from . import gpu, vecio


def _dev(x):
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("dataset preparation runs on the GPU (liberq.so); none is visible")
    return torch.as_tensor(np.ascontiguousarray(x, np.float32)).cuda()


def ivf_cluster(x, K, iters=20, seed=0):
    """(centroids [K, d] float32, cluster ids [N] int32, distance to the centroid [N] float32)."""
    xt = _dev(x)
    if K > xt.shape[0]:
        raise ValueError("more clusters than vectors")
    c, assign, d2 = gpu.kmeans_dev(xt, K, iters, gpu.kmeans_init_rows(xt.shape[0], K, seed))
    return (c.cpu().numpy(), assign.cpu().numpy().astype(np.int32),
            np.sqrt(d2.cpu().numpy()).astype(np.float32))


def compute_gt(base, queries, k=1000):
    """[nq, k] int32 ids of the exact L2 nearest neighbours, nearest first."""
    b, q = _dev(base), _dev(queries)
    k = min(k, b.shape[0])
    ids, _ = gpu.knn_dev(b, q, k)
    return ids.cpu().numpy().astype(np.int32)


def _path(root, name, suffix):
    return os.path.join(root, f"{name}_{suffix}")


def ivf_files(root, name, K, **kw):
    """python/ivf.py over {root}/{name}_base.fvecs; writes the three clustering files."""
    x = vecio.read_fvecs(_path(root, name, "base.fvecs"), mmap=True)
    c, cid, dist = ivf_cluster(np.asarray(x), K, **kw)
    vecio.write_fvecs(_path(root, name, f"dist_to_centroid_{K}.fvecs"), dist[:, None])
    vecio.write_ivecs(_path(root, name, f"cluster_id_{K}.ivecs"), cid[:, None])
    vecio.write_fvecs(_path(root, name, f"centroid_{K}.fvecs"), c)


def gt_files(root, name, k=1000, **kw):
    """python/compute_gt.py: writes {name}_groundtruth.ivecs."""
    b = vecio.read_fvecs(_path(root, name, "base.fvecs"), mmap=True)
    q = vecio.read_fvecs(_path(root, name, "query.fvecs"))
    vecio.write_ivecs(_path(root, name, "groundtruth.ivecs"), compute_gt(np.asarray(b), q, k, **kw))


def create_index_files(root, name, K, bits, **build_kw):
    """src/create_index.cpp: builds {root}/ivf_exhaf{bits}.index from the dataset's files."""
    out = os.path.join(root, f"ivf_exhaf{bits}.index")
    return vecio.build_from_files(_path(root, name, "base.fvecs"),
                                  _path(root, name, f"centroid_{K}.fvecs"),
                                  _path(root, name, f"cluster_id_{K}.ivecs"), bits, out,
                                  **build_kw)
