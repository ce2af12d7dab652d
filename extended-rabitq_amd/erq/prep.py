"""Dataset preparation of the reference's python/ scripts, on torch (GPU when one is there).

ivf_cluster      python/ivf.py:8-35 (faiss IVF{K},Flat training + nearest-centroid search):
                 Lloyd k-means here (erq.data.kmeans_torch), then each vector's nearest
                 centroid and the distance to it.  Clusters differ from faiss's (another
                 trainer, another seed); the files have the same meaning and format.
compute_gt       python/compute_gt.py:6-18: the ids of each query's k nearest base vectors
                 by L2, nearest first.
The *_files variants read and write the dataset directory as the scripts and
src/create_index.cpp / src/test_search.cpp name its files:
  {name}_base.fvecs, {name}_query.fvecs, {name}_groundtruth.ivecs,
  {name}_centroid_{K}.fvecs, {name}_cluster_id_{K}.ivecs, {name}_dist_to_centroid_{K}.fvecs.
This is setup work, not the search path: it runs on whatever device torch offers.
"""
import os

import numpy as np

from . import data, vecio


def _device(device):
    import torch
    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    return torch.device(device)


def ivf_cluster(x, K, iters=20, seed=0, device=None, chunk=65536):
    """(centroids [K, d] float32, cluster ids [N] int32, distance to the centroid [N] float32)."""
    import torch
    dev = _device(device)
    xt = torch.as_tensor(np.ascontiguousarray(x, np.float32)).to(dev)
    if K > xt.shape[0]:
        raise ValueError("more clusters than vectors")
    c, assign = data.kmeans_torch(xt, K, iters, seed, chunk=chunk)
    dist = torch.empty(xt.shape[0], dtype=torch.float32, device=dev)
    for s in range(0, xt.shape[0], chunk):
        diff = xt[s:s + chunk] - c[assign[s:s + chunk]]
        dist[s:s + chunk] = diff.norm(dim=1)
    return (c.cpu().numpy().astype(np.float32), assign.cpu().numpy().astype(np.int32),
            dist.cpu().numpy())


def compute_gt(base, queries, k=1000, device=None):
    """[nq, k] int32 ids of the exact L2 nearest neighbours, nearest first (FP32)."""
    import torch
    dev = _device(device)
    b = torch.as_tensor(np.ascontiguousarray(base, np.float32)).to(dev)
    q = torch.as_tensor(np.ascontiguousarray(queries, np.float32)).to(dev)
    k = min(k, b.shape[0])
    ids, _ = data.exact_knn_torch(b, q, k)
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
