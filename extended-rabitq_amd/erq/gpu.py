"""Host-side mirror of the reference's IVF class (inc/index/IVF.hpp:26-91) over liberq.so.

``IVF.load(path)`` / ``IVF.search(query, k, nprobe)`` / ``padded_dim()`` / ``k()`` keep the
reference's names and argument meaning; the ``*_dev`` methods take torch CUDA tensors
(device memory and streams only -- every computation runs in liberq.so's HIP kernels).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class IVF:
    def __init__(self, handle):
        self._h = handle

    # ---- construction (IVF::load, IVF.hpp:254-293) ----
    @staticmethod
    def load(path):
        h = ctypes.c_void_p()
        check(_lib.lib().erq_index_load(str(path).encode(), ctypes.byref(h)))
        return IVF(h)

    @staticmethod
    def from_index(idx):
        """Upload an in-memory erq.IVFIndex (same arrays IVF::load reads)."""
        idx.validate()
        arrs = [np.ascontiguousarray(idx.cluster_sizes, dtype=np.uint64),
                np.ascontiguousarray(idx.P, dtype=np.float32),
                np.ascontiguousarray(idx.centroids, dtype=np.float32),
                np.ascontiguousarray(idx.short_data, dtype=np.uint8),
                np.ascontiguousarray(idx.long_code, dtype=np.uint8),
                np.ascontiguousarray(idx.ex_factor, dtype=np.float32),
                np.ascontiguousarray(idx.ids, dtype=np.uint32)]
        h = ctypes.c_void_p()
        create = (_lib.lib().erq_index_create if getattr(idx, "factors", 1) == 1
                  else _lib.lib().erq_index_create_nonha)
        check(create(idx.N, idx.DIM, idx.K, idx.EX, *[_ptr(a) for a in arrs], ctypes.byref(h)))
        return IVF(h)

    def close(self):
        if self._h:
            _lib.lib().erq_index_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    PLAIN_AUTO, PLAIN_LUT, PLAIN_MFMA = 0, 1, 2

    def set_plain_kernel(self, mode):
        """Integer kernel of the per-query scan (erq_set_option ERQ_OPT_PLAIN_KERNEL)."""
        check(_lib.lib().erq_set_option(self._h, 1, mode))

    def option(self, opt):
        return _lib.lib().erq_get_option(self._h, opt)

    def set_option(self, opt, value):
        check(_lib.lib().erq_set_option(self._h, opt, value))

    def autotune(self, q_rot, k, nprobe):
        """erq_autotune: returns (best per-query ms, best batched ms) and keeps the shapes."""
        out = np.zeros(2, np.float32)
        check(_lib.lib().erq_autotune(self._h, _ptr(q_rot), q_rot.shape[0], k, nprobe, _ptr(out),
                                      _stream()))
        return float(out[0]), float(out[1])

    STAGES = ("select", "replay", "prep", "scan", "reest", "merge", "redo")

    def timing_begin(self, max_records=4096):
        """Record HIP events around every search stage from now on (erq_timing_begin)."""
        check(_lib.lib().erq_timing_begin(self._h, max_records))

    def timing_end(self):
        """{stage: (summed ms, recorded count)} since timing_begin (erq_timing_end)."""
        ms = np.zeros(len(self.STAGES), np.float64)
        cnt = np.zeros(len(self.STAGES), np.uint64)
        check(_lib.lib().erq_timing_end(self._h, _ptr(ms), _ptr(cnt)))
        return {n: (float(ms[i]), int(cnt[i])) for i, n in enumerate(self.STAGES)}

    # ---- header (IVF.hpp:86-90) ----
    def padded_dim(self):
        return _lib.lib().erq_padded_dim(self._h)

    def k(self):
        return _lib.lib().erq_num_clusters(self._h)

    def dim(self):
        return _lib.lib().erq_dim(self._h)

    def ex_bits(self):
        return _lib.lib().erq_ex_bits(self._h)

    def high_acc(self):
        """True for the HA block format (HIGH_ACC_FAST_SCAN), False for the non-HA one."""
        return _lib.lib().erq_short_factors(self._h) == 1

    # ---- host-pointer path (Rotator::rotate, IVF::search) ----
    def rotate(self, q):
        q = np.ascontiguousarray(q, dtype=np.float32)
        if q.ndim == 1:
            q = q[None]
        out = np.empty((q.shape[0], self.padded_dim()), np.float32)
        check(_lib.lib().erq_rotate(self._h, _ptr(q), q.shape[0], _ptr(out)))
        return out

    def search(self, query, k, nprobe, with_dists=False):
        """IVF::search(query, data, k, nprobe, results) on one rotated query (D floats)."""
        q = np.ascontiguousarray(query, dtype=np.float32).reshape(-1)
        if q.size != self.padded_dim():
            raise ValueError("query must be the rotated, padded D-vector")
        ids = np.empty(k, np.uint32)
        d = np.empty(k, np.float32)
        check(_lib.lib().erq_search(self._h, _ptr(q), k, nprobe, _ptr(ids), _ptr(d)))
        return (ids, d) if with_dists else ids

    def search_batch(self, q_rot, k, nprobe):
        q = np.ascontiguousarray(q_rot, dtype=np.float32)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        check(_lib.lib().erq_search_batch(self._h, _ptr(q), nq, k, nprobe, _ptr(ids), _ptr(d)))
        return ids, d

    def search_batch_nprobe(self, q_rot, k, nprobe_per_q, batched=False):
        """IVF::search over many queries, each with its own nprobe (host arrays); batched
        selects the cluster-major rounds instead of the exact per-query replay."""
        q = np.ascontiguousarray(q_rot, dtype=np.float32)
        npq = np.ascontiguousarray(nprobe_per_q, dtype=np.uint32)
        nq = q.shape[0]
        if npq.shape != (nq,):
            raise ValueError("one nprobe per query")
        ids = np.empty((nq, k), np.uint32)
        d = np.empty((nq, k), np.float32)
        check(_lib.lib().erq_search_batch_nprobe(self._h, _ptr(q), nq, k, _ptr(npq), _ptr(ids),
                                                 _ptr(d), 1 if batched else 0))
        return ids, d

    # ---- device path (torch CUDA tensors; enqueued on torch's current stream) ----
    def rotate_dev(self, q):
        import torch
        out = torch.empty((q.shape[0], self.padded_dim()), dtype=torch.float32, device=q.device)
        check(_lib.lib().erq_rotate_dev(self._h, _ptr(q.contiguous()), q.shape[0], _ptr(out),
                                        _stream()))
        return out

    def centroid_dist_dev(self, q_rot):
        import torch
        out = torch.empty((q_rot.shape[0], self.k()), dtype=torch.float32, device=q_rot.device)
        check(_lib.lib().erq_centroid_dist_dev(self._h, _ptr(q_rot), q_rot.shape[0], _ptr(out),
                                               _stream()))
        return out

    def probe_dev(self, q_rot, nprobe):
        import torch
        nq = q_rot.shape[0]
        pid = torch.empty((nq, nprobe), dtype=torch.int32, device=q_rot.device)
        pd = torch.empty((nq, nprobe), dtype=torch.float32, device=q_rot.device)
        check(_lib.lib().erq_probe_dev(self._h, _ptr(q_rot), nq, nprobe, _ptr(pid), _ptr(pd),
                                       _stream()))
        return pid, pd

    def scan_dev(self, q_rot, k, probe_off, probe_ids, probe_d, stats=None):
        import torch
        nq = q_rot.shape[0]
        ids = torch.empty((nq, k), dtype=torch.int32, device=q_rot.device)
        d = torch.empty((nq, k), dtype=torch.float32, device=q_rot.device)
        check(_lib.lib().erq_scan_dev(self._h, _ptr(q_rot), nq, k, _ptr(probe_off),
                                      _ptr(probe_ids), _ptr(probe_d), _ptr(ids), _ptr(d),
                                      _ptr(stats), _stream()))
        return ids, d

    def scan_batched_dev(self, q_rot, k, probe_off, probe_ids, probe_d, max_nprobe, stats=None):
        import torch
        nq = q_rot.shape[0]
        ids = torch.empty((nq, k), dtype=torch.int32, device=q_rot.device)
        d = torch.empty((nq, k), dtype=torch.float32, device=q_rot.device)
        check(_lib.lib().erq_scan_batched_dev(self._h, _ptr(q_rot), nq, k, _ptr(probe_off),
                                              _ptr(probe_ids), _ptr(probe_d), max_nprobe,
                                              _ptr(ids), _ptr(d), _ptr(stats), _stream()))
        return ids, d

    def search_batched_dev(self, q_rot, k, nprobe, stats=None):
        import torch
        nq = q_rot.shape[0]
        ids = torch.empty((nq, k), dtype=torch.int32, device=q_rot.device)
        d = torch.empty((nq, k), dtype=torch.float32, device=q_rot.device)
        check(_lib.lib().erq_search_batched_dev(self._h, _ptr(q_rot), nq, k, nprobe, _ptr(ids),
                                                _ptr(d), _ptr(stats), _stream()))
        return ids, d

    def probe_learned_dev(self, model, q_rot, mass, min_nprobe, max_nprobe, fallback_nprobe,
                          min_conf):
        """CSR probe lists chosen by the learned predictor (erq_probe_learned_dev)."""
        import torch
        nq = q_rot.shape[0]
        dev = q_rot.device
        off = torch.empty(nq + 1, dtype=torch.int32, device=dev)
        pid = torch.empty(nq * max_nprobe, dtype=torch.int32, device=dev)
        pd = torch.empty(nq * max_nprobe, dtype=torch.float32, device=dev)
        nfb = torch.zeros(1, dtype=torch.int32, device=dev)
        check(_lib.lib().erq_probe_learned_dev(self._h, model.handle, _ptr(q_rot), nq, mass,
                                               min_nprobe, max_nprobe, fallback_nprobe, min_conf,
                                               _ptr(off), _ptr(pid), _ptr(pd), _ptr(nfb),
                                               _stream()))
        return off, pid, pd, nfb

    def search_learned_dev(self, model, q_rot, k, mass, min_nprobe, max_nprobe, fallback_nprobe,
                           min_conf, stats=None):
        import torch
        nq = q_rot.shape[0]
        ids = torch.empty((nq, k), dtype=torch.int32, device=q_rot.device)
        d = torch.empty((nq, k), dtype=torch.float32, device=q_rot.device)
        check(_lib.lib().erq_search_learned_dev(self._h, model.handle, _ptr(q_rot), nq, k, mass,
                                                min_nprobe, max_nprobe, fallback_nprobe, min_conf,
                                                _ptr(ids), _ptr(d), _ptr(stats), _stream()))
        return ids, d

    def search_dev(self, q_rot, k, nprobe, stats=None):
        import torch
        nq = q_rot.shape[0]
        ids = torch.empty((nq, k), dtype=torch.int32, device=q_rot.device)
        d = torch.empty((nq, k), dtype=torch.float32, device=q_rot.device)
        check(_lib.lib().erq_search_dev(self._h, _ptr(q_rot), nq, k, nprobe, _ptr(ids), _ptr(d),
                                        _ptr(stats), _stream()))
        return ids, d


# ---------------------------------------------------------------- dataset preparation
# python/ivf.py and python/compute_gt.py on liberq.so's kernels (erq_knn.hip); inputs are
# torch CUDA float32 tensors of shape [rows, dim].

def _rows_dev(x):
    import torch
    if not (torch.is_tensor(x) and x.is_cuda and x.dtype == torch.float32 and x.dim() == 2):
        raise TypeError("expected a 2-D float32 CUDA tensor")
    return x.contiguous()


def knn_dev(base, queries, k):
    """(ids int32 [nq, k], squared distances float32 [nq, k]): exact k nearest base rows of
    each query, ordered by (distance, id) (erq_knn_dev)."""
    import torch
    b, q = _rows_dev(base), _rows_dev(queries)
    if b.shape[1] != q.shape[1]:
        raise ValueError("dimension mismatch")
    ids = torch.empty((q.shape[0], k), dtype=torch.int32, device=q.device)
    d = torch.empty((q.shape[0], k), dtype=torch.float32, device=q.device)
    check(_lib.lib().erq_knn_dev(_ptr(b), b.shape[0], _ptr(q), q.shape[0], b.shape[1], k,
                                 _ptr(ids), _ptr(d), _stream()))
    return ids, d


def kmeans_assign_dev(x, centroids):
    """(assign int32 [n], squared distance float32 [n]): L2Sqr16-nearest centroid, ties by id."""
    import torch
    x, c = _rows_dev(x), _rows_dev(centroids)
    a = torch.empty(x.shape[0], dtype=torch.int32, device=x.device)
    d = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
    check(_lib.lib().erq_kmeans_assign_dev(_ptr(x), x.shape[0], x.shape[1], _ptr(c), c.shape[0],
                                           _ptr(a), _ptr(d), _stream()))
    return a, d


def kmeans_update_dev(x, assign, centroids):
    """One Lloyd update of `centroids` in place; returns the member counts (int32 [K])."""
    import torch
    x = _rows_dev(x)
    if not centroids.is_contiguous():
        raise ValueError("centroids must be contiguous")
    a = assign.to(torch.int32).contiguous()
    cnt = torch.empty(centroids.shape[0], dtype=torch.int32, device=x.device)
    check(_lib.lib().erq_kmeans_update_dev(_ptr(x), x.shape[0], x.shape[1], _ptr(a),
                                           centroids.shape[0], _ptr(centroids), _ptr(cnt),
                                           _stream()))
    return cnt


def kmeans_dev(x, K, iters, init_rows):
    """(centroids [K, dim], assign int32 [n], squared distance [n]) after `iters` Lloyd steps
    from the rows init_rows (erq_kmeans_dev)."""
    import torch
    x = _rows_dev(x)
    init = torch.as_tensor(np.ascontiguousarray(init_rows, np.int32)).to(x.device)
    if init.numel() != K:
        raise ValueError("init_rows must hold K row indices")
    c = torch.empty((K, x.shape[1]), dtype=torch.float32, device=x.device)
    a = torch.empty(x.shape[0], dtype=torch.int32, device=x.device)
    d = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
    check(_lib.lib().erq_kmeans_dev(_ptr(x), x.shape[0], x.shape[1], K, iters, _ptr(init),
                                    _ptr(c), _ptr(a), _ptr(d), _stream()))
    return c, a, d


def kmeans_init_rows(n, K, seed):
    """K distinct row indices drawn by a seeded generator (the initial centroids)."""
    return np.sort(np.random.default_rng(seed).choice(n, K, replace=False)).astype(np.int32)


# ---------------------------------------------------------------- index building
# IVF::construct on liberq.so's kernels (erq_build.hip).

def build_quantize_dev(resid, onorm, ex_bits):
    """(long_code uint8 [n, D*EX/8], xipnorm float32 [n]) of rotated residuals [n, D] and
    their norms (erq_build_quantize_dev)."""
    import torch
    r = _rows_dev(resid)
    on = onorm.contiguous()
    n, D = r.shape
    lc = torch.empty((n, D * ex_bits // 8), dtype=torch.uint8, device=r.device)
    xip = torch.empty(n, dtype=torch.float32, device=r.device)
    check(_lib.lib().erq_build_quantize_dev(_ptr(r), _ptr(on), n, D, ex_bits, _ptr(lc),
                                            _ptr(xip), _stream()))
    return lc, xip


def build_pack_dev(resid, onorm, cluster_sizes):
    """Short blocks (uint8) of residuals in cluster order (erq_build_pack_dev)."""
    import torch
    r = _rows_dev(resid)
    on = onorm.contiguous()
    sizes = np.ascontiguousarray(cluster_sizes, np.uint64)
    D = r.shape[1]
    nb = int(((sizes + 31) // 32).sum())
    out = torch.empty(nb * (4 * D + 128), dtype=torch.uint8, device=r.device)
    check(_lib.lib().erq_build_pack_dev(_ptr(r), _ptr(on), _ptr(sizes), sizes.size, D,
                                        _ptr(out), _stream()))
    return out


def build_index_dev(data, centroids, assign, bits, P, high_acc=True):
    """IVF::construct on the GPU (erq_build_index_dev): raw vectors, un-rotated centroids and
    cluster ids (torch CUDA tensors) and the rotation P (numpy [D, D]) -> an IVFIndex with
    host arrays, EX = bits - 1.  high_acc=False builds the non-HA format
    (erq_build_index_dev_nonha: four factors per vector)."""
    import torch
    from .index import IVFIndex, padded_dim
    x, c = _rows_dev(data), _rows_dev(centroids)
    N, DIM = x.shape
    K = c.shape[0]
    EX = bits - 1
    D = padded_dim(DIM)
    a = assign.to(torch.int32).contiguous()
    if a.numel() != N:
        raise ValueError("one cluster id per vector")
    sizes = np.bincount(a.cpu().numpy(), minlength=K).astype(np.uint64)
    if sizes.size != K:
        raise ValueError("cluster id out of range")
    Pd = torch.as_tensor(np.ascontiguousarray(P, np.float32)).to(x.device)
    nb = int(((sizes + 31) // 32).sum())
    crot = torch.empty((K, D), dtype=torch.float32, device=x.device)
    F = 1 if high_acc else 4
    short = torch.empty(nb * (4 * D + 128 * F), dtype=torch.uint8, device=x.device)
    lc = torch.empty(N * D * EX // 8, dtype=torch.uint8, device=x.device)
    xip = torch.empty(N, dtype=torch.float32, device=x.device)
    ids = torch.empty(N, dtype=torch.int32, device=x.device)
    got = np.zeros(K, np.uint64)
    fn = _lib.lib().erq_build_index_dev if high_acc else _lib.lib().erq_build_index_dev_nonha
    check(fn(_ptr(x), N, DIM, _ptr(c), K, _ptr(a), _ptr(Pd), EX, _ptr(got), _ptr(crot),
             _ptr(short), _ptr(lc), _ptr(xip), _ptr(ids), _stream()))
    if not np.array_equal(got, sizes):
        raise RuntimeError("cluster sizes disagree")
    return IVFIndex(N, DIM, K, EX, got, np.ascontiguousarray(P, np.float32), crot.cpu().numpy(),
                    short.cpu().numpy(), lc.cpu().numpy(), xip.cpu().numpy(),
                    ids.cpu().numpy().view(np.uint32), F)
