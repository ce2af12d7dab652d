"""HA-format IVF index: arrays, file I/O and the build orchestration.

File layout written by IVF::save / read by IVF::load (reference inc/index/IVF.hpp:219-293),
all little-endian:
  N, DIM, K, EX_BITS            4 x size_t
  cluster_sizes[K]              size_t
  P[D x D]                      float32 row-major (Rotator.hpp:34-57)
  centroids[K x D]              float32, rotated (Initializer.hpp:70-76)
  SHORT_DATA                    sum_k ceil(n_k/32) blocks of (4D + 128 F) bytes: F = 1 factor
                                per vector (HA, HIGH_ACC_FAST_SCAN) or 4 (the non-HA Searcher,
                                Quantizer.hpp:107-121); the file does not record F, so load()
                                infers it from the file size
  LONG_CODE                     N * D*EX/8 bytes
  EX_FACTOR                     N float32 (xipnorm)
  IDs                           N uint32
D = DIM rounded up to a multiple of 64 (IVF.hpp:96).
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _lib

EX_BITS_SUPPORTED = (2, 3, 4, 6, 7, 8)


def padded_dim(dim):
    return (dim + 63) // 64 * 64


@dataclass
class IVFIndex:
    N: int
    DIM: int
    K: int
    EX: int
    cluster_sizes: np.ndarray  # uint64 [K]
    P: np.ndarray              # float32 [D, D]
    centroids: np.ndarray      # float32 [K, D] (rotated)
    short_data: np.ndarray     # uint8
    long_code: np.ndarray      # uint8 [N * D*EX/8]
    ex_factor: np.ndarray      # float32 [N]
    ids: np.ndarray            # uint32 [N]
    factors: int = 1           # short factors per vector: 1 (HA) or 4 (non-HA)

    @property
    def D(self):
        return padded_dim(self.DIM)

    @property
    def high_acc(self):
        return self.factors == 1

    @property
    def block_bytes(self):
        return 4 * self.D + 128 * self.factors

    @property
    def code_len(self):
        return self.D * self.EX // 8

    def block_start(self):
        nb = (self.cluster_sizes.astype(np.uint64) + 31) // 32
        return np.concatenate([[0], np.cumsum(nb)]).astype(np.uint64)

    def vec_start(self):
        return np.concatenate([[0], np.cumsum(self.cluster_sizes)]).astype(np.uint64)

    def validate(self):
        D = self.D
        if self.EX not in EX_BITS_SUPPORTED:
            raise ValueError(f"EX_BITS={self.EX} unsupported")
        if self.factors not in (1, 4):
            raise ValueError("factors per vector must be 1 (HA) or 4 (non-HA)")
        if int(self.cluster_sizes.sum()) != self.N:
            raise ValueError("cluster sizes do not sum to N")
        nblocks = int(self.block_start()[-1])
        if self.short_data.size != nblocks * self.block_bytes:
            raise ValueError("short data size mismatch")
        if self.long_code.size != self.N * self.code_len:
            raise ValueError("long code size mismatch")
        if self.P.shape != (D, D) or self.centroids.shape != (self.K, D):
            raise ValueError("rotation/centroid shape mismatch")
        if self.ex_factor.size != self.N or self.ids.size != self.N:
            raise ValueError("factor/id count mismatch")

    # ---- file I/O (IVF.hpp:219-293) ----
    def save(self, path):
        self.validate()
        with open(path, "wb") as f:
            np.array([self.N, self.DIM, self.K, self.EX], dtype="<u8").tofile(f)
            self.cluster_sizes.astype("<u8").tofile(f)
            self.P.astype("<f4").tofile(f)
            self.centroids.astype("<f4").tofile(f)
            self.short_data.tofile(f)
            self.long_code.tofile(f)
            self.ex_factor.astype("<f4").tofile(f)
            self.ids.astype("<u4").tofile(f)

    @staticmethod
    def load(path):
        size = os.path.getsize(path)
        with open(path, "rb") as f:
            hdr = np.fromfile(f, dtype="<u8", count=4)
            if hdr.size != 4:
                raise ValueError("truncated header")
            N, DIM, K, EX = (int(x) for x in hdr)
            D = padded_dim(DIM)
            cs = np.fromfile(f, dtype="<u8", count=K)
            nblocks = int(((cs + 31) // 32).sum())
            rest = 32 + 8 * K + 4 * D * D + 4 * K * D + N * D * EX // 8 + 8 * N
            fits = [F for F in (1, 4) if rest + nblocks * (4 * D + 128 * F) == size]
            if len(fits) != 1:
                raise ValueError("file size matches neither the HA nor the non-HA block size")
            F = fits[0]
            P = np.fromfile(f, dtype="<f4", count=D * D).reshape(D, D)
            C = np.fromfile(f, dtype="<f4", count=K * D).reshape(K, D)
            sd = np.fromfile(f, dtype=np.uint8, count=nblocks * (4 * D + 128 * F))
            lc = np.fromfile(f, dtype=np.uint8, count=N * D * EX // 8)
            ef = np.fromfile(f, dtype="<f4", count=N)
            ids = np.fromfile(f, dtype="<u4", count=N)
        idx = IVFIndex(N, DIM, K, EX, cs, P, C, sd, lc, ef, ids, F)
        idx.validate()
        return idx


def random_rotation(D, seed):
    """P = Q^T for Q from the QR of a uniform random D x D matrix (Rotator.hpp:14-19)."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1.0, 1.0, size=(D, D))
    Q, R = np.linalg.qr(A)
    Q = Q * np.sign(np.diag(R))[None, :]
    return np.ascontiguousarray(Q.T.astype(np.float32))


def _np_matmul(a, b):
    return a @ b


def build_index(data, centroids, assign, bits, P=None, seed=3, matmul=None, threads=0,
                chunk=65536, high_acc=True):
    """Build an HA index (IVF::construct, IVF.hpp:116-153) from raw vectors; high_acc=False
    writes the non-HA format (four factors per vector, Quantizer.hpp:386-437).

    data [N, DIM] float32, centroids [K, DIM] float32 (un-rotated), assign [N] cluster ids,
    bits = total bits per dimension B (EX = B - 1).  matmul(a, b) computes a @ b in float32
    for [rows, D] x [D, D] (numpy by default; callers pass a GPU routine for large N).
    """
    data = np.ascontiguousarray(data, dtype=np.float32)
    N, DIM = data.shape
    K = centroids.shape[0]
    EX = bits - 1
    if EX not in EX_BITS_SUPPORTED:
        raise ValueError(f"B={bits} unsupported (EX_BITS must be one of {EX_BITS_SUPPORTED})")
    D = padded_dim(DIM)
    if P is None:
        P = random_rotation(D, seed)
    matmul = matmul or _np_matmul
    assign = np.asarray(assign, dtype=np.int64)
    if assign.min() < 0 or assign.max() >= K:
        raise ValueError("bad cluster id")
    order = np.argsort(assign, kind="stable")            # ids ascending within a cluster
    sizes = np.bincount(assign, minlength=K).astype(np.uint64)
    cpad = np.zeros((K, D), np.float32)
    cpad[:, :DIM] = centroids
    crot = np.ascontiguousarray(matmul(cpad, P), dtype=np.float32)
    resid = np.empty((N, D), np.float32)
    onorm = np.empty(N, np.float32)
    onorm2 = np.empty(N, np.float32)
    xp = np.zeros((chunk, D), np.float32)
    for s in range(0, N, chunk):
        sel = order[s:s + chunk]
        n = sel.size
        xp[:n, :DIM] = data[sel]
        resid[s:s + n] = matmul(xp[:n], P) - crot[assign[sel]]
        diff = data[sel] - centroids[assign[sel]]
        onorm2[s:s + n] = np.einsum("ij,ij->i", diff, diff, dtype=np.float32)
        onorm[s:s + n] = np.sqrt(onorm2[s:s + n])
    nblocks = int(((sizes + 31) // 32).sum())
    F = 1 if high_acc else 4
    short = np.zeros(nblocks * (4 * D + 128 * F), np.uint8)
    longc = np.zeros(N * D * EX // 8, np.uint8)
    xip = np.zeros(N, np.float32)
    fn = _lib.builder().erqb_quantize_all if high_acc else _lib.builder().erqb_quantize_all_nonha
    # the non-HA blocks keep ||o - c||^2 (the HA ones ||o - c||)
    fac = onorm if high_acc else np.ascontiguousarray(onorm2, np.float32)
    rc = fn(
        resid.ctypes.data_as(ctypes.c_void_p), fac.ctypes.data_as(ctypes.c_void_p),
        sizes.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(K), ctypes.c_size_t(D),
        ctypes.c_int(EX), short.ctypes.data_as(ctypes.c_void_p),
        longc.ctypes.data_as(ctypes.c_void_p), xip.ctypes.data_as(ctypes.c_void_p),
        ctypes.c_int(threads))
    if rc != 0:
        raise ValueError("index build failed: non-finite vectors or unsupported shape")
    return IVFIndex(N, DIM, K, EX, sizes, P, crot, short, longc, xip,
                    order.astype(np.uint32), F)


def index_bytes(idx):
    return (idx.short_data.nbytes + idx.long_code.nbytes + idx.ex_factor.nbytes +
            idx.ids.nbytes + idx.centroids.nbytes)


def save_default_path(tmpdir, name):
    return os.path.join(tmpdir, f"{name}.index")
