"""Vector file formats of the reference's data pipeline.

fvecs / ivecs (inc/utils/IO.hpp:26-81, python/utils/io.py:5-30): one record per row, an int32
dimension d followed by d float32 / int32 values, all little-endian; every row has the same d.
fbin / ibin (inc/utils/IO.hpp:83-110, python/utils/io.py:33-41): a header of two 32-bit
integers (rows, cols) followed by rows x cols float32 / int32 values.

Readers memory-map the file (mmap=True) or read it whole, and check what the reference's
readers assume without checking: a positive dimension, a size that is a whole number of
rows, the same dimension in every row's prefix.  Writers emit the same bytes as the
reference's writers, one vectorised write instead of a struct.pack per row.
"""
import os

import numpy as np

_I32 = np.dtype("<i4")


def _read_vecs(path, dtype, mmap):
    size = os.path.getsize(path)
    if size == 0:
        return np.zeros((0, 0), dtype)
    if size < 4:
        raise ValueError(f"{path}: truncated record header")
    d = int(np.fromfile(path, dtype=_I32, count=1)[0])
    if d <= 0:
        raise ValueError(f"{path}: bad dimension {d}")
    row = 4 * (d + 1)
    if size % row:
        raise ValueError(f"{path}: size {size} is not a multiple of the record size {row}")
    n = size // row
    if mmap:
        raw = np.memmap(path, dtype=_I32, mode="r", shape=(n, d + 1))
    else:
        raw = np.fromfile(path, dtype=_I32).reshape(n, d + 1)
    if not np.all(raw[:, 0] == d):
        raise ValueError(f"{path}: rows of different dimensions")
    return raw[:, 1:].view(dtype)


def _write_vecs(path, m, dtype):
    m = np.asarray(m)
    if m.ndim != 2:
        raise ValueError("expected a 2-D array")
    n, d = m.shape
    out = np.empty((n, d + 1), _I32)
    out[:, 0] = d
    out[:, 1:] = np.ascontiguousarray(m, dtype=dtype).view(_I32)
    out.tofile(path)


def read_fvecs(path, mmap=False):
    """[n, d] float32 rows of an fvecs file (read_fvecs, python/utils/io.py:13-14)."""
    return _read_vecs(path, np.dtype("<f4"), mmap)


def read_ivecs(path, mmap=False):
    """[n, d] int32 rows of an ivecs file (read_ivecs, python/utils/io.py:5-10)."""
    return _read_vecs(path, _I32, mmap)


def write_fvecs(path, m):
    """Write float32 rows as fvecs (write_fvecs, python/utils/io.py:29-31)."""
    _write_vecs(path, m, np.dtype("<f4"))


def write_ivecs(path, m):
    """Write int32 rows as ivecs (write_ivecs, python/utils/io.py:17-26)."""
    _write_vecs(path, m, _I32)


def _read_bin(path, dtype, mmap):
    size = os.path.getsize(path)
    if size < 8:
        raise ValueError(f"{path}: truncated header")
    n, d = (int(x) for x in np.fromfile(path, dtype="<u4", count=2))
    if size != 8 + 4 * n * d:
        raise ValueError(f"{path}: header says {n} x {d} but the file holds {size - 8} bytes")
    if mmap:
        return np.memmap(path, dtype=dtype, mode="r", offset=8, shape=(n, d))
    return np.fromfile(path, dtype=dtype, offset=8).reshape(n, d)


def _write_bin(path, m, dtype):
    m = np.asarray(m)
    if m.ndim != 2:
        raise ValueError("expected a 2-D array")
    with open(path, "wb") as f:
        np.array(m.shape, "<u4").tofile(f)
        np.ascontiguousarray(m, dtype=dtype).tofile(f)


def read_fbin(path, mmap=False):
    """[rows, cols] float32 of an fbin file (load_bin, IO.hpp:83-110; read_fbin, io.py:40-41)."""
    return _read_bin(path, np.dtype("<f4"), mmap)


def read_ibin(path, mmap=False):
    """[rows, cols] int32 of an ibin file (read_ibin, python/utils/io.py:33-37)."""
    return _read_bin(path, _I32, mmap)


def write_fbin(path, m):
    _write_bin(path, m, np.dtype("<f4"))


def write_ibin(path, m):
    _write_bin(path, m, _I32)


def build_from_files(base_path, centroids_path, cluster_id_path, bits, index_path=None,
                     **build_kw):
    """The index build of src/create_index.cpp:10-57 from its three input files.

    base vectors (fvecs), centroids (fvecs) and one cluster id per vector (ivecs, one
    column), as python/ivf.py:20-35 writes them; bits = B (EX_BITS = B - 1).  Returns the
    IVFIndex and, given index_path, saves it in the IVF::save format.  build_kw passes on
    to erq.index.build_index (rotation seed, matmul routine, threads).
    """
    from .index import build_index

    data = read_fvecs(base_path, mmap=True)
    centroids = read_fvecs(centroids_path)
    cids = read_ivecs(cluster_id_path)
    if cids.shape[1] != 1 or cids.shape[0] != data.shape[0]:
        raise ValueError("cluster id file must hold one id per base vector")
    if centroids.shape[1] != data.shape[1]:
        raise ValueError("centroid and base dimensions differ")
    idx = build_index(np.asarray(data), centroids, cids[:, 0], bits, **build_kw)
    if index_path is not None:
        idx.save(index_path)
    return idx
