"""fvecs / ivecs / fbin / ibin files (inc/utils/IO.hpp:26-110, python/utils/io.py) and the
file-driven index build of src/create_index.cpp."""
import struct

import numpy as np
import pytest

from erq import vecio


def test_fvecs_bytes(tmp_path):
    # two rows of d = 2: int32 2, then the float32 values, little-endian
    p = tmp_path / "a.fvecs"
    vecio.write_fvecs(p, np.array([[1.0, 2.0], [-0.5, 3.25]], np.float32))
    want = (struct.pack("<i", 2) + struct.pack("<2f", 1.0, 2.0) +
            struct.pack("<i", 2) + struct.pack("<2f", -0.5, 3.25))
    assert p.read_bytes() == want
    np.testing.assert_array_equal(vecio.read_fvecs(p), [[1.0, 2.0], [-0.5, 3.25]])


def test_ivecs_bytes(tmp_path):
    p = tmp_path / "g.ivecs"
    p.write_bytes(struct.pack("<4i", 3, 7, -1, 2 ** 31 - 1))
    got = vecio.read_ivecs(p)
    assert got.dtype == np.int32 and got.shape == (1, 3)
    assert got.tolist() == [[7, -1, 2 ** 31 - 1]]
    vecio.write_ivecs(tmp_path / "h.ivecs", got)
    assert (tmp_path / "h.ivecs").read_bytes() == p.read_bytes()


def test_fbin_bytes(tmp_path):
    p = tmp_path / "a.fbin"
    m = np.arange(6, dtype=np.float32).reshape(3, 2) / 4
    vecio.write_fbin(p, m)
    assert p.read_bytes() == struct.pack("<2I", 3, 2) + struct.pack("<6f", *m.ravel())
    np.testing.assert_array_equal(vecio.read_fbin(p), m)
    np.testing.assert_array_equal(vecio.read_fbin(p, mmap=True), m)
    q = tmp_path / "a.ibin"
    vecio.write_ibin(q, [[5, -6]])
    assert q.read_bytes() == struct.pack("<2I2i", 1, 2, 5, -6)
    assert vecio.read_ibin(q).tolist() == [[5, -6]]


@pytest.mark.parametrize("mmap", [False, True])
def test_round_trip(tmp_path, mmap):
    rng = np.random.default_rng(0)
    m = rng.standard_normal((257, 1536)).astype(np.float32)
    vecio.write_fvecs(tmp_path / "m.fvecs", m)
    got = vecio.read_fvecs(tmp_path / "m.fvecs", mmap=mmap)
    np.testing.assert_array_equal(got.view(np.uint32), m.view(np.uint32))
    # fvecs -> fbin conversion of python/cvt_data.py, and back
    vecio.write_fbin(tmp_path / "m.fbin", got)
    np.testing.assert_array_equal(vecio.read_fbin(tmp_path / "m.fbin"), m)


def test_bad_files(tmp_path):
    p = tmp_path / "e.fvecs"
    p.write_bytes(b"")
    assert vecio.read_fvecs(p).shape == (0, 0)
    p.write_bytes(struct.pack("<i2f", 2, 1.0, 2.0) + b"\x00")        # trailing byte
    with pytest.raises(ValueError):
        vecio.read_fvecs(p)
    p.write_bytes(struct.pack("<i2f", 2, 1.0, 2.0) + struct.pack("<i2f", 3, 1.0, 2.0))
    with pytest.raises(ValueError):                                   # ragged rows
        vecio.read_fvecs(p)
    p.write_bytes(struct.pack("<i", 0))
    with pytest.raises(ValueError):
        vecio.read_fvecs(p)
    b = tmp_path / "e.fbin"
    b.write_bytes(struct.pack("<2I3f", 2, 2, 1.0, 2.0, 3.0))          # one value short
    with pytest.raises(ValueError):
        vecio.read_fbin(b)


def test_build_from_files(tmp_path, small_data):
    from erq import build_index, IVFIndex
    x, q, c, assign = small_data[:4]
    vecio.write_fvecs(tmp_path / "base.fvecs", x)
    vecio.write_fvecs(tmp_path / "centroid_64.fvecs", c)
    vecio.write_ivecs(tmp_path / "cluster_id_64.ivecs", np.asarray(assign, np.int32)[:, None])
    out = tmp_path / "ivf_exhaf4.index"
    idx = vecio.build_from_files(tmp_path / "base.fvecs", tmp_path / "centroid_64.fvecs",
                                 tmp_path / "cluster_id_64.ivecs", 4, out, seed=5)
    ref = build_index(x, c, assign, 4, seed=5)
    ref.save(tmp_path / "direct.index")
    assert out.read_bytes() == (tmp_path / "direct.index").read_bytes()
    back = IVFIndex.load(out)
    assert back.N == idx.N and back.EX == 3
    vecio.write_ivecs(tmp_path / "bad.ivecs", np.zeros((x.shape[0] - 1, 1), np.int32))
    with pytest.raises(ValueError):
        vecio.build_from_files(tmp_path / "base.fvecs", tmp_path / "centroid_64.fvecs",
                               tmp_path / "bad.ivecs", 4)
