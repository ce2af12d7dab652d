"""C-ABI checks that need no GPU: the library loads, exports every declared symbol, and
rejects bad files before touching the device."""
import os
import re

import numpy as np
import pytest

from conftest import REPO


def _declared():
    src = open(os.path.join(REPO, "include", "erq.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(erq_[a-z_]+)\s*\(", src)))


def test_header_symbols_exported():
    from erq import _lib
    lib = _lib.lib()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), f"liberq.so does not export {n}"
        assert n in _lib.SIGNATURES, f"ctypes signature missing for {n}"


def test_builder_symbols_exported():
    from erq import _lib
    b = _lib.builder()
    for n in _lib.BUILDER_SIGNATURES:
        assert hasattr(b, n)


def _load_rc(path):
    import ctypes
    from erq import _lib
    h = ctypes.c_void_p()
    rc = _lib.lib().erq_index_load(path.encode(), ctypes.byref(h))
    return rc, _lib.lib().erq_last_error().decode()


def test_missing_file_is_eio(tmp_path):
    rc, msg = _load_rc(str(tmp_path / "nope.index"))
    assert rc == 2 and "cannot open" in msg


def test_truncated_header_is_eio(tmp_path):
    p = tmp_path / "t.index"
    p.write_bytes(b"\x01\x02\x03")
    rc, msg = _load_rc(str(p))
    assert rc == 2 and "truncated" in msg


def test_bad_ex_bits_is_eformat(tmp_path):
    p = tmp_path / "b.index"
    np.array([10, 64, 2, 5], dtype="<u8").tofile(str(p))
    rc, msg = _load_rc(str(p))
    assert rc == 3


def test_sizes_not_summing_is_eformat(tmp_path):
    p = tmp_path / "s.index"
    with open(p, "wb") as f:
        np.array([10, 64, 2, 3], dtype="<u8").tofile(f)
        np.array([4, 5], dtype="<u8").tofile(f)
    rc, msg = _load_rc(str(p))
    assert rc == 3 and "sum" in msg


def test_truncated_body_is_eio(small_data, tmp_path):
    from erq import index
    x, q, c, assign = small_data
    idx = index.build_index(x[:2000], c, assign[:2000] if assign.size >= 2000 else assign,
                            bits=4, seed=3)
    p = tmp_path / "i.index"
    idx.save(str(p))
    data = p.read_bytes()
    p.write_bytes(data[:-100])
    rc, msg = _load_rc(str(p))
    assert rc == 2 and "truncated" in msg


def test_null_arguments_rejected():
    import ctypes
    from erq import _lib
    L = _lib.lib()
    assert L.erq_index_load(None, None) == 1
    assert L.erq_search(None, None, 10, 1, None, None) == 1
    assert L.erq_padded_dim(None) == 0


def test_huge_header_is_rejected_before_allocating(tmp_path):
    """A header whose sections would need ~10^13 bytes, with cluster sizes that do sum to N:
    the file-size check returns ERQ_EIO before any buffer is sized from the header (an
    allocation that size would throw across the C ABI)."""
    p = tmp_path / "h.index"
    N = (1 << 32) - 1
    with open(p, "wb") as f:
        np.array([N, 4096, 2, 8], dtype="<u8").tofile(f)
        np.array([N - 7, 7], dtype="<u8").tofile(f)
        f.write(b"\0" * 4096)
    rc, msg = _load_rc(str(p))
    assert rc == 2 and "truncated" in msg


def test_trailing_bytes_are_eformat(small_data, tmp_path):
    from erq import index
    x, q, c, assign = small_data
    idx = index.build_index(x[:2000], c, assign[:2000], bits=4, seed=3)
    p = tmp_path / "t.index"
    idx.save(str(p))
    with open(p, "ab") as f:
        f.write(b"\0" * 8)
    rc, msg = _load_rc(str(p))
    assert rc == 3 and "after the index body" in msg


def test_batch_nprobe_rejects_bad_counts():
    import ctypes
    from erq import _lib
    L = _lib.lib()
    npq = np.array([4, 0], np.uint32)
    # a null handle is rejected before anything else is read
    assert L.erq_search_batch_nprobe(None, None, 2, 10, npq.ctypes.data_as(ctypes.c_void_p),
                                     None, None, 0) == 1


def test_nonha_file_size_checks(small_data, tmp_path):
    """The file does not record its block format: erq_index_load takes the one whose size the
    file has.  A non-HA file cut short, or one with bytes appended, matches neither and is
    refused before anything is uploaded."""
    from erq import index
    x, q, c, assign = small_data
    idx = index.build_index(x[:2000], c, assign[:2000], bits=4, seed=3, high_acc=False)
    p = tmp_path / "n.index"
    idx.save(str(p))
    data = p.read_bytes()
    p.write_bytes(data[:-100])
    rc, msg = _load_rc(str(p))
    assert rc == 3 and "neither block format" in msg
    p.write_bytes(data + b"\0" * 4)
    rc, msg = _load_rc(str(p))
    assert rc == 3 and "after the index body" in msg
