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
