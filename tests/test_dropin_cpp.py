"""The C++ drop-in (extended-rabitq_amd/host/ivf_gpu.hpp) compiled with the reference's
test_search call pattern: load -> rotator().rotate -> search per query."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO, gpu_available


def _compile(tmp_path):
    exe = tmp_path / "test_search_dropin"
    lib = os.path.join(REPO, "extended-rabitq_amd", "lib")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(REPO, "include"),
                    "-I", os.path.join(REPO, "extended-rabitq_amd", "host"),
                    os.path.join(REPO, "tests", "cpp", "test_search_dropin.cpp"),
                    "-L", lib, "-lerq", f"-Wl,-rpath,{lib}", "-o", str(exe)], check=True)
    return exe


def test_dropin_compiles_and_reports_missing_index(tmp_path):
    exe = _compile(tmp_path)
    r = subprocess.run([str(exe), str(tmp_path / "none.index"), "q", "1", "10", "1", "o"],
                       capture_output=True, text=True)
    assert r.returncode == 3 and "cannot open" in r.stderr


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a GPU")
def test_dropin_search_matches_python_path(tmp_path, small_data):
    from erq import index
    from erq.gpu import IVF
    x, q, c, assign = small_data
    idx = index.build_index(x, c, assign, bits=4, seed=3)
    path = tmp_path / "i.index"
    idx.save(str(path))
    qf = tmp_path / "q.f32"
    q[:8].astype(np.float32).tofile(str(qf))
    exe = _compile(tmp_path)
    out = tmp_path / "r.u32"
    r = subprocess.run([str(exe), str(path), str(qf), "8", "100", "8", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(str(out), np.uint32).reshape(8, 100)
    g = IVF.load(str(path))
    qr = g.rotate(q[:8])
    want, _ = g.search_batch(qr, 100, 8)
    np.testing.assert_array_equal(got, want)
    g.close()


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a GPU")
def test_dropin_timing_file(tmp_path, small_data):
    """The timing mode searches every query `rounds` times with the same results and writes
    test_search's QPS (queries over summed search time plus rotation time) per round."""
    import json
    from erq import index
    x, q, c, assign = small_data
    idx = index.build_index(x, c, assign, bits=4, seed=3)
    path = tmp_path / "i.index"
    idx.save(str(path))
    qf = tmp_path / "q.f32"
    q[:16].astype(np.float32).tofile(str(qf))
    exe = _compile(tmp_path)
    out1, out2, tj = tmp_path / "a.u32", tmp_path / "b.u32", tmp_path / "t.json"
    r = subprocess.run([str(exe), str(path), str(qf), "16", "50", "6", str(out1)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), str(path), str(qf), "16", "50", "6", str(out2), str(tj), "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(np.fromfile(str(out1), np.uint32), np.fromfile(str(out2), np.uint32))
    t = json.loads(tj.read_text())
    assert t["nq"] == 16 and t["nprobe"] == 6 and len(t["rounds"]) == 2
    for rd in t["rounds"]:
        assert rd["search_us"] > 0
        assert abs(rd["qps"] - 16 / ((rd["search_us"] + t["rotate_us"]) / 1e6)) <= 1e-3 * rd["qps"] + 0.1
    assert 0 < t["median_us"] and 0 < t["mean_us"]
