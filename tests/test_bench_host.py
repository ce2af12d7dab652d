"""bench.py's host-side contract, on the CPU: the one-GPU refusal, the roofline arithmetic and
the traffic figure it reads from the committed counter profile."""
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(monkeypatch, argv=()):
    sys.path.insert(0, REPO)
    import bench
    monkeypatch.setattr(sys, "argv", ["bench.py", *argv])
    return bench, bench.parse()


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=REPO)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_roofline_arithmetic(monkeypatch):
    bench, _ = _bench(monkeypatch)
    # the per-query scan: blocks x (4D + 128) + survivors x (D*EX/8 + 8), over its time
    nbytes = 68441459 * 6272 + 11944812 * 584
    assert nbytes == 436240601056
    r = bench.roofline("replay", nbytes, 73.598)
    assert r["achieved"] == pytest.approx(436240601056 / 0.073598 / 1e9, abs=0.1)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0, abs=1e-4)
    assert (r["bound"], r["unit"], r["peak"], r["kernel"]) == ("hbm", "GB/s", 8000.0,
                                                              "scan_mfma_kernel")
    # a batched stage of 4 round dispatches: bytes and time per dispatch, same fraction
    s = bench.roofline("scan", 18046216368, 5.6057, launches=4)
    assert s["algorithmic_bytes"] == 18046216368 // 4
    assert s["ms_per_launch"] == pytest.approx(5.6057 / 4, abs=1e-4)
    assert s["frac"] == pytest.approx(18046216368 / 5.6057e-3 / 1e9 / 8000.0, abs=1e-4)
    assert (s["kernel"], s["launches_per_search"]) == ("batch_scan_kernel", 4)


def test_interpolation_brackets_the_target():
    sys.path.insert(0, REPO)
    import bench
    sweep = [[64, 0.90], [128, 0.94], [192, 0.96], [256, 0.97]]
    lo, hi, f = bench.interpolate(sweep, 0.95)
    assert (lo[0], hi[0]) == (128, 192) and f == pytest.approx(0.5)
    lo, hi, f = bench.interpolate(sweep, 0.85)   # below the first point: no lower bracket
    assert lo is None and hi[0] == 64 and f == 1.0
    assert bench.interpolate(sweep, 0.99) is None


def test_profiled_traffic_reads_committed_profile(monkeypatch):
    bench, a = _bench(monkeypatch)
    name = "scan_mfma_kernel_plain_d1536_n1m_b4_nprobe192.md"
    traffic, src = bench.profiled_traffic(a, name)
    assert src == "profiles/" + name
    text = open(os.path.join(REPO, src)).read()
    assert f"traffic_bytes_per_launch: {traffic}" in text
    # FETCH_SIZE (KiB, x2 copy calibration) + WRITE_SIZE (KiB, x1) of the same launch
    assert traffic > 436240601056  # at least the algorithmic bytes
    # another workload has no committed profile
    monkeypatch.setattr(a, "nq", 1000)
    assert bench.profiled_traffic(a, name) == (None, None)


def test_host_info_names_the_cpu():
    sys.path.insert(0, REPO)
    import bench
    h = bench.host_info()
    assert h["logical_cpus"] >= 1 and h["affinity_cpus"] >= 1 and h["cpu_model"]
