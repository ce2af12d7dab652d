"""Benchmark: QPS at recall@100 = 0.95 on the embedding-shaped workload
(N = 1M, D = 1536 unit-norm power-law mixture, 4096 IVF clusters, B = 4, K = 100,
20,000 queries per batch), plus the scan kernel's HBM roofline and a CPU baseline.

A "step" is one pass of the search over the whole query batch, inputs resident in HBM:
rotation (MFMA GEMM) -> centroid contraction + exact top-nprobe -> per-query scan.
Protocol (reference src/test_search.cpp:83-150): sweep nprobe, recall against exact
ground truth (GPU FP32, FP64 spot check), QPS at recall 0.95 by linear interpolation
between the two bracketing nprobes; each bracketing nprobe is timed over --steps steps.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "extended-rabitq_amd"))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=1536)
    ap.add_argument("--clusters", type=int, default=4096)
    ap.add_argument("--bits", type=int, default=4)
    ap.add_argument("--nq", type=int, default=20000)
    ap.add_argument("--k", type=int, default=100)
    ap.add_argument("--target", type=float, default=0.95)
    ap.add_argument("--kmeans-iters", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--nprobes", type=str, default="8,16,32,64,96,128,160,192,224,256,320,384,512")
    ap.add_argument("--sigma", type=float, default=1.15,
                    help="within-component spread relative to the component spread")
    ap.add_argument("--train-queries", type=int, default=200000)
    ap.add_argument("--label-nprobe", type=int, default=0,
                    help="nprobe of the label search (0: 256) and the learned lists' maximum")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB device copy first (HBM counter calibration)")
    ap.add_argument("--profile-plain", type=int, default=0, metavar="NPROBE",
                    help="only time the plain path (rotate, probe, scan) at this nprobe and "
                         "print its roofline: the command the kernel profiles are taken with")
    ap.add_argument("--profile-learned", type=float, default=0.0, metavar="MASS",
                    help="train the probe predictor, then only time the learned batched search "
                         "at this mass (the command its kernel profile is taken with)")
    ap.add_argument("--save-mlp", type=str, default="",
                    help="save the trained probe predictor to this .npz file")
    ap.add_argument("--load-mlp", type=str, default="",
                    help="load the probe predictor from this .npz file instead of training it")
    ap.add_argument("--sweep-only", action="store_true",
                    help="print the recall sweep of both paths and exit (no timing)")
    return ap.parse_args()


def build_workload(a, torch):
    from erq import data, index
    t0 = time.time()
    ncomp = a.clusters
    x, (basis, centres) = data.mixture_powerlaw_torch(a.n, a.dim, ncomp, a.sigma, seed=1,
                                                      alpha=1.0)
    q, _ = data.mixture_powerlaw_torch(a.nq, a.dim, ncomp, a.sigma, seed=2, alpha=1.0,
                                       centres=centres, basis=basis)
    log(f"data generated N={a.n} D={a.dim} nq={a.nq} in {time.time() - t0:.1f}s")
    t0 = time.time()
    c, assign = data.kmeans_torch(x, a.clusters, a.kmeans_iters, seed=4)
    log(f"k-means {a.clusters} clusters, {a.kmeans_iters} iters in {time.time() - t0:.1f}s")
    t0 = time.time()
    D = index.padded_dim(a.dim)
    P = index.random_rotation(D, 3)
    Pt = torch.from_numpy(P).cuda()

    def matmul(m, b):
        return (torch.from_numpy(np.ascontiguousarray(m)).cuda() @ Pt).cpu().numpy()

    idx = index.build_index(x.cpu().numpy(), c.cpu().numpy(), assign.cpu().numpy(), a.bits,
                            P=P, matmul=matmul, threads=16, chunk=131072)
    log(f"HA index built (B={a.bits}) in {time.time() - t0:.1f}s; "
        f"short {idx.short_data.nbytes / 1e6:.0f} MB, long {idx.long_code.nbytes / 1e6:.0f} MB")
    return x, q, idx, (basis, centres)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the search runs on one GPU (no multi-GPU path); --gpus must be 1"}))
        sys.exit(2)
    import torch

    from erq import data
    from erq.gpu import IVF
    torch.cuda.set_device(0)
    if a.calib_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 1073741824 bytes read, 1073741824 bytes written")
        del src, dst
    x, q, idx, mix = build_workload(a, torch)
    g = IVF.from_index(idx)
    D, EX = idx.D, idx.EX
    if a.profile_plain:
        del x
        profile_plain(a, g, q, idx, torch)
        g.close()
        return
    if a.profile_learned:
        del x
        profile_learned(a, g, q, idx, mix, torch)
        g.close()
        return

    t0 = time.time()
    gt, _ = data.exact_knn_torch(x, q, a.k)
    gt = gt.cpu().numpy()
    # FP64 spot check of the FP32 ground truth on 4 queries
    xs = x.cpu().numpy()
    qs = q[:4].cpu().numpy()
    g64, _ = data.exact_knn_np(xs, qs, a.k)
    spot = data.recall_at_k(gt[:4], g64, a.k)
    del xs
    log(f"ground truth (GPU FP32) in {time.time() - t0:.1f}s, FP64 spot-check agreement {spot:.4f}")
    del x
    torch.cuda.empty_cache()

    stream = torch.cuda.current_stream()
    stats = torch.zeros(4, dtype=torch.int64, device="cuda")

    stats = torch.zeros(8, dtype=torch.int64, device="cuda")

    def step(nprobe, path, timed=False):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record(stream)
        qr = g.rotate_dev(q)
        ev[1].record(stream)
        pid, pd = g.probe_dev(qr, nprobe)
        off = torch.arange(0, (a.nq + 1) * nprobe, nprobe, dtype=torch.int32, device="cuda")
        ev[2].record(stream)
        st = stats if timed else None
        if path == "plain":
            ids, dist = g.scan_dev(qr, a.k, off, pid, pd, stats=st)
        else:
            ids, dist = g.scan_batched_dev(qr, a.k, off, pid, pd, nprobe, stats=st)
        ev[3].record(stream)
        return ids, ev

    def measure(path):
        sweep = []
        for npb in [int(s) for s in a.nprobes.split(",")]:
            if npb > idx.K:
                break
            ids, _ = step(npb, path)
            torch.cuda.synchronize()
            rec = data.recall_at_k(ids.cpu().numpy().view(np.uint32), gt, a.k)
            sweep.append((npb, rec))
            log(f"[{path}] nprobe={npb:4d} recall@{a.k}={rec:.4f}")
            if rec >= a.target:
                break
        above = [s for s in sweep if s[1] >= a.target]
        below = [s for s in sweep if s[1] < a.target]
        if not above:
            raise SystemExit(f"recall {a.target} not reached by nprobe {sweep[-1][0]}")
        hi = above[0]
        lo = below[-1] if below else None

        def timed(npb):
            for _ in range(a.warmup):
                step(npb, path)
            torch.cuda.synchronize()
            stats.zero_()
            t_rot = t_probe = t_scan = 0.0
            evs = []
            t0 = time.perf_counter()
            for _ in range(a.steps):
                _, ev = step(npb, path, timed=True)
                evs.append(ev)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            for ev in evs:
                t_rot += ev[0].elapsed_time(ev[1])
                t_probe += ev[1].elapsed_time(ev[2])
                t_scan += ev[2].elapsed_time(ev[3])
            st = (stats.cpu().numpy() // a.steps).tolist()
            return dict(nprobe=npb, wall_ms=wall * 1e3 / a.steps, rot_ms=t_rot / a.steps,
                        probe_ms=t_probe / a.steps, scan_ms=t_scan / a.steps, stats=st)

        th = timed(hi[0])
        th.update(scan_volume(g, q, idx, hi[0], th, torch))
        log(f"[{path}] timed nprobe={hi[0]}: {th}")
        qps_hi = a.nq / (th["wall_ms"] / 1e3)
        tl = None
        if lo is not None:
            tl = timed(lo[0])
            log(f"[{path}] timed nprobe={lo[0]}: {tl}")
            qps_lo = a.nq / (tl["wall_ms"] / 1e3)
            f = (a.target - lo[1]) / (hi[1] - lo[1])
            qps = qps_lo + f * (qps_hi - qps_lo)
        else:
            qps = qps_hi
        return dict(qps=qps, hi=hi, lo=lo, th=th, tl=tl, sweep=sweep)

    if a.sweep_only:
        for path in ("plain", "batched"):
            for npb in [int(s) for s in a.nprobes.split(",")]:
                if npb > idx.K:
                    break
                ids, _ = step(npb, path)
                torch.cuda.synchronize()
                rec = data.recall_at_k(ids.cpu().numpy().view(np.uint32), gt, a.k)
                log(f"[{path}] sigma={a.sigma} nprobe={npb:4d} recall@{a.k}={rec:.4f}")
        return

    plain = measure("plain")
    try:
        batched = measure("batched")
    except Exception as e:  # reported, never hidden
        log(f"batched path failed: {e!r}")
        batched = None
    learned = None
    if not a.no_learned:
        try:
            learned = measure_learned(a, g, idx, q, gt, mix, step, stats, stream, torch,
                                      fixed_np=(batched or plain)["hi"][0])
        except Exception as e:  # reported, never hidden
            log(f"learned path failed: {e!r}")
    uneven = None
    try:
        uneven = measure_uneven(a, g, q, gt, torch)
    except Exception as e:  # reported, never hidden
        log(f"uneven-nprobe leg failed: {e!r}")
    best = batched if batched is not None and batched["qps"] > plain["qps"] else plain
    qps = best["qps"]
    hi, lo, th = plain["hi"], plain["lo"], plain["th"]
    ms_per_step = a.nq / qps * 1e3

    # roofline of the plain path's dominant kernel (the scan) at its upper bracketing nprobe
    roofline = scan_roofline(th["stats"], th["scan_ms"], D, EX)
    roofline["traffic"], roofline["traffic_source"] = profiled_traffic(a, hi[0])

    # the drop-in's host-pointer entry (erq_search_batch): rotated queries from host memory,
    # results back to host memory, device buffers allocated per call -- PCIe-inclusive
    qr_host = g.rotate_dev(q).cpu().numpy()
    g.search_batch(qr_host, a.k, hi[0])
    t0 = time.perf_counter()
    g.search_batch(qr_host, a.k, hi[0])
    host_qps = a.nq / (time.perf_counter() - t0)
    log(f"[host] erq_search_batch at nprobe {hi[0]}: {host_qps:.0f} queries/s")
    del qr_host

    cpu = cpu_baseline(idx, q, a, hi[0], torch)
    out = {
        "metric": "QPS at recall@100=0.95, D1536 N1M B4: learned-probe vs fixed nprobe",
        "value": round(learned["qps"] if learned else qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 codes / int32 FastScan sums / f32 distances",
        "data": "synthetic power-law Gaussian mixture (seeded), unit-norm",
        "config": {"workload": "embedding-shaped IVF search", "N": a.n, "D": D, "clusters": idx.K,
                   "bits": a.bits, "nq": a.nq, "k": a.k,
                   "path": "learned-probe batched" if learned else
                   ("batched" if best is batched else "plain"),
                   "sigma": a.sigma},
        "fixed_nprobe_qps_at_target": round(qps, 1),
        "host_pointer_plain_qps": round(host_qps, 1),
        "uneven_nprobe_graph": uneven,
        "learned": None if learned is None else {
            "qps_at_target": round(learned["qps"], 1),
            "vs_fixed": round(learned["qps"] / qps, 3),
            "mass_bracket": learned["bracket"], "train": learned["train"],
            "sweep": learned["sweep"]},
        "paths": {name: None if r is None else {
            "qps_at_target": round(r["qps"], 1),
            "nprobe_bracket": [r["lo"][0] if r["lo"] else None, r["hi"][0]],
            "recall_bracket": [round(r["lo"][1], 4) if r["lo"] else None, round(r["hi"][1], 4)],
            "stages_ms_at_hi": {"rotate": round(r["th"]["rot_ms"], 3),
                                "probe": round(r["th"]["probe_ms"], 3),
                                "scan": round(r["th"]["scan_ms"], 3)},
            "stats_at_hi": r["th"]["stats"],
            "scan_algorithmic_gbs_at_hi": r["th"].get("algorithmic_gbs"),
            "scan_unique_bytes_at_hi": r["th"].get("unique_bytes"),
            "sweep": [[n, round(x, 4)] for n, x in r["sweep"]]}
            for name, r in (("plain", plain), ("batched", batched))},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    g.close()


def scan_roofline(stats, scan_ms, D, EX):
    """Algorithmic bytes of the per-query scan (every probed block's 4D code bytes and 32
    norms, plus each survivor's EX-bit code and two factors) over its measured time."""
    scan_bytes = stats[0] * (4 * D + 128) + stats[2] * (D * EX // 8 + 8)
    achieved = scan_bytes / (scan_ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "scan_mfma_kernel", "algorithmic_bytes": int(scan_bytes)}


def scan_volume(g, q, idx, nprobe, th, torch):
    """Algorithmic scan bytes of one step at this nprobe (every (query, cluster) pair's
    blocks plus each survivor's EX-bit code and factors), the same bytes counted once per
    distinct probed cluster, and the algorithmic rate over the measured scan time."""
    D, EX = idx.D, idx.EX
    pid, _ = g.probe_dev(g.rotate_dev(q), nprobe)
    nblk = torch.from_numpy(((np.diff(idx.vec_start()) + 31) // 32).astype(np.int64)).cuda()
    p = pid.long().reshape(-1)
    blocks = int(nblk[p].sum())
    unique_blocks = int(nblk[torch.unique(p)].sum())
    alg = blocks * (4 * D + 128) + th["stats"][2] * (D * EX // 8 + 8)
    return {"pair_blocks": blocks, "algorithmic_bytes": alg,
            "algorithmic_gbs": round(alg / (th["scan_ms"] / 1e3) / 1e9, 1),
            "unique_bytes": unique_blocks * (4 * D + 128)}


def measure_uneven(a, g, q, gt, torch):
    """Per-query nprobe drawn uniformly from 8..256 (seed 6), the whole search (rotation,
    probe lists, batched scan) captured once in a HIP graph and replayed; launch shapes from
    the autotuner."""
    from erq import data
    rng = np.random.default_rng(6)
    npq = torch.from_numpy(rng.integers(8, 257, a.nq).astype(np.int64)).cuda()
    qr0 = g.rotate_dev(q)
    tuned = g.autotune(qr0, a.k, 128)
    cols = torch.arange(256, device="cuda")
    off = torch.zeros(a.nq + 1, dtype=torch.int64, device="cuda")
    off[1:] = torch.cumsum(npq, 0)
    off32 = off.to(torch.int32)
    keep = torch.nonzero((cols[None, :] < npq[:, None]).reshape(-1)).reshape(-1)

    def run():  # capturable: fixed shapes, no host synchronisation
        qr = g.rotate_dev(q)
        pid, pd = g.probe_dev(qr, 256)
        return g.scan_batched_dev(qr, a.k, off32, pid.reshape(-1).index_select(0, keep),
                                  pd.reshape(-1).index_select(0, keep), 256)

    ids, _ = run()  # eager: sizes the workspace and the boolean-mask outputs
    torch.cuda.synchronize()
    rec = data.recall_at_k(ids.cpu().numpy().view(np.uint32), gt, a.k)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        gids, _ = run()
    graph.replay()
    torch.cuda.synchronize()
    same = bool(torch.equal(gids, ids))
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(a.steps):
        graph.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    log(f"[uneven] nprobe U[8,256] recall@{a.k}={rec:.4f} graph replay {ms:.2f} ms "
        f"({a.nq / ms * 1e3:.0f} q/s), graph == eager: {same}")
    return {"qps": round(a.nq / ms * 1e3, 1), "recall": round(rec, 4), "ms": round(ms, 3),
            "mean_nprobe": round(float(npq.float().mean()), 1), "graph_equals_eager": same,
            "autotune_ms": [round(x, 3) for x in tuned],
            "plain_waves": g.option(2), "batch_wg_per_cu": g.option(3)}


def profiled_traffic(a, nprobe):
    """HBM bytes per scan launch from the committed rocprofv3 counter summary of this exact
    workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None if there is none."""
    if a.nq != 20000 or a.n % 1_000_000:
        return None, None
    name = (f"scan_mfma_kernel_plain_d{a.dim}_n{a.n // 1_000_000}m_b{a.bits}_"
            f"nprobe{nprobe}.md")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        m = re.search(r"traffic_bytes_per_launch: (\d+)", open(path).read())
    except OSError:
        return None, None
    return (int(m.group(1)), "profiles/" + name) if m else (None, None)


def profile_plain(a, g, q, idx, torch):
    """The plain path alone at one nprobe: W untimed and K timed steps, HIP events around
    each stage on the stream the kernels run on."""
    nprobe = a.profile_plain
    stream = torch.cuda.current_stream()
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    off = torch.arange(0, (a.nq + 1) * nprobe, nprobe, dtype=torch.int32, device="cuda")

    def one(st):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record(stream)
        qr = g.rotate_dev(q)
        ev[1].record(stream)
        pid, pd = g.probe_dev(qr, nprobe)
        ev[2].record(stream)
        g.scan_dev(qr, a.k, off, pid, pd, stats=st)
        ev[3].record(stream)
        return ev

    for _ in range(a.warmup):
        one(None)
    torch.cuda.synchronize()
    evs = [one(stats) for _ in range(a.steps)]
    torch.cuda.synchronize()
    ms = [sum(e[i].elapsed_time(e[i + 1]) for e in evs) / a.steps for i in range(3)]
    st = (stats.cpu().numpy() // a.steps).tolist()
    D, EX = idx.D, idx.EX
    flop_rot = 2.0 * a.nq * D * D
    flop_cen = 2.0 * a.nq * idx.K * D
    print(json.dumps({"profile": "plain", "nprobe": nprobe, "nq": a.nq, "steps": a.steps,
                      "rotate_ms": round(ms[0], 3), "probe_ms": round(ms[1], 3),
                      "scan_ms": round(ms[2], 3), "stats": st,
                      "rotation_tflops": round(flop_rot / ms[0] / 1e9, 1),
                      "roofline": scan_roofline(st, ms[2], D, EX),
                      "centroid_gemm_flop": flop_cen}), flush=True)


def train_probe_model(a, g, idx, mix, torch):
    """Labels from the fixed-nprobe batched search over separate training queries, then the
    probe MLP trained on them.  Returns (model, label_nprobe, losses, label_s, train_s)."""
    from erq import data
    from erq.probe_model import ProbeMLP, cluster_of_ids
    basis, centres = mix
    if a.load_mlp:
        m = ProbeMLP.load(a.load_mlp)
        log(f"probe MLP loaded from {a.load_mlp}")
        return m, min(a.label_nprobe or 256, idx.K, 2048), [], 0.0, 0.0
    t0 = time.time()
    qt, _ = data.mixture_powerlaw_torch(a.train_queries, a.dim, a.clusters, a.sigma, seed=5,
                                        alpha=1.0, centres=centres, basis=basis)
    qt_r = g.rotate_dev(qt)
    del qt
    c_of = torch.from_numpy(cluster_of_ids(idx)).cuda()
    label_np = min(a.label_nprobe or 256, idx.K, 2048)
    lab = torch.empty((a.train_queries, a.k), dtype=torch.int32, device="cuda")
    for s0 in range(0, a.train_queries, 20000):
        ids, _ = g.search_batched_dev(qt_r[s0:s0 + 20000], a.k, label_np)
        lab[s0:s0 + 20000] = c_of[ids.long()]
    torch.cuda.synchronize()
    t_lab = time.time() - t0
    log(f"labels from fixed nprobe={label_np} over {a.train_queries} queries in {t_lab:.1f}s")
    t0 = time.time()
    # inputs scaled to unit variance per coordinate (rotated unit-norm queries)
    m = ProbeMLP(idx.D, idx.K, hidden=512, max_batch=1024, seed=7,
                 input_scale=float(np.sqrt(idx.D)))
    losses = m.train(qt_r, lab, epochs=a.epochs, batch=1024, lr=2e-3, pos_weight=4.0, log=log,
                     schedule="cosine")
    torch.cuda.synchronize()
    if a.save_mlp:
        m.save(a.save_mlp)
    return m, label_np, losses, t_lab, time.time() - t0


def profile_learned(a, g, q, idx, mix, torch):
    m, label_np, _, _, _ = train_probe_model(a, g, idx, mix, torch)
    stream = torch.cuda.current_stream()
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")

    def one(st):
        return g.search_learned_dev(m, g.rotate_dev(q), a.k, a.profile_learned, 4, label_np,
                                    192, 0.5, stats=st)

    for _ in range(a.warmup):
        one(None)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(a.steps):
        one(stats)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    print(json.dumps({"profile": "learned", "mass": a.profile_learned, "nq": a.nq,
                      "steps": a.steps, "ms_per_search": round(ms, 3),
                      "qps": round(a.nq / ms * 1e3, 1),
                      "stats": (stats.cpu().numpy() // a.steps).tolist()}), flush=True)
    m.close()


def measure_learned(a, g, idx, q, gt, mix, step, stats, stream, torch, fixed_np):
    """Train the probe predictor on labels from fixed-nprobe search over separate training
    queries, then sweep the mass threshold and report QPS at the recall target."""
    from erq import data
    m, label_np, losses, t_lab, t_train = train_probe_model(a, g, idx, mix, torch)
    max_np = label_np

    def lstep(mass, timed=False):
        qr = g.rotate_dev(q)
        return g.search_learned_dev(m, qr, a.k, mass, 4, max_np, min(fixed_np, max_np),
                                    0.5, stats=stats if timed else None)

    sweep = []
    for mass in (0.9, 0.95, 0.97, 0.975, 0.98, 0.985, 0.99, 0.995, 0.998):
        ids, _ = lstep(mass)
        torch.cuda.synchronize()
        rec = data.recall_at_k(ids.cpu().numpy().view(np.uint32), gt, a.k)
        sweep.append((mass, rec))
        log(f"[learned] mass={mass} recall@{a.k}={rec:.4f}")
        if rec >= a.target:
            break
    above = [s for s in sweep if s[1] >= a.target]
    below = [s for s in sweep if s[1] < a.target]
    if not above:
        raise RuntimeError("learned path did not reach the recall target")

    def timed(mass):
        for _ in range(a.warmup):
            lstep(mass)
        torch.cuda.synchronize()
        stats.zero_()
        t1 = time.perf_counter()
        for _ in range(a.steps):
            lstep(mass, timed=True)
        torch.cuda.synchronize()
        st = (stats.cpu().numpy() // a.steps).tolist()
        return (time.perf_counter() - t1) * 1e3 / a.steps, st

    hi = above[0]
    ms_hi, st_hi = timed(hi[0])
    qps = a.nq / (ms_hi / 1e3)
    lo = below[-1] if below else None
    if lo is not None:
        ms_lo, _ = timed(lo[0])
        qlo = a.nq / (ms_lo / 1e3)
        f = (a.target - lo[1]) / (hi[1] - lo[1])
        qps = qlo + f * (qps - qlo)
    log(f"[learned] qps at target {qps:.0f} (mass {lo[0] if lo else None}..{hi[0]}), "
        f"stats {st_hi}")
    m.close()
    return dict(qps=qps, bracket=[lo[0] if lo else None, hi[0]],
                sweep=[[x, round(r, 4)] for x, r in sweep],
                train={"queries": a.train_queries, "label_nprobe": label_np,
                       "fallback_nprobe": fixed_np,
                       "epochs": a.epochs, "loss": [round(x, 3) for x in losses],
                       "label_s": round(t_lab, 1), "train_s": round(t_train, 1),
                       "fallbacks_at_hi": st_hi[7]})


def cpu_baseline(idx, q, a, nprobe, torch):
    """The CPU oracle (a scalar C port of the reference search), single thread, timed on
    a bounded sample of the same queries at the same nprobe: QPS on this host."""
    try:
        from oracle import oracle as orc
    except Exception as e:  # pragma: no cover
        return {"error": str(e)}
    oi = orc.OracleIndex(idx)
    qr = q[:512].cpu().numpy().astype(np.float64)
    qp = np.zeros((qr.shape[0], idx.D))
    qp[:, :qr.shape[1]] = qr
    qr = (qp @ idx.P.astype(np.float64)).astype(np.float32)
    n = 0
    t0 = time.perf_counter()
    while n < qr.shape[0] and time.perf_counter() - t0 < a.cpu_seconds:
        oi.search(qr[n], a.k, nprobe)
        n += 1
    dt = time.perf_counter() - t0
    log(f"cpu baseline: {n} queries in {dt:.1f}s (1 thread)")
    out = {"value": round(n / dt, 2), "unit": "queries/s", "cores": 1, "kind": "port",
           "sample": f"{n} queries of the same batch, nprobe={nprobe}, k={a.k}, single thread, "
                     f"scalar C oracle"}
    # all host cores (OpenMP over queries) on a bounded sample sized from the 1-thread rate
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or 10**6)
    m = int(min(qr.shape[0], max(threads, n / dt * threads * 2.0)))
    t0 = time.perf_counter()
    oi.search_batch(qr[:m], a.k, nprobe, threads=threads)
    dt = time.perf_counter() - t0
    log(f"cpu baseline: {m} queries in {dt:.1f}s ({threads} threads)")
    out["all_cores"] = {"value": round(m / dt, 2), "cores": threads,
                        "sample": f"{m} queries, OpenMP over queries"}
    return out


if __name__ == "__main__":
    main()
