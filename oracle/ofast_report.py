"""How far the reference binary's -Ofast float rewrites (F1..F6 in erq_oracle.c) move the
search results from the source's literal semantics, measured with the oracle on data shaped
like the headline workload (unit-norm power-law mixture, D = 1536, B = 4, clusters of ~244
vectors, the same probed fraction of the index as nprobe 192 of 4096).

For every probed (query, cluster) pair and every vector of it: the re-estimated distance
under both semantics, |delta| relative to |o - c|^2 + y^2; the lower-bound mask against the
literal search's final k-th distance; and over whole searches, the survivors and the top-k
ids.  Test infrastructure: runs on the CPU (oracle), prints one JSON line.

  python oracle/ofast_report.py [--n 250000] [--clusters 1024] [--nprobe 48] [--nq 200]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
# run as a script, this folder is first on sys.path, where oracle.py would shadow the package
sys.path[:] = [p for p in sys.path if os.path.abspath(p or ".") != HERE]
sys.path[:0] = [REPO, os.path.join(REPO, "extended-rabitq_amd")]


def build(n, dim, clusters, bits, nq, sigma, seed=1):
    from erq import data, index
    x, (basis, centres) = data.mixture_powerlaw(n + nq, dim, clusters, sigma, seed=seed,
                                                alpha=1.0)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    q, x = x[n:], x[:n]
    c, assign = data.kmeans(x, clusters, 6, seed=seed + 3)
    idx = index.build_index(x, c, assign, bits=bits, seed=seed + 2)
    qp = np.zeros((nq, idx.D), np.float32)
    qp[:, :dim] = q
    return idx, (qp.astype(np.float64) @ idx.P.astype(np.float64)).astype(np.float32)


def measure(idx, qr, k, nprobe):
    from oracle import oracle as orc
    oi = orc.OracleIndex(idx)
    onorm_of = {}
    bs = idx.block_start()
    rel_max, n_pairs, n_vec, mask_flips = 0.0, 0, 0, 0
    top_diff, surv = 0, [0, 0]
    for i in range(qr.shape[0]):
        pid, pd = oi.centroid_topn(qr[i], nprobe)
        ids_l, d_l, st_l = oi.search(qr[i], k, nprobe, stats=True)
        with orc.fast_math():
            ids_f, d_f, st_f = oi.search(qr[i], k, nprobe, stats=True)
        surv[0] += st_l[2]
        surv[1] += st_f[2]
        top_diff += k - len(np.intersect1d(ids_l, ids_f))
        distk = float(d_l[-1])
        for c, sy in zip(pid, pd):
            n = int(idx.cluster_sizes[c])
            if n == 0:
                continue
            if c not in onorm_of:
                nb = (n + 31) // 32
                off = int(bs[c]) * idx.block_bytes
                blocks = idx.short_data[off:off + nb * idx.block_bytes].reshape(nb, -1)
                onorm_of[c] = blocks[:, 4 * idx.D:].copy().view(np.float32).reshape(-1)[:n]
            lo_l, ex_l = oi.pair_estimates(qr[i], c, sy)
            with orc.fast_math():
                lo_f, ex_f = oi.pair_estimates(qr[i], c, sy)
            scale = onorm_of[c].astype(np.float64) ** 2 + float(sy)
            rel = np.abs(ex_f.astype(np.float64) - ex_l) / scale
            rel_max = max(rel_max, float(rel.max()))
            mask_flips += int(np.count_nonzero((lo_l < distk) != (lo_f < distk)))
            n_pairs += 1
            n_vec += n
    return dict(queries=int(qr.shape[0]), pairs=n_pairs, vectors=n_vec,
                max_rel_ex_dist_delta=rel_max, mask_flips_at_final_kth=mask_flips,
                survivors_literal=int(surv[0]), survivors_fast=int(surv[1]),
                topk_ids_changed=int(top_diff), topk_total=int(k * qr.shape[0]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=250000)
    ap.add_argument("--dim", type=int, default=1536)
    ap.add_argument("--clusters", type=int, default=1024)
    ap.add_argument("--bits", type=int, default=4)
    ap.add_argument("--nq", type=int, default=200)
    ap.add_argument("--k", type=int, default=100)
    ap.add_argument("--nprobe", type=int, default=48)
    ap.add_argument("--sigma", type=float, default=1.15)
    a = ap.parse_args()
    t0 = time.time()
    idx, qr = build(a.n, a.dim, a.clusters, a.bits, a.nq, a.sigma)
    out = measure(idx, qr, a.k, a.nprobe)
    out.update(workload=dict(N=a.n, D=idx.D, clusters=a.clusters, bits=a.bits, k=a.k,
                             nprobe=a.nprobe, sigma=a.sigma), seconds=round(time.time() - t0, 1))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
