"""GPU parity of the cluster-major batched search (erq_scan_batched_dev /
erq_search_batched_dev) against the oracle's round-synchronous schedule, bit for bit;
plus determinism and recall against the per-query path."""
import numpy as np
import pytest

from conftest import build_cached, gpu_available, rotate_np

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _build(key, n, dim, ncl, bits, seed=1, nq=48):
    from erq import data, index

    def make():
        x, cen = data.mixture_isotropic(n, dim, ncl, 0.3, seed=seed)
        q, _ = data.mixture_isotropic(nq, dim, ncl, 0.3, seed=seed + 1, centres=cen)
        c, assign = data.kmeans(x, ncl, 5, seed=seed + 3)
        idx = index.build_index(x, c, assign, bits=bits, seed=seed + 2)
        return idx, q, x
    return build_cached(("batched",) + key, make)


def _oracle_rounds(idx, qr, k, probe_lists):
    from oracle import oracle as orc
    oi = orc.OracleIndex(idx)
    ids = np.empty((qr.shape[0], k), np.uint32)
    ds = np.empty((qr.shape[0], k), np.float32)
    tot = np.zeros(3, np.int64)  # rank-0 blocks, survivors, later rounds' candidates
    for i in range(qr.shape[0]):
        pid, pd = probe_lists[i]
        ids[i], ds[i], st = oi.search_rounds(qr[i], k, pid, pd, stats=True)
        tot += np.asarray(st, np.int64)
    return ids, ds, tot


def _probe_lists(idx, qr, nprobes):
    from oracle import oracle as orc
    oi = orc.OracleIndex(idx)
    return [oi.centroid_topn(qr[i], int(nprobes[i])) for i in range(qr.shape[0])]


def _gpu_batched(idx, qr, k, probe_lists, max_nprobe, survivor_cap=None, pair_budget=None,
                 options=()):
    import torch
    from erq.gpu import IVF
    g = IVF.from_index(idx)
    for opt, val in options:
        g.set_option(opt, val)
        assert g.option(opt) == val
    if survivor_cap is not None:
        g.set_option(4, survivor_cap)   # ERQ_OPT_BATCH_SURVIVOR_CAP
    if pair_budget is not None:
        g.set_option(6, pair_budget)    # ERQ_OPT_BATCH_PAIR_BUDGET
        assert g.option(6) == pair_budget
    npq = np.array([len(p[0]) for p in probe_lists])
    off = np.concatenate([[0], np.cumsum(npq)]).astype(np.uint32)
    pid = np.concatenate([p[0] for p in probe_lists]).astype(np.uint32)
    pd = np.concatenate([p[1] for p in probe_lists]).astype(np.float32)
    t = lambda a: torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).cuda()
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    ids, d = g.scan_batched_dev(torch.from_numpy(qr).cuda(), k, t(off), t(pid), t(pd),
                                max_nprobe, stats=stats)
    torch.cuda.synchronize()
    out = ids.cpu().numpy().view(np.uint32), d.cpu().numpy(), stats.cpu().numpy()
    g.close()
    return out


def _check(idx, q, k, nprobes, label, survivor_cap=None, pair_budget=None, redo=0, options=()):
    qr = rotate_np(idx, q)
    pl = _probe_lists(idx, qr, nprobes)
    oid, od, ost = _oracle_rounds(idx, qr, k, pl)
    print(f"[{label}] k={k} nq={qr.shape[0]} max_nprobe={int(max(nprobes))}", flush=True)
    gid, gd, st = _gpu_batched(idx, qr, k, pl, int(max(nprobes)), survivor_cap, pair_budget,
                               options)
    print(f"[{label}] stats={st.tolist()} oracle blocks/survivors/candidates={ost.tolist()}",
          flush=True)
    # [5]: (query, round) rescans after a candidate overflow
    assert st[5] == redo, f"rescans {st[5]}, expected {redo}"
    # the work counters match the schedule's: rank-0 blocks, survivors, round candidates
    assert (st[0], st[2], st[4]) == tuple(int(v) for v in ost)
    np.testing.assert_array_equal(gid, oid, err_msg=f"{label}: ids")
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32), err_msg=f"{label}: dists")
    return gid, gd


@pytest.mark.parametrize("bits", [3, 4, 7, 9])
def test_batched_matches_round_oracle_d128(bits):
    idx, q, _ = _build(("d128", bits), 20000, 128, 64, bits)
    _check(idx, q, 100, np.full(q.shape[0], 24), f"D128 B{bits}")


@pytest.mark.parametrize("cap", [0, 1000])
def test_batched_survivor_list_overflow(cap):
    """With the survivor list capped (none, or part of a round's survivors), workgroups
    re-estimate the overflow themselves; the result stays bit-exact."""
    idx, q, _ = _build(("d128", 4), 20000, 128, 64, 4)
    _check(idx, q, 100, np.full(q.shape[0], 24), f"cap {cap}", survivor_cap=cap)


def test_batched_query_chunks():
    """A pair budget of 1024 with rounds up to 48 wide takes the 48 queries in chunks of
    21; each query's rounds are its own, so the result is the same."""
    idx, q, _ = _build(("d128", 4), 20000, 128, 64, 4)
    _check(idx, q, 100, np.full(q.shape[0], 64), "chunks", pair_budget=1024)


def test_batched_rounds_past_rank_192():
    """256 probes: rounds [4,16), [16,64), [64,192) and [192,256), all prepared at once."""
    idx, q, _ = _build(("d128", "k256"), 20000, 128, 256, 4, nq=24)
    _check(idx, q, 100, np.full(q.shape[0], 256), "rank 192+")


def test_batched_duplicate_vectors():
    """Every vector 24 times over (ties everywhere), rounds past the exact ranks."""
    from erq import data, index

    def make():
        base, cen = data.mixture_isotropic(300, 128, 8, 0.3, seed=31)
        x = np.repeat(base, 24, axis=0)
        x = x[np.random.default_rng(32).permutation(x.shape[0])]
        c, assign = data.kmeans(x, 8, 4, seed=33)
        idx = index.build_index(x, c, assign, bits=4, seed=34)
        q, _ = data.mixture_isotropic(24, 128, 8, 0.3, seed=35, centres=cen)
        return idx, q, x
    idx, q, _ = build_cached(("batched", "dup"), make)
    _check(idx, q, 100, np.full(q.shape[0], 8), "duplicates")


def test_batched_uneven_probe_counts():
    idx, q, _ = _build(("d128", 4), 20000, 128, 64, 4)
    rng = np.random.default_rng(5)
    _check(idx, q, 50, rng.integers(1, 64, q.shape[0]), "uneven")


def test_batched_d1024_b8():
    idx, q, _ = _build(("d960", 8), 8000, 960, 48, 8, nq=24)
    _check(idx, q, 100, np.full(q.shape[0], 20), "D960 B8")


def test_batched_d2048_b3():
    """Four 512-dim chunks (the largest MFMA shape: 88 KB of LDS per workgroup)."""
    idx, q, _ = _build(("d2048", 3), 4000, 2048, 24, 3, nq=16)
    _check(idx, q, 50, np.full(q.shape[0], 16), "D2048 B3")


@pytest.mark.parametrize("cap", [None, 0])
def test_batched_d1536_b4(cap):
    idx, q, _ = _build(("d1536", 4), 8000, 1536, 48, 4, nq=24)
    _check(idx, q, 100, np.full(q.shape[0], 20), f"D1536 B4 cap {cap}", survivor_cap=cap)


def test_batched_exact_ranks_equal_plain_and_is_deterministic():
    import torch
    from erq.gpu import IVF
    idx, q, _ = _build(("d128", 4), 20000, 128, 64, 4)
    qr = torch.from_numpy(rotate_np(idx, q)).cuda()
    g = IVF.from_index(idx)
    pi, pdist = g.search_dev(qr, 100, 1)
    bi, bd = g.search_batched_dev(qr, 100, 1)
    p4i, p4d = g.search_dev(qr, 100, 4)       # ranks [0,4) are the exact replay
    b4i, b4d = g.search_batched_dev(qr, 100, 4)
    b2i, b2d = g.search_batched_dev(qr, 100, 16)
    b3i, b3d = g.search_batched_dev(qr, 100, 16)
    torch.cuda.synchronize()
    assert torch.equal(pi, bi) and torch.equal(pdist, bd)
    assert torch.equal(p4i, b4i) and torch.equal(p4d, b4d)
    assert torch.equal(b2i, b3i) and torch.equal(b2d, b3d)
    g.close()


def test_batched_recall_close_to_plain():
    import torch
    from erq import data
    from erq.gpu import IVF
    idx, q, x = _build(("d128", 4), 20000, 128, 64, 4)
    gt, _ = data.exact_knn_np(x, q, 100)
    qr = torch.from_numpy(rotate_np(idx, q)).cuda()
    g = IVF.from_index(idx)
    for nprobe in (4, 16):
        pi, _ = g.search_dev(qr, 100, nprobe)
        bi, _ = g.search_batched_dev(qr, 100, nprobe)
        torch.cuda.synchronize()
        rp = data.recall_at_k(pi.cpu().numpy(), gt, 100)
        rb = data.recall_at_k(bi.cpu().numpy(), gt, 100)
        print(f"nprobe={nprobe} recall plain={rp:.4f} batched={rb:.4f}")
        assert rb >= rp - 0.005
    g.close()


def test_batched_search_in_hip_graph_replays_match_eager():
    """The whole batched search (probe GEMM + selection + rounds + merges) captured in a
    HIP graph; every replay, with eager work in between, equals the eager result."""
    import torch
    from erq.gpu import IVF
    idx, q, _ = _build(("d128", 4), 20000, 128, 64, 4)
    qr = torch.from_numpy(rotate_np(idx, q)).cuda()
    g = IVF.from_index(idx)
    ei, ed = g.search_batched_dev(qr, 100, 24)        # eager; also sizes the workspace
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        gi, gd = g.search_batched_dev(qr, 100, 24)
    for rep in range(4):
        gi.fill_(-1)
        gd.fill_(0)
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(gi, ei) and torch.equal(gd, ed), f"replay {rep}"
        xi, xd = g.search_batched_dev(qr, 100, 24)    # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(xi, ei)
    g.close()


@pytest.mark.parametrize("tiles,prep_wg", [(8, 6), (16, 1), (8, 3)])
def test_batched_item_tiles_and_prep_grid_keep_results(tiles, prep_wg):
    """ERQ_OPT_BATCH_ITEM_TILES (9) and ERQ_OPT_PREP_WG_PER_CU (10) change only how the
    rounds are split into work items and how the preparation is spread: the results and the
    work counters stay the oracle's."""
    idx, q, _ = _build(("d128", 4), 20000, 128, 64, 4)
    rng = np.random.default_rng(7)
    nprobes = rng.integers(1, 64, q.shape[0])
    _check(idx, q, 100, nprobes, f"tiles {tiles} prep {prep_wg}", options=((9, tiles), (10, prep_wg)))


def test_autotuner_keeps_results():
    import torch
    from erq.gpu import IVF
    idx, q, _ = _build(("d128", 4), 20000, 128, 64, 4)
    qr = torch.from_numpy(rotate_np(idx, q)).cuda()
    g = IVF.from_index(idx)
    pi, pd = g.search_dev(qr, 100, 16)
    bi, bd = g.search_batched_dev(qr, 100, 16)
    ms = g.autotune(qr, 100, 16)
    print("autotune best ms", ms, "waves", g.option(2), "wg/cu", g.option(3), "tiles",
          g.option(9), "prep wg/cu", g.option(10))
    assert 1 <= g.option(2) <= 4 and 1 <= g.option(3) <= 4
    assert g.option(9) in (8, 16) and g.option(10) in (2, 4, 6)
    pi2, pd2 = g.search_dev(qr, 100, 16)
    bi2, bd2 = g.search_batched_dev(qr, 100, 16)
    torch.cuda.synchronize()
    assert torch.equal(pi, pi2) and torch.equal(pd, pd2)
    assert torch.equal(bi, bi2) and torch.equal(bd, bd2)
    g.close()


def _skewed_index(small, large, n_small, n_large, dup_large=False, seed=31):
    """Clusters at growing distance along one axis (cluster c centred at 3c e_0), the first
    `small` of n_small vectors, the next `large` of n_large; with dup_large every vector of a
    large cluster is the same vector.  Queries sit at the first centre, so they probe the
    clusters in order."""
    from erq import index
    rng = np.random.default_rng(seed)
    dim = 128
    K = small + large
    cen = np.zeros((K, dim), np.float32)
    cen[:, 0] = 3.0 * np.arange(K)
    xs, lab = [], []
    for c in range(K):
        n = n_small if c < small else n_large
        if dup_large and c >= small:
            one = cen[c] + 0.3 * rng.standard_normal(dim).astype(np.float32)
            xs.append(np.repeat(one[None], n, axis=0))
        else:
            xs.append(cen[c] + 0.3 * rng.standard_normal((n, dim)).astype(np.float32))
        lab += [c] * n
    x = np.concatenate(xs)
    idx = index.build_index(x, cen, np.asarray(lab), bits=4, seed=seed + 1)
    q = cen[0] + 0.3 * rng.standard_normal((6, dim)).astype(np.float32)
    return idx, q


def test_batched_candidate_overflow_is_rescanned():
    """k = 1024 and 4 small nearest clusters (200 vectors): the pool is not full after the
    exact ranks, so every vector of the next 12 clusters (12 x 800 > 8192) is a candidate of
    round [4,16); each query's round is rescanned exactly, and the later round runs normally."""
    idx, q = _skewed_index(4, 16, 50, 800)
    _check(idx, q, 1024, np.full(q.shape[0], 20), "candidate overflow", redo=q.shape[0])


def test_batched_ties_beyond_merge_buffer_are_rescanned():
    """5000 copies of one vector in the fifth cluster: its 5000 equal distances sit at the
    k-th key (k = 1024), more than the merge keeps; the rescan orders them by id."""
    idx, q = _skewed_index(4, 1, 50, 5000, dup_large=True)
    _check(idx, q, 1024, np.full(q.shape[0], 5), "tied candidates", redo=q.shape[0])


def test_calls_on_different_streams_are_ordered():
    """A batched search enqueued on a side stream, then at once a host-pointer search on the
    handle's own stream (which reuses the same workspace): the second call waits for the
    first call's kernels, so the side-stream result equals the one computed alone."""
    import torch
    from erq.gpu import IVF
    idx, q, _ = _build(("d128", 4), 20000, 128, 64, 4, nq=48)
    qr = rotate_np(idx, q)
    qd = torch.from_numpy(np.tile(qr, (40, 1))).cuda()     # enough work to still be running
    g = IVF.from_index(idx)
    ref_ids, ref_d = g.search_batched_dev(qd, 100, 48)
    torch.cuda.synchronize()
    ref_h, _ = g.search_batch(qr[:8], 100, 48)
    side = torch.cuda.Stream()
    for _ in range(3):
        with torch.cuda.stream(side):
            ids, d = g.search_batched_dev(qd, 100, 48)
        hids, _ = g.search_batch(qr[:8], 100, 48)           # no synchronisation in between
        torch.cuda.synchronize()
        assert torch.equal(ids, ref_ids) and torch.equal(d, ref_d)
        np.testing.assert_array_equal(hids, ref_h)
    g.close()
