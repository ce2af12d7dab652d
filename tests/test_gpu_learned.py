"""Learned probe predictor on the GPU: BF16 MFMA forward/backward against the FP64 MLP
oracle (with and without the GPU's BF16 rounding steps), training, adaptive probing and
the fixed-nprobe fallback."""
import numpy as np
import pytest

from conftest import build_cached, gpu_available, rotate_np

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _mlp_case(din=64, hid=64, dout=40, B=64, seed=0):
    rng = np.random.default_rng(seed)
    n = hid * din + hid + hid * hid + hid + dout * hid + dout
    theta = (rng.standard_normal(n) * 0.2).astype(np.float32)
    x = rng.standard_normal((B, din)).astype(np.float32)
    y = (rng.random((B, dout)) < 0.15).astype(np.float32)
    return theta, x, y


@pytest.mark.parametrize("xs", [1.0, 3.7])
def test_mlp_forward_and_gradients_match_oracle(xs):
    import torch
    from erq.probe_model import ProbeMLP
    from oracle import mlp_oracle as mo
    din, hid, dout, B = 64, 64, 40, 64
    theta, x, y = _mlp_case(din, hid, dout, B)
    if xs != 1.0:
        x = x / xs   # keep the pre-activations at the unscaled case's size
    m = ProbeMLP(din, dout, hidden=hid, max_batch=B, seed=1, input_scale=xs)
    m.set_params(theta)
    xt, yt = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    p = m.forward_dev(xt)
    loss = m.loss_grad_dev(xt, yt, 3.0)
    torch.cuda.synchronize()
    g = m.get_grads()
    L_ref, g_ref, p_ref = mo.grad_bf16(theta.astype(np.float64), x, y, din, hid, dout, 3.0, xs)
    # same precision steps: only FP32 accumulation order differs
    np.testing.assert_allclose(p.cpu().numpy(), p_ref, rtol=2e-3, atol=1e-5)
    assert abs(float(loss[0]) - L_ref) <= 1e-3 * abs(L_ref)
    scale = np.abs(g_ref).max()
    assert np.abs(g - g_ref).max() <= 2e-2 * scale, np.abs(g - g_ref).max() / scale
    # against the exact FP64 network: the BF16 path keeps the gradient's direction
    g64 = mo.grad(theta.astype(np.float64), x, y, din, hid, dout, 3.0, xs)
    assert g @ g64 / np.linalg.norm(g) / np.linalg.norm(g64) > 0.995
    m.close()


def test_mlp_configured_size_matches_oracle():
    """The configured network, 1536 -> 512 -> 512 -> 4096 (inputs scaled by sqrt(1536)), on a
    batch of 32: forward and gradients against the FP64 oracle with the GPU's BF16 steps."""
    import torch
    from erq.probe_model import ProbeMLP
    from oracle import mlp_oracle as mo
    din, hid, dout, B = 1536, 512, 4096, 32
    rng = np.random.default_rng(21)
    sizes = [(hid, din), (hid,), (hid, hid), (hid,), (dout, hid), (dout,)]
    parts = []
    for shp in sizes:   # He-uniform weights, small random biases
        fan = shp[1] if len(shp) == 2 else 1
        a = np.sqrt(6.0 / fan) if len(shp) == 2 else 0.1
        parts.append(rng.uniform(-a, a, int(np.prod(shp))))
    theta = np.concatenate(parts).astype(np.float32)
    xs = float(np.sqrt(din))
    x = rng.standard_normal((B, din)).astype(np.float32) / xs   # unit-norm-like queries
    y = (rng.random((B, dout)) < 0.02).astype(np.float32)
    m = ProbeMLP(din, dout, hidden=hid, max_batch=B, seed=1, input_scale=xs)
    m.set_params(theta)
    xt, yt = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    p = m.forward_dev(xt)
    loss = m.loss_grad_dev(xt, yt, 4.0)
    torch.cuda.synchronize()
    g = m.get_grads()
    L_ref, g_ref, p_ref = mo.grad_bf16(theta.astype(np.float64), x, y, din, hid, dout, 4.0, xs)
    np.testing.assert_allclose(p.cpu().numpy(), p_ref, rtol=5e-3, atol=1e-5)
    assert abs(float(loss[0]) - L_ref) <= 1e-3 * abs(L_ref)
    # per parameter block: FP32 accumulation order is the only difference from the oracle
    o = 0
    for shp in sizes:
        n = int(np.prod(shp))
        gb, rb = g[o:o + n], g_ref[o:o + n]
        assert np.abs(gb - rb).max() <= 2e-2 * np.abs(rb).max(), (shp, np.abs(gb - rb).max())
        o += n
    g64 = mo.grad(theta.astype(np.float64), x, y, din, hid, dout, 4.0, xs)
    cos = g @ g64 / np.linalg.norm(g) / np.linalg.norm(g64)
    print(f"configured MLP: loss {float(loss[0]):.4f} (oracle {L_ref:.4f}), cos(g, g64) {cos:.5f}")
    assert cos > 0.99
    m.close()


def test_mlp_non_multiple_batch_forward_and_chunking():
    import torch
    from erq.probe_model import ProbeMLP
    from oracle import mlp_oracle as mo
    din, hid, dout = 64, 64, 40
    theta, x, _ = _mlp_case(din, hid, dout, 100, seed=2)
    m = ProbeMLP(din, dout, hidden=hid, max_batch=32, seed=1)
    m.set_params(theta)
    p = m.forward_dev(torch.from_numpy(x).cuda()).cpu().numpy()
    _, _, p_ref = mo.grad_bf16(theta.astype(np.float64), x, np.zeros((100, dout)), din, hid,
                               dout, 1.0)
    np.testing.assert_allclose(p, p_ref, rtol=2e-3, atol=1e-5)
    # batches above max_batch run through the inference buffers; every row is computed
    # as in a batch of at most max_batch rows
    xt = torch.from_numpy(x).cuda()
    parts = [m.forward_dev(xt[i:i + 32]) for i in range(0, 100, 32)]
    assert torch.equal(torch.cat(parts), torch.from_numpy(p).cuda())
    m.close()


def test_mlp_training_reduces_loss():
    import torch
    from erq.probe_model import ProbeMLP
    rng = np.random.default_rng(5)
    din, dout, n = 64, 32, 2048
    W = rng.standard_normal((dout, din)).astype(np.float32)
    x = rng.standard_normal((n, din)).astype(np.float32)
    top = np.argsort(-(x @ W.T), axis=1)[:, :3].astype(np.int32)   # learnable labels
    m = ProbeMLP(din, dout, hidden=64, max_batch=256, seed=3)
    losses = m.train(torch.from_numpy(x).cuda(), torch.from_numpy(top).cuda(), epochs=6,
                     batch=256, lr=3e-3, pos_weight=2.0)
    print("losses", losses)
    assert losses[-1] < 0.6 * losses[0]
    m.close()


def _small_index():
    from erq import data, index

    def make():
        x, cen = data.mixture_isotropic(30000, 128, 64, 0.3, seed=1)
        qtr, _ = data.mixture_isotropic(4096, 128, 64, 0.3, seed=5, centres=cen)
        qte, _ = data.mixture_isotropic(256, 128, 64, 0.3, seed=6, centres=cen)
        c, assign = data.kmeans(x, 64, 6, seed=4)
        idx = index.build_index(x, c, assign, bits=4, seed=3)
        return idx, x, qtr, qte
    return build_cached(("learned-d128",), make)


def test_learned_probe_search_end_to_end():
    import torch
    from erq import data
    from erq.gpu import IVF
    from erq.probe_model import ProbeMLP, cluster_of_ids
    idx, x, qtr, qte = _small_index()
    g = IVF.from_index(idx)
    c_of = torch.from_numpy(cluster_of_ids(idx)).cuda()
    qtr_r = torch.from_numpy(rotate_np(idx, qtr)).cuda()
    qte_r = torch.from_numpy(rotate_np(idx, qte)).cuda()
    ids, _ = g.search_batched_dev(qtr_r, 100, 32)
    labels = c_of[ids.long()]
    m = ProbeMLP(idx.D, idx.K, hidden=128, max_batch=512, seed=0)
    losses = m.train(qtr_r, labels, epochs=8, batch=512, lr=2e-3, pos_weight=4.0)
    assert losses[-1] < losses[0]
    gt, _ = data.exact_knn_np(x, qte, 100)
    # the gain: the recall the learned lists reach at mass 0.8 costs fixed-nprobe search at
    # least twice as many probes per query (interpolating its recall curve)
    fixed = []
    for npb in range(1, 9):
        fi, _ = g.search_batched_dev(qte_r, 100, npb)
        fixed.append((npb, data.recall_at_k(fi.cpu().numpy(), gt, 100)))
    lg, _ = g.search_learned_dev(m, qte_r, 100, mass=0.8, min_nprobe=1, max_nprobe=48,
                                 fallback_nprobe=16, min_conf=0.0)
    off8, _, _, _ = g.probe_learned_dev(m, qte_r, 0.8, 1, 48, 16, 0.0)
    torch.cuda.synchronize()
    n_learned = float(np.diff(off8.cpu().numpy()).mean())
    r_learned = data.recall_at_k(lg.cpu().numpy(), gt, 100)
    n_fixed = None
    for (n0, r0), (n1, r1) in zip(fixed, fixed[1:]):
        if r0 < r_learned <= r1:
            n_fixed = n0 + (r_learned - r0) / (r1 - r0) * (n1 - n0)
            break
    print(f"learned (mass 0.8): recall {r_learned:.4f} at {n_learned:.2f} probes per query; "
          f"fixed needs {n_fixed} probes; fixed curve {[(n, round(r, 4)) for n, r in fixed]}")
    assert r_learned > fixed[0][1], "the learned lists must beat one probe per query"
    assert n_fixed is not None and n_fixed >= 2.0 * n_learned
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    li, ld = g.search_learned_dev(m, qte_r, 100, mass=0.95, min_nprobe=2, max_nprobe=48,
                                  fallback_nprobe=16, min_conf=0.0, stats=stats)
    off, pid, pd, nfb = g.probe_learned_dev(m, qte_r, 0.95, 2, 48, 16, 0.0)
    torch.cuda.synchronize()
    offs = off.cpu().numpy()
    # probe lists are ordered by exact centroid distance and hold distinct clusters
    pid_h, pd_h = pid.cpu().numpy(), pd.cpu().numpy()
    from oracle import oracle as orc
    oi = orc.OracleIndex(idx)
    qte_np = rotate_np(idx, qte)
    for q in range(0, 256, 37):
        seg = slice(offs[q], offs[q + 1])
        assert np.all(np.diff(pd_h[seg]) >= 0)
        assert len(set(pid_h[seg].tolist())) == offs[q + 1] - offs[q]
        # the distances are the oracle's L2Sqr16, bit for bit
        all_i, all_d = oi.centroid_topn(qte_np[q], idx.K)
        want = np.empty(idx.K, np.float32)
        want[all_i] = all_d
        np.testing.assert_array_equal(pd_h[seg].view(np.uint32),
                                      want[pid_h[seg]].view(np.uint32))
    # no confidence -> every query falls back to the fixed list = batched fixed search
    fb_i, fb_d = g.search_learned_dev(m, qte_r, 100, mass=0.95, min_nprobe=2, max_nprobe=48,
                                      fallback_nprobe=16, min_conf=1e30)
    bi, bd = g.search_batched_dev(qte_r, 100, 16)
    torch.cuda.synchronize()
    assert torch.equal(fb_i, bi) and torch.equal(fb_d, bd)
    m.close()
    g.close()


def _kept_by_rule(p, mass, min_np, max_np):
    """The mass-prefix rule of erq_probe_learned_dev restated in numpy with the selection
    kernel's float32 sums: clusters by probability (descending, ties by id), the first
    T = min(max_np, K); thread t of 256 sums entries [t per, (t + 1) per) in order, the
    thread sums are scanned (Hillis-Steele), the goal is float32(mass) times the total, and
    n is the first prefix (thread's base plus its own entries, in order) reaching it; then
    n is clamped to [max(min_np, 1), T]."""
    K = p.size
    order = np.lexsort((np.arange(K), -p.astype(np.float64)))
    T = min(max_np, K)
    top = order[:T]
    pv = p[top].astype(np.float32)
    L = T
    per = (L + 255) // 256
    own = np.zeros(256, np.float32)
    for t in range(256):
        acc = np.float32(0)
        for i in range(t * per, min(L, (t + 1) * per)):
            acc = np.float32(acc + pv[i])
        own[t] = acc
    s = own.copy()
    o = 1
    while o < 256:
        s = np.concatenate([s[:o], (s[o:] + s[:-o]).astype(np.float32)])
        o <<= 1
    goal = np.float32(np.float32(mass) * s[255])
    n = L
    for t in range(256):
        acc = np.float32(s[t - 1]) if t else np.float32(0)
        for i in range(t * per, min(L, (t + 1) * per)):
            acc = np.float32(acc + pv[i])
            if acc >= goal:
                n = min(n, i + 1)
                break
    n = max(min(max(n, min_np), T), 1)
    return top[:n]


@pytest.mark.parametrize("mass,min_np,max_np", [(0.5, 1, 48), (0.9, 2, 48), (0.99, 4, 30),
                                                (1.0, 1, 64)])
def test_learned_kept_set_is_the_mass_prefix_rule(mass, min_np, max_np):
    """erq_probe_learned_dev keeps exactly the clusters the mass-prefix rule names (restated
    in numpy from the model's own probabilities), ordered by (exact distance, id)."""
    import torch
    from erq.gpu import IVF
    from erq.probe_model import ProbeMLP
    from oracle import oracle as orc
    idx, x, qtr, qte = _small_index()
    g = IVF.from_index(idx)
    m = ProbeMLP(idx.D, idx.K, hidden=128, max_batch=512, seed=4)
    qte_np = rotate_np(idx, qte)
    q = torch.from_numpy(qte_np).cuda()
    prob = m.forward_dev(q).cpu().numpy()
    off, pid, pd, nfb = g.probe_learned_dev(m, q, mass, min_np, max_np, 8, 0.0)
    torch.cuda.synchronize()
    off, pid, pd = off.cpu().numpy(), pid.cpu().numpy().view(np.uint32), pd.cpu().numpy()
    oi = orc.OracleIndex(idx)
    lens = []
    for i in range(0, qte_np.shape[0], 5):
        kept = _kept_by_rule(prob[i], mass, min_np, max_np)
        all_i, all_d = oi.centroid_topn(qte_np[i], idx.K)
        dist = np.empty(idx.K, np.float32)
        dist[all_i] = all_d
        want = kept[np.lexsort((kept, dist[kept]))]
        seg = slice(off[i], off[i + 1])
        np.testing.assert_array_equal(pid[seg], want.astype(np.uint32), err_msg=f"query {i}")
        np.testing.assert_array_equal(pd[seg].view(np.uint32), dist[want].view(np.uint32))
        lens.append(off[i + 1] - off[i])
    print(f"mass {mass}: kept {np.mean(lens):.1f} clusters on average")
    m.close()
    g.close()


def test_learned_lists_of_many_queries():
    """More queries than one 8192-length tile of the offsets scan: the CSR offsets are the
    running sum of the kept lengths, and lists past the first tile hold what the rule keeps."""
    import torch
    from erq.gpu import IVF
    from erq.probe_model import ProbeMLP
    from oracle import oracle as orc
    idx, x, qtr, qte = _small_index()
    g = IVF.from_index(idx)
    m = ProbeMLP(idx.D, idx.K, hidden=128, max_batch=512, seed=4)
    qr = rotate_np(idx, np.concatenate([qtr, qtr, qtr, qte]))   # 12544 queries
    q = torch.from_numpy(qr).cuda()
    prob = m.forward_dev(q).cpu().numpy()
    mass, min_np, max_np = 0.9, 2, 48
    off, pid, pd, nfb = g.probe_learned_dev(m, q, mass, min_np, max_np, 8, 0.0)
    torch.cuda.synchronize()
    off, pid = off.cpu().numpy().astype(np.int64), pid.cpu().numpy().view(np.uint32)
    lens = np.diff(off)
    assert off[0] == 0 and np.all(lens >= min_np) and np.all(lens <= max_np)
    assert off[-1] <= pid.size
    oi = orc.OracleIndex(idx)
    for i in list(range(0, qr.shape[0], 997)) + [8191, 8192, 8193, qr.shape[0] - 1]:
        kept = _kept_by_rule(prob[i], mass, min_np, max_np)
        all_i, all_d = oi.centroid_topn(qr[i], idx.K)
        dist = np.empty(idx.K, np.float32)
        dist[all_i] = all_d
        want = kept[np.lexsort((kept, dist[kept]))]
        np.testing.assert_array_equal(pid[off[i]:off[i + 1]], want.astype(np.uint32),
                                      err_msg=f"query {i}")
    m.close()
    g.close()


def test_mlp_large_batch_tiles_match_small_batches():
    """A batch large enough for the GEMM's 128-row tiles (layer 3 at 4096 rows x 4096 outputs)
    gives every row the probabilities that batches of 64 rows (64-row tiles) give."""
    import torch
    from erq.probe_model import ProbeMLP
    rng = np.random.default_rng(11)
    din, hid, dout, rows = 256, 128, 4096, 4096
    m = ProbeMLP(din, dout, hidden=hid, max_batch=64, seed=3, input_scale=16.0)
    x = torch.from_numpy((rng.standard_normal((rows, din)) / 16.0).astype(np.float32)).cuda()
    p_big = m.forward_dev(x)
    p_small = torch.cat([m.forward_dev(x[i:i + 64]) for i in range(0, rows, 64)])
    torch.cuda.synchronize()
    assert torch.equal(p_big, p_small)
    m.close()


def test_mlp_forward_rejects_unaligned_buffers():
    """Inference reads its input and writes its probabilities 16 bytes at a time: a buffer
    off that alignment is refused with ERQ_EINVAL before anything is launched."""
    import ctypes
    import torch
    from erq import _lib
    from erq.probe_model import ProbeMLP
    m = ProbeMLP(64, 40, hidden=64, max_batch=32, seed=1)
    x = torch.zeros(8 * 64 + 1, dtype=torch.float32, device="cuda")
    p = torch.zeros(8 * 40 + 1, dtype=torch.float32, device="cuda")
    L = _lib.lib()
    rc = L.erq_mlp_forward_dev(m.handle, ctypes.c_void_p(x.data_ptr() + 4), 8,
                               ctypes.c_void_p(p.data_ptr()), None)
    assert rc != 0 and b"16-byte aligned" in L.erq_last_error()
    rc = L.erq_mlp_forward_dev(m.handle, ctypes.c_void_p(x.data_ptr()), 8,
                               ctypes.c_void_p(p.data_ptr() + 4), None)
    assert rc != 0
    assert L.erq_mlp_forward_dev(m.handle, ctypes.c_void_p(x.data_ptr()), 8,
                                 ctypes.c_void_p(p.data_ptr()), None) == 0
    torch.cuda.synchronize()
    m.close()


def test_probe_model_save_load_round_trip(tmp_path):
    import torch
    from erq.probe_model import ProbeMLP
    theta, x, _ = _mlp_case(64, 64, 40, 32, seed=9)
    m = ProbeMLP(64, 40, hidden=64, max_batch=32, seed=1, input_scale=2.5)
    m.set_params(theta)
    path = str(tmp_path / "probe.npz")
    m.save(path)
    m2 = ProbeMLP.load(path, max_batch=32)
    assert (m2.in_dim, m2.out_dim, m2.hidden) == (64, 40, 64) and m2.input_scale() == 2.5
    xt = torch.from_numpy(x).cuda()
    assert torch.equal(m.forward_dev(xt), m2.forward_dev(xt))
    m.close()
    m2.close()


def test_learned_mixed_fallback_lists():
    """min_conf between the queries' total predicted masses: the queries below it take the
    fixed list (their fb_np nearest clusters by exact distance, ties by id), computed across
    the GPU after the selection; the others keep the mass prefix.  Both against the oracle."""
    import torch
    from erq.gpu import IVF
    from erq.probe_model import ProbeMLP
    from oracle import oracle as orc
    idx, x, qtr, qte = _small_index()
    g = IVF.from_index(idx)
    m = ProbeMLP(idx.D, idx.K, hidden=128, max_batch=512, seed=4)
    qte_np = rotate_np(idx, qte)
    q = torch.from_numpy(qte_np).cuda()
    prob = m.forward_dev(q).cpu().numpy()
    tot = prob.astype(np.float64).sum(1)
    min_conf = float(np.median(tot))
    fb_np, mass, min_np, max_np = 12, 0.9, 2, 48
    off, pid, pd, nfb = g.probe_learned_dev(m, q, mass, min_np, max_np, fb_np, min_conf)
    torch.cuda.synchronize()
    off, pid, pd = off.cpu().numpy(), pid.cpu().numpy().view(np.uint32), pd.cpu().numpy()
    oi = orc.OracleIndex(idx)
    nf = 0
    for i in range(qte_np.shape[0]):
        all_i, all_d = oi.centroid_topn(qte_np[i], idx.K)
        seg = slice(off[i], off[i + 1])
        # the kernel's float32 total decides; keep clear of the boundary
        if abs(tot[i] - min_conf) < 1e-3 * max(1.0, min_conf):
            continue
        if tot[i] < min_conf:
            nf += 1
            np.testing.assert_array_equal(pid[seg], all_i[:fb_np], err_msg=f"query {i}")
            np.testing.assert_array_equal(pd[seg].view(np.uint32), all_d[:fb_np].view(np.uint32))
        else:
            kept = _kept_by_rule(prob[i], mass, min_np, max_np)
            dist = np.empty(idx.K, np.float32)
            dist[all_i] = all_d
            want = kept[np.lexsort((kept, dist[kept]))]
            np.testing.assert_array_equal(pid[seg], want.astype(np.uint32), err_msg=f"query {i}")
    assert nf > 0 and int(nfb[0]) >= nf
    m.close()
    g.close()
