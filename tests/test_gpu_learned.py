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
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    li, ld = g.search_learned_dev(m, qte_r, 100, mass=0.95, min_nprobe=2, max_nprobe=48,
                                  fallback_nprobe=16, min_conf=0.0, stats=stats)
    off, pid, pd, nfb = g.probe_learned_dev(m, qte_r, 0.95, 2, 48, 16, 0.0)
    torch.cuda.synchronize()
    offs = off.cpu().numpy()
    nprobe_mean = float(np.diff(offs).mean())
    r_learned = data.recall_at_k(li.cpu().numpy(), gt, 100)
    fi, _ = g.search_batched_dev(qte_r, 100, max(1, int(round(nprobe_mean))))
    r_fixed = data.recall_at_k(fi.cpu().numpy(), gt, 100)
    print(f"mean nprobe {nprobe_mean:.1f}: recall learned {r_learned:.4f} fixed {r_fixed:.4f} "
          f"fallbacks {int(nfb[0])}")
    assert r_learned >= r_fixed - 0.01
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
