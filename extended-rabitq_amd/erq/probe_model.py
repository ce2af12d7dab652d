"""Learned probe predictor over liberq.so's erq_mlp_* entry points (include/erq.h).

The network maps a rotated query (D floats) to per-cluster probabilities of holding one of
its top-k neighbours; erq_probe_learned_dev turns them into an adaptive probe list.  torch
is used only for device buffers, batching and label scatter.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check
from .gpu import _ptr, _stream


class ProbeMLP:
    def __init__(self, in_dim, out_dim, hidden=512, max_batch=1024, seed=0, input_scale=1.0):
        h = ctypes.c_void_p()
        check(_lib.lib().erq_mlp_create(in_dim, hidden, out_dim, max_batch, seed,
                                        ctypes.byref(h)))
        self._h = h
        if input_scale != 1.0:
            check(_lib.lib().erq_mlp_set_input_scale(h, input_scale))
        self.in_dim, self.out_dim, self.hidden, self.max_batch = in_dim, out_dim, hidden, max_batch

    def close(self):
        if self._h:
            _lib.lib().erq_mlp_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def num_params(self):
        return _lib.lib().erq_mlp_num_params(self._h)

    def get_params(self):
        out = np.empty(self.num_params(), np.float32)
        check(_lib.lib().erq_mlp_get_params(self._h, _ptr(out), None))
        return out

    def get_grads(self):
        out = np.empty(self.num_params(), np.float32)
        check(_lib.lib().erq_mlp_get_params(self._h, None, _ptr(out)))
        return out

    def set_params(self, theta):
        theta = np.ascontiguousarray(theta, np.float32)
        assert theta.size == self.num_params()
        check(_lib.lib().erq_mlp_set_params(self._h, _ptr(theta)))

    def forward_dev(self, x):
        import torch
        p = torch.empty((x.shape[0], self.out_dim), dtype=torch.float32, device=x.device)
        check(_lib.lib().erq_mlp_forward_dev(self._h, _ptr(x.contiguous()), x.shape[0], _ptr(p),
                                             _stream()))
        return p

    def loss_grad_dev(self, x, y, pos_weight):
        import torch
        loss = torch.zeros(1, dtype=torch.float32, device=x.device)
        check(_lib.lib().erq_mlp_loss_grad_dev(self._h, _ptr(x.contiguous()), _ptr(y.contiguous()),
                                               x.shape[0], pos_weight, _ptr(loss), _stream()))
        return loss

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        check(_lib.lib().erq_mlp_adam_step_dev(self._h, lr, beta1, beta2, eps, _stream()))

    def train(self, x, label_clusters, epochs=3, batch=512, lr=1e-3, pos_weight=4.0, seed=0,
              log=None, calibrate=True, schedule="constant", soft=False):
        """x: [n, D] CUDA float tensor; label_clusters: [n, k] CUDA int tensor of the cluster
        of each of the query's top-k results (-1 = none).  Returns the per-epoch mean loss.

        The positive-weighted loss is minimised by sigmoid(z) = w p / (w p + 1 - p), i.e.
        z = ln w + logit p.  With calibrate, ln w is taken off the output bias afterwards so
        that the network's outputs estimate p itself, which the probability-mass criterion
        of erq_probe_learned_dev assumes.  schedule "cosine" decays the learning rate from lr
        to 0 over the run (half a cosine); "constant" keeps lr.  soft: the target of a cluster
        is the share of the query's top-k it holds (instead of 1 if it holds any), so the
        outputs rank clusters by how many neighbours they hold."""
        import torch
        n = x.shape[0]
        g = torch.Generator(device=x.device)
        g.manual_seed(seed)
        losses = []
        steps = max(1, epochs * (n // batch))
        step = 0
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g, device=x.device)
            tot, cnt = torch.zeros((), device=x.device), 0
            for s in range(0, n - batch + 1, batch):
                sel = perm[s:s + batch]
                y = labels_dense(label_clusters[sel], self.out_dim, soft=soft)
                tot = tot + self.loss_grad_dev(x[sel], y, pos_weight)[0]
                lr_t = lr if schedule == "constant" else 0.5 * lr * (1 + np.cos(np.pi * step / steps))
                self.adam_step(float(lr_t))
                step += 1
                cnt += 1
            losses.append(float(tot) / max(cnt, 1))
            if log:
                log(f"probe MLP epoch {ep}: loss {losses[-1]:.4f}")
        if calibrate and pos_weight != 1.0:
            self.shift_output_bias(-float(np.log(pos_weight)))
        return losses

    def save(self, path):
        """Weights and shape to an .npz file (plain arrays; loads with allow_pickle=False)."""
        np.savez(path, params=self.get_params(),
                 shape=np.array([self.in_dim, self.hidden, self.out_dim], np.int64),
                 input_scale=np.array([self.input_scale()], np.float32))

    @classmethod
    def load(cls, path, max_batch=1024):
        with np.load(path, allow_pickle=False) as f:
            din, hid, dout = (int(v) for v in f["shape"])
            m = cls(din, dout, hidden=hid, max_batch=max_batch,
                    input_scale=float(f["input_scale"][0]))
            m.set_params(f["params"])
        return m

    def input_scale(self):
        return float(_lib.lib().erq_mlp_input_scale(self._h))

    def shift_output_bias(self, delta):
        """Adds delta to every output logit (the last out_dim parameters are the output bias)."""
        import torch
        torch.cuda.synchronize()   # training steps run on the caller's stream
        theta = self.get_params()
        theta[-self.out_dim:] += np.float32(delta)
        self.set_params(theta)


def labels_dense(label_clusters, K, soft=False):
    """[n, K] float labels from per-query cluster lists (-1 entries ignored): 1 where the
    cluster holds any listed neighbour, or with soft, the share of the list it holds."""
    import torch
    n = label_clusters.shape[0]
    y = torch.zeros((n, K + 1), dtype=torch.float32, device=label_clusters.device)
    lc = label_clusters.long().clone()
    lc[lc < 0] = K
    if soft:
        y.scatter_add_(1, lc, torch.full(lc.shape, 1.0 / lc.shape[1], device=lc.device))
    else:
        y.scatter_(1, lc, 1.0)
    return y[:, :K].contiguous()


def cluster_of_ids(idx):
    """cluster id of every vector id of an erq.IVFIndex (host numpy)."""
    out = np.empty(idx.N, np.int32)
    vs = idx.vec_start()
    for k in range(idx.K):
        out[idx.ids[vs[k]:vs[k + 1]]] = k
    return out
