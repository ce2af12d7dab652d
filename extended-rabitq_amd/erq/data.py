"""Synthetic workloads (Gaussian mixtures) and k-means IVF partitioning.

Shapes follow the benchmark configurations:
  mixture_isotropic   centres ~ N(0, I), x = centre + sigma * N(0, I)
  mixture_decay       per-dim std s_i = i^-0.25: centre = s*z, x = centre + sigma * s*z'
  mixture_powerlaw    eigen-spectrum lambda_i ~ i^-alpha in a random orthonormal basis U:
                      centre = U(sqrt(lambda)*z), x = centre + sigma*U(sqrt(lambda)*z'),
                      then L2-normalised
numpy generators are used for CPU-sized cases; the *_torch variants generate on the GPU
(torch is used for device memory and RNG only).
"""
import numpy as np


def mixture_isotropic(n, dim, ncomp, sigma, seed, centres=None):
    rng = np.random.default_rng(seed)
    if centres is None:
        centres = rng.standard_normal((ncomp, dim)).astype(np.float32)
    lab = rng.integers(0, ncomp, size=n)
    x = centres[lab] + sigma * rng.standard_normal((n, dim)).astype(np.float32)
    return x.astype(np.float32), centres


def mixture_decay(n, dim, ncomp, sigma, seed, centres=None):
    rng = np.random.default_rng(seed)
    s = (np.arange(1, dim + 1, dtype=np.float64) ** -0.25).astype(np.float32)
    if centres is None:
        centres = (rng.standard_normal((ncomp, dim)).astype(np.float32) * s)
    lab = rng.integers(0, ncomp, size=n)
    x = centres[lab] + sigma * s * rng.standard_normal((n, dim)).astype(np.float32)
    return x.astype(np.float32), centres


def powerlaw_basis(dim, alpha, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
    lam = np.arange(1, dim + 1, dtype=np.float64) ** -alpha
    lam = lam / lam.sum() * dim
    return U.astype(np.float32), np.sqrt(lam).astype(np.float32)


def mixture_powerlaw(n, dim, ncomp, sigma, seed, alpha=1.0, basis=None, centres=None):
    rng = np.random.default_rng(seed)
    if basis is None:
        basis = powerlaw_basis(dim, alpha, seed + 1000)
    U, sl = basis
    if centres is None:
        centres = (rng.standard_normal((ncomp, dim)).astype(np.float32) * sl) @ U.T
    lab = rng.integers(0, ncomp, size=n)
    x = centres[lab] + sigma * ((rng.standard_normal((n, dim)).astype(np.float32) * sl) @ U.T)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    return x.astype(np.float32), (basis, centres)


def kmeans(x, k, iters, seed, chunk=65536):
    """Lloyd k-means (numpy): init from k distinct sampled points."""
    rng = np.random.default_rng(seed)
    n = x.shape[0]
    c = x[rng.choice(n, size=k, replace=False)].copy()
    assign = np.zeros(n, np.int64)
    for _ in range(iters):
        cn = (c * c).sum(1)
        for s in range(0, n, chunk):
            xs = x[s:s + chunk]
            d = cn[None, :] - 2.0 * (xs @ c.T)
            assign[s:s + chunk] = d.argmin(1)
        sums = np.zeros_like(c, dtype=np.float64)
        np.add.at(sums, assign, x)
        cnt = np.bincount(assign, minlength=k)
        nz = cnt > 0
        c[nz] = (sums[nz] / cnt[nz, None]).astype(np.float32)
    cn = (c * c).sum(1)
    for s in range(0, n, chunk):
        xs = x[s:s + chunk]
        assign[s:s + chunk] = (cn[None, :] - 2.0 * (xs @ c.T)).argmin(1)
    return c.astype(np.float32), assign


# ---------------- GPU (torch) variants for the 1M-vector configurations ----------------

def torch_generator(seed, device):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return g


def mixture_powerlaw_torch(n, dim, ncomp, sigma, seed, alpha=1.0, device="cuda", centres=None,
                           basis=None):
    import torch
    g = torch_generator(seed, device)
    if basis is None:
        gb = torch_generator(seed + 1000, device)
        A = torch.randn(dim, dim, generator=gb, device=device, dtype=torch.float32)
        U, _ = torch.linalg.qr(A)
        lam = torch.arange(1, dim + 1, device=device, dtype=torch.float64) ** -alpha
        lam = lam / lam.sum() * dim
        basis = (U.contiguous(), lam.sqrt().float())
    U, sl = basis
    if centres is None:
        centres = (torch.randn(ncomp, dim, generator=g, device=device) * sl) @ U.T
    lab = torch.randint(0, ncomp, (n,), generator=g, device=device)
    x = torch.empty(n, dim, device=device, dtype=torch.float32)
    step = 131072
    for s in range(0, n, step):
        e = min(n, s + step)
        z = torch.randn(e - s, dim, generator=g, device=device) * sl
        xs = centres[lab[s:e]] + sigma * (z @ U.T)
        x[s:e] = xs / xs.norm(dim=1, keepdim=True)
    return x, (basis, centres)


def mixture_decay_torch(n, dim, ncomp, sigma, seed, device="cuda", centres=None):
    import torch
    g = torch_generator(seed, device)
    s = torch.arange(1, dim + 1, device=device, dtype=torch.float32) ** -0.25
    if centres is None:
        centres = torch.randn(ncomp, dim, generator=g, device=device) * s
    lab = torch.randint(0, ncomp, (n,), generator=g, device=device)
    x = torch.empty(n, dim, device=device, dtype=torch.float32)
    step = 131072
    for st in range(0, n, step):
        e = min(n, st + step)
        x[st:e] = centres[lab[st:e]] + sigma * s * torch.randn(e - st, dim, generator=g,
                                                               device=device)
    return x, centres


def exact_knn_np(base, queries, k):
    """Exact L2 top-k in FP64 on the CPU (spot checks)."""
    b = base.astype(np.float64)
    q = queries.astype(np.float64)
    d = (q * q).sum(1)[:, None] + (b * b).sum(1)[None, :] - 2.0 * q @ b.T
    idx = np.argsort(d, axis=1, kind="stable")[:, :k]
    return idx, np.take_along_axis(d, idx, 1)


def recall_at_k(result_ids, gt_ids, k):
    """test_search.cpp:113-123: |result[0:k] intersect gt[0:k]| / k, averaged."""
    tot = 0
    for r, g in zip(np.asarray(result_ids), np.asarray(gt_ids)):
        tot += len(set(r[:k].tolist()) & set(g[:k].tolist()))
    return tot / (len(result_ids) * k)


def distance_ratio(gt_sqr, ann_sqr):
    """test_search.cpp's average distance ratio (get_ratio, tools.hpp:16-51; the harness divides
    the summed per-query values by NQ * k): per query, its k true and k returned squared
    distances each sorted and paired rank by rank, the mean over pairs whose true distance
    exceeds 1e-5 of sqrt(returned / true) (1 when no pair counts), averaged over queries.
    gt_sqr, ann_sqr: [nq, k] torch tensors (any device); returns a float."""
    import torch
    g = torch.sort(gt_sqr.double(), dim=1).values
    a = torch.sort(ann_sqr.double(), dim=1).values
    valid = g > 1e-5
    r = torch.where(valid, torch.sqrt(a / torch.where(valid, g, torch.ones_like(g))),
                    torch.zeros_like(g))
    n = valid.sum(1)
    per_q = torch.where(n > 0, r.sum(1) / n.clamp(min=1), torch.ones_like(r[:, 0]))
    return float(per_q.mean())
