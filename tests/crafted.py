"""Hand-built HA indexes for cases natural data does not reach (test data, no GPU needed).

wrap_index(): D = 3072, where the reference's u16 FastScan lanes can wrap
(FastScan.hpp:264-290).  Every (query, cluster) pair quantizes to the same digits
q_d = s_d * (255 + 256 j_d) (s_d = +-1, j_d in 0..31, |q_0| = 8191), because every residual
q - c_k is a positive multiple of one direction.  Each such |q_d| has low byte 255, so a
codebook whose shifted LUT entry selects n >= 1 of its four |q| has low byte 256 - n, and a
parity half of 384 codebooks wraps once at least ~259 of them select something.  A vector
selects dim d when its code bit equals [q_d > 0]; the selection rate is drawn per vector
and per codebook parity, so some vectors wrap in neither half, some in one, some in both.
"""
import numpy as np


def wrap_index(n_per_cluster=70, K=6, ex=3, seed=11, nq=3):
    from erq import index
    from oracle import oracle as orc

    rng = np.random.default_rng(seed)
    D = 3072
    j = rng.integers(0, 32, D)
    j[0] = 31
    s = np.where(rng.random(D) < 0.7, 1, -1)
    s[0] = 1
    pattern = (s * (255 + 256 * j)).astype(np.float64)
    direction = (pattern / np.linalg.norm(pattern)).astype(np.float32)
    eps = 0.1 * np.arange(K)
    centroids = (-eps[:, None] * direction[None, :]).astype(np.float32)
    t = np.array([1.0, 0.7, 1.6][:nq])
    queries = (t[:, None] * direction[None, :]).astype(np.float32)

    N = n_per_cluster * K
    sign_bit = (pattern > 0).astype(np.uint8)
    bits = np.empty((N, D), np.uint8)
    rates = np.array([0.05, 0.15, 0.25, 0.35, 0.6, 0.9])
    even = (np.arange(D) // 4) % 2 == 0          # dims of even codebooks
    for v in range(N):
        r_even, r_odd = rng.choice(rates, 2)
        rate = np.where(even, r_even, r_odd)
        match = rng.random(D) < rate
        bits[v] = np.where(match, sign_bit, 1 - sign_bit)
    onorm = rng.uniform(0.2, 1.2, N).astype(np.float32)
    raw = rng.integers(0, 1 << ex, (N, D)).astype(np.uint8)
    xip = rng.uniform(0.004, 0.012, N).astype(np.float32)

    sizes = np.full(K, n_per_cluster, np.uint64)
    nb = (n_per_cluster + 31) // 32
    BB = 4 * D + 128
    short = np.zeros(K * nb * BB, np.uint8)
    longc = np.empty(N * D * ex // 8, np.uint8)
    clen = D * ex // 8
    for k in range(K):
        for b in range(nb):
            v0 = k * n_per_cluster + 32 * b
            n = min(32, n_per_cluster - 32 * b)
            blk = (k * nb + b) * BB
            short[blk:blk + 4 * D] = orc.pack_block_codes(bits[v0:v0 + n], D)
            f = np.zeros(32, np.float32)
            f[:n] = onorm[v0:v0 + n]
            short[blk + 4 * D:blk + BB] = f.view(np.uint8)
    for v in range(N):
        longc[v * clen:(v + 1) * clen] = orc.ex_encode(raw[v], ex)
    ids = rng.permutation(N).astype(np.uint32) + 1000
    P = np.eye(D, dtype=np.float32)
    idx = index.IVFIndex(N, D, K, ex, sizes, P, centroids, short, longc, xip, ids)
    idx.validate()
    return idx, queries, bits


def wrap_index_nonha(n_per_cluster=70, K=6, ex=3, seed=12, nq=3):
    """A non-HA index (Searcher.hpp, four factors per vector) at D = 3072 whose FastScan
    half-sums wrap for some vectors and not for others (accumulate_robust, FastScan.hpp:
    136-212).  Every residual q - c_k is a positive multiple of one all-positive direction, so
    every digit is (about) 16383 and a codebook whose nibble selects n >= 1 dims has LUT
    entry n * 16383: low byte 256 - n, high byte 64 n - 1.  Per parity (384 codebooks) the
    low-byte sum wraps once about 2/3 of the codebooks select something, the high-byte sum
    once the selection rate passes 2/3; rates are drawn per vector and per parity.  Factors
    are drawn in plausible ranges (sumxb is the vector's popcount)."""
    from erq import index
    from oracle import oracle as orc

    rng = np.random.default_rng(seed)
    D = 3072
    direction = rng.uniform(0.5, 1.0, D)
    direction = (direction / np.linalg.norm(direction)).astype(np.float32)
    eps = 0.1 * np.arange(K)
    centroids = (-eps[:, None] * direction[None, :]).astype(np.float32)
    t = np.array([1.0, 0.7, 1.6][:nq])
    queries = (t[:, None] * direction[None, :]).astype(np.float32)

    N = n_per_cluster * K
    bits = np.empty((N, D), np.uint8)
    rates = np.array([0.05, 0.15, 0.25, 0.35, 0.6, 0.75, 0.9])
    even = (np.arange(D) // 4) % 2 == 0
    for v in range(N):
        r_even, r_odd = rng.choice(rates, 2)
        bits[v] = rng.random(D) < np.where(even, r_even, r_odd)
    x2 = rng.uniform(0.05, 1.5, N).astype(np.float32)
    fip = rng.uniform(0.5, 2.0, N).astype(np.float32)
    ferr = rng.uniform(0.0, 0.05, N).astype(np.float32)
    sumxb = bits.sum(1).astype(np.float32)
    raw = rng.integers(0, 1 << ex, (N, D)).astype(np.uint8)
    xip = rng.uniform(0.004, 0.012, N).astype(np.float32)

    sizes = np.full(K, n_per_cluster, np.uint64)
    nb = (n_per_cluster + 31) // 32
    BB = 4 * D + 512
    short = np.zeros(K * nb * BB, np.uint8)
    clen = D * ex // 8
    longc = np.empty(N * clen, np.uint8)
    for k in range(K):
        for b in range(nb):
            v0 = k * n_per_cluster + 32 * b
            n = min(32, n_per_cluster - 32 * b)
            blk = (k * nb + b) * BB
            short[blk:blk + 4 * D] = orc.pack_block_codes(bits[v0:v0 + n], D)
            f = np.zeros((4, 32), np.float32)
            for r, arr in enumerate((x2, fip, sumxb, ferr)):
                f[r, :n] = arr[v0:v0 + n]
            short[blk + 4 * D:blk + BB] = f.reshape(-1).view(np.uint8)
    for v in range(N):
        longc[v * clen:(v + 1) * clen] = orc.ex_encode(raw[v], ex)
    ids = rng.permutation(N).astype(np.uint32) + 1000
    P = np.eye(D, dtype=np.float32)
    idx = index.IVFIndex(N, D, K, ex, sizes, P, centroids, short, longc, xip, ids, 4)
    idx.validate()
    return idx, queries, bits
