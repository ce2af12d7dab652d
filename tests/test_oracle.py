"""Pinning the CPU oracle (no GPU).

The reference has no tests or golden vectors and cannot be compiled here (it needs
Eigen 3.4, which is absent), so each oracle function is pinned by known-answer inputs
derived from the reference source, by closed forms / invariants, or by an independent
restatement at a different level (the AVX-512 register dataflow of
accumulate_one_block_high_acc emulated lane by lane in numpy).
"""
import bisect

import numpy as np
import pytest

from oracle import oracle as orc

RNG = np.random.default_rng(12345)


# ---------------------------------------------------------------- reductions
def test_reduce16_tree_known_answer():
    # _mm512_reduce_add_ps pairs lanes (i, i+8), then (j, j+4), (k, k+2), (0, 1).
    # p[0]=1e8, p[8]=-1e8, others 1: pairs (0,8) cancel exactly -> 14 ones survive.
    # A left-to-right sum loses the ones next to 1e8 (ulp 8) and gives 7; adjacent
    # pairing gives 12.
    p = np.ones(16, np.float32)
    p[0], p[8] = 1e8, -1e8
    assert orc.reduce16(p) == 14.0


def test_l2sqr16_lane_chains_known_answer():
    # (x-y) = [2^12, 1, 1, 1, 0...] over D=16: lane sums [2^24, 1, 1, 1]; the tree adds
    # lanes 1 and 3 first (=2) and 2^24 + 2 is exact; a sequential sum stays at 2^24.
    x = np.zeros(16, np.float32)
    x[0], x[1], x[2], x[3] = 4096.0, 1.0, 1.0, 1.0
    assert orc.l2sqr16(x, np.zeros(16, np.float32)) == 16777218.0


def test_l2sqr16_exact_on_integers_and_bound():
    for D in (16, 128, 1536):
        x = RNG.integers(-20, 20, D).astype(np.float32)
        y = RNG.integers(-20, 20, D).astype(np.float32)
        assert orc.l2sqr16(x, y) == float(((x.astype(np.int64) - y) ** 2).sum())
        xf = RNG.standard_normal(D).astype(np.float32)
        yf = RNG.standard_normal(D).astype(np.float32)
        ref = float(((xf.astype(np.float64) - yf) ** 2).sum())
        assert abs(orc.l2sqr16(xf, yf) - ref) <= (D / 16 + 4) * 2 ** -23 * ref


# ---------------------------------------------------------------- query prep
def test_normalize_query_unit_and_sum():
    D = 256
    q = RNG.standard_normal(D).astype(np.float32)
    c = RNG.standard_normal(D).astype(np.float32)
    y = float(np.sqrt(orc.l2sqr16(q, c)))
    u, s = orc.normalize_query(q, c, y)
    fac = np.float32(1.0 / np.float64(np.float32(y)))
    np.testing.assert_array_equal(u, (q - c) * fac)
    assert abs(np.linalg.norm(u.astype(np.float64)) - 1) < 1e-5
    assert abs(s - u.astype(np.float64).sum()) < 1e-4
    # degenerate residual (norm <= 1e-5): constant 1/sqrt(D), sum = D * value
    u0, s0 = orc.normalize_query(q, q, 0.0)
    v = np.float32(1.0 / np.float64(np.sqrt(np.float32(D))))
    assert np.all(u0 == v) and s0 == np.float32(D) * v


def test_normalize_query_sum_order_known_answer():
    # normalize_query16 (space.hpp:749-778) adds u into 16 lane accumulators (lane l takes
    # dims l, l+16, l+32, ... in order), then _mm512_reduce_add_ps.  With q - c = u exactly
    # (c = 0, norm = 1 so fac = 1): lane 0 holds 2^24, +1, +1, -2^24 (dims 0, 16, 32, 48):
    # 2^24 + 1 rounds back to 2^24 twice, so the lane ends at 0; lane 8 holds 3 (dim 8);
    # the tree gives 3.  A sequential sum over dims meets 3 while at 2^24 (2^24 + 3 -> 2^24 + 4),
    # keeps 2^24 + 4 through both +1 (ties to even), ends at 4.
    D = 64
    u = np.zeros(D, np.float32)
    u[0], u[16], u[32], u[48], u[8] = 2.0 ** 24, 1.0, 1.0, -(2.0 ** 24), 3.0
    unit, sumq = orc.normalize_query(u, np.zeros(D, np.float32), 1.0)
    np.testing.assert_array_equal(unit, u)
    assert sumq == 3.0
    seq = np.float32(0)
    for v in u:
        seq = np.float32(seq + v)
    assert seq == 4.0                                   # the order the known answer rules out


@pytest.mark.parametrize("ex", [2, 3, 4, 6, 7, 8])
def test_ip_ex_lane_order_and_fma_known_answers(ex):
    # IP*_fxu* (space.hpp:109-497): per 16-dim group an FMA into 16 lane accumulators, then
    # _mm512_reduce_add_ps.  Codes 0..3 fit every EX.
    D = 64
    # (a) lane order: lane 0 gets 2^24*1, 1*1, 1*1, -2^24*1 -> 0 (each +1 rounds away);
    #     lane 8 gets 1*3 -> 3; total 3.  Summed over dims in order: 2^24 + 3 -> 2^24 + 4,
    #     +1, +1 stay (ties to even), -2^24 -> 4.
    u = np.zeros(D, np.float32)
    e = np.zeros(D, np.uint8)
    u[0], u[16], u[32], u[48], u[8] = 2.0 ** 24, 1.0, 1.0, -(2.0 ** 24), 1.0
    e[0] = e[16] = e[32] = e[48] = 1
    e[8] = 3
    assert orc.ip_ex(u, orc.ex_encode(e, ex), ex) == 3.0
    # (b) fused multiply-add: lane 1 gets (-1)*3 then (1 + 2^-23)*3.  fma(1 + 2^-23, 3, -3)
    #     is exactly 3 * 2^-23; a rounded product (3 + 2^-21) minus 3 would give 2^-21.
    u = np.zeros(D, np.float32)
    e = np.zeros(D, np.uint8)
    u[1], u[17] = -1.0, np.float32(1.0 + 2.0 ** -23)
    e[1] = e[17] = 3
    got = orc.ip_ex(u, orc.ex_encode(e, ex), ex)
    assert got == np.float32(3 * 2.0 ** -23)
    assert got != np.float32(2.0 ** -21)


def test_quantize_query_rounding_and_range():
    D = 512
    u = RNG.standard_normal(D).astype(np.float32)
    q, delta = orc.quantize_query(u)
    vmax = np.float32(np.abs(u).max())
    assert delta == vmax / np.float32(8191)
    assert np.abs(q).max() == 8191 and q.dtype == np.int16
    tmp = u / np.float32(delta)                    # float32 quotient, as in the source
    expect = np.trunc(tmp.astype(np.float64) + 0.5 - (tmp < 0)).astype(np.int16)
    np.testing.assert_array_equal(q, expect)
    # half-way cases round away from zero
    u2 = np.array([8191, 2.5, -2.5, 0.5, -0.5, -8191] + [0] * 10, np.float32)
    q2, d2 = orc.quantize_query(u2)
    assert d2 == 1.0
    assert q2[:6].tolist() == [8191, 3, -3, 1, -1, -8191]


def test_lut_closed_form():
    D = 128
    q = RNG.integers(-8191, 8192, D).astype(np.int16)
    lut, shift = orc.pack_lut(q)
    M = D // 4
    assert lut.shape == (M, 16)
    for m in range(M):
        w = q[4 * m:4 * m + 4].astype(np.int64)
        full = np.array([sum(((j >> (3 - t)) & 1) * w[t] for t in range(4)) for j in range(16)])
        vmin = min(0, full.min())
        np.testing.assert_array_equal(lut[m].astype(np.int64), full - vmin)
        assert lut[m].min() == 0 or vmin == 0
    assert shift == int(np.minimum(q.astype(np.int64), 0).sum())
    assert lut.max() <= 4 * 8191


# ---------------------------------------------------------------- FastScan block
PERM0 = [0, 8, 1, 9, 2, 10, 3, 11, 4, 12, 5, 13, 6, 14, 7, 15]  # pack_codes.hpp:44


def ref_pack_block(bits):
    """pack_binary + byte/nibble swaps + pack_codes (Quantizer.hpp:371-384,
    pack_codes.hpp:33-98) restated from the reference source for one 32-vector block."""
    n, D = bits.shape
    M = D // 4
    # pack_binary: dim col+i at bit 63-i of a u64 per 64 dims, stored little-endian
    words = np.zeros((32, D // 64), np.uint64)
    for v in range(n):
        for w in range(D // 64):
            cur = 0
            for i in range(64):
                cur |= int(bits[v, 64 * w + i]) << (63 - i)
            words[v, w] = cur
    b8 = words.view(np.uint8).reshape(32, D // 8).copy()
    for v in range(32):                       # swap bytes k <-> 7-k in each 8-byte group
        for j in range(D // 64):
            g = b8[v, 8 * j:8 * j + 8].copy()
            b8[v, 8 * j:8 * j + 8] = g[::-1]
    b8 = ((b8 & 15) << 4) | (b8 >> 4)         # swap nibbles
    out = []
    for m in range(0, M, 2):                  # pack_codes(u8)
        c = b8[:, m // 2]
        c0, c1 = c & 15, c >> 4
        d0 = [c0[PERM0[j]] | (c0[PERM0[j] + 16] << 4) for j in range(16)]
        d1 = [c1[PERM0[j]] | (c1[PERM0[j] + 16] << 4) for j in range(16)]
        out += d0 + d1
    return np.array(out, np.uint8)


def emulate_avx512_accumulate(codes, lut_u16, D):
    """accumulate_one_block_high_acc's integer dataflow (FastScan.hpp:237-293) emulated
    on 512-bit registers as numpy arrays: per-128-bit-lane byte shuffles, u16 lane adds
    with wraparound, then the 256-bit halves combined with permute2f128(0x21) / blend."""
    M = D // 4
    # HC_LUT layout (HASearcher.hpp:169-179): per 4 codebooks 128 bytes: lo x4, hi x4
    lut = np.zeros(M * 32, np.uint8)
    for m in range(M):
        base = (m // 4) * 128 + (m % 4) * 16
        lut[base:base + 16] = lut_u16[m] & 255
        lut[base + 64:base + 80] = lut_u16[m] >> 8
    accu = np.zeros((2, 4, 32), np.uint16)
    for it in range(M // 4):
        c = codes[64 * it:64 * it + 64]
        lo = c & 15
        hi = (c >> 4) & 15
        for h in range(2):
            L = lut[128 * it + 64 * h:128 * it + 64 * h + 64]
            res_lo = np.concatenate([L[16 * l + lo[16 * l:16 * l + 16]] for l in range(4)])
            res_hi = np.concatenate([L[16 * l + hi[16 * l:16 * l + 16]] for l in range(4)])
            w_lo = res_lo.view(np.uint16)
            w_hi = res_hi.view(np.uint16)
            accu[h, 0] += w_lo
            accu[h, 1] += w_lo >> 8
            accu[h, 2] += w_hi
            accu[h, 3] += w_hi >> 8

    def combine(a0, a1):
        t0 = a0[:16] + a0[16:]
        t1 = a1[:16] + a1[16:]
        t0 = t0 - (t1 << 8)
        perm = np.concatenate([t0[8:], t1[:8]]).astype(np.int64)   # permute2f128 0x21
        blend = np.concatenate([t0[:8], t1[8:]]).astype(np.int64)  # blend_epi32 0xF0
        return perm + blend

    with np.errstate(over="ignore"):
        d0 = [combine(accu[h, 0], accu[h, 1]) for h in range(2)]
        d1 = [combine(accu[h, 2], accu[h, 3]) for h in range(2)]
    return np.concatenate([d0[0] + (d0[1] << 8), d1[0] + (d1[1] << 8)]).astype(np.int32)


@pytest.mark.parametrize("D", [64, 128, 1024])
def test_fastscan_matches_register_emulation_and_inner_product(D):
    n = 32
    bits = RNG.integers(0, 2, (n, D)).astype(np.uint8)
    q = RNG.integers(-8191, 8192, D).astype(np.int16)
    lut, shift = orc.pack_lut(q)
    codes = orc.pack_block_codes(bits, D)
    np.testing.assert_array_equal(codes, ref_pack_block(bits))
    _, _, sums, _ = orc.fastscan_block(codes, lut, np.ones(32, np.float32), 0.0, 1.0, 1.0, shift,
                                       0.1, 0.0, D)
    np.testing.assert_array_equal(sums, emulate_avx512_accumulate(codes, lut, D))
    # no wrap for D <= 2048: S + shift is the exact integer inner product
    np.testing.assert_array_equal(sums.astype(np.int64) + shift,
                                  bits.astype(np.int64) @ q.astype(np.int64))


def test_fastscan_single_hot_vectors():
    D = 128
    q = RNG.integers(-8191, 8192, D).astype(np.int16)
    lut, shift = orc.pack_lut(q)
    for v, d in [(0, 0), (8, 3), (15, 127), (16, 4), (17, 64), (31, 126)]:
        bits = np.zeros((32, D), np.uint8)
        bits[v, d] = 1
        codes = orc.pack_block_codes(bits, D)
        _, _, sums, _ = orc.fastscan_block(codes, lut, np.ones(32, np.float32), 0.0, 1.0, 1.0,
                                           shift, 0.1, 0.0, D)
        assert sums[v] + shift == q[d]
        others = np.delete(np.arange(32), v)
        assert np.all(sums[others] + shift == 0)


def test_fastscan_u16_wrap_at_d3072():
    # all-ones codes against q = +8191 everywhere: each LUT entry 15 is 32764 (low byte
    # 252); the 384 even codebooks sum to 96768 low-byte units, which wraps mod 2^16.
    D = 3072
    bits = np.ones((32, D), np.uint8)
    bits[5] = RNG.integers(0, 2, D)
    q = np.full(D, 8191, np.int16)
    lut, shift = orc.pack_lut(q)
    codes = orc.pack_block_codes(bits, D)
    _, _, sums, _ = orc.fastscan_block(codes, lut, np.ones(32, np.float32), 0.0, 1.0, 1.0, shift,
                                       0.1, 0.0, D)
    emu = emulate_avx512_accumulate(codes, lut, D)
    np.testing.assert_array_equal(sums, emu)
    exact = 8191 * D
    assert sums[0] == exact - 2 * 65536           # both low-byte halves wrapped once
    assert sums[0] != exact


def test_fastscan_float_bound_and_mask():
    D = 256
    bits = RNG.integers(0, 2, (32, D)).astype(np.uint8)
    u = RNG.standard_normal(D).astype(np.float32)
    u /= np.linalg.norm(u)
    q, delta = orc.quantize_query(u)
    lut, shift = orc.pack_lut(q)
    codes = orc.pack_block_codes(bits, D)
    onorm = RNG.uniform(0.5, 2.0, 32).astype(np.float32)
    sumq, y = float(u.sum()), 1.3
    oosd = orc.one_over_sqrtD(D)
    for distk in (1e30, 3.0, 2.0):
        mask, ipq, sums, lower = orc.fastscan_block(codes, lut, onorm, sumq, y, delta, shift,
                                                    oosd, distk, D)
        ip64 = (sums.astype(np.float64) + shift) * delta
        np.testing.assert_allclose(ipq, ip64, rtol=1e-6)
        ref = ((ip64 - (0.5 * sumq - 0.58)) * (-5 * y / np.sqrt(D)) * onorm + y * y +
               onorm.astype(np.float64) ** 2)
        np.testing.assert_allclose(lower, ref, rtol=1e-5, atol=1e-5)
        assert mask == sum(1 << v for v in range(32) if lower[v] < np.float32(distk))


# ---------------------------------------------------------------- ex codes
def layout_position(ex, d, b):
    """(byte, bit) where bit b of the code of dim d (0..63 within a 64-dim chunk, or
    0..31 for EX=4) lands, read off store_compacted_code (Quantizer.hpp:491-631)."""
    if ex == 8:
        return d, b
    if ex == 4:  # per 32 dims: b_j = e_j | e_{j+16} << 4
        return d % 16, b + 4 * (d // 16)
    if ex in (2, 3):
        if b < 2:  # b_j = e_j | e_{j+16}<<2 | e_{j+32}<<4 | e_{j+48}<<6
            return d % 16, b + 2 * (d // 16)
        return 16 + d % 8, d // 8     # top bit: byte (d mod 8), bit (d / 8) of u64
    if ex in (6, 7):
        j, s = d % 16, d // 16
        if b == 6:
            return 48 + d % 8, d // 8
        if s < 2:
            if b < 6:
                return 16 * s + j, b  # e_j & 63 / e_{16+j} & 63
            raise AssertionError
        if b < 4:                     # bytes 32..47: e_{32+j} & 15 | (e_{48+j} & 15) << 4
            return 32 + j, b + 4 * (s - 2)
        return 16 * (s - 2) + j, 6 + (b - 4)   # bits 4,5 of e_{32+j}/e_{48+j} -> top of b0/b1
    raise AssertionError


@pytest.mark.parametrize("ex", [2, 3, 4, 6, 7, 8])
def test_ex_layout_single_hot(ex):
    D = 64
    clen = D * ex // 8
    for d in range(D):
        for b in range(ex):
            raw = np.zeros(D, np.uint8)
            raw[d] = 1 << b
            enc = orc.ex_encode(raw, ex)
            assert enc.size == clen
            nz = np.flatnonzero(enc)
            assert nz.size == 1, (ex, d, b)
            byte = int(nz[0])
            bit = int(enc[byte]).bit_length() - 1
            assert enc[byte] == 1 << bit
            eb, ebit = layout_position(ex, d if ex != 4 else d % 32, b)
            if ex == 4:
                eb += 16 * (d // 32)
            assert (byte, bit) == (eb, ebit), (ex, d, b)
            np.testing.assert_array_equal(orc.ex_decode(enc, D, ex), raw)


@pytest.mark.parametrize("ex", [2, 3, 4, 6, 7, 8])
def test_ex_round_trip_and_ip(ex):
    D = 1536
    raw = RNG.integers(0, 1 << ex, D).astype(np.uint8)
    enc = orc.ex_encode(raw, ex)
    np.testing.assert_array_equal(orc.ex_decode(enc, D, ex), raw)
    u = RNG.integers(-50, 50, D).astype(np.float32)  # exact partial sums
    assert orc.ip_ex(u, enc, ex) == float(u.astype(np.int64) @ raw.astype(np.int64))
    uf = RNG.standard_normal(D).astype(np.float32)
    ref = float(uf.astype(np.float64) @ raw.astype(np.float64))
    assert abs(orc.ip_ex(uf, enc, ex) - ref) <= D * 2 ** -23 * float(np.abs(uf) @ raw)


# ---------------------------------------------------------------- result pool
class PoolModel:
    """Sorted list with bisect_left insertion (new before equal), capacity cap."""

    def __init__(self, cap):
        self.cap, self.d, self.i = cap, [], []

    def insert(self, pid, dist):
        if len(self.d) == self.cap and dist > self.d[-1]:
            return
        pos = bisect.bisect_left(self.d, dist)
        self.d.insert(pos, dist)
        self.i.insert(pos, pid)
        del self.d[self.cap:], self.i[self.cap:]


@pytest.mark.parametrize("cap", [1, 5, 100])
def test_pool_matches_model_with_ties(cap):
    p, m = orc.Pool(cap), PoolModel(cap)
    vals = RNG.integers(0, 30, 500).astype(np.float32)   # many exact ties
    for t, v in enumerate(vals):
        p.insert(t, float(v))
        m.insert(t, float(v))
        dk = m.d[-1] if len(m.d) == cap else np.finfo(np.float32).max
        assert p.distk() == np.float32(dk)
    ids, d = p.contents()
    assert ids.tolist() == m.i and d.tolist() == m.d


def test_pool_full_tie_at_kth_replaces_last():
    p = orc.Pool(3)
    for i, v in enumerate([1.0, 2.0, 3.0]):
        p.insert(i, v)
    p.insert(9, 3.0)          # equal to the k-th: accepted, goes before the old 3.0
    ids, d = p.contents()
    assert ids.tolist() == [0, 1, 9] and d.tolist() == [1.0, 2.0, 3.0]
    p.insert(10, 3.5)         # larger than the k-th: rejected
    assert p.contents()[0].tolist() == [0, 1, 9]


# ---------------------------------------------------------------- end to end
def test_centroid_topn_is_sorted_l2sqr16(small_data):
    x, q, c, assign = small_data
    from erq import index
    idx = index.build_index(x, c, assign, bits=4, seed=3)
    oi = orc.OracleIndex(idx)
    from conftest import rotate_np
    qr = rotate_np(idx, q[:4])
    for i in range(4):
        ids, d = oi.centroid_topn(qr[i], 10)
        all_d = np.array([orc.l2sqr16(qr[i], idx.centroids[j]) for j in range(idx.K)],
                         np.float32)
        order = np.lexsort((np.arange(idx.K), all_d))[:10]
        np.testing.assert_array_equal(ids, order)
        np.testing.assert_array_equal(d, all_d[order])


def test_search_recall_monotone_and_exhaustive(small_data):
    from conftest import rotate_np
    from erq import data, index
    x, q, c, assign = small_data
    idx = index.build_index(x, c, assign, bits=9, seed=3)
    oi = orc.OracleIndex(idx)
    qr = rotate_np(idx, q[:32])
    gt, _ = data.exact_knn_np(x, q[:32], 100)
    recalls = []
    for nprobe in (1, 4, 16, 64):
        res = np.array([oi.search(qr[i], 100, nprobe)[0] for i in range(32)])
        recalls.append(data.recall_at_k(res, gt, 100))
    assert all(b >= a - 0.01 for a, b in zip(recalls, recalls[1:]))
    assert recalls[-1] > 0.97      # 8-bit ex codes, every cluster probed
    ids, d = oi.search(qr[0], 100, 64)
    assert np.all(np.diff(d) >= 0)


def test_round_schedule_oracle(small_data):
    """The round-synchronous schedule: one probe equals the per-query search; with more
    probes the result is sorted by (distance, id) and its recall tracks the per-query one."""
    from conftest import rotate_np
    from erq import data, index
    x, q, c, assign = small_data
    idx = index.build_index(x, c, assign, bits=4, seed=3)
    oi = orc.OracleIndex(idx)
    qr = rotate_np(idx, q[:32])
    gt, _ = data.exact_knn_np(x, q[:32], 100)
    res_p, res_r = [], []
    for i in range(32):
        pid, pd = oi.centroid_topn(qr[i], 1)
        a = oi.search_probes(qr[i], 100, pid, pd)
        b = oi.search_rounds(qr[i], 100, pid, pd)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
        pid, pd = oi.centroid_topn(qr[i], 16)
        ids, d = oi.search_rounds(qr[i], 100, pid, pd)
        assert all((d[j], ids[j]) <= (d[j + 1], ids[j + 1]) for j in range(99))
        res_r.append(ids)
        res_p.append(oi.search_probes(qr[i], 100, pid, pd)[0])
    assert data.recall_at_k(res_r, gt, 100) >= data.recall_at_k(res_p, gt, 100) - 0.01
