"""The oracle's non-HA searcher (Searcher.hpp, IVF::search without HIGH_ACC_FAST_SCAN) pinned:
the residual sum's order, scalar_quantize16's reciprocal multiply and truncation, the u16 LUT,
accumulate_robust against an emulation of its AVX-512 register dataflow (with the u16
wraparound above D = 2048), and whole searches against exact distances."""
import numpy as np
import pytest

from oracle import oracle as orc
from test_oracle import ref_pack_block

RNG = np.random.default_rng(11)


def test_minus_sum16_residual_and_order():
    D = 48
    q = np.zeros(D, np.float32)
    c = np.zeros(D, np.float32)
    # lane 0: 2^24, then +1 twice (each rounds away); lane 1: three ones -> 3.  Lane chains +
    # butterfly: 2^24 + 3 -> 2^24 + 4.  A sequential sum over the dims loses every 1: 2^24.
    q[0] = 2.0 ** 24
    q[16] = q[32] = 1.0
    q[1] = q[17] = q[33] = 1.0
    s, r = orc.minus_sum16(q, c)
    np.testing.assert_array_equal(r, q - c)
    assert s == 2.0 ** 24 + 4
    seq = np.float32(0)
    for v in q:
        seq = np.float32(seq + v)
    assert seq == 2.0 ** 24


def test_scalar_quantize16_reciprocal_multiply_and_truncation():
    # (r - vl) * f32(1 / width) + 0.5, truncated: find inputs where this differs from the
    # IEEE quotient (r - vl) / width and from rounding half to even
    width = np.float32(3.0)
    o = np.float32(1.0 / np.float64(width))
    # r next to (k - 1/2) width, where the quotient sits next to a rounding boundary
    base = ((np.arange(1, 4000) - 0.5) * np.float64(width)).astype(np.float32)
    cand = np.concatenate([np.nextafter(base, np.float32(-np.inf)), base,
                           np.nextafter(base, np.float32(np.inf))])
    by_mul = np.trunc((cand * o) + np.float32(0.5)).astype(np.int64)
    by_div = np.trunc((cand / width) + np.float32(0.5)).astype(np.int64)
    diff = np.nonzero(by_mul != by_div)[0]
    assert diff.size > 0
    r = cand[diff[:64]]
    got = orc.scalar_quantize16(r, 0.0, float(width)).astype(np.int64)
    np.testing.assert_array_equal(got, by_mul[diff[:64]])
    assert np.all(got != by_div[diff[:64]])
    # truncation of x + 0.5, not round-half-even: 2.5 -> 3, 4.5 -> 5
    got = orc.scalar_quantize16(np.array([2.0, 4.0], np.float32), 0.0, 1.0)
    np.testing.assert_array_equal(got, [2, 4])
    got = orc.scalar_quantize16(np.array([2.5, 4.5], np.float32), 0.0, 1.0)
    np.testing.assert_array_equal(got, [3, 5])
    # the range starts at vl: r = vl -> 0
    assert orc.scalar_quantize16(np.array([-1.5], np.float32), -1.5, 0.25)[0] == 0


def test_pack_lut_u16_closed_form_and_wrap():
    D = 64
    q = RNG.integers(0, 16384, D).astype(np.uint16)
    lut = orc.pack_lut_u16(q).reshape(D // 4, 16)
    for m in range(D // 4):
        w = q[4 * m:4 * m + 4].astype(np.int64)
        full = np.array([sum(((j >> (3 - t)) & 1) * w[t] for t in range(4)) for j in range(16)])
        np.testing.assert_array_equal(lut[m].astype(np.int64), full % 65536)
    q = np.full(4, 16384, np.uint16)       # 4 x 16384 = 65536 wraps to 0 in u16
    assert orc.pack_lut_u16(q)[15] == 0


def emulate_accumulate_robust(codes, lut8, D):
    """accumulate_robust (FastScan.hpp:136-212) on 512-bit registers as numpy arrays: per
    128-bit lane byte shuffles of a u8 LUT, u16 lane adds with wraparound, the 256-bit halves
    combined with permute2f128(0x21) / blend, widened to u32 and added."""
    accu = np.zeros((4, 32), np.uint16)
    for it in range(D // 16):             # 64 code bytes (4 codebooks) per step
        c = codes[64 * it:64 * it + 64]
        L = lut8[64 * it:64 * it + 64]
        lo = c & 15
        hi = (c >> 4) & 15
        res_lo = np.concatenate([L[16 * l + lo[16 * l:16 * l + 16]] for l in range(4)])
        res_hi = np.concatenate([L[16 * l + hi[16 * l:16 * l + 16]] for l in range(4)])
        w_lo = res_lo.view(np.uint16)
        w_hi = res_hi.view(np.uint16)
        with np.errstate(over="ignore"):
            accu[0] += w_lo
            accu[1] += w_lo >> 8
            accu[2] += w_hi
            accu[3] += w_hi >> 8

    def combine(a0, a1):
        with np.errstate(over="ignore"):
            t0 = a0[:16] + a0[16:]
            t1 = a1[:16] + a1[16:]
            t0 = t0 - (t1 << 8)
        perm = np.concatenate([t0[8:], t1[:8]]).astype(np.int64)
        blend = np.concatenate([t0[:8], t1[8:]]).astype(np.int64)
        return perm + blend

    return np.concatenate([combine(accu[0], accu[1]), combine(accu[2], accu[3])])


def _ints_reference(codes, lut, D):
    lower = emulate_accumulate_robust(codes, (lut & 255).astype(np.uint8), D)
    upper = emulate_accumulate_robust(codes, (lut >> 8).astype(np.uint8), D)
    return ((upper << 8) + lower).astype(np.int32)


@pytest.mark.parametrize("D", [64, 1024, 2048])
def test_block_ints_match_emulation_and_inner_product(D):
    bits = RNG.integers(0, 2, (32, D)).astype(np.uint8)
    q = RNG.integers(0, 16384, D).astype(np.uint16)
    lut = orc.pack_lut_u16(q)
    codes = orc.pack_block_codes(bits, D)
    np.testing.assert_array_equal(codes, ref_pack_block(bits))
    ints = orc.nonha_block_ints(codes, lut, D)
    np.testing.assert_array_equal(ints, _ints_reference(codes, lut, D))
    # no wrap up to D = 2048: the exact inner product of the bits with the digits
    np.testing.assert_array_equal(ints.astype(np.int64), bits.astype(np.int64) @ q.astype(np.int64))


def test_block_ints_wrap_at_d3072():
    D = 3072
    bits = np.ones((32, D), np.uint8)
    q = np.full(D, 16383, np.uint16)      # every LUT byte at its largest: both byte sums wrap
    lut = orc.pack_lut_u16(q)
    codes = orc.pack_block_codes(bits, D)
    ints = orc.nonha_block_ints(codes, lut, D)
    np.testing.assert_array_equal(ints, _ints_reference(codes, lut, D))
    assert ints[0] != D * 16383


def test_nonha_search_recall_and_exhaustive():
    from erq import data, index
    x, cen = data.mixture_isotropic(8000, 128, 32, 0.3, seed=1)
    q, _ = data.mixture_isotropic(40, 128, 32, 0.3, seed=2, centres=cen)
    c, assign = data.kmeans(x, 32, 5, seed=4)
    idx = index.build_index(x, c, assign, 5, seed=3, high_acc=False)
    assert idx.factors == 4 and idx.block_bytes == 4 * 128 + 512
    oi = orc.OracleIndex(idx)
    gt, _ = data.exact_knn_np(x, q, 50)
    qr = q @ idx.P
    rec = []
    for npb in (1, 4, 32):
        ids, d = oi.search_batch(qr, 50, npb)
        rec.append(data.recall_at_k(ids, gt, 50))
    assert rec[0] <= rec[1] <= rec[2]
    # every cluster probed: the recall of the HA searcher on the same data and bits
    ha = index.build_index(x, c, assign, 5, seed=3)
    ids_ha, _ = orc.OracleIndex(ha).search_batch(qr, 50, 32)
    assert abs(rec[2] - data.recall_at_k(ids_ha, gt, 50)) <= 0.01
    # the returned distances estimate the true squared distances of the returned ids
    idx9 = index.build_index(x, c, assign, 9, seed=3, high_acc=False)
    ids, d = orc.OracleIndex(idx9).search_batch(qr, 50, 32)
    assert data.recall_at_k(ids, gt, 50) > 0.99
    true = ((x[ids.astype(np.int64)] - q[:, None, :]) ** 2).sum(-1)
    assert np.median(np.abs(d - true) / true) < 1e-3


def test_nonha_index_file_round_trip(tmp_path):
    from erq import data, index
    x, _ = data.mixture_isotropic(3000, 64, 8, 0.3, seed=5)
    c, assign = data.kmeans(x, 8, 4, seed=6)
    idx = index.build_index(x, c, assign, 4, seed=3, high_acc=False)
    idx.save(tmp_path / "nonha.index")
    back = index.IVFIndex.load(tmp_path / "nonha.index")
    assert back.factors == 4
    np.testing.assert_array_equal(back.short_data, idx.short_data)
    ha = index.build_index(x, c, assign, 4, seed=3)
    ha.save(tmp_path / "ha.index")
    assert index.IVFIndex.load(tmp_path / "ha.index").factors == 1
