"""The hand-built D = 3072 wrap index (tests/crafted.py) on the CPU oracle: its digits are
the intended pattern, the reference's u16 wraparound hits a mix of vectors (none, one or
both low-byte halves), the oracle's integer equals the AVX-512 register-dataflow emulation
on its blocks, and the search results include wrapped vectors, so a kernel that ignored the
wraparound would return other distances."""
import numpy as np

from crafted import wrap_index
from oracle import oracle as orc
from test_oracle import emulate_avx512_accumulate


def _wraps_per_vector(idx, q_rot, bits):
    oi = orc.OracleIndex(idx)
    pid, pd = oi.centroid_topn(q_rot, idx.K)
    u, _ = orc.normalize_query(q_rot, idx.centroids[pid[0]], float(np.sqrt(pd[0])))
    qq, _ = orc.quantize_query(u)
    lut, shift = orc.pack_lut(qq)
    exact = bits.astype(np.int64) @ qq.astype(np.int64)
    out, blk, D = [], idx.block_start(), idx.D
    n = int(idx.cluster_sizes[0])
    for k in range(idx.K):
        for b in range((n + 31) // 32):
            off = int(blk[k] + b) * idx.block_bytes
            codes = idx.short_data[off:off + 4 * D]
            _, _, sums, _ = orc.fastscan_block(codes, lut, np.ones(32, np.float32), 0.0, 1.0,
                                               1.0, shift, 0.1, 0.0, D)
            if k == 0 and b == 0:
                np.testing.assert_array_equal(sums, emulate_avx512_accumulate(codes, lut, D))
            m = min(32, n - 32 * b)
            v0 = k * n + 32 * b
            out.append((exact[v0:v0 + m] - (sums[:m] + shift)) // 65536)
    return qq, np.concatenate(out)


def test_crafted_digits_and_wrap_mix():
    idx, q, bits = wrap_index()
    qq, wr = _wraps_per_vector(idx, q[0], bits)
    assert np.abs(qq).max() == 8191
    assert np.all(np.abs(qq.astype(np.int64)) % 256 == 255)
    counts = np.bincount(wr, minlength=3)
    assert counts.size == 3 and np.all(counts > 20), counts   # 0, 1 and 2 wraps all occur
    oi = orc.OracleIndex(idx)
    by_id = np.empty(int(idx.ids.max()) + 1, np.int64)
    by_id[idx.ids] = np.arange(idx.N)
    for qi in range(q.shape[0]):
        ids, _, st = oi.search(q[qi], 20, idx.K, stats=True)
        assert 0 < st[2] < idx.N                          # the lower bound prunes some
        assert np.count_nonzero(wr[by_id[ids]]) > 0       # wrapped vectors are in the result
