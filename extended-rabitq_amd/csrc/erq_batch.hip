// erq_batch.hip -- cluster-major batched ExRaBitQ search on gfx950.
//
// Many queries with uneven probe lists are searched together.  Probe ranks are split into
// rounds ([0,4) [4,16) [16,64) [64,192), then 128 wide).  Round 0 is the exact per-query
// scan (erq_search.hip).  The later rounds:
//   1. bucket every round's (query, rank) pairs by (round, cluster) at once (count / bin
//      scans / scatter) and prepare every pair's quantized residual (prep_kernel),
//   2. per round, run one persistent kernel over (cluster, <= 16 entries) work items, so
//      each cluster's codes are read once per round for all the queries that probe it,
//      then re-estimate the survivors (reest_kernel),
//   3. merge each query's new candidates into its top-k pool.
// Within a round the pruning threshold of a query is fixed (its k-th distance at round
// start), so the candidate set -- and the merged result -- does not depend on scheduling.
//
// The FastScan integer I = <x_b, q> (q: the 14-bit quantized query residual of
// high_acc_quantize16, space.hpp:695-711; q = 128 hi + lo, lo in [0,127]) on int8 MFMA:
//   * at load, the 1-bit codes are re-laid out per 512-dim chunk as 64 bytes per vector,
//     byte B bit p = dim 512c + 8B + p ("bit-plane" layout);
//   * main pass, v_mfma_i32_32x32x32_i8: B = 16 code bytes of one vector per lane taken as
//     bit-plane p ((byte >> p) & 1), A = the digits (16 entries x 2 rows) of the dims that
//     plane covers; all planes and chunks accumulate into one integer tile;
//   * inline pass, v_mfma_i32_16x16x64_i8 (8 entries x 2 rows): B = the bytes masked in
//     place (one v_and per dword, values 0 or 2^p), I = sum_p (acc_p >> p), plane 7 with
//     the sign flipped (0x80 is -128).
// The integer equals the reference's S + shift whenever its u16 lanes do not wrap
// (always for D <= 2048).  Lower bounds and EX-bit re-estimates use the same formulas,
// in the same order, as the per-query kernel.
#include "erq_batch.h"
#include "erq_device.h"
#include "erq_internal.h"

namespace erq {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kItemGroup = 16; // entries per work item (32 rows of the main pass's MFMA)
constexpr int kGroup = 8;      // entries per inline pass over an item half (16 MFMA rows)
constexpr int kThreads = 256;  // 4 waves
constexpr int kItemTiles = 16; // 16-vector tiles (8 blocks of 32) per work item, at most
                               // (BatchArgs::item_tiles: 8 or 16)
constexpr int kQSurv = kItemTiles * 16;  // survivors of one query in one work item, at most
constexpr int kMainSurv = 128; // of them kept in LDS by the main pass (more: inline pass)

// ---------------------------------------------------------------- relayout
// One thread per output byte: (block, chunk, vector, byte B).  Blocks of BB bytes: 4D code
// bytes and the factors (128 bytes per factor per 32 vectors, one factor in the HA format,
// four in the non-HA one).
__global__ void relayout_kernel(const uint8_t* __restrict__ short_data, size_t nblocks, int D,
                                int nchunk, size_t BB, uint8_t* __restrict__ bp) {
    const size_t per_block = (size_t)nchunk * 32 * 64;
    const size_t total = nblocks * per_block;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        size_t blk = i / per_block;
        int rem = (int)(i % per_block);
        int c = rem / (32 * 64);
        int v = (rem / 64) % 32;
        int B = rem % 64;
        int d0 = 512 * c + 8 * B;
        uint32_t out = 0;
        if (d0 < D) {
            const uint8_t* codes = short_data + blk * BB;
            int j = byte_of_vec(v);
            for (int h = 0; h < 2; h++) {
                int m = (d0 >> 2) + h;  // codebook: dims 4m..4m+3, dim 4m at nibble bit 3
                uint8_t b = codes[(m >> 1) * 32 + (m & 1) * 16 + j];
                uint32_t nib = v < 16 ? (b & 15u) : (b >> 4);
                for (int k = 0; k < 4; k++) out |= ((nib >> (3 - k)) & 1u) << (4 * h + k);
            }
        }
        bp[i] = (uint8_t)out;
    }
}

// exl (lane-major EX-bit codes, erq_device.h) from the reference's compact layout: one
// thread per (vector, lane), values decoded with decode4 and packed in increasing i.
template <int EX>
__global__ void exl_relayout_kernel(const uint8_t* __restrict__ long_code, size_t N, int D, int W,
                                    uint32_t* __restrict__ exl) {
    constexpr int vpd = exl_vpd(EX), sp = exl_sp(EX);
    const size_t CL = (size_t)D * EX / 8;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < N * 16;
         t += (size_t)gridDim.x * blockDim.x) {
        const size_t v = t >> 4;
        const int l = (int)(t & 15);
        const uint8_t* code = long_code + v * CL;
        uint32_t* out = exl + (v * 16 + l) * (size_t)W;
        uint32_t cur = 0;
        int w = 0;
        for (int ch = 0; ch < D / 64; ch++) {
            float e[4];
            decode4<EX>(code + ch * 8 * EX, l, e);
            for (int s4 = 0; s4 < 4; s4++) {
                const int i = 4 * ch + s4;
                if (i / vpd != w) {
                    out[w++] = cur;
                    cur = 0;
                }
                const int r = i % vpd;  // value r of the dword: byte r % 4, slot r / 4
                cur |= (uint32_t)e[s4] << (8 * (r & 3) + sp * (r >> 2));
            }
        }
        out[w++] = cur;
        for (; w < W; w++) out[w] = 0;
    }
}

// ---------------------------------------------------------------- work lists
__global__ void zero_u32_kernel(uint32_t* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// The round of a probe rank past the exact ones: [4,16), [16,64), [64,192), then 128 wide
// (the host's round schedule).
__device__ __forceinline__ uint32_t round_of_rank(uint32_t rank) {
    if (rank < 16) return 0;
    if (rank < 64) return 1;
    if (rank < 192) return 2;
    return 3 + (rank - 192) / 128;
}

// Entries of all rounds at once, binned by (round, cluster): bin = round * K + cluster.
// One wave per query (workgroups of 256 = 4 queries); the lanes take its probes, so the
// atomics of one query are in flight together.
__global__ void count_kernel(const uint32_t* __restrict__ off, const uint32_t* __restrict__ pid,
                             int nq, int r_first, int r_end, int K,
                             uint32_t* __restrict__ counts) {
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (q >= nq) return;
    const uint32_t s = off[q], e = min(off[q + 1], s + (uint32_t)r_end);
    for (uint32_t p = s + (uint32_t)r_first + lane; p < e; p += 64) {
        const uint32_t c = pid[p];
        if (c < (uint32_t)K) atomicAdd(&counts[round_of_rank(p - s) * K + c], 1u);
    }
}

// Work items from the bin counts, by workgroups of 256 (small enough to start beside other
// kernels).  A work item is (cluster, up to kItemGroup consecutive entries, up to
// item_tiles tiles of the cluster), three words: cluster, (first entry << 5) | entry count,
// first tile.  Items and entries are in (round, cluster) order; rinfo[4r + 0..3] = the
// round's items, first item, entries, items queued for the inline pass (0); rinfo[4R + 0],
// [4R + 2] = the totals.
__device__ __forceinline__ uint32_t bin_items(const uint32_t* csize, int K, int bin, uint32_t cnt,
                                              uint32_t& nsplit, int item_tiles) {
    const uint32_t ntile = (csize[bin % K] + 15) / 16;
    nsplit = (ntile + item_tiles - 1) / item_tiles;  // 0: empty cluster
    return (cnt + kItemGroup - 1) / kItemGroup * nsplit;
}

// Inclusive scan of (a, b) over the 256 threads of a workgroup.
__device__ __forceinline__ void block_scan2(uint32_t& a, uint32_t& b, uint32_t* wa, uint32_t* wb) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t x = __shfl_up(a, o), y = __shfl_up(b, o);
        if (lane >= o) {
            a += x;
            b += y;
        }
    }
    if (lane == 63) {
        wa[wv] = a;
        wb[wv] = b;
    }
    __syncthreads();
    for (int w = 0; w < wv; w++) {
        a += wa[w];
        b += wb[w];
    }
}

// Per bin: entries and items before it within its workgroup of 256 bins; per workgroup
// the totals.
__global__ void __launch_bounds__(256) bin_scan_kernel(const uint32_t* __restrict__ counts,
                                                      const uint32_t* __restrict__ csize, int K,
                                                      int nb, int item_tiles,
                                                      uint32_t* __restrict__ local,
                                                      uint32_t* __restrict__ btot) {
    __shared__ uint32_t wa[4], wb[4];
    const int bin = blockIdx.x * 256 + threadIdx.x;
    uint32_t ns, cnt = bin < nb ? counts[bin] : 0;
    uint32_t it = bin < nb ? bin_items(csize, K, bin, cnt, ns, item_tiles) : 0;
    uint32_t ic = cnt, ii = it;
    block_scan2(ic, ii, wa, wb);
    if (bin < nb) {
        local[2 * bin] = ic - cnt;
        local[2 * bin + 1] = ii - it;
    }
    if (threadIdx.x == 255) {
        btot[2 * blockIdx.x] = ic;
        btot[2 * blockIdx.x + 1] = ii;
    }
}

// One workgroup: exclusive prefix of the workgroup totals (in place) and the round table.
__global__ void __launch_bounds__(256) bin_total_kernel(uint32_t* __restrict__ btot, int nblk,
                                                       const uint32_t* __restrict__ local, int K,
                                                       int R, uint32_t* __restrict__ rinfo) {
    __shared__ uint32_t wa[4], wb[4];
    __shared__ uint32_t s_item[kMaxRounds + 1], s_ent[kMaxRounds + 1];
    const int t = threadIdx.x, per = (nblk + 255) / 256;
    const int i_lo = min(nblk, t * per), i_hi = min(nblk, (t + 1) * per);
    uint32_t oc = 0, oi = 0;
    for (int i = i_lo; i < i_hi; i++) {
        oc += btot[2 * i];
        oi += btot[2 * i + 1];
    }
    uint32_t ic = oc, ii = oi;
    block_scan2(ic, ii, wa, wb);
    uint32_t c = ic - oc, it = ii - oi;  // exclusive prefix of this thread's first total
    for (int i = i_lo; i < i_hi; i++) {
        const uint32_t a = btot[2 * i], b = btot[2 * i + 1];
        btot[2 * i] = c;
        btot[2 * i + 1] = it;
        c += a;
        it += b;
    }
    if (t == 255) {  // the totals
        s_ent[R] = ic;
        s_item[R] = ii;
    }
    __syncthreads();  // every workgroup offset is in btot
    if (t < R) {  // round t starts at bin t K: its workgroup's offset + the bin's local prefix
        const int bin = t * K, blk = bin / 256;
        s_ent[t] = btot[2 * blk] + local[2 * bin];
        s_item[t] = btot[2 * blk + 1] + local[2 * bin + 1];
    }
    __syncthreads();
    if (t <= R) {
        rinfo[4 * t + 0] = t < R ? s_item[t + 1] - s_item[t] : s_item[R];
        rinfo[4 * t + 1] = t < R ? s_item[t] : 0;
        rinfo[4 * t + 2] = t < R ? s_ent[t + 1] - s_ent[t] : s_ent[R];
        rinfo[4 * t + 3] = 0;
    }
}

// Per bin, one wave each (the lanes take its items, so a bin with many entries does not
// hold its workgroup): the cursor of its entries and its items.
__global__ void __launch_bounds__(256) bin_items_kernel(const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ csize, int K,
                                                       int nb, int item_tiles,
                                                       const uint32_t* __restrict__ local,
                                                       const uint32_t* __restrict__ btot,
                                                       uint32_t* __restrict__ cursor,
                                                       uint32_t* __restrict__ items) {
    const int bin = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (bin >= nb) return;
    const uint32_t cnt = counts[bin];
    uint32_t ns;
    const uint32_t nit = bin_items(csize, K, bin, cnt, ns, item_tiles);
    const int blk = bin >> 8;  // the bin's workgroup of bin_scan_kernel
    const uint32_t c0 = btot[2 * blk] + local[2 * bin];
    const uint32_t i0 = btot[2 * blk + 1] + local[2 * bin + 1];
    if (lane == 0) cursor[bin] = c0;
    for (uint32_t j = lane; j < nit; j += 64) {
        const uint32_t grp = j / ns, sp = j % ns, w = i0 + j;
        items[3 * w] = (uint32_t)(bin % K);
        items[3 * w + 1] = ((c0 + grp * kItemGroup) << 5) |
                           min((uint32_t)kItemGroup, cnt - grp * kItemGroup);
        items[3 * w + 2] = sp * (uint32_t)item_tiles;
    }
}

// One wave per query, as count_kernel.
__global__ void scatter_kernel(const uint32_t* __restrict__ off, const uint32_t* __restrict__ pid,
                               int nq, int r_first, int r_end, int K,
                               uint32_t* __restrict__ cursor, uint32_t* __restrict__ entries) {
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (q >= nq) return;
    const uint32_t s = off[q], e = min(off[q + 1], s + (uint32_t)r_end);
    for (uint32_t p = s + (uint32_t)r_first + lane; p < e; p += 64) {
        const uint32_t c = pid[p];
        if (c >= (uint32_t)K) continue;
        const uint32_t pos = atomicAdd(&cursor[round_of_rank(p - s) * K + c], 1u);
        reinterpret_cast<uint2*>(entries)[pos] = make_uint2(p, (uint32_t)q);
    }
}

// The same counts and scatter with the atomics of a workgroup's entries gathered in LDS:
// kCountGroups workgroups of 1024 threads, workgroup g taking queries [nq g / G,
// nq (g + 1) / G) (a wave per query).  count_priv_kernel counts its entries per bin in LDS,
// then adds each nonzero count to the bin's total with one atomic and keeps what the atomic
// returns (the entries of the bin placed before its own) in wgbase; scatter_priv_kernel,
// after the bins' starts are known (cursor, from bin_items_kernel), places its entries
// from there with LDS atomics.  A bin's entries come out grouped by workgroup, in any
// order within it, as the global-atomic kernels leave them in any order.
__global__ void __launch_bounds__(1024) count_priv_kernel(const uint32_t* __restrict__ off,
                                                          const uint32_t* __restrict__ pid,
                                                          int nq, int r_first, int r_end, int K,
                                                          int nb, uint32_t* __restrict__ counts,
                                                          uint32_t* __restrict__ wgbase) {
    extern __shared__ uint32_t hcnt[];  // nb
    for (int i = threadIdx.x; i < nb; i += 1024) hcnt[i] = 0;
    __syncthreads();
    const int g = blockIdx.x, G = gridDim.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int q0 = (int)((long long)nq * g / G), q1 = (int)((long long)nq * (g + 1) / G);
    for (int q = q0 + wv; q < q1; q += 16) {
        const uint32_t s = off[q], e = min(off[q + 1], s + (uint32_t)r_end);
        for (uint32_t p = s + (uint32_t)r_first + lane; p < e; p += 64) {
            const uint32_t c = pid[p];
            if (c < (uint32_t)K) atomicAdd(&hcnt[round_of_rank(p - s) * K + c], 1u);
        }
    }
    __syncthreads();
    for (int bin = threadIdx.x; bin < nb; bin += 1024) {
        const uint32_t cnt = hcnt[bin];
        wgbase[(size_t)g * nb + bin] = cnt ? atomicAdd(&counts[bin], cnt) : 0u;
    }
}

__global__ void __launch_bounds__(1024) scatter_priv_kernel(const uint32_t* __restrict__ off,
                                                            const uint32_t* __restrict__ pid,
                                                            int nq, int r_first, int r_end, int K,
                                                            int nb,
                                                            const uint32_t* __restrict__ cursor,
                                                            const uint32_t* __restrict__ wgbase,
                                                            uint32_t* __restrict__ entries) {
    extern __shared__ uint32_t cur[];  // nb
    const int g = blockIdx.x, G = gridDim.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int bin = threadIdx.x; bin < nb; bin += 1024)
        cur[bin] = cursor[bin] + wgbase[(size_t)g * nb + bin];
    __syncthreads();
    const int q0 = (int)((long long)nq * g / G), q1 = (int)((long long)nq * (g + 1) / G);
    for (int q = q0 + wv; q < q1; q += 16) {
        const uint32_t s = off[q], e = min(off[q + 1], s + (uint32_t)r_end);
        for (uint32_t p = s + (uint32_t)r_first + lane; p < e; p += 64) {
            const uint32_t c = pid[p];
            if (c >= (uint32_t)K) continue;
            const uint32_t pos = atomicAdd(&cur[round_of_rank(p - s) * K + c], 1u);
            reinterpret_cast<uint2*>(entries)[pos] = make_uint2(p, (uint32_t)q);
        }
    }
}

// ---------------------------------------------------------------- per-pair preparation
// One wave per (query, cluster) pair of the round: the reference's query preparation
// (normalize_query16 + high_acc_quantize16, space.hpp:695-778) and the int8 digits of the
// quantized residual in the MFMA operand order, [chunk][plane p][lane group][digit][16 B]
// (digit 0: q mod 128, digit 1: floor(q / 128)), plus the scalars the scan needs.
struct QParam {
    float y, sqr_y, fac, sumq, delta, c1, c2, qs, thr;
    int q;
    uint32_t cid;  // the pair's cluster
};

__global__ void __launch_bounds__(256) prep_kernel(BatchArgs a, int per_wave_bytes) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ float s_sum[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int D = a.D, nchunk = a.nchunk, Dp = nchunk * 512, S = ul_stride(D);
    float* uq = reinterpret_cast<float*>(smem + (size_t)wid * per_wave_bytes);
    // A fixed grid (a multiple of 8 workgroups) walks the entries, four at a time per
    // workgroup (one per wave).  XCD-aware: workgroup b runs on XCD b % 8; entries are
    // grouped by cluster, so giving each XCD a contiguous eighth of them keeps a centroid's
    // reads in one L2.
    const uint32_t n = a.n_items[2];
    const uint32_t x = blockIdx.x & 7, nw = (gridDim.x >> 3) * 4;
    const uint32_t lo = (uint32_t)((uint64_t)n * x / 8), hi = (uint32_t)((uint64_t)n * (x + 1) / 8);
    uint32_t step = 0;
    for (uint32_t base = lo + (blockIdx.x >> 3) * 4; base < hi; base += nw, step++) {  // uniform
        const uint32_t e = base + wid;
        const bool valid = e < hi;
        QParam P;
        float delta = 0.f;
        bool res = true;
        if (valid) {
            const uint2 ent = reinterpret_cast<const uint2*>(a.entries)[e];
            const uint32_t slot = ent.x, cid = a.probe_ids[slot];
            P.q = (int)ent.y;
            P.cid = cid;
            P.sqr_y = a.probe_sqr_y[slot];
            P.y = sqrtf(P.sqr_y);
            // three slices in flight: the entry's rows arrive in two waits, not six (D = 1536)
            res = residual_delta<3>(a.q_rot + (size_t)P.q * D, a.centroids + (size_t)cid * D,
                                    P.y, D, S, uq, lane, P.fac, delta);
        }
        __syncthreads();
        // the four residual sums by one wave (16 lanes per entry; the wave rotates), while
        // every wave quantizes its own entry
        if (wid == (int)(step & 3)) {
            const float* ur =
                reinterpret_cast<const float*>(smem + (size_t)(lane >> 4) * per_wave_bytes);
            const float part = residual_sum16(ur, S, D, lane & 15);
            if ((lane & 15) == 0) s_sum[lane >> 4] = part;
        }
        // the digits go straight to the entry's operand rows in HBM
        if (valid)
            quantize_digits(uq, S, D, nchunk, 1.0 / (double)delta,
                            reinterpret_cast<uint32_t*>(a.adig + (size_t)e * 2 * Dp), lane);
        __syncthreads();
        if (valid && lane == 0) {
            const float sumq = res ? s_sum[wid] : (float)D * (float)(1.0 / (double)sqrtf((float)D));
            P.sumq = sumq;
            P.delta = delta;
            P.c1 = (float)(0.5 * (double)sumq - (double)0.58f);
            P.c2 = (-5.0f * P.y) * a.one_over_sqrtD;
            P.qs = P.y * P.y;
            P.thr = 0.f;  // the round's threshold: batch_scan_kernel writes it
            reinterpret_cast<QParam*>(a.qparams)[e] = P;
        }
    }
}

// ---------------------------------------------------------------- batched scan
// One workgroup per work item.  Each wave loads the codes of its (up to 4) tiles of the
// item while the item's digits are staged in LDS, runs the bit-plane MFMAs, and appends
// every vector whose lower bound is below a query's threshold to that query's LDS list.
// The lists are then written to the round's survivor list (one record per (query,
// vector): entry, vector, scaled FastScan product, |o - c|) for reest_kernel.  If that list
// is full, the workgroup re-estimates its survivors itself (same formulas, same result).

// The EX-bit re-estimate of one survivor by a 16-lane group, in the reference's order:
// 16 lane-strided FMA chains over u * code (ex_chain on the lane-major codes), reduced by
// the xor butterfly, then the distance in double as C promotes it (HASearcher.hpp:183-231).
// uf(i) gives u for the lane's dim 16 i + l.
template <int EX>
__device__ __forceinline__ float reest_distance(const BatchArgs& a, const QParam& P, float part,
                                                bool act, float ipq, float on, uint32_t vi) {
    const float ip = butterfly16(part);
    if (!act) return FLT_MAX;
    const float R = (float)(1 << EX);
    const float sqr_x = on * on;
    const float Aa = sqr_x + P.sqr_y;
    const float Bb = a.ex_factor[vi] * P.y;
    const float Cc = R * ipq + ip;
    const double G = (double)Cc - ((double)R - 0.5) * (double)P.sumq;
    const double H = (double)Bb * G;
    return (float)((double)Aa - H);
}

template <int EX>
__device__ __forceinline__ float reestimate(const BatchArgs& a, const QParam& P, const float* ur,
                                            const uint32_t* lanecode, bool act, float ipq,
                                            float on, uint32_t vi) {
    const float part = act ? ex_chain<EX>(ur, lanecode, a.D >> 4, a.exl_w) : 0.f;
    return reest_distance<EX>(a, P, part, act, ipq, on, vi);
}

// Above D = 2048 (nchunk > 4): a survivor found on the exact FastScan integer I gets the
// reference's integer (lo_wraps, on the pair's digits qv) and the lower-bound test again,
// against the threshold in P.  Every lane of the 16-lane group calls it (inactive groups with
// any valid vector).  Returns whether the vector survives; ipq = its scaled product.
__device__ __forceinline__ bool wrap_retest(const BatchArgs& a, const QParam& P,
                                            const int16_t* qv, int vloc, int I, float on, int l,
                                            float& ipq) {
    const uint32_t blk = a.bstart[P.cid] + (uint32_t)(vloc >> 5);
    const int vb = vloc & 31;
    const int wr = lo_wraps(qv, a.nchunk, l, [&](int c) {
        return *reinterpret_cast<const uint32_t*>(
            a.bp + (((size_t)blk * a.nchunk + c) * 32 + vb) * 64 + 4 * l);
    });
    ipq = (float)(I - 65536 * wr) * P.delta;
    float tl = ipq - P.c1;
    tl = tl * P.c2;
    tl = tl * on;
    tl = tl + P.qs;
    tl = tl + on * on;
    return tl < P.thr;
}

__device__ __forceinline__ void add_candidate(const BatchArgs& a, int q, float exd, uint32_t id) {
    const uint32_t cs = atomicAdd(&a.cand_cnt[q], 1u);
    if (cs < (uint32_t)a.cand_cap) {
        a.cand_d[(size_t)q * a.cand_cap + cs] = exd;
        a.cand_i[(size_t)q * a.cand_cap + cs] = id;
    }
}

template <int NCH>
__device__ __forceinline__ void load_tile_codes(const uint8_t* bp, uint32_t blk, int vib, int lg,
                                                uint4 (&w)[NCH]) {
#pragma unroll
    for (int c = 0; c < NCH; c++)
        w[c] = *reinterpret_cast<const uint4*>(bp + (((size_t)blk * NCH + c) * 32 + vib) * 64 +
                                               16 * lg);
}

// The main pass: one workgroup per work item (cluster, up to 16 entries, up to 8 blocks of
// 32 vectors; each wave takes blocks w and w + 4).  The item's digits are staged in LDS as
// the 32-row A operand of v_mfma_i32_32x32x32_i8 (row 2 qq + d = digit d of entry qq);
// B is bit-plane p of a block's code bytes as 0/1 bytes ((byte >> p) & 1), so all planes
// and chunks accumulate into one 32 x 32 integer tile: lane (h, n) ends with the sums of
// vector n for rows 8 b + 4 h + i.  Every vector whose lower bound is below its entry's
// threshold goes to that entry's LDS list; the lists are written to the round's survivor
// list (one record per (entry, vector): entry, vector, FastScan product scaled by delta --
// above 4 chunks the exact integer instead, for the wrap correction --, |o - c|)
// for reest_kernel.  An item with a list beyond kMainSurv, or whose records do not fit in
// the workgroup's slice, is queued (both halves) for batch_inline_kernel instead.  Up to
// 4 chunks the next item's digits and both blocks' codes are held in registers ahead of
// use; above (D > 2048) they are loaded when needed.
template <int EX, int NCH>
__global__ void __launch_bounds__(kThreads) batch_scan_kernel(BatchArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int Dp = NCH * 512;
    constexpr bool kAhead = NCH <= 4;
    const int D = a.D;
    int8_t* A = reinterpret_cast<int8_t*>(smem);  // [NCH][8 p][4 g][32 rows][16 B]
    uint32_t* s_vec = reinterpret_cast<uint32_t*>(smem + (size_t)NCH * 16384);
    int* s_int = reinterpret_cast<int*>(s_vec + kItemGroup * kMainSurv);
    float* s_on = reinterpret_cast<float*>(s_int + kItemGroup * kMainSurv);
    __shared__ QParam qp[kItemGroup];
    __shared__ uint32_t s_cnt[kItemGroup];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int h = lane >> 5, n32 = lane & 31;
    const size_t BB = 4 * (size_t)D + 128;
    unsigned long long st_tiles = 0;
    // this workgroup's slice of the survivor list
    const size_t slice_cap = a.surv_cap / gridDim.x;
    uint4* slice = a.surv + (size_t)blockIdx.x * slice_cap;
    size_t used = 0;
    const uint32_t nitems = a.n_items[0], ifirst = a.n_items[1];

    // XCD-aware order: workgroup b runs on XCD b % 8, so each XCD takes a contiguous eighth
    // of the items (clusters in order) and the items of one cluster share that XCD's L2.
    uint32_t it_lo = 0, it_hi = nitems, it_step = gridDim.x, it_first = blockIdx.x;
    if ((gridDim.x & 7) == 0) {
        const uint32_t x = blockIdx.x & 7;
        it_lo = (uint32_t)((uint64_t)nitems * x / 8);
        it_hi = (uint32_t)((uint64_t)nitems * (x + 1) / 8);
        it_step = gridDim.x >> 3;
        it_first = it_lo + (blockIdx.x >> 3);
    }
    // The item's digits (16-byte units (c, p, g) x 32 rows; rows of missing entries are 0)
    // pass through registers: the next item's are loaded while this one computes.
    constexpr int kStage = NCH * 1024 / kThreads;
    constexpr int kHeld = kAhead ? kStage : 1;
    uint4 dreg[kHeld];
    auto digit_unit = [&](uint32_t item, int i) {
        const uint32_t f = a.items[3 * item + 1], de0 = f >> 5, delen = f & 31u;
        const int u = tid + kThreads * i, row = u & 31, cpg = u >> 5;
        const int qq = row >> 1, dgt = row & 1;
        return qq < (int)delen ? reinterpret_cast<const uint4*>(
                                     a.adig + (size_t)(de0 + qq) * 2 * Dp)[cpg * 2 + dgt]
                               : make_uint4(0, 0, 0, 0);
    };
    auto load_digits = [&](uint32_t item) {
        if constexpr (kAhead) {
#pragma unroll
            for (int i = 0; i < kStage; i++) dreg[i] = digit_unit(item, i);
        }
    };
    // block blk's codes (16 bytes per (chunk, K half) per lane) and its |o - c| of vector n32
    auto load_codes = [&](uint32_t blk, uint4 (&wc)[NCH][2], float& on) {
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const uint8_t* src = a.bp + (((size_t)blk * NCH + c) * 32 + n32) * 64 + 16 * h;
            wc[c][0] = *reinterpret_cast<const uint4*>(src);
            wc[c][1] = *reinterpret_cast<const uint4*>(src + 32);
        }
        on = reinterpret_cast<const float*>(a.short_data + (size_t)blk * BB + 4 * (size_t)D)[n32];
    };
    // this wave's blocks of an item (bfirst + wid and bfirst + wid + 4): up to 4 chunks their
    // codes and factors are loaded one item ahead, while the previous item's vectors are
    // tested and its lists written
    uint4 w[kAhead ? 2 : 1][NCH][2];
    float onp[2] = {0.f, 0.f};
    auto fetch_codes = [&](uint32_t item) {
        const uint32_t c = a.items[3 * item];
        const int bf = (int)a.items[3 * item + 2] >> 1;
        const int be = min((int)((a.csize[c] + 31) >> 5), bf + a.item_tiles / 2);
        const uint32_t bb = a.bstart[c];
#pragma unroll
        for (int j = 0; j < (kAhead ? 2 : 0); j++) {
            const int bl = bf + wid + 4 * j;
            if (bl < be) load_codes(bb + (uint32_t)bl, w[j], onp[j]);
        }
    };
    if (it_first < it_hi) {
        load_digits(ifirst + it_first);
        if constexpr (kAhead) fetch_codes(ifirst + it_first);
    }
    for (uint32_t it = it_first; it < it_hi; it += it_step) {
        __syncthreads();  // the previous item's LDS is no longer read
        const uint32_t item = ifirst + it;  // global index
        const uint32_t cid = a.items[3 * item];
        const uint32_t e0 = a.items[3 * item + 1] >> 5, elen = a.items[3 * item + 1] & 31u;
        const int bfirst = (int)a.items[3 * item + 2] >> 1;
        const int n = (int)a.csize[cid];
        const uint32_t b0 = a.bstart[cid], v0 = a.vstart[cid];
        const int bend = min((n + 31) >> 5, bfirst + a.item_tiles / 2);
        if (tid < kItemGroup) {
            s_cnt[tid] = 0;
            if (tid < (int)elen) {
                QParam P = reinterpret_cast<const QParam*>(a.qparams)[e0 + tid];
                P.thr = a.pool_d[(size_t)P.q * a.k + a.k - 1];  // k-th distance at round start
                qp[tid] = P;
                // for reest_kernel, which then needs no dependent read of the pool: written
                // by the entry's first item only (the pool does not change within the round)
                if (bfirst == 0) reinterpret_cast<QParam*>(a.qparams)[e0 + tid].thr = P.thr;
            }
        }
        if constexpr (kAhead) {
#pragma unroll
            for (int i = 0; i < kStage; i++)
                reinterpret_cast<uint4*>(A)[tid + kThreads * i] = dreg[i];
        } else {
            for (int i = 0; i < kStage; i++)
                reinterpret_cast<uint4*>(A)[tid + kThreads * i] = digit_unit(item, i);
        }
        __syncthreads();
        if (it + it_step < it_hi) load_digits(item + it_step);

        // the lower-bound test of block bl's 32 vectors against the item's entries, from the
        // block's integer tile (lane (h, n): rows 8 b + 4 h + i of vector n)
        auto test_block = [&](int bl, const v16i& acc, float on) {
            const int vloc = 32 * bl + n32;
            if (lane == 0) st_tiles += (n - 32 * bl > 16) ? 2 : 1;
#pragma unroll
            for (int b = 0; b < 4; b++) {
#pragma unroll
                for (int i2 = 0; i2 < 2; i2++) {
                    const int qq = 4 * b + 2 * h + i2;
                    if (qq >= (int)elen) continue;
                    const QParam& P = qp[qq];
                    const int I = 128 * acc[4 * b + 2 * i2 + 1] + acc[4 * b + 2 * i2];
                    const float ipq = (float)I * P.delta;
                    float tl = ipq - P.c1;
                    tl = tl * P.c2;
                    tl = tl * on;
                    tl = tl + P.qs;
                    tl = tl + on * on;
                    if (vloc < n && tl < P.thr) {
                        const uint32_t sidx = atomicAdd(&s_cnt[qq], 1u);
                        if (sidx < (uint32_t)kMainSurv) {
                            s_vec[qq * kMainSurv + sidx] = (uint32_t)vloc;
                            // the exact integer where the wrap correction needs it, else the
                            // scaled product the re-estimate uses
                            s_int[qq * kMainSurv + sidx] = NCH > 4 ? I : __float_as_int(ipq);
                            s_on[qq * kMainSurv + sidx] = on;
                        }
                    }
                }
            }
        };
        // the A fragment of (chunk c, K half sk, plane p) and bit-plane p of a code dword set
        auto afrag = [&](int c, int sk, int p) {
            return *reinterpret_cast<const v4i*>(
                A + ((((c * 8 + p) * 4 + 2 * sk + h) * 32) + n32) * 16);
        };
        auto plane = [](const uint4& cw, int p) {
            return v4i{(int)((cw.x >> p) & 0x01010101u), (int)((cw.y >> p) & 0x01010101u),
                       (int)((cw.z >> p) & 0x01010101u), (int)((cw.w >> p) & 0x01010101u)};
        };
        if constexpr (kAhead) {
            const int bl0 = bfirst + wid, bl1 = bl0 + 4;
            if (bl0 < bend) {
                if (bl1 < bend) {
                    // both of the wave's blocks: each A fragment is read from LDS once and
                    // feeds one MFMA per block (two independent accumulation chains)
                    v16i acc0 = {}, acc1 = {};
#pragma unroll
                    for (int c = 0; c < NCH; c++) {
#pragma unroll
                        for (int sk = 0; sk < 2; sk++) {
#pragma unroll 2
                            for (int p = 0; p < 8; p++) {
                                const v4i av = afrag(c, sk, p);
                                acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                                    av, plane(w[0][c][sk], p), acc0, 0, 0, 0);
                                acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                                    av, plane(w[1][c][sk], p), acc1, 0, 0, 0);
                            }
                        }
                    }
                    const float on0 = onp[0], on1 = onp[1];
                    if (it + it_step < it_hi) fetch_codes(item + it_step);
                    test_block(bl0, acc0, on0);
                    test_block(bl1, acc1, on1);
                } else {
                    // one block: even and odd planes in two chains, summed after (integer
                    // addition, so the order does not matter)
                    v16i acc4[2] = {};
#pragma unroll
                    for (int c = 0; c < NCH; c++) {
#pragma unroll
                        for (int sk = 0; sk < 2; sk++) {
#pragma unroll
                            for (int p = 0; p < 8; p++)
                                acc4[p & 1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                                    afrag(c, sk, p), plane(w[0][c][sk], p), acc4[p & 1], 0, 0, 0);
                        }
                    }
                    const float on0 = onp[0];
                    if (it + it_step < it_hi) fetch_codes(item + it_step);
                    test_block(bl0, acc4[0] + acc4[1], on0);
                }
            } else if (it + it_step < it_hi) {
                fetch_codes(item + it_step);
            }
        } else {
            for (int j = 0; j < 2; j++) {
                const int bl = bfirst + wid + 4 * j;
                if (bl >= bend) break;
                uint4 (&wj)[NCH][2] = w[0];
                float on;
                load_codes(b0 + (uint32_t)bl, wj, on);
                v16i acc4[2] = {};
#pragma unroll
                for (int c = 0; c < NCH; c++) {
#pragma unroll
                    for (int sk = 0; sk < 2; sk++) {
#pragma unroll
                        for (int p = 0; p < 8; p++)
                            acc4[p & 1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                                afrag(c, sk, p), plane(wj[c][sk], p), acc4[p & 1], 0, 0, 0);
                    }
                }
                test_block(bl, acc4[0] + acc4[1], on);
            }
        }
        __syncthreads();

        int qoff[kItemGroup + 1];
        qoff[0] = 0;
        bool full = false;
#pragma unroll
        for (int qq = 0; qq < kItemGroup; qq++) {
            const int c = qq < (int)elen ? (int)s_cnt[qq] : 0;
            full |= c > kMainSurv;
            qoff[qq + 1] = qoff[qq] + c;
        }
        const int tot = qoff[kItemGroup];
        if (tot == 0) continue;  // uniform: every thread read the same counts
        if (!full && used + (size_t)tot <= slice_cap) {
            for (int i = tid; i < tot; i += kThreads) {
                int qq = 0;
#pragma unroll
                for (int jq = 1; jq < kItemGroup; jq++) qq += i >= qoff[jq] ? 1 : 0;
                const int si = qq * kMainSurv + (i - qoff[qq]);
                slice[used + i] = make_uint4(e0 + (uint32_t)qq, v0 + s_vec[si],
                                             (uint32_t)s_int[si], __float_as_uint(s_on[si]));
            }
            used += (size_t)tot;
        } else if (tid == 0) {  // lists or slice full: both halves go to the inline pass
            const uint32_t at = atomicAdd(&a.n_items[3], 2u);
            a.ovf[at] = 2 * item;
            a.ovf[at + 1] = 2 * item + 1;
        }
    }
    if (tid == 0) a.surv_count[blockIdx.x] = (uint32_t)used;
    if (a.stats) wave_stat_add(&a.stats[6], st_tiles);
}

// The inline pass: the halves (up to kGroup entries) of the items the main pass queued
// (survivors beyond its LDS lists or its slice), scanned again with the 16 x 16 x 64 MFMA
// and their survivors re-estimated in the workgroup (the same formulas), so every item's
// survivors are handled exactly once either way.  Queue entry = 2 x item + half.
template <int EX, int NCH>
__global__ void __launch_bounds__(kThreads) batch_inline_kernel(BatchArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int Dp = NCH * 512;
    constexpr int kWaveTiles = kItemTiles / 4;
    const int D = a.D;
    // LDS: A operands [NCH][8 p][4 g][16 rows][16 B] (row 2qq+d = digit d of query qq),
    // then per-query survivor lists (vector, scaled FastScan product).  When the
    // workgroup re-estimates survivors itself, the A region holds each wave's residual
    // (NCH * 8192 >= 4 waves * D floats).
    int8_t* A = reinterpret_cast<int8_t*>(smem);
    // the A region also holds the inline pass's residuals: 4 waves x 16 x S floats; above
    // 4 chunks each wave's digits for the wrap correction follow (Dp int16 per wave)
    constexpr bool kWrap = NCH > 4;
    const int S = ul_stride(D);
    const size_t abytes = max((size_t)NCH * 8192, (size_t)256 * S);
    int16_t* qvw = reinterpret_cast<int16_t*>(smem + abytes) + (size_t)(threadIdx.x >> 6) * Dp;
    uint32_t* s_vec = reinterpret_cast<uint32_t*>(smem + abytes + (kWrap ? 8 * (size_t)Dp : 0));
    int* s_int = reinterpret_cast<int*>(s_vec + kGroup * kQSurv);
    float* s_on = reinterpret_cast<float*>(s_int + kGroup * kQSurv);
    __shared__ QParam qp[kGroup];
    __shared__ uint32_t s_cnt[kGroup];

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int ln = lane & 15, lg = lane >> 4;
    const size_t BB = 4 * (size_t)D + 128;
    const size_t CL = (size_t)D * EX / 8;
    unsigned long long st_surv = 0, st_cand = 0;
    const uint32_t nq_items = a.n_items[3];  // queued item halves
    for (uint32_t it = blockIdx.x; it < nq_items; it += gridDim.x) {
        __syncthreads();  // the previous item's LDS is no longer read
        const uint32_t qe = a.ovf[it], item = qe >> 1, half = qe & 1u;  // global item index
        const uint32_t cid = a.items[3 * item];
        const uint32_t ecnt = a.items[3 * item + 1] & 31u;
        if (ecnt <= kGroup * half) continue;  // uniform: no entries in this half
        const uint32_t e0 = (a.items[3 * item + 1] >> 5) + kGroup * half;
        const uint32_t elen = min((uint32_t)kGroup, ecnt - kGroup * half);
        const int tfirst = (int)a.items[3 * item + 2];
        const int n = (int)a.csize[cid];
        const uint32_t b0 = a.bstart[cid], v0 = a.vstart[cid];
        const int tend = min((n + 15) >> 4, tfirst + a.item_tiles);
        // this wave's tiles: tfirst + wid + 4j, j < kWaveTiles; codes loaded before staging
        uint4 w[kWaveTiles][NCH];
#pragma unroll
        for (int j = 0; j < kWaveTiles; j++) {
            const int t = tfirst + wid + 4 * j;
            if (t < tend) load_tile_codes<NCH>(a.bp, b0 + (uint32_t)(t >> 1), 16 * (t & 1) + ln, lg, w[j]);
        }
        if (tid < kGroup) {
            s_cnt[tid] = 0;
            if (tid < (int)elen) {
                QParam P = reinterpret_cast<const QParam*>(a.qparams)[e0 + tid];
                P.thr = a.pool_d[(size_t)P.q * a.k + a.k - 1];  // k-th distance at round start
                qp[tid] = P;
            }
        }
        // stage the digits: 16-byte units (c, p, g, row); rows of missing queries are 0
        for (int u = tid; u < NCH * 512; u += kThreads) {
            const int row = u & 15, cpg = u >> 4;
            const int qq = row >> 1, dgt = row & 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (qq < (int)elen)
                v = reinterpret_cast<const uint4*>(a.adig + (size_t)(e0 + qq) * 2 * Dp)[cpg * 2 + dgt];
            reinterpret_cast<uint4*>(A)[u] = v;
        }
        __syncthreads();

#pragma unroll
        for (int j = 0; j < kWaveTiles; j++) {
            const int t = tfirst + wid + 4 * j;
            if (t >= tend) break;
            const uint32_t blk = b0 + (uint32_t)(t >> 1);
            const int vib = 16 * (t & 1) + ln;
            v4i acc[8];
#pragma unroll
            for (int p = 0; p < 8; p++) acc[p] = v4i{0, 0, 0, 0};
#pragma unroll
            for (int c = 0; c < NCH; c++) {
#pragma unroll
                for (int p = 0; p < 8; p++) {
                    const uint32_t m = 0x01010101u << p;
                    v4i bv = v4i{(int)(w[j][c].x & m), (int)(w[j][c].y & m), (int)(w[j][c].z & m),
                                 (int)(w[j][c].w & m)};
                    const v4i av = *reinterpret_cast<const v4i*>(
                        A + (((c * 8 + p) * 4 + lg) * 16 + ln) * 16);
                    acc[p] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bv, acc[p], 0, 0, 0);
                }
            }
            int X[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int sacc = -(acc[7][r] >> 7);
#pragma unroll
                for (int p = 0; p < 7; p++) sacc += acc[p][r] >> p;
                X[r] = sacc;
            }
            const int vloc = 16 * t + ln;
            const float on = reinterpret_cast<const float*>(
                a.short_data + (size_t)blk * BB + 4 * (size_t)D)[vib];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int qq = 2 * lg + h;
                if (qq >= (int)elen) continue;
                const QParam& P = qp[qq];
                const int I = 128 * X[2 * h + 1] + X[2 * h];
                const float ipq = (float)I * P.delta;
                float tl = ipq - P.c1;
                tl = tl * P.c2;
                tl = tl * on;
                tl = tl + P.qs;
                tl = tl + on * on;
                if (vloc < n && tl < P.thr) {
                    const uint32_t sidx = atomicAdd(&s_cnt[qq], 1u);
                    s_vec[qq * kQSurv + sidx] = (uint32_t)vloc;
                    s_int[qq * kQSurv + sidx] = I;
                    s_on[qq * kQSurv + sidx] = on;
                }
            }
        }
        __syncthreads();

        int qoff[kGroup + 1];
        qoff[0] = 0;
#pragma unroll
        for (int qq = 0; qq < kGroup; qq++)
            qoff[qq + 1] = qoff[qq] + (qq < (int)elen ? (int)s_cnt[qq] : 0);
        const int tot = qoff[kGroup];
        if (tot == 0) continue;  // uniform: every thread read the same counts
        // inline pass: re-estimate this item's survivors here
        __syncthreads();  // every wave is done with the digits in A
        const float value = (float)(1.0 / (double)sqrtf((float)D));
        float* uw = reinterpret_cast<float*>(A) + (size_t)wid * 16 * S;
        const float* cen = a.centroids + (size_t)cid * D;
        int cstart[kGroup + 1];
        cstart[0] = 0;
#pragma unroll
        for (int qq = 0; qq < kGroup; qq++)
            cstart[qq + 1] = cstart[qq] + ((qoff[qq + 1] - qoff[qq] + 15) >> 4);
        int staged = -1;
        for (int ci = wid; ci < cstart[kGroup]; ci += 4) {
            int qq = 0;
#pragma unroll
            for (int j = 1; j < kGroup; j++) qq += ci >= cstart[j] ? 1 : 0;
            const QParam& P = qp[qq];
            if (qq != staged) {
                wave_sync();
                const float* qv = a.q_rot + (size_t)P.q * D;
                const bool use_res = P.y > 1e-5f;
#pragma unroll 4
                for (int d = lane; d < D; d += 64)
                    uw[ul_index(d, S)] = use_res ? (qv[d] - cen[d]) * P.fac : value;
                wave_sync();
                if constexpr (kWrap) {
                    store_qv(uw, S, D, NCH, 1.0 / (double)P.delta, qvw, lane);
                    wave_sync();
                }
                staged = qq;
            }
            const int cnt = qoff[qq + 1] - qoff[qq];
            const int cb = (ci - cstart[qq]) * 16;
            for (int it = 0; it < 4 && cb + 4 * it < cnt; it++) {
                const int si = cb + 4 * it + lg;
                bool act = si < cnt;
                const int vloc = act ? (int)s_vec[qq * kQSurv + si] : 0;
                const uint32_t vi = v0 + (uint32_t)vloc;
                const float on = act ? s_on[qq * kQSurv + si] : 0.f;
                const int I = act ? s_int[qq * kQSurv + si] : 0;
                float ipq = (float)I * P.delta;
                if constexpr (kWrap) act = wrap_retest(a, P, qvw, vloc, I, on, ln, ipq) && act;
                const float exd = reestimate<EX>(
                    a, P, uw + ln * S,
                    a.exl + ((size_t)vi * 16 + ln) * a.exl_w, act, ipq, on, vi);
                if (act && ln == 0) {
                    if (exd <= P.thr) {
                        add_candidate(a, P.q, exd, a.ids[vi]);
                        st_cand++;
                    }
                    st_surv++;
                }
            }
        }
    }
    if (a.stats) {
        wave_stat_add(&a.stats[2], st_surv);
        wave_stat_add(&a.stats[4], st_cand);
    }
}

// The round's survivor list, all slices as one index space (exclusive prefix of the slice
// counts in LDS), each wave taking an equal contiguous share.  Records of one (query,
// cluster) pair come in runs; the wave walks its share run by run: it stages the pair's
// residual u once in its LDS slice (lane-major, from 16-byte loads of the rotated query and
// the centroid) and re-estimates the run's records four at a time, one per 16-lane group.
constexpr int kReestWG = 4;  // re-estimation workgroups per scan workgroup
constexpr int kReestSlices = 3;  // 256-dim slices of a pair's rows in flight while staging
template <int EX, bool WRAP>
__global__ void __launch_bounds__(256) reest_kernel(BatchArgs a, int scan_grid) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int D = a.D;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int l16 = lane & 15, lg = lane >> 4;
    const int S = ul_stride(D);  // lane-major residual rows (erq_device.h)
    float* uw = reinterpret_cast<float*>(smem) + (size_t)wid * 16 * S;
    // above 4 chunks (WRAP): each wave's digits of the staged pair, for the wrap correction
    constexpr bool wrap = WRAP;
    const size_t qv_bytes = wrap ? (size_t)a.nchunk * 1024 : 0;
    int16_t* qvw = reinterpret_cast<int16_t*>(smem + (size_t)256 * S + (size_t)wid * qv_bytes);
    uint32_t* pre = reinterpret_cast<uint32_t*>(smem + (size_t)256 * S + 4 * qv_bytes);
    __shared__ uint32_t wsum[4];
    // XCD-aware: workgroup b runs on XCD b % 8, like scan workgroup g on XCD g % 8; the
    // workgroups of one XCD share the slices of that XCD's scan workgroups (clusters of one
    // contiguous range, whose centroid rows then stay in that XCD's L2).  ns slices
    // x, x + xs, x + 2 xs, ...; their exclusive count prefix goes to pre[0..ns].
    const bool xcd = (gridDim.x & 7) == 0 && (scan_grid & 7) == 0;
    const int x = xcd ? (int)(blockIdx.x & 7) : 0, xs = xcd ? 8 : 1;
    const int ns = scan_grid / xs;
    const int per_t = (ns + 255) / 256;
    uint32_t own = 0;
    for (int i = tid * per_t; i < min(ns, (tid + 1) * per_t); i++) own += a.surv_count[x + xs * i];
    uint32_t incl = own;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    uint32_t base = incl - own;
    for (int w = 0; w < wid; w++) base += wsum[w];
    for (int i = tid * per_t; i < min(ns, (tid + 1) * per_t); i++) {
        pre[i] = base;
        base += a.surv_count[x + xs * i];
    }
    if (tid == 255) pre[ns] = base;
    __syncthreads();
    const size_t total = pre[ns];
    const size_t slice_cap = a.surv_cap / scan_grid;
    const size_t nw = (size_t)(gridDim.x / xs) * 4;
    const size_t wv = (size_t)(blockIdx.x / xs) * 4 + wid;
    const size_t r_begin = total * wv / nw, r_end = total * (wv + 1) / nw;
    const float value = (float)(1.0 / (double)sqrtf((float)D));
    unsigned long long st_surv = 0, st_cand = 0;
    uint32_t staged = 0xFFFFFFFFu;
    QParam P{}, Pn{};
    uint32_t pn_e = 0xFFFFFFFFu;  // entry whose parameters are in Pn
    int sl = -1;  // slice of this group's record (records only move forward)
    auto load_rec = [&](size_t rr) {
        uint4 v = make_uint4(0xFFFFFFFFu, 0, 0, 0);
        if (rr < r_end) {
            if (sl < 0) {  // the last i with pre[i] <= rr
                int lo = 0, hi = ns - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (pre[mid] <= (uint32_t)rr) lo = mid; else hi = mid - 1;
                }
                sl = lo;
            }
            while (pre[sl + 1] <= (uint32_t)rr) sl++;
            v = a.surv[(size_t)(x + xs * sl) * slice_cap + (rr - pre[sl])];
        }
        return v;
    };
    int nfd, rem, nqw;  // shape of a lane's exl codes
    ex_chain_shape<EX>(D >> 4, nfd, rem, nqw);
    if (!wrap && nqw <= 4) {
        // Every lane's codes fit four 16-byte words: a step's code words are requested before
        // its pair's residual is staged, so the two waits overlap.  Records and the next
        // run's parameters are loaded one step ahead, as below.
        uint4 rec = load_rec(r_begin + lg);
        for (size_t r = r_begin; r < r_end;) {  // wave-uniform
            const bool inr = r + lg < r_end;
            const uint32_t e0 = __shfl(rec.x, 0);
            const uint64_t same = __ballot(inr && rec.x == e0);
            const int m = same == ~0ull ? 4 : __builtin_ctzll(~same) >> 4;
            const bool act = lg < m;
            uint4 cw[4];
            if (act) {
                const uint4* src =
                    reinterpret_cast<const uint4*>(a.exl + ((size_t)rec.y * 16 + l16) * a.exl_w);
#pragma unroll
                for (int j4 = 0; j4 < 4; j4++)
                    if (j4 < nqw) cw[j4] = src[j4];
            }
            const uint4 rec_next = load_rec(r + m + lg);
            const uint32_t e_next = m < 4 ? __shfl(rec.x, 16 * m) : 0xFFFFFFFFu;
            if (e0 != staged) {
                if (e0 != pn_e) P = reinterpret_cast<const QParam*>(a.qparams)[e0];
                else P = Pn;
                const float* qv = a.q_rot + (size_t)P.q * D;
                const float* cen = a.centroids + (size_t)P.cid * D;
                wave_sync();
                if (P.y > 1e-5f) {
                    stage_residual<kReestSlices>(qv, cen, P.fac, D, S, uw, lane);
                } else {
                    for (int d = lane; d < D; d += 64) uw[ul_index(d, S)] = value;
                }
                wave_sync();
                staged = e0;
            }
            if (e_next != 0xFFFFFFFFu) {  // the next run's first record is in this step
                Pn = reinterpret_cast<const QParam*>(a.qparams)[e_next];
                pn_e = e_next;
            }
            const uint32_t vi = act ? rec.y : 0u;
            const float on = act ? __uint_as_float(rec.w) : 0.f;
            const float ipq = act ? __int_as_float((int)rec.z) : 0.f;  // the scaled product
            float part = 0.f;
            if (act)
                ex_chain_words<EX>(reinterpret_cast<const float4*>(uw + l16 * S), cw, 0, nqw,
                                   nfd, rem, part);
            const float exd = reest_distance<EX>(a, P, part, act, ipq, on, vi);
            if (act && l16 == 0) {
                if (exd <= P.thr) {
                    add_candidate(a, P.q, exd, a.ids[vi]);
                    st_cand++;
                }
                st_surv++;
            }
            rec = rec_next;
            r += m;
        }
        if (a.stats) {
            wave_stat_add(&a.stats[2], st_surv);
            wave_stat_add(&a.stats[4], st_cand);
        }
        return;
    }
    // Otherwise each step's records, and the next run's parameters when that run starts
    // inside this step, are loaded one step ahead.
    uint4 rec = load_rec(r_begin + lg);
    for (size_t r = r_begin; r < r_end;) {  // wave-uniform
        const bool inr = r + lg < r_end;
        // the run of group 0's pair: the leading groups whose record has its entry
        const uint32_t e0 = __shfl(rec.x, 0);  // group 0 is always in range
        const uint64_t same = __ballot(inr && rec.x == e0);
        const int m = same == ~0ull ? 4 : __builtin_ctzll(~same) >> 4;
        const bool act = lg < m;
        const uint4 rec_next = load_rec(r + m + lg);
        const uint32_t e_next = m < 4 ? __shfl(rec.x, 16 * m) : 0xFFFFFFFFu;
        if (e0 != staged) {
            // the parameters carry the round's threshold (batch_scan_kernel wrote it)
            if (e0 != pn_e) P = reinterpret_cast<const QParam*>(a.qparams)[e0];
            else P = Pn;
            const float* qv = a.q_rot + (size_t)P.q * D;
            const float* cen = a.centroids + (size_t)P.cid * D;
            wave_sync();
            if (P.y > 1e-5f) {
                stage_residual<kReestSlices>(qv, cen, P.fac, D, S, uw, lane);
            } else {
                for (int d = lane; d < D; d += 64) uw[ul_index(d, S)] = value;
            }
            wave_sync();
            if constexpr (wrap) {
                store_qv(uw, S, D, a.nchunk, 1.0 / (double)P.delta, qvw, lane);
                wave_sync();
            }
            staged = e0;
        }
        if (e_next != 0xFFFFFFFFu) {  // the next run's first record is in this step
            Pn = reinterpret_cast<const QParam*>(a.qparams)[e_next];
            pn_e = e_next;
        }
        const uint32_t vi = act ? rec.y : (wrap ? a.vstart[P.cid] : 0u);
        const float on = act ? __uint_as_float(rec.w) : 0.f;
        const int I = act ? (int)rec.z : 0;  // the integer (wrap) or the scaled product
        float ipq = wrap ? (float)I * P.delta : __int_as_float(I);
        bool keep = act;
        if constexpr (wrap)
            keep = wrap_retest(a, P, qvw, (int)(vi - a.vstart[P.cid]), I, on, l16, ipq) && act;
        const float exd = reestimate<EX>(a, P, uw + l16 * S,
                                         a.exl + ((size_t)vi * 16 + l16) * a.exl_w, keep, ipq,
                                         on, vi);
        if (keep && l16 == 0) {
            if (exd <= P.thr) {
                add_candidate(a, P.q, exd, a.ids[vi]);
                st_cand++;
            }
            st_surv++;
        }
        rec = rec_next;
        r += m;
    }
    if (a.stats) {
        wave_stat_add(&a.stats[2], st_surv);
        wave_stat_add(&a.stats[4], st_cand);
    }
}

// ---------------------------------------------------------------- merge
// Per query: pool (k sorted, FLT_MAX-padded) + candidates -> the k smallest by (dist, id).
// The k-th smallest distance key is found by an 8-bit radix select over both lists read
// from global memory; entries up to it are gathered in LDS and bitonic-sorted.
constexpr int kMergeBuf = 4096;

__device__ __forceinline__ void merge_entry(int i, int k, int q, const float* pool_d,
                                            const uint32_t* pool_i, const float* cand_d,
                                            const uint32_t* cand_i, int cap, float& d,
                                            uint32_t& id) {
    if (i < k) {
        d = pool_d[(size_t)q * k + i];
        id = d < FLT_MAX ? pool_i[(size_t)q * k + i] : 0xFFFFFFFFu;
    } else {
        d = cand_d[(size_t)q * cap + (i - k)];
        id = cand_i[(size_t)q * cap + (i - k)];
    }
}

// A query whose candidates did not all fit (more than `cap`, or more than kMergeBuf entries
// at or below the k-th key) keeps its pool unchanged and goes to the round's redo list: the
// round-mode scan then rescans the round for it with an LDS pool, exactly.
__device__ __forceinline__ void merge_redo(int q, uint32_t* cand_cnt, uint32_t* redo,
                                           unsigned long long* stats) {
    const uint32_t at = atomicAdd(&redo[0], 1u);
    redo[1 + at] = (uint32_t)q;
    cand_cnt[q] = 0;
    if (stats) atomicAdd(&stats[5], 1ull);
}

// The general merge of one query by a workgroup (merge_fast_kernel hands it the queries it
// does not take): a radix select of the k-th key over pool and candidates, then a bitonic
// sort of the entries up to it.
__device__ void merge_query(int q, float* __restrict__ pool_d, uint32_t* __restrict__ pool_i,
                            int k, uint32_t* __restrict__ cand_cnt,
                            const float* __restrict__ cand_d, const uint32_t* __restrict__ cand_i,
                            int cap, uint32_t* __restrict__ redo, unsigned long long* stats,
                            float* sd, uint32_t* si, uint32_t* hist, uint32_t& sh_sel,
                            uint32_t& sh_rank, uint32_t& sh_cnt) {
    const int tid = threadIdx.x;
    const uint32_t raw = cand_cnt[q];
    if (raw > (uint32_t)cap) {  // uniform
        if (tid == 0) merge_redo(q, cand_cnt, redo, stats);
        return;
    }
    const int nc = (int)raw;
    if (nc == 0) {
        // no candidates: the pool (sorted by distance) is already in (distance, id) order
        // unless two entries share a distance (a pool from the per-query replay puts the
        // later-inserted of equal distances first); then it is re-sorted like any merge
        if (tid == 0) sh_cnt = 0;
        __syncthreads();
        for (int i = tid; i + 1 < k; i += 256) {
            const float a = pool_d[(size_t)q * k + i];
            if (a < FLT_MAX && a == pool_d[(size_t)q * k + i + 1]) sh_cnt = 1;
        }
        __syncthreads();
        if (sh_cnt == 0) return;
        __syncthreads();
    }
    const int n = k + nc;
    // radix select of the k-th smallest distance key
    uint32_t prefix = 0, pmask = 0;
    if (tid == 0) {
        sh_rank = (uint32_t)min(k, n);
        sh_cnt = 0;
    }
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = tid; i < 256; i += 256) hist[i] = 0;
        __syncthreads();
        for (int i = tid; i < n; i += 256) {
            float d;
            uint32_t id;
            merge_entry(i, k, q, pool_d, pool_i, cand_d, cand_i, cap, d, id);
            uint32_t kv = float_key(d);
            if ((kv & pmask) == prefix) atomicAdd(&hist[(kv >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid < 64) {  // the digit holding rank sh_rank: four bins per lane, shuffle scan
            const uint32_t h0 = hist[4 * tid], h1 = hist[4 * tid + 1], h2 = hist[4 * tid + 2],
                           h3 = hist[4 * tid + 3];
            const uint32_t own = h0 + h1 + h2 + h3;
            uint32_t incl = own;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t v = __shfl_up(incl, o);
                if (tid >= o) incl += v;
            }
            const uint32_t r = sh_rank, excl = incl - own;
            if ((excl < r && r <= incl) || (tid == 63 && r > incl)) {
                uint32_t cum = excl, sel = 4 * tid + 3;
                const uint32_t hs[4] = {h0, h1, h2, h3};
                for (int j = 0; j < 4; j++) {
                    if (cum + hs[j] >= r) {
                        sel = 4 * tid + j;
                        break;
                    }
                    cum += hs[j];
                }
                sh_sel = sel;
                sh_rank = r - cum;
            }
        }
        __syncthreads();
        prefix |= sh_sel << shift;
        pmask |= 255u << shift;
        __syncthreads();
    }
    for (int i = tid; i < n; i += 256) {
        float d;
        uint32_t id;
        merge_entry(i, k, q, pool_d, pool_i, cand_d, cand_i, cap, d, id);
        if (float_key(d) <= prefix) {
            uint32_t slot = atomicAdd(&sh_cnt, 1u);
            if (slot < (uint32_t)kMergeBuf) {
                sd[slot] = d;
                si[slot] = id;
            }
        }
    }
    __syncthreads();
    if (sh_cnt > (uint32_t)kMergeBuf) {  // uniform: ties at the k-th key beyond the buffer
        if (tid == 0) merge_redo(q, cand_cnt, redo, stats);
        return;
    }
    const int m = (int)sh_cnt;
    int P2 = 1;
    while (P2 < max(m, k)) P2 <<= 1;
    for (int i = m + tid; i < P2; i += 256) {
        sd[i] = FLT_MAX;
        si[i] = 0xFFFFFFFFu;
    }
    __syncthreads();
    for (int size = 2; size <= P2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = tid; t < (P2 >> 1); t += 256) {
                int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
                bool up = (lo & size) == 0;
                float dl = sd[lo], dh = sd[hi];
                uint32_t il = si[lo], ih = si[hi];
                bool lt_hl = dh < dl || (dh == dl && ih < il);
                bool lt_lh = dl < dh || (dl == dh && il < ih);
                if (up ? lt_hl : lt_lh) {
                    sd[lo] = dh;
                    sd[hi] = dl;
                    si[lo] = ih;
                    si[hi] = il;
                }
            }
            __syncthreads();
        }
    for (int i = tid; i < k; i += 256) {
        pool_d[(size_t)q * k + i] = sd[i];
        pool_i[(size_t)q * k + i] = sd[i] < FLT_MAX ? si[i] : 0u;
    }
    if (tid == 0) cand_cnt[q] = 0;
}

// merge_query over the queries of list (list[0] = count, then the queries), a fixed grid.
__global__ void __launch_bounds__(256) merge_list_kernel(float* __restrict__ pool_d,
                                                         uint32_t* __restrict__ pool_i, int k,
                                                         uint32_t* __restrict__ cand_cnt,
                                                         const float* __restrict__ cand_d,
                                                         const uint32_t* __restrict__ cand_i,
                                                         int cap, const uint32_t* __restrict__ list,
                                                         uint32_t* __restrict__ redo,
                                                         unsigned long long* stats) {
    __shared__ float sd[kMergeBuf];
    __shared__ uint32_t si[kMergeBuf];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sh_sel, sh_rank, sh_cnt;
    const uint32_t n = list[0];
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {  // uniform
        merge_query((int)list[1 + i], pool_d, pool_i, k, cand_cnt, cand_d, cand_i, cap, redo,
                    stats, sd, si, hist, sh_sel, sh_rank, sh_cnt);
        __syncthreads();
    }
}

// (d, id) order of the merges.
__device__ __forceinline__ bool did_less(float da, uint32_t ia, float db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

constexpr int kFastCand = 64;  // candidates a one-wave merge takes

// One wave per query: with at most kFastCand candidates and a pool in (distance, id) order
// (no two finite entries of equal distance), the candidates are bitonic-sorted across the
// wave's lanes and both sorted lists are merged by rank: a candidate ends at its index plus
// the pool entries before it, a pool entry at its index plus the candidates before it;
// positions below k are the new pool -- the result of merge_query.  Any other query goes
// to `slow` for merge_list_kernel (or to the redo list when its candidates overflowed).
__global__ void __launch_bounds__(256) merge_fast_kernel(float* __restrict__ pool_d,
                                                         uint32_t* __restrict__ pool_i, int k,
                                                         int nq, uint32_t* __restrict__ cand_cnt,
                                                         const float* __restrict__ cand_d,
                                                         const uint32_t* __restrict__ cand_i,
                                                         int cap, uint32_t* __restrict__ slow,
                                                         uint32_t* __restrict__ redo,
                                                         unsigned long long* stats) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int q = blockIdx.x * 4 + wid;
    if (q >= nq) return;  // wave-uniform; no workgroup barrier below
    const int kcap = (k + 63) & ~63;
    float* pd = reinterpret_cast<float*>(smem) + (size_t)wid * (2 * kcap + 2 * kFastCand);
    uint32_t* pi = reinterpret_cast<uint32_t*>(pd + kcap);
    float* cd = pd + 2 * kcap;
    uint32_t* ci = reinterpret_cast<uint32_t*>(cd + kFastCand);
    const uint32_t raw = cand_cnt[q];
    if (raw > (uint32_t)cap) {
        if (lane == 0) merge_redo(q, cand_cnt, redo, stats);
        return;
    }
    const int nc = (int)raw;
    // the pool: ids of FLT_MAX padding read as 0xFFFFFFFF (last in (d, id) order)
    bool tie = false;
    for (int i = lane; i < kcap; i += 64) {
        const float d = i < k ? pool_d[(size_t)q * k + i] : FLT_MAX;
        pd[i] = d;
        pi[i] = d < FLT_MAX ? pool_i[(size_t)q * k + i] : 0xFFFFFFFFu;
    }
    wave_sync();
    for (int i = lane; i + 1 < k; i += 64) tie = tie || (pd[i] < FLT_MAX && pd[i] == pd[i + 1]);
    if (nc == 0 && !ballot(tie)) return;  // already in (distance, id) order
    if (nc > kFastCand || ballot(tie)) {
        if (lane == 0) slow[1 + atomicAdd(&slow[0], 1u)] = (uint32_t)q;
        return;
    }
    float d = FLT_MAX;
    uint32_t id = 0xFFFFFFFFu;
    if (lane < nc) {
        d = cand_d[(size_t)q * cap + lane];
        id = cand_i[(size_t)q * cap + lane];
    }
    // bitonic sort of the 64 lanes' (d, id), ascending in lane order
    for (int size = 2; size <= 64; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const float od = __shfl_xor(d, stride);
            const uint32_t oi = __shfl_xor(id, stride);
            const bool up = (lane & size) == 0, low = (lane & stride) == 0;
            const bool other_less = did_less(od, oi, d, id);
            // the lower lane of a pair keeps the smaller entry when ascending
            if ((low == up) ? other_less : !other_less && !(od == d && oi == id)) {
                d = od;
                id = oi;
            }
        }
    cd[lane] = d;
    ci[lane] = id;
    wave_sync();
    // candidate lane (< nc): pool entries before it
    int at = k;
    if (lane < nc) {
        int lo = 0, hi = k;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (did_less(pd[mid], pi[mid], d, id)) lo = mid + 1; else hi = mid;
        }
        at = lane + lo;
    }
    // pool entries: candidates before each
    float od[16];
    uint32_t oi[16];
    int opos[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int i = lane + 64 * r;
        opos[r] = k;
        if (i < k) {
            od[r] = pd[i];
            oi[r] = pi[i];
            int lo = 0, hi = nc;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (did_less(cd[mid], ci[mid], od[r], oi[r])) lo = mid + 1; else hi = mid;
            }
            opos[r] = i + lo;
        }
    }
    float* gd = pool_d + (size_t)q * k;
    uint32_t* gi = pool_i + (size_t)q * k;
#pragma unroll
    for (int r = 0; r < 16; r++)
        if (opos[r] < k) {
            gd[opos[r]] = od[r];
            gi[opos[r]] = od[r] < FLT_MAX ? oi[r] : 0u;
        }
    if (at < k) {
        gd[at] = d;
        gi[at] = d < FLT_MAX ? id : 0u;
    }
    if (lane == 0) cand_cnt[q] = 0;
}

}  // namespace

int bp_nchunk(int D) {
    const int c = (D + 511) / 512;
    return c == 5 ? 6 : (c == 7 ? 8 : c);
}

hipError_t launch_relayout(const uint8_t* short_data, size_t nblocks, int D, int factors,
                           uint8_t* bp, hipStream_t s) {
    if (nblocks == 0) return hipSuccess;
    const int nchunk = bp_nchunk(D);
    size_t total = nblocks * (size_t)nchunk * 32 * 64;
    int blocks = (int)std::min<size_t>((total + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(relayout_kernel, dim3(blocks), dim3(256), 0, s, short_data, nblocks, D,
                       nchunk, 4 * (size_t)D + 128 * (size_t)factors, bp);
    return hipGetLastError();
}

size_t batch_lds_bytes(int D) {  // main pass
    return (size_t)bp_nchunk(D) * 16384 + (size_t)kItemGroup * kMainSurv * 12;
}
size_t batch_inline_lds_bytes(int D) {
    const size_t nch = (size_t)bp_nchunk(D);
    return std::max(nch * 8192, (size_t)256 * ul_stride(D)) + (nch > 4 ? 8 * nch * 512 : 0) +
           (size_t)kGroup * kQSurv * 12;
}

size_t batch_items_bound(size_t pairs, size_t K, size_t N, size_t max_csize, int item_tiles) {
    const size_t tiles_per_item = (size_t)item_tiles * 16;
    const size_t max_split = (max_csize + tiles_per_item - 1) / tiles_per_item;
    return (pairs / kItemGroup + 1) * max_split + N / tiles_per_item + K + 8;
}

size_t prep_wave_bytes(int D) { return 4 * 16 * (size_t)ul_stride(D); }

hipError_t launch_exl_relayout(const uint8_t* long_code, size_t N, int D, int EX, uint32_t* exl,
                               hipStream_t s) {
    if (N == 0) return hipSuccess;
    const int W = exl_words(D, EX);
    const int blocks = (int)std::min<size_t>((N * 16 + 255) / 256, 1 << 16);
    switch (EX) {
    case 2: hipLaunchKernelGGL(exl_relayout_kernel<2>, dim3(blocks), dim3(256), 0, s, long_code, N, D, W, exl); break;
    case 3: hipLaunchKernelGGL(exl_relayout_kernel<3>, dim3(blocks), dim3(256), 0, s, long_code, N, D, W, exl); break;
    case 4: hipLaunchKernelGGL(exl_relayout_kernel<4>, dim3(blocks), dim3(256), 0, s, long_code, N, D, W, exl); break;
    case 6: hipLaunchKernelGGL(exl_relayout_kernel<6>, dim3(blocks), dim3(256), 0, s, long_code, N, D, W, exl); break;
    case 7: hipLaunchKernelGGL(exl_relayout_kernel<7>, dim3(blocks), dim3(256), 0, s, long_code, N, D, W, exl); break;
    case 8: hipLaunchKernelGGL(exl_relayout_kernel<8>, dim3(blocks), dim3(256), 0, s, long_code, N, D, W, exl); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_zero(uint32_t* p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

template <typename K>
static hipError_t set_lds(K kernel, size_t bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <int EX, int NCH>
static hipError_t launch_batch_scan_nch(const BatchArgs& ba, int grid, hipStream_t s) {
    const size_t lds = batch_lds_bytes(ba.D), ilds = batch_inline_lds_bytes(ba.D);
    hipError_t e = hipSuccess;
    if (lds + 2048 > 65536)  // dynamic plus static LDS beyond the 64 KB default
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&batch_scan_kernel<EX, NCH>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess && ilds + 2048 > 65536)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&batch_inline_kernel<EX, NCH>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)ilds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((batch_scan_kernel<EX, NCH>), dim3(grid), dim3(kThreads), lds, s, ba);
    // the queued item halves (none unless a list or a workgroup's slice filled up)
    hipLaunchKernelGGL((batch_inline_kernel<EX, NCH>), dim3(grid), dim3(kThreads), ilds, s, ba);
    return hipGetLastError();
}

template <int EX>
static hipError_t launch_batch_scan_ex(const BatchArgs& ba, int grid, hipStream_t s) {
    switch (ba.nchunk) {
    case 1: return launch_batch_scan_nch<EX, 1>(ba, grid, s);
    case 2: return launch_batch_scan_nch<EX, 2>(ba, grid, s);
    case 3: return launch_batch_scan_nch<EX, 3>(ba, grid, s);
    case 4: return launch_batch_scan_nch<EX, 4>(ba, grid, s);
    case 6: return launch_batch_scan_nch<EX, 6>(ba, grid, s);
    case 8: return launch_batch_scan_nch<EX, 8>(ba, grid, s);
    default: return hipErrorInvalidValue;
    }
}

static hipError_t launch_batch_scan(const BatchArgs& ba, int grid, hipStream_t s) {
    switch (ba.EX) {
    case 2: return launch_batch_scan_ex<2>(ba, grid, s);
    case 3: return launch_batch_scan_ex<3>(ba, grid, s);
    case 4: return launch_batch_scan_ex<4>(ba, grid, s);
    case 6: return launch_batch_scan_ex<6>(ba, grid, s);
    case 7: return launch_batch_scan_ex<7>(ba, grid, s);
    case 8: return launch_batch_scan_ex<8>(ba, grid, s);
    default: return hipErrorInvalidValue;
    }
}

template <bool WRAP>
static hipError_t launch_reest_w(const BatchArgs& ba, int scan_grid, hipStream_t s) {
    // residuals, digits (wrap correction above 4 chunks), slice prefix
    const size_t lds = (size_t)256 * ul_stride(ba.D) +
                       (WRAP ? 4 * (size_t)ba.nchunk * 1024 : 0) + 4 * ((size_t)scan_grid + 1);
    if (lds > 65536) {  // above the 64 KB default (large D): raise the kernel's limit
        hipError_t e = hipSuccess;
        switch (ba.EX) {
        case 2: e = set_lds(reest_kernel<2, WRAP>, lds); break;
        case 3: e = set_lds(reest_kernel<3, WRAP>, lds); break;
        case 4: e = set_lds(reest_kernel<4, WRAP>, lds); break;
        case 6: e = set_lds(reest_kernel<6, WRAP>, lds); break;
        case 7: e = set_lds(reest_kernel<7, WRAP>, lds); break;
        case 8: e = set_lds(reest_kernel<8, WRAP>, lds); break;
        default: return hipErrorInvalidValue;
        }
        if (e != hipSuccess) return e;
    }
    // as many workgroups as stay resident together (bound by LDS or registers), a multiple of
    // 8: each takes a fixed share of the records, so a second partial wave of workgroups
    // would be a tail
    int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, 163840 / (lds + 1024)));
    {
        int occ = 0;
        hipError_t e = hipSuccess;
        switch (ba.EX) {
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reest_kernel<2, WRAP>, 256, lds); break;
        case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reest_kernel<3, WRAP>, 256, lds); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reest_kernel<4, WRAP>, 256, lds); break;
        case 6: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reest_kernel<6, WRAP>, 256, lds); break;
        case 7: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reest_kernel<7, WRAP>, 256, lds); break;
        case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reest_kernel<8, WRAP>, 256, lds); break;
        default: return hipErrorInvalidValue;
        }
        if (e != hipSuccess) return e;
        if (occ > 0) per_cu = std::min(per_cu, occ);
    }
    const dim3 grid(ba.num_cu > 0 ? (unsigned)(per_cu * ((ba.num_cu + 7) / 8 * 8))
                                  : (unsigned)(scan_grid * kReestWG));
    switch (ba.EX) {
    case 2: hipLaunchKernelGGL((reest_kernel<2, WRAP>), grid, dim3(256), lds, s, ba, scan_grid); break;
    case 3: hipLaunchKernelGGL((reest_kernel<3, WRAP>), grid, dim3(256), lds, s, ba, scan_grid); break;
    case 4: hipLaunchKernelGGL((reest_kernel<4, WRAP>), grid, dim3(256), lds, s, ba, scan_grid); break;
    case 6: hipLaunchKernelGGL((reest_kernel<6, WRAP>), grid, dim3(256), lds, s, ba, scan_grid); break;
    case 7: hipLaunchKernelGGL((reest_kernel<7, WRAP>), grid, dim3(256), lds, s, ba, scan_grid); break;
    case 8: hipLaunchKernelGGL((reest_kernel<8, WRAP>), grid, dim3(256), lds, s, ba, scan_grid); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

static hipError_t launch_reest(const BatchArgs& ba, int scan_grid, hipStream_t s) {
    return ba.nchunk > 4 ? launch_reest_w<true>(ba, scan_grid, s)
                         : launch_reest_w<false>(ba, scan_grid, s);
}

hipError_t launch_prep_all(const BatchArgs& ba, const RoundBufs& rb, int R, int r_end,
                           int prep_grid, hipStream_t s) {
    if (prep_grid < 8 || prep_grid % 8) return hipErrorInvalidValue;
    const int nq = ba.nq, K = rb.K;
    if (R < 1 || R > kMaxRounds) return hipErrorInvalidValue;
    hipError_t e = launch_zero(rb.counts, (size_t)K * R, s);
    if (e != hipSuccess) return e;
    dim3 qg((nq + 3) / 4), qb(256);  // a wave per query
    const int nb = K * R, nblk = (nb + 255) / 256;
    const size_t priv_lds = 4 * (size_t)nb;
    const bool priv = rb.wgbase && priv_lds <= kCountLds;
    if (priv) {
        e = set_lds(count_priv_kernel, priv_lds);
        if (e == hipSuccess) e = set_lds(scatter_priv_kernel, priv_lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(count_priv_kernel, dim3(kCountGroups), dim3(1024), priv_lds, s,
                           ba.probe_off, rb.probe_ids, nq, (int)kFirstRoundRank, r_end, K, nb,
                           rb.counts, rb.wgbase);
    } else {
        hipLaunchKernelGGL(count_kernel, qg, qb, 0, s, ba.probe_off, rb.probe_ids, nq,
                           (int)kFirstRoundRank, r_end, K, rb.counts);
    }
    hipLaunchKernelGGL(bin_scan_kernel, dim3(nblk), dim3(256), 0, s, rb.counts, ba.csize, K, nb,
                       ba.item_tiles, rb.local, rb.btot);
    hipLaunchKernelGGL(bin_total_kernel, dim3(1), dim3(256), 0, s, rb.btot, nblk, rb.local, K, R,
                       ba.n_items);
    hipLaunchKernelGGL(bin_items_kernel, dim3((nb + 3) / 4), dim3(256), 0, s, rb.counts, ba.csize,
                       K, nb,
                       ba.item_tiles, rb.local, rb.btot, rb.cursor, (uint32_t*)ba.items);
    if (priv) {
        hipLaunchKernelGGL(scatter_priv_kernel, dim3(kCountGroups), dim3(1024), priv_lds, s,
                           ba.probe_off, rb.probe_ids, nq, (int)kFirstRoundRank, r_end, K, nb,
                           rb.cursor, rb.wgbase, (uint32_t*)ba.entries);
    } else {
        hipLaunchKernelGGL(scatter_kernel, qg, qb, 0, s, ba.probe_off, rb.probe_ids, nq,
                           (int)kFirstRoundRank, r_end, K, rb.cursor, (uint32_t*)ba.entries);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // every entry of every round at once: ba.n_items points at the totals (rinfo + 4 R)
    BatchArgs all = ba;
    all.n_items = ba.n_items + 4 * R;
    const int wpb = 4;
    const size_t pw = prep_wave_bytes(ba.D);
    if (pw * wpb > 65536) {
        e = set_lds(prep_kernel, pw * wpb);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(prep_kernel, dim3((unsigned)prep_grid), dim3(64 * wpb), pw * wpb, s, all,
                       (int)pw);
    return hipGetLastError();
}

hipError_t launch_round_stage(const BatchArgs& ba, const RoundBufs& rb, int grid, int stage,
                              hipStream_t s) {
    switch (stage) {
    case 0: return launch_batch_scan(ba, grid, s);
    case 1: return launch_reest(ba, grid, s);
    case 2: {
        const size_t kcap = (size_t)((ba.k + 63) & ~63);
        const size_t lds = 4 * (2 * kcap + 2 * kFastCand) * 4;
        hipLaunchKernelGGL(merge_fast_kernel, dim3((ba.nq + 3) / 4), dim3(256), lds, s, ba.pool_d,
                           rb.pool_i, ba.k, ba.nq, ba.cand_cnt, ba.cand_d, ba.cand_i,
                           ba.cand_cap, ba.slow, ba.redo, ba.stats);
        const unsigned lgrid = (unsigned)std::max(8, std::min(4 * std::max(1, ba.num_cu), ba.nq));
        hipLaunchKernelGGL(merge_list_kernel, dim3(lgrid), dim3(256), 0, s, ba.pool_d, rb.pool_i,
                           ba.k, ba.cand_cnt, ba.cand_d, ba.cand_i, ba.cand_cap, ba.slow, ba.redo,
                           ba.stats);
        return hipGetLastError();
    }
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_round_scan(const BatchArgs& ba, const RoundBufs& rb, int grid, hipStream_t s) {
    for (int stage = 0; stage < 3; stage++) {
        hipError_t e = launch_round_stage(ba, rb, grid, stage, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace erq
