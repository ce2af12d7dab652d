// erq_search.hip -- plain (per-query) ExRaBitQ HA scan on gfx950.
//
// One wavefront owns one query and replays the reference's sequential semantics
// exactly: clusters in the given (ascending-distance) order, blocks in storage order,
// each block's candidate mask taken against the k-th distance at block start, and
// candidates inserted in ascending lane order (HASearcher.hpp:82-112, :187-237;
// IVF.hpp:313-323).  Per cluster the wave builds, in its LDS slice:
//   unit_q[D]      (q - c)/||q - c||                 normalize_query16, space.hpp:749-778
//   lut[M][16]     u16 shifted subset sums           pack_high_acc_LUT, HASearcher.hpp:128-185
// then scans 32-vector blocks straight from the HA layout in HBM (4D code bytes +
// 32 float factors per block) and re-ranks survivors from their EX-bit long codes.
#include "erq_batch.h"
#include "erq_device.h"
#include "erq_internal.h"
#include "erq_pool.h"

namespace erq {

namespace {

// The query a wave scans: slot blockIdx.x * waves + wid of the order when there is one
// (queries by scan cost, heaviest first: workgroups start in index order, so the light
// queries fill the last, partial generation of waves), else the query of that index.
// -1: none.
__device__ __forceinline__ int query_of_wave(const ScanArgs& a, int wid) {
    const int slot = blockIdx.x * (blockDim.x >> 6) + wid;
    if (slot >= a.nq) return -1;
    return a.qorder ? (int)a.qorder[slot] : slot;
}

// The order's bin of query q: its scan cost -- the 32-vector blocks of its probed clusters
// (the first max_probes of them when max_probes > 0) -- shifted right by `shift`, taken
// from K - 1 so that the heaviest queries come first.
__device__ __forceinline__ uint32_t order_bin(const uint32_t* off, const uint32_t* pid,
                                              const uint32_t* csize, int q, int K, int max_probes,
                                              int shift) {
    const uint32_t s = off[q];
    uint32_t e = off[q + 1];
    if (max_probes > 0) e = min(e, s + (uint32_t)max_probes);
    uint32_t blocks = 0;
    for (uint32_t p = s; p < e; p++) {
        const uint32_t c = pid[p];
        if (c < (uint32_t)K) blocks += (csize[c] + 31) >> 5;
    }
    return (uint32_t)K - 1 - min(blocks >> shift, (uint32_t)K - 1);
}

__global__ void order_zero_kernel(uint32_t* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void order_count_kernel(const uint32_t* __restrict__ off, const uint32_t* __restrict__ pid,
                                   const uint32_t* __restrict__ csize, int nq, int K,
                                   int max_probes, int shift, uint32_t* __restrict__ counts) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq) atomicAdd(&counts[order_bin(off, pid, csize, q, K, max_probes, shift)], 1u);
}

// Single workgroup: counts[0..n) -> exclusive prefix sums, in place.
__global__ void __launch_bounds__(1024) order_scan_kernel(uint32_t* counts, int n) {
    __shared__ uint32_t sc[1024];
    __shared__ uint32_t carry;
    const int t = threadIdx.x;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const uint32_t v = base + t < n ? counts[base + t] : 0;
        sc[t] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const uint32_t u = t >= o ? sc[t - o] : 0;
            __syncthreads();
            sc[t] += u;
            __syncthreads();
        }
        if (base + t < n) counts[base + t] = carry + sc[t] - v;
        __syncthreads();
        if (t == 1023) carry += sc[1023];
        __syncthreads();
    }
}

__global__ void order_scatter_kernel(const uint32_t* __restrict__ off,
                                     const uint32_t* __restrict__ pid,
                                     const uint32_t* __restrict__ csize, int nq, int K,
                                     int max_probes, int shift, uint32_t* __restrict__ cursor,
                                     uint32_t* __restrict__ order) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq)
        order[atomicAdd(&cursor[order_bin(off, pid, csize, q, K, max_probes, shift)], 1u)] =
            (uint32_t)q;
}


template <int EX, bool WRAP>
__global__ void __launch_bounds__(256) scan_kernel(ScanArgs a, int per_wave_bytes) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int qi = query_of_wave(a, wid);
    if (qi < 0) return;
    const int D = a.D, M = D >> 2, k = a.k;
    const int kcap = (k + 63) & ~63;
    unsigned char* base = smem + (size_t)wid * per_wave_bytes;
    WaveLds w;
    w.uq = reinterpret_cast<float*>(base);
    w.lut = reinterpret_cast<uint16_t*>(base + 4 * D);
    w.nd = nullptr;
    w.nid = nullptr;
    w.sums = reinterpret_cast<int*>(base + 12 * D);
    w.pd = reinterpret_cast<float*>(base + 12 * D + 128);
    w.pid = reinterpret_cast<uint32_t*>(w.pd + kcap);
    for (int i = lane; i < kcap; i += kWave) {
        w.pd[i] = FLT_MAX;
        w.pid[i] = 0;
    }
    int psize = 0;
    wave_sync();

    const float* q = a.q_rot + (size_t)qi * D;
    const size_t BB = 4 * (size_t)D + 128;
    const size_t CL = (size_t)D * EX / 8;
    const int R = 1 << EX;
    const float const_bound = 0.58f;  // FastScan.hpp:295
    unsigned long long st_blocks = 0, st_vec = 0, st_surv = 0;
    const int t4 = lane & 3, s16 = lane >> 2;
    const int g = lane >> 4, l16 = lane & 15;

    const uint32_t p0 = a.probe_off[qi];
    uint32_t p1 = a.probe_off[qi + 1];
    if (a.max_probes > 0 && p1 - p0 > (uint32_t)a.max_probes) p1 = p0 + (uint32_t)a.max_probes;
    for (uint32_t p = p0; p < p1; ++p) {
        const uint32_t cid = a.probe_ids[p];
        if (cid >= (uint32_t)a.K) continue;  // not a cluster of this index
        const float sqr_y = a.probe_sqr_y[p];
        const float y = sqrtf(sqr_y);  // HASearcher.hpp:85
        const float* c = a.centroids + (size_t)cid * D;

        // ---- normalize_query16: unit residual + reference-ordered sum ----
        float sumq;
        if (y > 1e-5f) {
            const float fac = (float)(1.0 / (double)y);
            for (int d = lane; d < D; d += kWave) w.uq[d] = (q[d] - c[d]) * fac;
            wave_sync();
            float part = 0.f;
            if (lane < 16)
                for (int d = lane; d < D; d += 16) part = part + w.uq[d];
            part = butterfly16(part);
            sumq = __shfl(part, 0);
        } else {
            const float value = (float)(1.0 / (double)sqrtf((float)D));
            for (int d = lane; d < D; d += kWave) w.uq[d] = value;
            wave_sync();
            sumq = (float)D * value;
        }

        // ---- high_acc_quantize16 + pack_high_acc_LUT ----
        float mx = 0.f, mn = 0.f;
        for (int d = lane; d < D; d += kWave) {
            float u = w.uq[d];
            mx = fmaxf(mx, u);
            mn = fminf(mn, u);
        }
        mx = wave_max(mx);
        mn = wave_min(mn);
        const float vmax = fmaxf(fabsf(mn), fabsf(mx));
        const float delta = vmax / 8191.0f;
        const double rdelta = 1.0 / (double)delta;
        int shift_part = 0;
        for (int m = lane; m < M; m += kWave) {
            int qv[4];
            for (int t = 0; t < 4; t++) qv[t] = quantize_digit(w.uq[4 * m + t], rdelta);
            // min over the 16 subset sums = sum of the negative terms
            const int vmin = min(0, qv[0]) + min(0, qv[1]) + min(0, qv[2]) + min(0, qv[3]);
            uint32_t wv[8];
            for (int j = 0; j < 16; j += 2) {
                // entry j: bit 3 <-> dim 4m, bit 0 <-> dim 4m+3 (HASearcher.hpp:130-158)
                int L0 = ((j & 8) ? qv[0] : 0) + ((j & 4) ? qv[1] : 0) + ((j & 2) ? qv[2] : 0) - vmin;
                int L1 = L0 + qv[3];
                wv[j >> 1] = (uint32_t)L0 | ((uint32_t)L1 << 16);
            }
            uint4* dst = reinterpret_cast<uint4*>(w.lut + m * 16);
            dst[0] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
            dst[1] = make_uint4(wv[4], wv[5], wv[6], wv[7]);
            shift_part += vmin;
        }
        const int shift = wave_sum_i(shift_part);
        wave_sync();

        // ---- block scan (accumulate_one_block_high_acc + scan_one_block) ----
        const int n = (int)a.csize[cid];
        const int nb = (n + 31) >> 5;
        const uint8_t* blk = a.short_data + (size_t)a.bstart[cid] * BB;
        const uint32_t v0 = a.vstart[cid];
        float distk = (psize == k) ? w.pd[k - 1] : FLT_MAX;
        const float c1 = (float)(0.5 * (double)sumq - (double)const_bound);
        const float c2 = (-5.0f * y) * a.one_over_sqrtD;
        const float qs = y * y;
        for (int b = 0; b < nb; ++b, blk += BB) {
            const int np = min(32, n - b * 32);
            // Lane (t4, s16) reads dword t4 (bytes 4t4..4t4+3, 8 vectors) of codebooks
            // s16, s16+16, ...; 64 lanes cover 256 contiguous code bytes per step.
            uint32_t acc[8], acch[8];
            for (int kk = 0; kk < 8; kk++) acc[kk] = acch[kk] = 0;
            const uint8_t* cbase = blk + t4 * 4;
            for (int i = 0; i < (M >> 4); ++i) {
                const int cb = s16 + 16 * i;
                const uint32_t wd =
                    *reinterpret_cast<const uint32_t*>(cbase + (cb >> 1) * 32 + (cb & 1) * 16);
                const uint16_t* L = w.lut + cb * 16;
                for (int kk = 0; kk < 4; kk++) {
                    const uint32_t byte = (wd >> (8 * kk)) & 0xFFu;
                    const uint32_t e0 = L[byte & 15u], e1 = L[byte >> 4];
                    acc[kk] += e0;
                    acc[4 + kk] += e1;
                    if (WRAP) {
                        acch[kk] += e0 >> 8;
                        acch[4 + kk] += e1 >> 8;
                    }
                }
            }
            if (WRAP) {
                // even and odd codebooks (lane bit 2) wrap separately, per byte plane
                for (int o = 8; o <= 32; o <<= 1)
                    for (int kk = 0; kk < 8; kk++) {
                        acc[kk] += __shfl_xor(acc[kk], o);
                        acch[kk] += __shfl_xor(acch[kk], o);
                    }
                for (int kk = 0; kk < 8; kk++) {
                    uint32_t lo = acc[kk] - (acch[kk] << 8);
                    acc[kk] = (lo & 0xFFFFu) + ((acch[kk] & 0xFFFFu) << 8);
                }
                for (int kk = 0; kk < 8; kk++) acc[kk] += __shfl_xor(acc[kk], 4);
            } else {
                for (int o = 4; o <= 32; o <<= 1)
                    for (int kk = 0; kk < 8; kk++) acc[kk] += __shfl_xor(acc[kk], o);
            }
            if (lane < 4)
                for (int kk = 0; kk < 4; kk++) {
                    int v = vec_of_byte(4 * lane + kk);
                    w.sums[v] = (int)acc[kk];
                    w.sums[v + 16] = (int)acc[4 + kk];
                }
            wave_sync();

            const float* onorm = reinterpret_cast<const float*>(blk + 4 * (size_t)D);
            bool cand = false;
            float ipq_l = 0.f;
            if (lane < 32) {
                const int I = w.sums[lane] + shift;
                ipq_l = (float)I * delta;
                const float on = onorm[lane];
                float t = ipq_l - c1;
                t = t * c2;
                t = t * on;
                t = t + qs;
                t = t + on * on;
                cand = (lane < np) && (t < distk);
            }
            uint64_t mask = ballot(cand);
            st_blocks++;
            st_vec += np;

            // ---- survivors, ascending lane order, 4 re-estimated at a time ----
            while (mask) {
                uint64_t mm = mask;
                for (int r = 0; r < g; r++) mm &= mm - 1;
                const int j = mm ? (__ffsll((unsigned long long)mm) - 1) : -1;
                const int jj = j < 0 ? 0 : j;
                const uint32_t vi = v0 + (uint32_t)(b * 32 + jj);
                float part = 0.f;
                if (j >= 0) {
                    const uint8_t* code = a.long_code + (size_t)vi * CL;
                    for (int ch = 0; ch < (D >> 6); ch++) {
                        float e[4];
                        decode4<EX>(code + ch * 8 * EX, l16, e);
                        for (int s4 = 0; s4 < 4; s4++)
                            part = fmaf(w.uq[64 * ch + 16 * s4 + l16], e[s4], part);
                    }
                }
                const float ip = butterfly16(part);
                const float ipq_j = __shfl(ipq_l, jj);
                float exd = 0.f;
                uint32_t idv = 0;
                if (j >= 0) {
                    // HASearcher.hpp:225-232 with C's float -> double promotions
                    const float on = onorm[jj];
                    const float sqr_x = on * on;
                    const float A = sqr_x + sqr_y;
                    const float B = a.ex_factor[vi] * y;
                    const float C = (float)R * ipq_j + ip;
                    const double G = (double)C - ((double)R - 0.5) * (double)sumq;
                    const double H = (double)B * G;
                    exd = (float)((double)A - H);
                    idv = a.ids[vi];
                }
                const int cnt = min(4, __popcll(mask));
                for (int gg = 0; gg < cnt; gg++) {
                    const float dd = __shfl(exd, 16 * gg);
                    const uint32_t ii = (uint32_t)__shfl((int)idv, 16 * gg);
                    pool_insert(w, psize, k, lane, dd, ii);
                }
                for (int r = 0; r < cnt; r++) mask &= mask - 1;
                st_surv += cnt;
            }
            distk = (psize == k) ? w.pd[k - 1] : FLT_MAX;
        }
    }
    uint32_t* oi = a.out_ids + (size_t)qi * k;
    float* od = a.out_dists ? a.out_dists + (size_t)qi * k : nullptr;
    for (int i = lane; i < k; i += kWave) {
        oi[i] = i < psize ? w.pid[i] : 0u;
        if (od) od[i] = i < psize ? w.pd[i] : FLT_MAX;
    }
    if (a.stats && lane == 0) {
        atomicAdd(&a.stats[0], st_blocks);
        atomicAdd(&a.stats[1], st_vec);
        atomicAdd(&a.stats[2], st_surv);
        atomicAdd(&a.stats[3], (unsigned long long)(p1 - p0));
    }
}

// The same replay with the FastScan integer taken from the bit-plane codes on
// v_mfma_i32_16x16x64_i8: A rows 0/1 hold the query's low/high int8 digits (q = 128*hi + lo),
// rows 2..15 are zero; B is bit-plane p of 16 vectors' code bytes (one v_and per dword);
// I = sum_p +-(acc_p >> p) is the exact inner product (LUT sum plus shift).  Above D = 2048
// the reference's u16 lanes can wrap: the vectors whose lower bound passes on the exact
// integer (a superset of the reference's, see lo_wraps) get the reference's integer from
// their low-byte half-sums before the mask is final.  Prep (normalize, quantize), lower
// bound, mask, survivor re-rank and pool are the LUT kernel's.
typedef int v4i __attribute__((ext_vector_type(4)));

// The re-estimated distance of a survivor in the reference's operation order
// (HASearcher.hpp:215-220, the double promotion as C performs it): on = |o| of the vector,
// Bb = ex_factor * y, ipq = scaled FastScan product, ip = <u, ex code>.
template <int EX>
__device__ __forceinline__ float ex_distance(float on, float sqr_y, float Bb, float ipq, float ip,
                                             float sumq) {
    constexpr int R = 1 << EX;
    const float sqr_x = on * on;
    const float Aa = sqr_x + sqr_y;
    const float Cc = (float)R * ipq + ip;
    const double G = (double)Cc - ((double)R - 0.5) * (double)sumq;
    const double H = (double)Bb * G;
    return (float)((double)Aa - H);
}

// (distance, id) order of the batched rounds' merge.
__device__ __forceinline__ bool did_less(float da, uint32_t ia, float db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

// ns new entries (nd, nid, ns <= 32, ids distinct and absent from the pool) merged into a
// pool sorted by (distance, id), keeping its k smallest: new entry t lands after the pool
// entries and the new entries below it, pool entry i moves up by the new entries below it.
// nd / nid are overwritten (sorted).
__device__ __forceinline__ void pool_merge_did(const WaveLds& w, int& psize, int k, int lane,
                                               float* nd, uint32_t* nid, int ns) {
    float dt = 0.f;
    uint32_t it = 0;
    int rt = 0, at = k;
    if (lane < ns) {
        dt = nd[lane];
        it = nid[lane];
        for (int u = 0; u < ns; u++) rt += did_less(nd[u], nid[u], dt, it) ? 1 : 0;
        int lo = 0, hi = psize;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (did_less(w.pd[mid], w.pid[mid], dt, it)) lo = mid + 1; else hi = mid;
        }
        at = lo + rt;
    }
    wave_sync();
    if (lane < ns) {
        nd[rt] = dt;
        nid[rt] = it;
    }
    wave_sync();
    for (int base = (psize - 1) & ~(kWave - 1); psize > 0 && base >= 0; base -= kWave) {
        const int i = base + lane;
        float vd = 0.f;
        uint32_t vid = 0;
        int dst = k;
        if (i < psize) {
            vd = w.pd[i];
            vid = w.pid[i];
            int lo = 0, hi = ns;  // new entries below (vd, vid)
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (did_less(nd[mid], nid[mid], vd, vid)) lo = mid + 1; else hi = mid;
            }
            dst = i + lo;
        }
        wave_sync();
        if (dst < k) {
            w.pd[dst] = vd;
            w.pid[dst] = vid;
        }
        wave_sync();
    }
    if (at < k) {
        w.pd[at] = dt;
        w.pid[at] = it;
    }
    wave_sync();
    psize = min(k, psize + ns);
}

enum { kModeReplay = 0, kModeRound = 1, kModeEstimate = 2 };
constexpr int kEstBlocks = 2;  // blocks per work item of the estimate mode
// 256-dim slices of q and c in flight while staging a pair's residual (fewer above 3 chunks,
// where the code registers leave no room for them)
template <int NCH> constexpr int kReplaySlices = NCH <= 3 ? 3 : 1;

// LDS of one wave of scan_mfma_kernel: the lane-major residual, the A operand and its zero
// twin, the digits for the wrap correction (NCH > 4), the block's integer sums, the pool and
// a block's new entries.
__host__ __device__ inline size_t mfma_wave_bytes(int D, int nchunk, int k) {
    const size_t kcap = (size_t)((k + 63) & ~63);
    return 4 * 16 * (size_t)ul_stride(D) + 2048 * (size_t)nchunk +
           (nchunk > 4 ? 1024 * (size_t)nchunk : 0) + 128 + 8 * kcap + 256;
}

template <int EX, int NCH, int MODE>
__device__ __forceinline__ void scan_query_mfma(const ScanArgs& a, int qi, unsigned char* base,
                                                int lane, int blo = 0, int bhi = 0x7FFFFFFF) {
    const int D = a.D, k = a.k;
    // estimate mode: qi is a (query, probe) pair; find its query
    const int pair = qi;
    if (MODE == kModeEstimate) {
        int lo = 0, hi = a.nq - 1;  // the last query whose probes start at or before pair
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.probe_off[mid] <= (uint32_t)pair) lo = mid; else hi = mid - 1;
        }
        qi = lo;
    }
    constexpr int nchunk = NCH;
    constexpr bool kWrap = NCH > 4;  // D > 2048
    const int kcap = (k + 63) & ~63;
    // the residual u, lane-major (u[16 i + l] at uq[l * S + i]) for the re-estimate's lanes
    const int S = ul_stride(D);
    float* uq = reinterpret_cast<float*>(base);
    const size_t ubytes = 4 * 16 * (size_t)S;
    int8_t* A = reinterpret_cast<int8_t*>(base + ubytes);                 // nchunk * 1 KB
    // nchunk * 1 KB of zeros laid out like A: lanes holding MFMA rows 2..15 read their
    // (zero) A operand from here at the same offsets, so every lane loads unconditionally
    int8_t* Z = A + 1024 * (size_t)nchunk;
    int16_t* qv = reinterpret_cast<int16_t*>(base + ubytes + 2048 * (size_t)nchunk);
    int* sums = reinterpret_cast<int*>(base + ubytes + 2048 * (size_t)nchunk +
                                       (kWrap ? 1024 * (size_t)nchunk : 0));
    WaveLds w;
    w.uq = uq;
    w.lut = nullptr;
    w.sums = sums;
    w.pd = reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(sums) + 128);
    w.pid = reinterpret_cast<uint32_t*>(w.pd + kcap);
    w.nd = reinterpret_cast<float*>(w.pid + kcap);
    w.nid = reinterpret_cast<uint32_t*>(w.nd + 32);
    for (int i = lane; i < nchunk * 64; i += kWave)
        reinterpret_cast<uint4*>(Z)[i] = make_uint4(0, 0, 0, 0);
    int psize = 0;
    float thr = FLT_MAX;  // round mode: the pool's k-th distance at round start
    if (MODE == kModeReplay) {
        for (int i = lane; i < kcap; i += kWave) {
            w.pd[i] = FLT_MAX;
            w.pid[i] = 0;
        }
        wave_sync();
    } else if (MODE == kModeRound) {
        // the pool at round start (sorted by distance, FLT_MAX-padded); equal distances are
        // put in id order, as the round's merge orders its pool
        const float* gd = a.out_dists + (size_t)qi * k;
        const uint32_t* gi = a.out_ids + (size_t)qi * k;
        int cnt = 0;
        for (int i = lane; i < kcap; i += kWave) {
            const float d = i < k ? gd[i] : FLT_MAX;
            w.pd[i] = d;
            w.pid[i] = d < FLT_MAX ? gi[i] : 0xFFFFFFFFu;
            cnt += d < FLT_MAX ? 1 : 0;
        }
        psize = wave_sum_i(cnt);
        wave_sync();
        constexpr int kMaxOwn = 16;  // k <= 1024
        float od[kMaxOwn];
        uint32_t oid[kMaxOwn];
        int opos[kMaxOwn];
#pragma unroll
        for (int r = 0; r < kMaxOwn; r++) {
            const int i = lane + kWave * r;
            opos[r] = -1;
            if (i < psize) {
                const float d = w.pd[i];
                const uint32_t id = w.pid[i];
                int lo = 0, hi = psize;  // first entry with distance d
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (w.pd[mid] < d) lo = mid + 1; else hi = mid;
                }
                int pos = lo;
                for (int j = lo; j < psize && w.pd[j] == d; j++) pos += w.pid[j] < id ? 1 : 0;
                od[r] = d;
                oid[r] = id;
                opos[r] = pos;
            }
        }
        wave_sync();
#pragma unroll
        for (int r = 0; r < kMaxOwn; r++)
            if (opos[r] >= 0) {
                w.pd[opos[r]] = od[r];
                w.pid[opos[r]] = oid[r];
            }
        wave_sync();
        thr = psize == k ? w.pd[k - 1] : FLT_MAX;
    }

    const float* q = a.q_rot + (size_t)qi * D;
    const size_t BB = 4 * (size_t)D + 128;
    const float const_bound = 0.58f;  // FastScan.hpp:295
    unsigned long long st_blocks = 0, st_vec = 0, st_surv = 0;
    const int g = lane >> 4, l16 = lane & 15;
    // A operand of this lane at (chunk, plane) = abase + (chunk * 8 + plane) * 128
    const int8_t* abase = l16 < 2 ? A + (g * 2 + l16) * 16 : Z + g * 32;

    uint32_t p0 = a.probe_off[qi];
    uint32_t p1 = a.probe_off[qi + 1];
    if (MODE == kModeReplay) {
        if (a.max_probes > 0 && p1 - p0 > (uint32_t)a.max_probes) p1 = p0 + (uint32_t)a.max_probes;
    } else if (MODE == kModeEstimate) {
        p0 = (uint32_t)pair;
        p1 = p0 + 1;
    } else {
        p1 = min(p1, p0 + (uint32_t)a.r1);
        p0 = p0 + (uint32_t)a.r0;
    }
    const uint32_t np_total = p1 > p0 ? p1 - p0 : 0;
    for (uint32_t p = p0; p < p1; ++p) {
        const uint32_t cid = a.probe_ids[p];
        if (cid >= (uint32_t)a.K) continue;
        const float sqr_y = a.probe_sqr_y[p];
        const float y = sqrtf(sqr_y);
        const float* c = a.centroids + (size_t)cid * D;
        const int n = (int)a.csize[cid];
        const int nb = (n + 31) >> 5;
        const uint32_t blk0 = a.bstart[cid];
        const uint32_t v0 = a.vstart[cid];
        // codes and norms of block b + 1 are loaded while block b is on the MFMA; block 0's
        // while the residual is prepared
        uint4 wn0[NCH], wn1[NCH];
        float on_n = 0.f;
        auto load_block = [&](uint32_t blk) {
#pragma unroll
            for (int ch = 0; ch < NCH; ch++) {
                const uint8_t* src = a.bp + (((size_t)blk * NCH + ch) * 32 + l16) * 64 + 16 * g;
                wn0[ch] = *reinterpret_cast<const uint4*>(src);
                wn1[ch] = *reinterpret_cast<const uint4*>(src + 16 * 64);
            }
            on_n = lane < 32 ? reinterpret_cast<const float*>(a.short_data + (size_t)blk * BB +
                                                              4 * (size_t)D)[lane]
                             : 0.f;
        };
        const int b_lo = min(blo, nb), b_hi = min(bhi, nb);
        if (b_lo < b_hi) load_block(blk0 + (uint32_t)b_lo);
        float fac, sumq, delta;
        prepare_residual<kReplaySlices<NCH>>(q, c, y, D, S, uq, lane, fac, sumq, delta);
        quantize_digits(uq, S, D, nchunk, 1.0 / (double)delta, reinterpret_cast<uint32_t*>(A),
                        lane, kWrap ? qv : nullptr);
        wave_sync();

        float distk = MODE == kModeReplay ? ((psize == k) ? w.pd[k - 1] : FLT_MAX)
                      : (MODE == kModeEstimate ? INFINITY : thr);
        const uint32_t est0 = MODE == kModeEstimate ? a.est_off[pair] : 0u;
        const float c1 = (float)(0.5 * (double)sumq - (double)const_bound);
        const float c2 = (-5.0f * y) * a.one_over_sqrtD;
        const float qs = y * y;
        for (int b = b_lo; b < b_hi; ++b) {
            const uint32_t blk = blk0 + (uint32_t)b;
            const int np = min(32, n - b * 32);
            uint4 wc0[NCH], wc1[NCH];
#pragma unroll
            for (int ch = 0; ch < NCH; ch++) {
                wc0[ch] = wn0[ch];
                wc1[ch] = wn1[ch];
            }
            const float on_l = on_n;  // lane < 32: norm of vector `lane` of this block
            if (b + 1 < b_hi) load_block(blk + 1);
            // plane by plane: the NCH-chunk MFMA chain of plane pp for both 16-vector tiles,
            // then folded into the running sums (plane 7 carries the sign: 0x80 is -128)
            int lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0;
#pragma unroll 2
            for (int pp = 0; pp < 8; pp++) {
                v4i acc0 = v4i{0, 0, 0, 0}, acc1 = v4i{0, 0, 0, 0};
                const uint32_t m = 0x01010101u << pp;
#pragma unroll
                for (int ch = 0; ch < NCH; ch++) {
                    const v4i av = *reinterpret_cast<const v4i*>(abase + (ch * 8 + pp) * 128);
                    const v4i b0 = v4i{(int)(wc0[ch].x & m), (int)(wc0[ch].y & m),
                                       (int)(wc0[ch].z & m), (int)(wc0[ch].w & m)};
                    const v4i b1 = v4i{(int)(wc1[ch].x & m), (int)(wc1[ch].y & m),
                                       (int)(wc1[ch].z & m), (int)(wc1[ch].w & m)};
                    acc0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b0, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b1, acc1, 0, 0, 0);
                }
                if (pp < 7) {
                    lo0 += acc0[0] >> pp;
                    hi0 += acc0[1] >> pp;
                    lo1 += acc1[0] >> pp;
                    hi1 += acc1[1] >> pp;
                } else {
                    lo0 -= acc0[0] >> 7;
                    hi0 -= acc0[1] >> 7;
                    lo1 -= acc1[0] >> 7;
                    hi1 -= acc1[1] >> 7;
                }
            }
            // rows 0 (low digits) and 1 (high digits) of column l16 live in lanes 0..15
            if (lane < 16) {
                sums[lane] = 128 * hi0 + lo0;
                sums[16 + lane] = 128 * hi1 + lo1;
            }
            wave_sync();
            bool cand = false;
            float ipq_l = 0.f, low_l = INFINITY;
            auto lower_test = [&]() {
                cand = false;
                if (lane < 32) {
                    ipq_l = (float)sums[lane] * delta;
                    const float on = on_l;
                    float t = ipq_l - c1;
                    t = t * c2;
                    t = t * on;
                    t = t + qs;
                    t = t + on * on;
                    low_l = lane < np ? t : INFINITY;
                    cand = (lane < np) && (t < distk);
                }
            };
            lower_test();
            uint64_t mask = ballot(cand);
            if (MODE == kModeEstimate) {
                if (lane < 32) a.est_lower[est0 + 32 * b + lane] = low_l;
                const float bm = wave_min(low_l);  // lanes past the block's vectors hold INFINITY
                if (lane == 0) a.est_bmin[(est0 >> 5) + b] = bm;
            }
            if constexpr (kWrap) {
                // the reference's integer for the vectors that pass on the exact one (four at
                // a time, one per 16-lane group), then the lower bound again
                if (mask) {
                    for (uint64_t mw = mask; mw;) {
                        uint64_t mm = mw;
                        for (int r = 0; r < g; r++) mm &= mm - 1;
                        const int j = mm ? (__ffsll((unsigned long long)mm) - 1) : -1;
                        const int jj = j < 0 ? 0 : j;
                        const int wr = lo_wraps(qv, NCH, l16, [&](int ch) {
                            return *reinterpret_cast<const uint32_t*>(
                                a.bp + (((size_t)blk * NCH + ch) * 32 + jj) * 64 + 4 * l16);
                        });
                        if (j >= 0 && l16 == 0 && wr) sums[j] -= 65536 * wr;
                        for (int r = 0; r < 4; r++) mw &= mw - 1;
                    }
                    wave_sync();
                    lower_test();
                    mask = ballot(cand);
                }
            }
            st_blocks++;
            st_vec += np;
            // the block's survivors, eight at a time (group g takes the g-th and (g + 4)-th
            // remaining set bits), re-estimated into the new-entry list in bit order, then
            // inserted at once
            const int ns = __popcll(mask);
            int nn = 0;  // round mode: new entries (estimate <= threshold)
            for (int s0 = 0; s0 < ns; s0 += 8) {
                uint64_t mm = mask;
                for (int r = 0; r < g; r++) mm &= mm - 1;
                const int j = mm ? (__ffsll((unsigned long long)mm) - 1) : -1;
                for (int r = 0; r < 4; r++) mm &= mm - 1;
                const int j2 = mm ? (__ffsll((unsigned long long)mm) - 1) : -1;
                const int jj = j < 0 ? 0 : j, jj2 = j2 < 0 ? 0 : j2;
                const uint32_t vi = v0 + (uint32_t)(b * 32 + jj);
                const uint32_t vi2 = v0 + (uint32_t)(b * 32 + jj2);
                float exf = 0.f, exf2 = 0.f;
                uint32_t idv = 0, idv2 = 0;
                float part = 0.f, part2 = 0.f;
                if (j >= 0) {
                    exf = a.ex_factor[vi];
                    idv = a.ids[vi];
                    if (j2 >= 0) {
                        exf2 = a.ex_factor[vi2];
                        idv2 = a.ids[vi2];
                    }
                    const float* ur = uq + l16 * S;
                    ex_chain2<EX>(ur,
                                  a.exl + ((size_t)vi * 16 + l16) * a.exl_w,
                                  a.exl + ((size_t)vi2 * 16 + l16) * a.exl_w, j2 >= 0, D >> 4,
                                  a.exl_w, part, part2);
                }
                const float ip = butterfly16(part), ip2 = butterfly16(part2);
                const float ipq_j = __shfl(ipq_l, jj), ipq_j2 = __shfl(ipq_l, jj2);
                const float on_j = __shfl(on_l, jj), on_j2 = __shfl(on_l, jj2);
                if (MODE == kModeEstimate) {
                    if (l16 == 0) {
                        if (j >= 0) {
                            a.est_exd[est0 + 32 * b + j] =
                                ex_distance<EX>(on_j, sqr_y, exf * y, ipq_j, ip, sumq);
                            a.est_id[est0 + 32 * b + j] = idv;
                        }
                        if (j2 >= 0) {
                            a.est_exd[est0 + 32 * b + j2] =
                                ex_distance<EX>(on_j2, sqr_y, exf2 * y, ipq_j2, ip2, sumq);
                            a.est_id[est0 + 32 * b + j2] = idv2;
                        }
                    }
                } else if (MODE == kModeReplay) {
                    if (l16 == 0) {
                        if (j >= 0) {
                            w.nd[s0 + g] = ex_distance<EX>(on_j, sqr_y, exf * y, ipq_j, ip, sumq);
                            w.nid[s0 + g] = idv;
                        }
                        if (j2 >= 0) {
                            w.nd[s0 + 4 + g] =
                                ex_distance<EX>(on_j2, sqr_y, exf2 * y, ipq_j2, ip2, sumq);
                            w.nid[s0 + 4 + g] = idv2;
                        }
                    }
                } else {
                    // candidates: estimate <= the round's threshold, kept in bit order
                    const float e1 = ex_distance<EX>(on_j, sqr_y, exf * y, ipq_j, ip, sumq);
                    const float e2 = ex_distance<EX>(on_j2, sqr_y, exf2 * y, ipq_j2, ip2, sumq);
                    const bool k1 = j >= 0 && e1 <= thr, k2 = j2 >= 0 && e2 <= thr;
                    const uint64_t b1 = ballot(k1 && l16 == 0), b2 = ballot(k2 && l16 == 0);
                    const int c1n = __popcll(b1), c2n = __popcll(b2);
                    if (l16 == 0) {
                        // group g's slot: after the groups before it (first survivors, then
                        // the second ones), in bit order of the survivors
                        const uint64_t below = (1ull << (16 * g)) - 1;
                        if (k1) {
                            const int at = nn + __popcll(b1 & below);
                            w.nd[at] = e1;
                            w.nid[at] = idv;
                        }
                        if (k2) {
                            const int at = nn + c1n + __popcll(b2 & below);
                            w.nd[at] = e2;
                            w.nid[at] = idv2;
                        }
                    }
                    nn += c1n + c2n;
                }
                for (int r = 0; r < 8; r++) mask &= mask - 1;
            }
            if (MODE == kModeReplay) {
                if (ns) {
                    wave_sync();
                    if (k <= 128) pool_insert_bulk128(w, psize, k, lane, w.nd, w.nid, ns);
                    else pool_insert_bulk(w, psize, k, lane, w.nd, w.nid, ns);
                }
                st_surv += ns;
                distk = (psize == k) ? w.pd[k - 1] : FLT_MAX;
            } else if (MODE == kModeRound && nn) {
                wave_sync();
                pool_merge_did(w, psize, k, lane, w.nd, w.nid, nn);
            }
        }
    }
    if (MODE == kModeEstimate) {
        // the blocks and vectors of the work item (the replay reads only candidate blocks)
        if (a.stats && lane == 0 && st_blocks) {
            atomicAdd(&a.stats[0], st_blocks);
            atomicAdd(&a.stats[1], st_vec);
        }
        return;
    }
    uint32_t* oi = a.out_ids + (size_t)qi * k;
    float* od = a.out_dists ? a.out_dists + (size_t)qi * k : nullptr;
    for (int i = lane; i < k; i += kWave) {
        oi[i] = i < psize ? w.pid[i] : 0u;
        if (od) od[i] = i < psize ? w.pd[i] : FLT_MAX;
    }
    if (MODE == kModeReplay && a.stats && lane == 0) {
        atomicAdd(&a.stats[0], st_blocks);
        atomicAdd(&a.stats[1], st_vec);
        atomicAdd(&a.stats[2], st_surv);
        atomicAdd(&a.stats[3], (unsigned long long)np_total);
    }
}

template <int EX, int NCH, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NCH > 4 ? 2 : 3)))
scan_mfma_kernel(ScanArgs a, int per_wave_bytes) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    unsigned char* base = smem + (size_t)wid * per_wave_bytes;
    if (MODE == kModeReplay) {
        const int qi = query_of_wave(a, wid);
        if (qi < 0) return;
        scan_query_mfma<EX, NCH, MODE>(a, qi, base, lane);
    } else if (MODE == kModeEstimate) {
        // items: kEstBlocks consecutive blocks of one (query, probe) pair; est_item (after
        // the pairs' slot offsets) holds each pair's first item
        const uint32_t np = a.probe_off[a.nq];
        const uint32_t* est_item = a.est_off + np + 1;
        const uint32_t ni = est_item[np];
        const uint32_t wpb = blockDim.x >> 6;
        for (uint32_t it = blockIdx.x * wpb + (uint32_t)wid; it < ni; it += gridDim.x * wpb) {
            uint32_t lo = 0, hi = np - 1;  // the last pair whose items start at or before it
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (est_item[mid] <= it) lo = mid; else hi = mid - 1;
            }
            const int g = (int)(it - est_item[lo]);
            scan_query_mfma<EX, NCH, MODE>(a, (int)lo, base, lane, g * kEstBlocks,
                                           (g + 1) * kEstBlocks);
            wave_sync();
        }
    } else {
        const uint32_t nl = a.qlist[0];
        const uint32_t wpb = blockDim.x >> 6;
        for (uint32_t i = blockIdx.x * wpb + (uint32_t)wid; i < nl; i += gridDim.x * wpb) {
            scan_query_mfma<EX, NCH, MODE>(a, (int)a.qlist[1 + i], base, lane);
            wave_sync();
        }
    }
}

template <int EX, int MODE>
hipError_t launch_mfma(const ScanArgs& a, dim3 grid, dim3 block, size_t lds, int per_wave,
                       hipStream_t s) {
    switch (a.nchunk) {
    case 1: hipLaunchKernelGGL((scan_mfma_kernel<EX, 1, MODE>), grid, block, lds, s, a, per_wave); break;
    case 2: hipLaunchKernelGGL((scan_mfma_kernel<EX, 2, MODE>), grid, block, lds, s, a, per_wave); break;
    case 3: hipLaunchKernelGGL((scan_mfma_kernel<EX, 3, MODE>), grid, block, lds, s, a, per_wave); break;
    case 4: hipLaunchKernelGGL((scan_mfma_kernel<EX, 4, MODE>), grid, block, lds, s, a, per_wave); break;
    case 6: hipLaunchKernelGGL((scan_mfma_kernel<EX, 6, MODE>), grid, block, lds, s, a, per_wave); break;
    case 8: hipLaunchKernelGGL((scan_mfma_kernel<EX, 8, MODE>), grid, block, lds, s, a, per_wave); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int EX>
hipError_t launch_ex(const ScanArgs& a, hipStream_t s) {
    int waves = a.waves_per_block > 0 ? a.waves_per_block : 1;
    if (a.bp) {
        const size_t per_wave = mfma_wave_bytes(a.D, a.nchunk, a.k);
        while (waves > 1 && per_wave * waves > 65536) waves--;
        dim3 grid((a.nq + waves - 1) / waves), block(64 * waves);
        return launch_mfma<EX, kModeReplay>(a, grid, block, per_wave * waves, (int)per_wave, s);
    }
    const size_t per_wave = (size_t)12 * a.D + 128 + (size_t)8 * ((a.k + 63) & ~63);
    while (waves > 1 && per_wave * waves > 65536) waves--;
    const size_t lds = per_wave * waves;
    dim3 grid((a.nq + waves - 1) / waves), block(64 * waves);
    if (a.D > 2048)
        hipLaunchKernelGGL((scan_kernel<EX, true>), grid, block, lds, s, a, (int)per_wave);
    else
        hipLaunchKernelGGL((scan_kernel<EX, false>), grid, block, lds, s, a, (int)per_wave);
    return hipGetLastError();
}

template <int EX>
hipError_t launch_round_ex(const ScanArgs& a, hipStream_t s) {
    const size_t per_wave = mfma_wave_bytes(a.D, a.nchunk, a.k);
    const int waves = per_wave * 2 <= 65536 ? 2 : 1;
    // a fixed grid: the listed queries are few (rounds whose candidates overflowed), and the
    // list's length is only known on the device
    dim3 grid(256), block(64 * waves);
    return launch_mfma<EX, kModeRound>(a, grid, block, per_wave * waves, (int)per_wave, s);
}

}  // namespace

hipError_t launch_query_order(const uint32_t* probe_off, const uint32_t* probe_ids,
                              const uint32_t* csize, int nq, int K, int max_probes, int shift,
                              uint32_t* scratch, uint32_t* qorder, hipStream_t s) {
    if (nq == 0) return hipSuccess;
    const int nb = K + 1;
    hipLaunchKernelGGL(order_zero_kernel, dim3((nb + 255) / 256), dim3(256), 0, s, scratch, nb);
    hipLaunchKernelGGL(order_count_kernel, dim3((nq + 255) / 256), dim3(256), 0, s, probe_off,
                       probe_ids, csize, nq, K, max_probes, shift, scratch);
    hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(1024), 0, s, scratch, nb);
    hipLaunchKernelGGL(order_scatter_kernel, dim3((nq + 255) / 256), dim3(256), 0, s, probe_off,
                       probe_ids, csize, nq, K, max_probes, shift, scratch, qorder);
    return hipGetLastError();
}

size_t scan_round_wave_bytes(int D, int k) { return mfma_wave_bytes(D, bp_nchunk(D), k); }

hipError_t launch_scan_round(const ScanArgs& a, hipStream_t s) {
    if (!a.bp || !a.qlist) return hipErrorInvalidValue;
    switch (a.EX) {
    case 2: return launch_round_ex<2>(a, s);
    case 3: return launch_round_ex<3>(a, s);
    case 4: return launch_round_ex<4>(a, s);
    case 6: return launch_round_ex<6>(a, s);
    case 7: return launch_round_ex<7>(a, s);
    case 8: return launch_round_ex<8>(a, s);
    default: return hipErrorInvalidValue;
    }
}

namespace {

// Slot offsets of the estimate mode: est_off[p] = sum over pairs p' < p of the probed
// cluster's size rounded up to 32 (one workgroup; pairs < probe_off[nq]).
__global__ void __launch_bounds__(1024) est_offsets_kernel(const uint32_t* __restrict__ probe_off,
                                                           const uint32_t* __restrict__ probe_ids,
                                                           const uint32_t* __restrict__ csize,
                                                           int nq, int K,
                                                           uint32_t* __restrict__ est_off) {
    __shared__ uint32_t wsum[16], isum[16];
    __shared__ uint32_t carry, icarry;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t np = probe_off[nq];
    uint32_t* est_item = est_off + np + 1;  // np + 1 more entries: each pair's first item
    if (tid == 0) {
        carry = 0;
        icarry = 0;
    }
    __syncthreads();
    for (uint32_t base = 0; base < np; base += 1024) {  // block-uniform
        const uint32_t p = base + (uint32_t)tid;
        uint32_t v = 0, u = 0;
        if (p < np) {
            const uint32_t cid = probe_ids[p];
            v = cid < (uint32_t)K ? ((csize[cid] + 31u) & ~31u) : 0u;
            u = (v / 32 + kEstBlocks - 1) / kEstBlocks;
        }
        uint32_t incl = v, iincl = u;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o), ti = __shfl_up(iincl, o);
            if (lane >= o) {
                incl += t;
                iincl += ti;
            }
        }
        if (lane == 63) {
            wsum[wid] = incl;
            isum[wid] = iincl;
        }
        __syncthreads();
        uint32_t pre = carry + incl - v, ipre = icarry + iincl - u;
        for (int w = 0; w < wid; w++) {
            pre += wsum[w];
            ipre += isum[w];
        }
        if (p < np) {
            est_off[p] = pre;
            est_item[p] = ipre;
        }
        __syncthreads();
        if (tid == 1023) {
            carry = pre + v;
            icarry = ipre + u;
        }
        __syncthreads();
    }
    if (tid == 0) {
        est_off[np] = carry;
        est_item[np] = icarry;
    }
}

// One wave per query: the query's blocks in probe order with the reference's per-block step
// of scan_query_mfma (the vectors whose lower bound is below the pool's k-th distance at the
// block's start, in slot order, into the pool).  A block whose smallest lower bound (est_bmin)
// is not below the k-th distance has no such vector, and the k-th distance only falls, so the
// replay reads the minima of 256 blocks at a time (the next 256 in flight) and fetches the
// slots of the candidate blocks only, eight blocks per request.  The block and vector
// counts come from the estimate kernel.
constexpr int kReplayWin = 256;  // block minima per step (4 per lane)
__global__ void __launch_bounds__(64) replay_est_kernel(ScanArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x, k = a.k, kcap = (k + 63) & ~63;
    const int qi = blockIdx.x;
    WaveLds w;
    w.uq = nullptr;
    w.lut = nullptr;
    w.sums = nullptr;
    w.pd = reinterpret_cast<float*>(smem);
    w.pid = reinterpret_cast<uint32_t*>(w.pd + kcap);
    w.nd = reinterpret_cast<float*>(w.pid + kcap);
    w.nid = reinterpret_cast<uint32_t*>(w.nd + 32);
    for (int i = lane; i < kcap; i += kWave) {
        w.pd[i] = FLT_MAX;
        w.pid[i] = 0;
    }
    wave_sync();
    int psize = 0;
    uint32_t p0 = a.probe_off[qi], p1 = a.probe_off[qi + 1];
    if (a.max_probes > 0 && p1 - p0 > (uint32_t)a.max_probes) p1 = p0 + (uint32_t)a.max_probes;
    const uint32_t B0 = a.est_off[p0] >> 5, B1 = a.est_off[p1] >> 5;  // slots come in 32s
    unsigned long long st_surv = 0;
    float distk = FLT_MAX;
    float bm[4], bn[4];
    auto minima = [&](uint32_t w0, float (&m)[4]) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint32_t b = w0 + (uint32_t)(64 * r + lane);
            m[r] = b < B1 ? a.est_bmin[b] : INFINITY;
        }
    };
    if (B0 < B1) minima(B0, bm);
    for (uint32_t w0 = B0; w0 < B1; w0 += kReplayWin) {  // wave-uniform
        if (w0 + kReplayWin < B1) minima(w0 + kReplayWin, bn);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint32_t base = w0 + 64 * r;
            uint64_t cand = ballot(bm[r] < distk);
            while (cand) {  // wave-uniform: up to 8 candidate blocks per request
                uint32_t blk[8];
                int nb = 0;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    blk[j] = base;
                    if (cand) {
                        blk[j] = base + (uint32_t)(__ffsll((unsigned long long)cand) - 1);
                        cand &= cand - 1;
                        nb++;
                    }
                }
                // lane l holds slot l % 32 of blocks 2 i + l / 32, i < 4
                float lv[4], ev[4];
                uint32_t iv[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int j = 2 * i + (lane >> 5);
                    const bool ok = j < nb;
                    const uint32_t bj = (lane >> 5) ? blk[2 * i + 1] : blk[2 * i];
                    const size_t sl = (size_t)bj * 32 + (lane & 31);
                    lv[i] = ok ? a.est_lower[sl] : INFINITY;
                    ev[i] = ok ? a.est_exd[sl] : 0.f;
                    iv[i] = ok ? a.est_id[sl] : 0u;
                }
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    if (j >= nb) break;
                    const int i = j >> 1, half = j & 1;
                    const uint64_t m64 = ballot(lv[i] < distk);
                    const uint32_t mask = half ? (uint32_t)(m64 >> 32) : (uint32_t)m64;
                    const int ns = __popc(mask);
                    if (ns) {
                        const int l32 = lane & 31;
                        if ((lane >> 5) == half && ((mask >> l32) & 1u)) {
                            const int t = __popc(mask & ((1u << l32) - 1u));
                            w.nd[t] = ev[i];
                            w.nid[t] = iv[i];
                        }
                        wave_sync();
                        if (k <= 128) pool_insert_bulk128(w, psize, k, lane, w.nd, w.nid, ns);
                        else pool_insert_bulk(w, psize, k, lane, w.nd, w.nid, ns);
                        st_surv += ns;
                        distk = (psize == k) ? w.pd[k - 1] : FLT_MAX;
                    }
                }
                // the rest of this group's candidates against the lowered k-th distance
                cand &= ballot(bm[r] < distk);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; r++) bm[r] = bn[r];
    }
    uint32_t* oi = a.out_ids + (size_t)qi * k;
    float* od = a.out_dists ? a.out_dists + (size_t)qi * k : nullptr;
    for (int i = lane; i < k; i += kWave) {
        oi[i] = i < psize ? w.pid[i] : 0u;
        if (od) od[i] = i < psize ? w.pd[i] : FLT_MAX;
    }
    if (a.stats && lane == 0) {
        atomicAdd(&a.stats[2], st_surv);
        atomicAdd(&a.stats[3], (unsigned long long)(p1 - p0));
    }
}

template <int EX>
hipError_t launch_estimate_ex(const ScanArgs& a, hipStream_t s) {
    const size_t per_wave = mfma_wave_bytes(a.D, a.nchunk, a.k);
    const int waves = per_wave * 2 <= 65536 ? 2 : 1;
    // a fixed grid: the pair count is only known on the device
    dim3 grid(1024), block(64 * waves);
    switch (a.nchunk) {
    case 1: hipLaunchKernelGGL((scan_mfma_kernel<EX, 1, kModeEstimate>), grid, block, per_wave * waves, s, a, (int)per_wave); break;
    case 2: hipLaunchKernelGGL((scan_mfma_kernel<EX, 2, kModeEstimate>), grid, block, per_wave * waves, s, a, (int)per_wave); break;
    case 3: hipLaunchKernelGGL((scan_mfma_kernel<EX, 3, kModeEstimate>), grid, block, per_wave * waves, s, a, (int)per_wave); break;
    case 4: hipLaunchKernelGGL((scan_mfma_kernel<EX, 4, kModeEstimate>), grid, block, per_wave * waves, s, a, (int)per_wave); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

size_t replay_est_lds(int k) {
    const size_t kcap = (size_t)((k + 63) & ~63);
    return 8 * kcap + 256 + 256;
}

hipError_t launch_scan_latency(const ScanArgs& a, hipStream_t s) {
    if (!a.bp || a.nchunk > 4 || !a.est_off || !a.est_lower || !a.est_exd || !a.est_id ||
        !a.est_bmin)
        return hipErrorInvalidValue;
    if (a.nq == 0) return hipSuccess;
    hipLaunchKernelGGL(est_offsets_kernel, dim3(1), dim3(1024), 0, s, a.probe_off, a.probe_ids,
                       a.csize, a.nq, a.K, a.est_off);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (a.EX) {
    case 2: e = launch_estimate_ex<2>(a, s); break;
    case 3: e = launch_estimate_ex<3>(a, s); break;
    case 4: e = launch_estimate_ex<4>(a, s); break;
    case 6: e = launch_estimate_ex<6>(a, s); break;
    case 7: e = launch_estimate_ex<7>(a, s); break;
    case 8: e = launch_estimate_ex<8>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(replay_est_kernel, dim3(a.nq), dim3(64), replay_est_lds(a.k), s, a);
    return hipGetLastError();
}

hipError_t launch_scan(const ScanArgs& a, hipStream_t s) {
    switch (a.EX) {
    case 2: return launch_ex<2>(a, s);
    case 3: return launch_ex<3>(a, s);
    case 4: return launch_ex<4>(a, s);
    case 6: return launch_ex<6>(a, s);
    case 7: return launch_ex<7>(a, s);
    case 8: return launch_ex<8>(a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace erq
