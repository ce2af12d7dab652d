// erq_nonha.hip -- the per-query scan of the non-HA index format on gfx950.
//
// The reference's IVF::search uses Searcher (inc/index/Searcher.hpp:14-236) instead of
// HASearcher when built without HIGH_ACC_FAST_SCAN.  Its blocks hold four factors per
// vector after the 4D code bytes (Quantizer.hpp:107-121): ||o-c||^2, fac_ip, sumxb and
// fac_err, 32 floats each.  Per probed cluster (Searcher::preparing, :118-124):
//   r = q - c, half_sum = 0.5 * sum(r)           minus_sum16, space.hpp:714-736
//   vl, vr = min(0, min r), max(0, max r)        data_range16, space.hpp:608-628
//   width = (vr - vl) / 16383
//   digit_d = u16(i32((r_d - vl) * f32(1/width) + 0.5))   scalar_quantize16, :644-686
//   LUT_m[j] = u16 sum of the digits of the dims whose bits j holds    pack_LUT, :126-172
// and per block (scan_one_block, :174-236) the integer of accumulate_robust (FastScan.hpp:
// 136-212) over the LUT's low and high byte planes: per plane, the even and the odd codebooks
// summed in u16 lanes (mod 2^16), the two sums added in 32 bits, then (upper << 8) + lower.
// The lower bound of vector j is
//   (||o-c||^2 + y^2) - ((I_j width + sumxb_j vl - half_sum) fac_ip_j + fac_err_j y)
// and, unlike the HA searcher, the pruning threshold is re-read after every insert: a vector
// is re-estimated iff its lower bound is not above the pool's current k-th distance.
//
// One wave per query.  Two integer kernels: a u16 LUT in LDS (every D; the half-sums of
// D > 2048 wrap as the reference's do), and for D <= 2048, where they cannot wrap, the int8
// MFMA over the bit-plane copy of the codes that the HA scan uses.  Every vector whose lower
// bound is within the k-th distance at block start is re-estimated speculatively (the
// threshold only falls), and the survivors are then inserted in lane order, each tested
// against the threshold left by the inserts before it.
#include "erq_device.h"
#include "erq_internal.h"
#include "erq_pool.h"

namespace erq {

namespace {

// x86 cvttps2dq: truncation toward zero; NaN and values outside int32 give INT_MIN.
__device__ __forceinline__ int32_t cvtt_i32(float t) {
    return (t >= -2147483648.0f && t < 2147483648.0f) ? (int32_t)t : INT32_MIN;
}

template <int EX, bool WRAP>
__device__ void nonha_lut_query(const ScanArgs& a, int qi, unsigned char* base, int lane) {
    const int D = a.D, M = D >> 2, k = a.k;
    const int kcap = (k + 63) & ~63;
    WaveLds w;
    w.uq = reinterpret_cast<float*>(base);                  // the residual r, D floats
    w.lut = reinterpret_cast<uint16_t*>(base + 4 * D);      // M x 16 u16
    w.sums = reinterpret_cast<int*>(base + 12 * D);         // 32 integers of a block
    w.pd = reinterpret_cast<float*>(base + 12 * D + 128);
    w.pid = reinterpret_cast<uint32_t*>(w.pd + kcap);
    w.nd = nullptr;
    w.nid = nullptr;
    for (int i = lane; i < kcap; i += kWave) {
        w.pd[i] = FLT_MAX;
        w.pid[i] = 0;
    }
    int psize = 0;
    wave_sync();

    const float* q = a.q_rot + (size_t)qi * D;
    const size_t BB = 4 * (size_t)D + 512;
    const size_t CL = (size_t)D * EX / 8;
    constexpr int R = 1 << EX;
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
        const float y = sqrtf(sqr_y);  // Searcher.hpp:93
        const float* c = a.centroids + (size_t)cid * D;

        // ---- minus_sum16: the residual and its reference-ordered sum ----
        float mx = 0.f, mn = 0.f;
        for (int d = lane; d < D; d += kWave) {
            const float r = q[d] - c[d];
            w.uq[d] = r;
            mx = fmaxf(mx, r);
            mn = fminf(mn, r);
        }
        wave_sync();
        float part = 0.f;
        if (lane < 16)
            for (int d = lane; d < D; d += 16) part = part + w.uq[d];
        part = butterfly16(part);
        const float half_sum = (float)(0.5 * (double)__shfl(part, 0));
        // ---- data_range16, width, scalar_quantize16, pack_LUT ----
        const float vr = wave_max(mx), vl = wave_min(mn);
        const float width = (vr - vl) / 16383.0f;
        const float o = (float)(1.0 / (double)width);
        for (int m = lane; m < M; m += kWave) {
            uint32_t qv[4];
            for (int t = 0; t < 4; t++) {
                float x = (w.uq[4 * m + t] - vl) * o;
                x = x + 0.5f;
                qv[t] = (uint32_t)cvtt_i32(x) & 0xFFFFu;
            }
            uint32_t wv[8];
            for (int j = 0; j < 16; j += 2) {
                // entry j: bit 3 <-> dim 4m, bit 0 <-> dim 4m+3; sums kept mod 2^16 as the
                // reference's u16 table stores them
                const uint32_t L0 =
                    (((j & 8) ? qv[0] : 0u) + ((j & 4) ? qv[1] : 0u) + ((j & 2) ? qv[2] : 0u)) &
                    0xFFFFu;
                const uint32_t L1 = (L0 + qv[3]) & 0xFFFFu;
                wv[j >> 1] = L0 | (L1 << 16);
            }
            uint4* dst = reinterpret_cast<uint4*>(w.lut + m * 16);
            dst[0] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
            dst[1] = make_uint4(wv[4], wv[5], wv[6], wv[7]);
        }
        wave_sync();

        // ---- blocks (scan_one_block) ----
        const int n = (int)a.csize[cid];
        const int nb = (n + 31) >> 5;
        const uint8_t* blk = a.short_data + (size_t)a.bstart[cid] * BB;
        const uint32_t v0 = a.vstart[cid];
        float distk = (psize == k) ? w.pd[k - 1] : FLT_MAX;
        for (int b = 0; b < nb; ++b, blk += BB) {
            const int np = min(32, n - b * 32);
            // Lane (t4, s16) reads dword t4 (bytes 4t4..4t4+3, 8 vectors) of codebooks
            // s16, s16 + 16, ...: the parity of a lane's codebooks is lane bit 2.
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
                // per parity: the high-plane sum and the low-plane sum (acc - 256 * high),
                // each taken mod 2^16; then the two parities added
                for (int o2 = 8; o2 <= 32; o2 <<= 1)
                    for (int kk = 0; kk < 8; kk++) {
                        acc[kk] += __shfl_xor(acc[kk], o2);
                        acch[kk] += __shfl_xor(acch[kk], o2);
                    }
                for (int kk = 0; kk < 8; kk++) {
                    const uint32_t lo = acc[kk] - (acch[kk] << 8);
                    acc[kk] = (lo & 0xFFFFu) + ((acch[kk] & 0xFFFFu) << 8);
                }
                for (int kk = 0; kk < 8; kk++) acc[kk] += __shfl_xor(acc[kk], 4);
            } else {
                for (int o2 = 4; o2 <= 32; o2 <<= 1)
                    for (int kk = 0; kk < 8; kk++) acc[kk] += __shfl_xor(acc[kk], o2);
            }
            if (lane < 4)
                for (int kk = 0; kk < 4; kk++) {
                    const int v = vec_of_byte(4 * lane + kk);
                    w.sums[v] = (int)acc[kk];
                    w.sums[v + 16] = (int)acc[4 + kk];
                }
            wave_sync();

            // Searcher.hpp:196-218, operation for operation
            const float* fx2 = reinterpret_cast<const float*>(blk + 4 * (size_t)D);
            float lower = 0.f, rip = 0.f, x2 = 0.f;
            bool cand = false;
            if (lane < 32) {
                x2 = fx2[lane];
                const float sum_sqr = x2 + sqr_y;
                const float xbvl = fx2[64 + lane] * vl;
                const float resf = (float)w.sums[lane];
                float ip = resf * width;
                ip = ip + xbvl;
                ip = ip - half_sum;
                rip = ip;
                ip = ip * fx2[32 + lane];
                const float err = fx2[96 + lane] * y;
                lower = sum_sqr - (ip + err);
                cand = (lane < np) && !(lower > distk);
            }
            uint64_t mask = ballot(cand);
            st_blocks++;
            st_vec += np;

            // ---- candidates in lane order, re-estimated 4 at a time ----
            while (mask) {
                uint64_t mm = mask;
                for (int r = 0; r < g; r++) mm &= mm - 1;
                const int j = mm ? (__ffsll((unsigned long long)mm) - 1) : -1;
                const int jj = j < 0 ? 0 : j;
                const uint32_t vi = v0 + (uint32_t)(b * 32 + jj);
                float pt = 0.f;
                if (j >= 0) {
                    // IP_FUNC(residual, long_code, D): 16 lane-strided FMA chains
                    const uint8_t* code = a.long_code + (size_t)vi * CL;
                    for (int ch = 0; ch < (D >> 6); ch++) {
                        float e[4];
                        decode4<EX>(code + ch * 8 * EX, l16, e);
                        for (int s4 = 0; s4 < 4; s4++)
                            pt = fmaf(w.uq[64 * ch + 16 * s4 + l16], e[s4], pt);
                    }
                }
                const float ipx = butterfly16(pt);
                const float rip_j = __shfl(rip, jj);
                const float x2_j = __shfl(x2, jj);
                const float low_j = __shfl(lower, jj);
                float exd = 0.f;
                uint32_t idv = 0;
                if (j >= 0) {
                    // Searcher.hpp:225-229, all in float
                    const float t1 = (float)R * rip_j;
                    const float t2 = t1 + ipx;
                    const float t3 = t2 - (float)(R - 1) * half_sum;
                    exd = (x2_j + sqr_y) - a.ex_factor[vi] * t3;
                    idv = a.ids[vi];
                }
                const int cnt = min(4, __popcll(mask));
                for (int gg = 0; gg < cnt; gg++) {
                    const float lj = __shfl(low_j, 16 * gg);
                    if (lj > distk) continue;  // wave-uniform
                    const float dd = __shfl(exd, 16 * gg);
                    const uint32_t ii = (uint32_t)__shfl((int)idv, 16 * gg);
                    pool_insert(w, psize, k, lane, dd, ii);
                    distk = (psize == k) ? w.pd[k - 1] : FLT_MAX;
                    st_surv++;
                }
                for (int r = 0; r < cnt; r++) mask &= mask - 1;
            }
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

template <int EX, bool WRAP>
__global__ void __launch_bounds__(256) nonha_scan_kernel(ScanArgs a, int per_wave_bytes) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int slot = blockIdx.x * (blockDim.x >> 6) + wid;
    if (slot >= a.nq) return;
    const int qi = a.qorder ? (int)a.qorder[slot] : slot;
    nonha_lut_query<EX, WRAP>(a, qi, smem + (size_t)wid * per_wave_bytes, lane);
}

// The queries the MFMA kernel handed back (a.nonha_redo[1 .. a.nonha_redo[0]]), by the LUT
// scan: a fixed grid of waves walks the list, whose length is only known on the device.
template <int EX>
__global__ void __launch_bounds__(256) nonha_redo_kernel(ScanArgs a, int per_wave_bytes) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const uint32_t cnt = a.nonha_redo[0];
    const uint32_t nw = gridDim.x * (blockDim.x >> 6);
    for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + wid; i < cnt; i += nw)
        nonha_lut_query<EX, false>(a, (int)a.nonha_redo[1 + i], smem + (size_t)wid * per_wave_bytes,
                                   lane);
}

typedef int v4i_t __attribute__((ext_vector_type(4)));

// LDS of one wave of nonha_mfma_kernel: the lane-major residual, the A operand and its zero
// twin, the block's integer sums, the pool.
__host__ __device__ inline size_t nonha_mfma_wave_bytes(int D, int nchunk, int k) {
    const size_t kcap = (size_t)((k + 63) & ~63);
    return 4 * 16 * (size_t)ul_stride(D) + 2048 * (size_t)nchunk + 128 + 8 * kcap;
}

// The unsigned 14-bit digits of scalar_quantize16 split for the int8 MFMA (digit 0 = q mod
// 128, digit 1 = q / 128), in the operand order of quantize_digits (erq_device.h): 16-byte
// units (chunk c, plane p, lane group g) at unit index 32 c + 4 p + g, low row then high row,
// byte 4 jw + b = dim 512 c + 128 g + 32 jw + 8 b + p.  Returns false on a lane that met a
// digit above 16383 (only a degenerate residual range gives one), which two int8 digits
// cannot carry.
__device__ __forceinline__ bool nonha_digits(const float* uq, int S, int D, int nchunk, float vl,
                                             float o, uint32_t* dst, int lane) {
    bool ok = true;
    for (int t = lane; t < nchunk * 64; t += kWave) {
        const int jw = t & 3, pp = (t >> 2) & 3, g = (t >> 4) & 3, c = t >> 6;
        const int base = 512 * c + 128 * g + 32 * jw + 2 * pp;
        uint32_t lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0;
        if (base < D) {
            uint32_t q0[4], q1[4];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int d = base + 8 * b;
                float x0 = (uq[ul_index(d, S)] - vl) * o;
                float x1 = (uq[ul_index(d + 1, S)] - vl) * o;
                x0 = x0 + 0.5f;
                x1 = x1 + 0.5f;
                q0[b] = (uint32_t)cvtt_i32(x0) & 0xFFFFu;
                q1[b] = (uint32_t)cvtt_i32(x1) & 0xFFFFu;
                ok = ok && q0[b] <= 16383u && q1[b] <= 16383u;
            }
            lo0 = pack_low_bytes(q0[0], q0[1], q0[2], q0[3]) & 0x7F7F7F7Fu;
            lo1 = pack_low_bytes(q1[0], q1[1], q1[2], q1[3]) & 0x7F7F7F7Fu;
            hi0 = pack_low_bytes(q0[0] >> 7, q0[1] >> 7, q0[2] >> 7, q0[3] >> 7);
            hi1 = pack_low_bytes(q1[0] >> 7, q1[1] >> 7, q1[2] >> 7, q1[3] >> 7);
        }
        const int un = 32 * c + 8 * pp + g;  // plane 2 pp; plane 2 pp + 1 is un + 4
        dst[8 * un + jw] = lo0;
        dst[8 * un + 4 + jw] = hi0;
        dst[8 * un + 32 + jw] = lo1;
        dst[8 * un + 36 + jw] = hi1;
    }
    return ok;
}

// The non-HA per-query scan with the FastScan integer on v_mfma_i32_16x16x64_i8 over the
// bit-plane codes (D <= 2048, where neither byte plane's u16 half-sums can wrap, so the
// integer is the exact inner product of the code bits with the digits): A rows 0 / 1 hold the
// query's low / high digits, rows 2..15 are zero; B is bit-plane p of 16 vectors' code bytes;
// I = sum_p (acc_p >> p), plane 7 with its sign.  One wave per query, everything else as the
// LUT kernel: the four-factor lower bound, candidates against the k-th distance at block
// start re-estimated eight at a time (two per 16-lane group, lane-major EX-bit codes), then
// inserted in bit order, each against the threshold the inserts before it left.  A query
// with a digit above 16383 goes to a.nonha_redo for the LUT kernel.
template <int EX, int NCH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
nonha_mfma_kernel(ScanArgs a, int per_wave_bytes) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int slot = blockIdx.x * (blockDim.x >> 6) + wid;
    if (slot >= a.nq) return;
    const int qi = a.qorder ? (int)a.qorder[slot] : slot;
    const int D = a.D, k = a.k;
    const int kcap = (k + 63) & ~63;
    const int S = ul_stride(D);
    unsigned char* base = smem + (size_t)wid * per_wave_bytes;
    float* uq = reinterpret_cast<float*>(base);  // r = q - c, lane-major
    const size_t ubytes = 4 * 16 * (size_t)S;
    int8_t* A = reinterpret_cast<int8_t*>(base + ubytes);  // NCH KB
    int8_t* Z = A + 1024 * NCH;                            // NCH KB of zeros
    int* sums = reinterpret_cast<int*>(Z + 1024 * NCH);
    WaveLds w;
    w.uq = uq;
    w.lut = nullptr;
    w.sums = sums;
    w.pd = reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(sums) + 128);
    w.pid = reinterpret_cast<uint32_t*>(w.pd + kcap);
    w.nd = nullptr;
    w.nid = nullptr;
    for (int i = lane; i < NCH * 64; i += kWave)
        reinterpret_cast<uint4*>(Z)[i] = make_uint4(0, 0, 0, 0);
    for (int i = lane; i < kcap; i += kWave) {
        w.pd[i] = FLT_MAX;
        w.pid[i] = 0;
    }
    int psize = 0;
    wave_sync();

    const float* q = a.q_rot + (size_t)qi * D;
    const size_t BB = 4 * (size_t)D + 512;
    constexpr int R = 1 << EX;
    unsigned long long st_blocks = 0, st_vec = 0, st_surv = 0;
    const int g = lane >> 4, l16 = lane & 15;
    const int8_t* abase = l16 < 2 ? A + (g * 2 + l16) * 16 : Z + g * 32;

    const uint32_t p0 = a.probe_off[qi];
    uint32_t p1 = a.probe_off[qi + 1];
    if (a.max_probes > 0 && p1 - p0 > (uint32_t)a.max_probes) p1 = p0 + (uint32_t)a.max_probes;
    for (uint32_t p = p0; p < p1; ++p) {
        const uint32_t cid = a.probe_ids[p];
        if (cid >= (uint32_t)a.K) continue;
        const float sqr_y = a.probe_sqr_y[p];
        const float y = sqrtf(sqr_y);  // Searcher.hpp:93
        const float* c = a.centroids + (size_t)cid * D;
        const int n = (int)a.csize[cid];
        const int nb = (n + 31) >> 5;
        const uint32_t blk0 = a.bstart[cid];
        const uint32_t v0 = a.vstart[cid];
        uint4 wn0[NCH], wn1[NCH];
        auto load_block = [&](uint32_t blk) {
#pragma unroll
            for (int ch = 0; ch < NCH; ch++) {
                const uint8_t* src = a.bp + (((size_t)blk * NCH + ch) * 32 + l16) * 64 + 16 * g;
                wn0[ch] = *reinterpret_cast<const uint4*>(src);
                wn1[ch] = *reinterpret_cast<const uint4*>(src + 16 * 64);
            }
        };
        if (nb > 0) load_block(blk0);
        // ---- preparing(): residual, its half sum, the range, the digits ----
        stage_residual<3>(q, c, 1.0f, D, S, uq, lane);
        wave_sync();
        float mx = 0.f, mn = 0.f;
        for (int d = lane; d < D; d += kWave) {
            const float r = uq[ul_index(d, S)];
            mx = fmaxf(mx, r);
            mn = fminf(mn, r);
        }
        const float half_sum = (float)(0.5 * (double)residual_sum16(uq, S, D, l16));
        const float vr = wave_max(mx), vl = wave_min(mn);
        const float width = (vr - vl) / 16383.0f;
        const float o = (float)(1.0 / (double)width);
        const bool ok = nonha_digits(uq, S, D, NCH, vl, o, reinterpret_cast<uint32_t*>(A), lane);
        if (ballot(!ok)) {  // wave-uniform: the LUT kernel redoes this query from the start
            if (lane == 0) a.nonha_redo[1 + atomicAdd(&a.nonha_redo[0], 1u)] = (uint32_t)qi;
            return;
        }
        wave_sync();

        float distk = (psize == k) ? w.pd[k - 1] : FLT_MAX;
        for (int b = 0; b < nb; ++b) {
            const uint32_t blk = blk0 + (uint32_t)b;
            const int np = min(32, n - b * 32);
            uint4 wc0[NCH], wc1[NCH];
#pragma unroll
            for (int ch = 0; ch < NCH; ch++) {
                wc0[ch] = wn0[ch];
                wc1[ch] = wn1[ch];
            }
            if (b + 1 < nb) load_block(blk + 1);
            int lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0;
#pragma unroll 2
            for (int pp = 0; pp < 8; pp++) {
                v4i_t acc0 = v4i_t{0, 0, 0, 0}, acc1 = v4i_t{0, 0, 0, 0};
                const uint32_t m = 0x01010101u << pp;
#pragma unroll
                for (int ch = 0; ch < NCH; ch++) {
                    const v4i_t av = *reinterpret_cast<const v4i_t*>(abase + (ch * 8 + pp) * 128);
                    const v4i_t b0 = v4i_t{(int)(wc0[ch].x & m), (int)(wc0[ch].y & m),
                                           (int)(wc0[ch].z & m), (int)(wc0[ch].w & m)};
                    const v4i_t b1 = v4i_t{(int)(wc1[ch].x & m), (int)(wc1[ch].y & m),
                                           (int)(wc1[ch].z & m), (int)(wc1[ch].w & m)};
                    acc0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b0, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b1, acc1, 0, 0, 0);
                }
                if (pp < 7) {
                    lo0 += acc0[0] >> pp;
                    hi0 += acc0[1] >> pp;
                    lo1 += acc1[0] >> pp;
                    hi1 += acc1[1] >> pp;
                } else {  // the code byte 0x80 is -128 as an int8
                    lo0 -= acc0[0] >> 7;
                    hi0 -= acc0[1] >> 7;
                    lo1 -= acc1[0] >> 7;
                    hi1 -= acc1[1] >> 7;
                }
            }
            if (lane < 16) {  // rows 0 (low digits) and 1 (high digits) of column l16
                sums[lane] = 128 * hi0 + lo0;
                sums[16 + lane] = 128 * hi1 + lo1;
            }
            wave_sync();
            // Searcher.hpp:196-218, operation for operation
            const float* fx2 = reinterpret_cast<const float*>(a.short_data + (size_t)blk * BB +
                                                              4 * (size_t)D);
            float lower = INFINITY, rip = 0.f, x2 = 0.f;
            bool cand = false;
            if (lane < 32) {
                x2 = fx2[lane];
                const float sum_sqr = x2 + sqr_y;
                const float xbvl = fx2[64 + lane] * vl;
                const float resf = (float)sums[lane];
                float ip = resf * width;
                ip = ip + xbvl;
                ip = ip - half_sum;
                rip = ip;
                ip = ip * fx2[32 + lane];
                const float err = fx2[96 + lane] * y;
                lower = sum_sqr - (ip + err);
                cand = (lane < np) && !(lower > distk);
            }
            uint64_t mask = ballot(cand);
            st_blocks++;
            st_vec += np;
            // candidates eight at a time (group g takes the g-th and (g + 4)-th remaining
            // set bits), each re-estimated, then inserted in bit order
            while (mask) {
                uint64_t mm = mask;
                for (int r = 0; r < g; r++) mm &= mm - 1;
                const int j = mm ? (__ffsll((unsigned long long)mm) - 1) : -1;
                for (int r = 0; r < 4; r++) mm &= mm - 1;
                const int j2 = mm ? (__ffsll((unsigned long long)mm) - 1) : -1;
                const int jj = j < 0 ? 0 : j, jj2 = j2 < 0 ? 0 : j2;
                const uint32_t vi = v0 + (uint32_t)(b * 32 + jj);
                const uint32_t vi2 = v0 + (uint32_t)(b * 32 + jj2);
                float part = 0.f, part2 = 0.f;
                if (j >= 0) {
                    const float* ur = uq + l16 * S;
                    ex_chain2<EX>(ur,
                                  a.exl + ((size_t)vi * 16 + l16) * a.exl_w,
                                  a.exl + ((size_t)vi2 * 16 + l16) * a.exl_w, j2 >= 0, D >> 4,
                                  a.exl_w, part, part2);
                }
                const float ipx = butterfly16(part), ipx2 = butterfly16(part2);
                // Searcher.hpp:225-229, all in float
                auto exd_of = [&](int jv, uint32_t vv, float ipv) {
                    const float rj = __shfl(rip, jv), xj = __shfl(x2, jv);
                    const float t1 = (float)R * rj;
                    const float t2 = t1 + ipv;
                    const float t3 = t2 - (float)(R - 1) * half_sum;
                    return (xj + sqr_y) - a.ex_factor[vv] * t3;
                };
                const float e1 = exd_of(jj, vi, ipx), e2 = exd_of(jj2, vi2, ipx2);
                const uint32_t id1 = a.ids[vi], id2 = a.ids[vi2];
                const float l1 = __shfl(lower, jj), l2 = __shfl(lower, jj2);
                for (int h = 0; h < 2; h++)
                    for (int gg = 0; gg < 4; gg++) {
                        const int jv = __shfl(h ? j2 : j, 16 * gg);  // wave-uniform
                        if (jv < 0) continue;
                        const float lj = __shfl(h ? l2 : l1, 16 * gg);
                        if (lj > distk) continue;
                        const float dd = __shfl(h ? e2 : e1, 16 * gg);
                        const uint32_t ii = (uint32_t)__shfl((int)(h ? id2 : id1), 16 * gg);
                        pool_insert(w, psize, k, lane, dd, ii);
                        distk = (psize == k) ? w.pd[k - 1] : FLT_MAX;
                        st_surv++;
                    }
                for (int r = 0; r < 8; r++) mask &= mask - 1;
            }
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

__global__ void nonha_zero_kernel(uint32_t* p) { *p = 0; }

template <int EX>
hipError_t launch_nonha_mfma(const ScanArgs& a, hipStream_t s) {
    int waves = a.waves_per_block > 0 ? a.waves_per_block : 1;
    const size_t pw = nonha_mfma_wave_bytes(a.D, a.nchunk, a.k);
    while (waves > 1 && pw * waves > 65536) waves--;
    dim3 grid((a.nq + waves - 1) / waves), block(64 * waves);
    hipLaunchKernelGGL(nonha_zero_kernel, dim3(1), dim3(1), 0, s, a.nonha_redo);
    switch (a.nchunk) {
    case 1: hipLaunchKernelGGL((nonha_mfma_kernel<EX, 1>), grid, block, pw * waves, s, a, (int)pw); break;
    case 2: hipLaunchKernelGGL((nonha_mfma_kernel<EX, 2>), grid, block, pw * waves, s, a, (int)pw); break;
    case 3: hipLaunchKernelGGL((nonha_mfma_kernel<EX, 3>), grid, block, pw * waves, s, a, (int)pw); break;
    case 4: hipLaunchKernelGGL((nonha_mfma_kernel<EX, 4>), grid, block, pw * waves, s, a, (int)pw); break;
    default: return hipErrorInvalidValue;
    }
    // the handed-back queries (normally none) by the LUT kernel, on a fixed grid
    const size_t lw = (size_t)12 * a.D + 128 + (size_t)8 * ((a.k + 63) & ~63);
    hipLaunchKernelGGL(nonha_redo_kernel<EX>, dim3(64), dim3(64), lw, s, a, (int)lw);
    return hipGetLastError();
}

template <int EX>
hipError_t launch_nonha_ex(const ScanArgs& a, hipStream_t s) {
    int waves = a.waves_per_block > 0 ? a.waves_per_block : 1;
    const size_t per_wave = (size_t)12 * a.D + 128 + (size_t)8 * ((a.k + 63) & ~63);
    while (waves > 1 && per_wave * waves > 65536) waves--;
    const size_t lds = per_wave * waves;
    dim3 grid((a.nq + waves - 1) / waves), block(64 * waves);
    if (a.D > 2048)
        hipLaunchKernelGGL((nonha_scan_kernel<EX, true>), grid, block, lds, s, a, (int)per_wave);
    else
        hipLaunchKernelGGL((nonha_scan_kernel<EX, false>), grid, block, lds, s, a, (int)per_wave);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_scan_nonha(const ScanArgs& a, hipStream_t s) {
    if (a.nq == 0) return hipSuccess;
    if (a.bp && a.exl && a.nonha_redo && a.nchunk <= 4) {
        switch (a.EX) {
        case 2: return launch_nonha_mfma<2>(a, s);
        case 3: return launch_nonha_mfma<3>(a, s);
        case 4: return launch_nonha_mfma<4>(a, s);
        case 6: return launch_nonha_mfma<6>(a, s);
        case 7: return launch_nonha_mfma<7>(a, s);
        case 8: return launch_nonha_mfma<8>(a, s);
        default: return hipErrorInvalidValue;
        }
    }
    switch (a.EX) {
    case 2: return launch_nonha_ex<2>(a, s);
    case 3: return launch_nonha_ex<3>(a, s);
    case 4: return launch_nonha_ex<4>(a, s);
    case 6: return launch_nonha_ex<6>(a, s);
    case 7: return launch_nonha_ex<7>(a, s);
    case 8: return launch_nonha_ex<8>(a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace erq
