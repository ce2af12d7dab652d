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
// One wave per query.  The integer comes from a u16 LUT in LDS (the half-sums of D > 2048
// wrap as the reference's do; at D <= 2048 they cannot); every vector whose lower bound is
// within the k-th distance at block start is re-estimated speculatively (the threshold only
// falls), four at a time by 16-lane groups, and the survivors are then inserted in lane
// order, each tested against the threshold left by the inserts before it.
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
__global__ void __launch_bounds__(256) nonha_scan_kernel(ScanArgs a, int per_wave_bytes) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int slot = blockIdx.x * (blockDim.x >> 6) + wid;
    if (slot >= a.nq) return;
    const int qi = a.qorder ? (int)a.qorder[slot] : slot;
    const int D = a.D, M = D >> 2, k = a.k;
    const int kcap = (k + 63) & ~63;
    unsigned char* base = smem + (size_t)wid * per_wave_bytes;
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
