// erq_cpu512.cpp -- the HA IVF search (IVF::search with HASearcher, IVF.hpp:295-324,
// HASearcher.hpp:82-237) on the host CPU with AVX-512: the CPU baseline that bench.py times
// beside the GPU figures, following test_search.cpp:83-150's protocol.
//
// TEST INFRASTRUCTURE: only tests/ and bench.py's cpu_baseline leg load this library; the HIP
// product never calls it.  Results are bit-identical to the scalar oracle (erq_oracle.c, the
// same float operations in the same order, FMA where the reference writes an FMA intrinsic;
// tests/test_cpu512.py checks it), so it measures the reference algorithm's speed on this host
// without being the reference binary (which needs Eigen and hnswlib, absent here).
//
// Built by oracle/Makefile with -O3 -march=x86-64-v4 (AVX-512 F/BW/DQ/VL) -ffp-contract=off;
// c512_available() reports whether the running CPU has those instructions.
#include <immintrin.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "erq_oracle.h"

namespace {

double now_us() {
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e6 + t.tv_nsec * 1e-3;
}

// space.hpp:18-31: 16 lane-strided FMA chains of (x - y)^2, then _mm512_reduce_add_ps.
inline float l2sqr16(const float* x, const float* y, size_t D) {
    __m512 acc = _mm512_setzero_ps();
    for (size_t i = 0; i < D; i += 16) {
        const __m512 t = _mm512_sub_ps(_mm512_loadu_ps(x + i), _mm512_loadu_ps(y + i));
        acc = _mm512_fmadd_ps(t, t, acc);
    }
    return _mm512_reduce_add_ps(acc);
}

// The 16-lane code values of dims 16 s + l (s = 0..3) of one 64-dim chunk of an EX-bit
// code in the store_compacted_code layouts (Quantizer.hpp:491-631).
template <int EX>
inline void decode_chunk(const uint8_t* c, __m512i e[4]) {
    auto bytes = [](const uint8_t* p) {
        return _mm512_cvtepu8_epi32(_mm_loadu_si128(reinterpret_cast<const __m128i*>(p)));
    };
    // bit t of the u64 at p, for dim t = 16 s + l: byte l & 7, bit (l >> 3) + 2 s
    auto top = [](const uint8_t* p) {
        const __m128i t8 = _mm_loadl_epi64(reinterpret_cast<const __m128i*>(p));
        const __m128i idx = _mm_set_epi8(7, 6, 5, 4, 3, 2, 1, 0, 7, 6, 5, 4, 3, 2, 1, 0);
        const __m512i b = _mm512_cvtepu8_epi32(_mm_shuffle_epi8(t8, idx));
        const __m512i sh = _mm512_set_epi32(1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0);
        return _mm512_srlv_epi32(b, sh);
    };
    const __m512i one = _mm512_set1_epi32(1);
    if constexpr (EX == 8) {
        for (int s = 0; s < 4; s++) e[s] = bytes(c + 16 * s);
    } else if constexpr (EX == 4) {
        const __m512i m = _mm512_set1_epi32(15);
        const __m512i b0 = bytes(c), b1 = bytes(c + 16);
        e[0] = _mm512_and_si512(b0, m);
        e[1] = _mm512_srli_epi32(b0, 4);
        e[2] = _mm512_and_si512(b1, m);
        e[3] = _mm512_srli_epi32(b1, 4);
    } else if constexpr (EX == 2 || EX == 3) {
        const __m512i m = _mm512_set1_epi32(3);
        const __m512i b = bytes(c);
        __m512i tb = _mm512_setzero_si512();
        if constexpr (EX == 3) tb = top(c + 16);
        for (int s = 0; s < 4; s++) {
            __m512i v = _mm512_and_si512(_mm512_srli_epi32(b, 2 * s), m);
            if constexpr (EX == 3)
                v = _mm512_or_si512(
                    v, _mm512_slli_epi32(_mm512_and_si512(_mm512_srli_epi32(tb, 2 * s), one), 2));
            e[s] = v;
        }
    } else {  // 6, 7
        const __m512i m6 = _mm512_set1_epi32(63), m4 = _mm512_set1_epi32(15);
        const __m512i b0 = bytes(c), b1 = bytes(c + 16), b2 = bytes(c + 32);
        e[0] = _mm512_and_si512(b0, m6);
        e[1] = _mm512_and_si512(b1, m6);
        e[2] = _mm512_or_si512(_mm512_and_si512(b2, m4),
                               _mm512_slli_epi32(_mm512_srli_epi32(b0, 6), 4));
        e[3] = _mm512_or_si512(_mm512_srli_epi32(b2, 4),
                               _mm512_slli_epi32(_mm512_srli_epi32(b1, 6), 4));
        if constexpr (EX == 7) {
            const __m512i tb = top(c + 48);
            for (int s = 0; s < 4; s++)
                e[s] = _mm512_or_si512(
                    e[s],
                    _mm512_slli_epi32(_mm512_and_si512(_mm512_srli_epi32(tb, 2 * s), one), 6));
        }
    }
}

// IP16_fxu8 / IP32_fxu4 / IP64_fxu{2,3,6,7} (space.hpp:109-497): lane l accumulates
// u[d] * e[d] by FMA over d = 16 i + l in increasing i, then _mm512_reduce_add_ps.
template <int EX>
float ip_ex(const float* u, const uint8_t* code, size_t D) {
    __m512 acc = _mm512_setzero_ps();
    for (size_t ch = 0; ch < D / 64; ch++, code += 8 * EX) {
        __m512i e[4];
        decode_chunk<EX>(code, e);
        for (int s = 0; s < 4; s++)
            acc = _mm512_fmadd_ps(_mm512_loadu_ps(u + 64 * ch + 16 * s), _mm512_cvtepi32_ps(e[s]),
                                  acc);
    }
    return _mm512_reduce_add_ps(acc);
}

struct Pool {  // ResultPool, Pool.hpp:10-50
    std::vector<uint32_t> ids;
    std::vector<float> dist;
    size_t size = 0, cap = 0;
    void reset(size_t k) {
        cap = k;
        size = 0;
        ids.assign(k + 1, 0);
        dist.assign(k + 1, 0.f);
    }
    float distk() const { return size == cap ? dist[size - 1] : FLT_MAX; }
    void insert(uint32_t id, float d) {
        if (size == cap && d > dist[size - 1]) return;
        const size_t lo = std::lower_bound(dist.begin(), dist.begin() + size, d) - dist.begin();
        memmove(&ids[lo + 1], &ids[lo], (size - lo) * sizeof(uint32_t));
        memmove(&dist[lo + 1], &dist[lo], (size - lo) * sizeof(float));
        ids[lo] = id;
        dist[lo] = d;
        size += size < cap;
    }
};

struct Searcher {
    const orc_index* idx;
    size_t D;
    std::vector<float, std::allocator<float>> uq;
    std::vector<int16_t> qq;
    std::vector<uint8_t> lut_lo, lut_hi;  // per codebook 16 bytes, 4 codebooks per 64 bytes
    std::vector<std::pair<float, uint32_t>> cand;
    Pool pool;
    float one_over_sqrtD;

    explicit Searcher(const orc_index* ix) : idx(ix), D(ix->D) {
        uq.resize(D);
        qq.resize(D);
        lut_lo.resize(4 * D);
        lut_hi.resize(4 * D);
        cand.resize(ix->K);
        one_over_sqrtD = (float)(1.0 / (double)std::sqrt((float)D));  // HASearcher.hpp:56
    }

    // normalize_query16 (space.hpp:749-778); returns sumq
    float normalize(const float* q, const float* c, float y) {
        if (y > 1e-5f) {
            const __m512 fac = _mm512_set1_ps((float)(1.0 / (double)y));
            __m512 acc = _mm512_setzero_ps();
            for (size_t i = 0; i < D; i += 16) {
                const __m512 u =
                    _mm512_mul_ps(_mm512_sub_ps(_mm512_loadu_ps(q + i), _mm512_loadu_ps(c + i)),
                                  fac);
                acc = _mm512_add_ps(acc, u);
                _mm512_storeu_ps(&uq[i], u);
            }
            return _mm512_reduce_add_ps(acc);
        }
        const float value = (float)(1.0 / (double)std::sqrt((float)D));
        for (size_t i = 0; i < D; i++) uq[i] = value;
        return (float)D * value;
    }

    // data_range16 + high_acc_quantize16 (space.hpp:608-628, 695-711); returns delta
    float quantize() {
        __m512 mx = _mm512_setzero_ps(), mn = _mm512_setzero_ps();
        for (size_t i = 0; i < D; i += 16) {
            const __m512 u = _mm512_loadu_ps(&uq[i]);
            mx = _mm512_max_ps(u, mx);
            mn = _mm512_min_ps(u, mn);
        }
        const float vr = _mm512_reduce_max_ps(mx), vl = _mm512_reduce_min_ps(mn);
        const float vmax = std::max(std::fabs(vl), std::fabs(vr));
        const float delta = vmax / (float)((1 << 13) - 1);
        const __m512 dv = _mm512_set1_ps(delta);
        const __m512d half = _mm512_set1_pd(0.5), one = _mm512_set1_pd(1.0);
        for (size_t i = 0; i < D; i += 16) {
            const __m512 t = _mm512_div_ps(_mm512_loadu_ps(&uq[i]), dv);
            const __mmask16 neg = _mm512_cmp_ps_mask(t, _mm512_setzero_ps(), _CMP_LT_OQ);
            __m512d a = _mm512_cvtps_pd(_mm512_castps512_ps256(t));
            __m512d b = _mm512_cvtps_pd(_mm512_extractf32x8_ps(t, 1));
            a = _mm512_mask_sub_pd(_mm512_add_pd(a, half), (__mmask8)(neg & 0xFF),
                                   _mm512_add_pd(a, half), one);
            b = _mm512_mask_sub_pd(_mm512_add_pd(b, half), (__mmask8)(neg >> 8),
                                   _mm512_add_pd(b, half), one);
            const __m256i ia = _mm512_cvttpd_epi32(a), ib = _mm512_cvttpd_epi32(b);
            const __m512i w = _mm512_inserti64x4(_mm512_castsi256_si512(ia), ib, 1);
            _mm256_storeu_si256(reinterpret_cast<__m256i*>(&qq[i]), _mm512_cvtepi32_epi16(w));
        }
        return delta;
    }

    // pack_high_acc_LUT (HASearcher.hpp:128-185): entry j of codebook m = the sum of the
    // digits whose bit j holds (bit 3 = dim 4m), minus the codebook's minimum; low and high
    // bytes in two tables.  Returns the summed minima (the shift).
    int32_t pack_lut() {
        // lane j of m3/m2/m1/m0: bit 3/2/1/0 of j
        const __mmask16 b3 = 0xFF00, b2 = 0xF0F0, b1 = 0xCCCC, b0 = 0xAAAA;
        int32_t shift = 0;
        for (size_t m = 0; m < D / 4; m++) {
            const int16_t* d = &qq[4 * m];
            __m512i L = _mm512_maskz_mov_epi32(b3, _mm512_set1_epi32(d[0]));
            L = _mm512_add_epi32(L, _mm512_maskz_mov_epi32(b2, _mm512_set1_epi32(d[1])));
            L = _mm512_add_epi32(L, _mm512_maskz_mov_epi32(b1, _mm512_set1_epi32(d[2])));
            L = _mm512_add_epi32(L, _mm512_maskz_mov_epi32(b0, _mm512_set1_epi32(d[3])));
            const int vmin = _mm512_reduce_min_epi32(L);  // includes entry 0 = 0
            L = _mm512_sub_epi32(L, _mm512_set1_epi32(vmin));
            _mm_storeu_si128(reinterpret_cast<__m128i*>(&lut_lo[16 * m]), _mm512_cvtepi32_epi8(L));
            _mm_storeu_si128(reinterpret_cast<__m128i*>(&lut_hi[16 * m]),
                             _mm512_cvtepi32_epi8(_mm512_srli_epi32(L, 8)));
            shift += vmin;
        }
        return shift;
    }

    // The FastScan integers of one block (accumulate_one_block_high_acc, FastScan.hpp:214-
    // 290): per LUT byte plane, codebooks of one parity (128-bit lanes 0,2 and 1,3) summed in
    // u16 (mod 2^16), the two parities added in 32 bits; low + 256 * high.
    void block_ints(const uint8_t* codes, int32_t out[32]) {
        const __m512i nib = _mm512_set1_epi8(0x0F);
        // [table lo/hi][nibble lo/hi]: u16 lanes summing byte pairs (A) and odd bytes (B)
        __m512i A[2][2], B[2][2];
        for (int t = 0; t < 2; t++)
            for (int n = 0; n < 2; n++) A[t][n] = B[t][n] = _mm512_setzero_si512();
        for (size_t i = 0; i < 4 * D; i += 64) {
            const __m512i c = _mm512_loadu_si512(codes + i);
            const __m512i lo = _mm512_and_si512(c, nib);
            const __m512i hi = _mm512_and_si512(_mm512_srli_epi16(c, 4), nib);
            const __m512i tl = _mm512_loadu_si512(&lut_lo[i]);
            const __m512i th = _mm512_loadu_si512(&lut_hi[i]);
            const __m512i r[2][2] = {{_mm512_shuffle_epi8(tl, lo), _mm512_shuffle_epi8(tl, hi)},
                                     {_mm512_shuffle_epi8(th, lo), _mm512_shuffle_epi8(th, hi)}};
            for (int t = 0; t < 2; t++)
                for (int n = 0; n < 2; n++) {
                    A[t][n] = _mm512_add_epi16(A[t][n], r[t][n]);
                    B[t][n] = _mm512_add_epi16(B[t][n], _mm512_srli_epi16(r[t][n], 8));
                }
        }
        // per (table, nibble): even-byte sums E = A - 256 B, odd-byte sums B, each folded
        // over the parity lanes (0 + 2, 1 + 3) in u16, then the parities widened and added
        __m256i S[2][2][2];  // [table][nibble][even/odd byte] -> 8 u32
        for (int t = 0; t < 2; t++)
            for (int n = 0; n < 2; n++) {
                const __m512i E = _mm512_sub_epi16(A[t][n], _mm512_slli_epi16(B[t][n], 8));
                const __m512i V[2] = {E, B[t][n]};
                for (int h = 0; h < 2; h++) {
                    const __m128i l0 = _mm512_extracti32x4_epi32(V[h], 0);
                    const __m128i l1 = _mm512_extracti32x4_epi32(V[h], 1);
                    const __m128i l2 = _mm512_extracti32x4_epi32(V[h], 2);
                    const __m128i l3 = _mm512_extracti32x4_epi32(V[h], 3);
                    const __m128i even_cb = _mm_add_epi16(l0, l2), odd_cb = _mm_add_epi16(l1, l3);
                    S[t][n][h] = _mm256_add_epi32(_mm256_cvtepu16_epi32(even_cb),
                                                  _mm256_cvtepu16_epi32(odd_cb));
                }
            }
        // u16 lane i of a codebook's 16 bytes: even byte 2i = vector i, odd byte = vector i+8
        // (perm0, pack_codes.hpp:44); the high nibble holds vector + 16
        for (int n = 0; n < 2; n++)
            for (int h = 0; h < 2; h++) {
                const __m256i v = _mm256_add_epi32(S[0][n][h], _mm256_slli_epi32(S[1][n][h], 8));
                _mm256_storeu_si256(reinterpret_cast<__m256i*>(out + 16 * n + 8 * h), v);
            }
    }

    template <int EX>
    void search_cluster(const float* q, size_t cid, float sqr_y) {
        const size_t n = (size_t)idx->cluster_sizes[cid];
        const float* c = idx->centroids + cid * D;
        const float y = std::sqrt(sqr_y);
        const float sumq = normalize(q, c, y);
        const float delta = quantize();
        const int32_t shift = pack_lut();
        constexpr int R = 1 << EX;
        const size_t BB = 4 * D + 128, CL = D * EX / 8;
        const uint8_t* blk = idx->short_data + idx->block_start[cid] * BB;
        const size_t v0 = idx->vec_start[cid];
        // FastScan.hpp:295-323
        const float c1 = (float)(0.5 * (double)sumq - (double)0.58f);
        const float c2 = (-5.0f * y) * one_over_sqrtD;
        const float qs = y * y;
        float distk = pool.distk();
        alignas(64) int32_t ints[32];
        alignas(64) float ipq[32];
        for (size_t b = 0; b * 32 < n; b++, blk += BB) {
            const size_t np = std::min<size_t>(32, n - b * 32);
            block_ints(blk, ints);
            const float* on = reinterpret_cast<const float*>(blk + 4 * D);
            uint32_t mask = 0;
            for (int h = 0; h < 2; h++) {
                const __m512i I = _mm512_add_epi32(_mm512_load_si512(ints + 16 * h),
                                                   _mm512_set1_epi32(shift));
                const __m512 p = _mm512_mul_ps(_mm512_cvtepi32_ps(I), _mm512_set1_ps(delta));
                _mm512_store_ps(ipq + 16 * h, p);
                const __m512 o = _mm512_loadu_ps(on + 16 * h);
                __m512 t = _mm512_sub_ps(p, _mm512_set1_ps(c1));
                t = _mm512_mul_ps(t, _mm512_set1_ps(c2));
                t = _mm512_mul_ps(t, o);
                t = _mm512_add_ps(t, _mm512_set1_ps(qs));
                t = _mm512_add_ps(t, _mm512_mul_ps(o, o));
                mask |= (uint32_t)_mm512_cmp_ps_mask(t, _mm512_set1_ps(distk), _CMP_LT_OQ)
                        << (16 * h);
            }
            if (np < 32) mask &= (1u << np) - 1u;  // HASearcher.hpp:216, intended semantics
            while (mask) {
                const int j = __builtin_ctz(mask);
                mask &= mask - 1;
                const size_t vi = v0 + b * 32 + (size_t)j;
                const float ip = ip_ex<EX>(uq.data(), idx->long_code + vi * CL, D);
                // HASearcher.hpp:225-232 with C's float -> double promotions
                const float a = on[j] * on[j] + sqr_y;
                const float bb = idx->ex_factor[vi] * y;
                const float cc = (float)R * ipq[j] + ip;
                const double g = (double)cc - ((double)R - 0.5) * (double)sumq;
                const double hh = (double)bb * g;
                pool.insert(idx->ids[vi], (float)((double)a - hh));
                distk = pool.distk();
            }
        }
    }

    // IVF::search (IVF.hpp:295-324): centroid distances, the nprobe nearest (ties by id),
    // clusters in ascending distance
    template <int EX>
    void search(const float* q, size_t k, size_t nprobe, uint32_t* ids, float* dists) {
        const size_t K = idx->K;
        for (size_t i = 0; i < K; i++) cand[i] = {l2sqr16(q, idx->centroids + i * D, D), (uint32_t)i};
        nprobe = std::min(nprobe, K);
        std::partial_sort(cand.begin(), cand.begin() + nprobe, cand.end());
        pool.reset(k);
        for (size_t p = 0; p < nprobe; p++) search_cluster<EX>(q, cand[p].second, cand[p].first);
        for (size_t i = 0; i < k; i++) {
            ids[i] = i < pool.size ? pool.ids[i] : 0u;
            if (dists) dists[i] = i < pool.size ? pool.dist[i] : FLT_MAX;
        }
    }

    void run(const float* q, size_t k, size_t nprobe, uint32_t* ids, float* dists) {
        switch (idx->EX) {
        case 2: search<2>(q, k, nprobe, ids, dists); break;
        case 3: search<3>(q, k, nprobe, ids, dists); break;
        case 4: search<4>(q, k, nprobe, ids, dists); break;
        case 6: search<6>(q, k, nprobe, ids, dists); break;
        case 7: search<7>(q, k, nprobe, ids, dists); break;
        default: search<8>(q, k, nprobe, ids, dists); break;
        }
    }
};

}  // namespace

extern "C" {

int c512_available(void) {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
           __builtin_cpu_supports("avx512dq") && __builtin_cpu_supports("avx512vl");
}

// Rotator::rotate on zero-padded queries (Rotator.hpp:60-63, test_search.cpp:83-92): out =
// q_pad[nq x D] * P[D x D], single thread (EIGEN_DONT_PARALLELIZE), four query rows and 64
// columns per register tile.  Returns the microseconds it took.
double c512_rotate(const float* q_pad, size_t nq, size_t D, const float* P, float* out) {
    const double t0 = now_us();
    for (size_t i0 = 0; i0 < nq; i0 += 4) {
        const size_t nr = std::min<size_t>(4, nq - i0);
        for (size_t j0 = 0; j0 < D; j0 += 64) {
            __m512 acc[4][4];
            for (int r = 0; r < 4; r++)
                for (int c = 0; c < 4; c++) acc[r][c] = _mm512_setzero_ps();
            for (size_t kk = 0; kk < D; kk++) {
                const float* prow = P + kk * D + j0;
                const __m512 p0 = _mm512_loadu_ps(prow), p1 = _mm512_loadu_ps(prow + 16),
                             p2 = _mm512_loadu_ps(prow + 32), p3 = _mm512_loadu_ps(prow + 48);
                for (size_t r = 0; r < nr; r++) {
                    const __m512 a = _mm512_set1_ps(q_pad[(i0 + r) * D + kk]);
                    acc[r][0] = _mm512_fmadd_ps(a, p0, acc[r][0]);
                    acc[r][1] = _mm512_fmadd_ps(a, p1, acc[r][1]);
                    acc[r][2] = _mm512_fmadd_ps(a, p2, acc[r][2]);
                    acc[r][3] = _mm512_fmadd_ps(a, p3, acc[r][3]);
                }
            }
            for (size_t r = 0; r < nr; r++)
                for (int c = 0; c < 4; c++)
                    _mm512_storeu_ps(out + (i0 + r) * D + j0 + 16 * c, acc[r][c]);
        }
    }
    return now_us() - t0;
}

// IVF::search for each of nq rotated queries.  threads <= 1: one thread, each search timed
// on its own as test_search.cpp:110-112 does, returning the summed microseconds; threads > 1:
// OpenMP over queries, returning the wall-clock microseconds.
double c512_search_batch(const orc_index* idx, const float* q_rot, size_t nq, size_t k,
                         size_t nprobe, uint32_t* ids, float* dists, int threads) {
    if (threads <= 1) {
        Searcher s(idx);
        double total = 0;
        for (size_t i = 0; i < nq; i++) {
            const double t0 = now_us();
            s.run(q_rot + i * idx->D, k, nprobe, ids + i * k, dists ? dists + i * k : nullptr);
            total += now_us() - t0;
        }
        return total;
    }
    const double t0 = now_us();
#pragma omp parallel num_threads(threads)
    {
        Searcher s(idx);
#pragma omp for schedule(dynamic, 4)
        for (long i = 0; i < (long)nq; i++)
            s.run(q_rot + i * idx->D, k, nprobe, ids + i * k, dists ? dists + i * k : nullptr);
    }
    return now_us() - t0;
}

// get_ratio (tools.hpp:16-51) of one query: the k true and the k returned squared distances,
// each sorted, paired largest with largest; the mean over pairs with a true distance above
// 1e-5 of sqrt(returned / true), times k (k when no pair counts).
double c512_ratio(const float* gt_d, const float* ann_d, size_t k) {
    std::vector<float> g(gt_d, gt_d + k), a(ann_d, ann_d + k);
    std::sort(g.begin(), g.end());
    std::sort(a.begin(), a.end());
    double ret = 0;
    size_t valid = 0;
    for (size_t i = k; i-- > 0;)
        if (g[i] > 1e-5f) {
            ret += std::sqrt((double)a[i] / g[i]);
            ++valid;
        }
    return valid ? ret / valid * k : 1.0 * k;
}

}  // extern "C"
