/*
 * erq_oracle.c -- CPU restatement of the Extended-RaBitQ HA IVF search path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the HIP product never calls into it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math), so every float
 * operation below is one IEEE-754 operation in the order written.  Citations are
 * reference file:line.  Pinning of each function is listed in oracle/README.md.
 */
#include "erq_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Fast-math variant (orc_set_fast_math(1)): the float rewrites GCC's -Ofast (the reference's
 * CMakeLists.txt:21) is free to apply to the path, applied here one by one so their effect on
 * the results can be measured against the literal source semantics (the default, 0):
 *   F1 1/norm in float instead of a double quotient (normalize_query16, space.hpp:763)
 *   F2 vmax / 8191 as vmax * (1/8191), u / width as u * (1/width) (space.hpp:705-708)
 *   F3 1/sqrt(D) as an approximate reciprocal square root (12 bits) plus one Newton step
 *   F4 (-5 * qnorm) * one_over_sqrtD reassociated as -5 * (qnorm * one_over_sqrtD)
 *   F5 the lower bound t*onorm + qs + onorm*onorm as fma(onorm, fma(t, c2, onorm), qs)
 *   F6 the re-estimated distance with its multiply-adds fused (HASearcher.hpp:225-232) */
static int g_fast = 0;
void orc_set_fast_math(int on) { g_fast = on; }
int orc_fast_math(void) { return g_fast; }

/* _mm512_reduce_add_ps (used by space.hpp:30, :120, :172, :774): halves added
 * pairwise -- lanes i and i+8, then j and j+4, then k and k+2, then 0 and 1. */
float orc_reduce16(const float p[16]) {
    float t3[8], t6[4], t8[2];
    for (int i = 0; i < 8; i++) t3[i] = p[i + 8] + p[i];
    for (int j = 0; j < 4; j++) t6[j] = t3[j + 4] + t3[j];
    t8[0] = t6[0] + t6[2];
    t8[1] = t6[1] + t6[3];
    return t8[0] + t8[1];
}

/* L2Sqr16, space.hpp:18-31: 16 lane-strided FMA chains, then reduce16. */
float orc_l2sqr16(const float* x, const float* y, size_t D) {
    float p[16] = {0};
    for (size_t i = 0; i < D; i += 16)
        for (int l = 0; l < 16; l++) {
            float t = x[i + l] - y[i + l];
            p[l] = fmaf(t, t, p[l]);
        }
    return orc_reduce16(p);
}

/* normalize_query16, space.hpp:749-778. */
float orc_normalize_query(float* unit_q, const float* q, const float* c, float norm, size_t D) {
    const float eps = 1e-5f;
    if (norm > eps) {
        float fac = g_fast ? 1.0f / norm : (float)(1.0 / (double)norm); /* F1 */
        float p[16] = {0};
        for (size_t i = 0; i < D; i += 16)
            for (int l = 0; l < 16; l++) {
                float u = (q[i + l] - c[i + l]) * fac;
                p[l] = p[l] + u;
                unit_q[i + l] = u;
            }
        return orc_reduce16(p);
    }
    float value = (float)(1.0 / (double)sqrtf((float)D));
    for (size_t i = 0; i < D; i++) unit_q[i] = value;
    return (float)D * value;
}

/* data_range16, space.hpp:608-628: running max/min start from 0. */
void orc_data_range(const float* q, size_t D, float* vl, float* vr) {
    float lo = 0.0f, hi = 0.0f;
    for (size_t i = 0; i < D; i++) {
        if (q[i] > hi) hi = q[i];
        if (q[i] < lo) lo = q[i];
    }
    *vl = lo;
    *vr = hi;
}

/* high_acc_quantize16, space.hpp:695-711: 14-bit signed, round half away from 0. */
void orc_quantize_query(int16_t* out, const float* u, size_t D, float* delta) {
    float vl, vr;
    orc_data_range(u, D, &vl, &vr);
    float a = fabsf(vl), b = fabsf(vr);
    float vmax = a < b ? b : a;
    float width = g_fast ? vmax * (1.0f / (float)((1 << 13) - 1)) /* F2 */
                         : vmax / (float)((1 << 13) - 1);
    const float rwidth = 1.0f / width;
    for (size_t i = 0; i < D; i++) {
        float tmp = g_fast ? u[i] * rwidth : u[i] / width; /* F2 */
        double r = (double)tmp + 0.5 - (double)(tmp < 0.0f);
        out[i] = (int16_t)(int)r;
    }
    *delta = width;
}

/* pack_high_acc_LUT, HASearcher.hpp:128-185.  Entry j of codebook m is
 * sum_k bit(j, 3-k) * q[4m+k] shifted by the codebook minimum so it is >= 0; the
 * reference stores its low and high bytes in two tables, here kept as one u16. */
int32_t orc_pack_lut(uint16_t* lut, const int16_t* q, size_t D) {
    /* pos[j]: dimension (within the 4) of the lowest set bit of j, HASearcher.hpp:130-147 */
    static const int pos[16] = {3, 3, 2, 3, 1, 3, 2, 3, 0, 3, 2, 3, 1, 3, 2, 3};
    size_t M = D >> 2;
    int32_t shift = 0;
    for (size_t m = 0; m < M; m++) {
        int L[16];
        int vmin = 0;
        L[0] = 0;
        for (int j = 1; j < 16; j++) {
            L[j] = L[j - (j & -j)] + q[4 * m + pos[j]];
            if (L[j] < vmin) vmin = L[j];
        }
        for (int j = 0; j < 16; j++) lut[m * 16 + j] = (uint16_t)(L[j] - vmin);
        shift += vmin;
    }
    return shift;
}

/* Byte position j (0..15) inside a codebook's 16 bytes that holds vector v%16:
 * pack_codes.hpp:44-65 writes byte j = code(perm0[j]) | code(perm0[j]+16) << 4 with
 * perm0 = {0,8,1,9,...,7,15}, so vector w<8 sits at j=2w and w+8 at j=2w+1. */
static inline int byte_of_vec(int v) {
    int w = v & 15;
    return w < 8 ? 2 * w : 2 * (w - 8) + 1;
}

/* Nibble of codebook m for vector v in a 4D-byte block (pack_codes.hpp:50-68). */
static inline int block_nibble(const uint8_t* codes, size_t m, int v) {
    const uint8_t* cb = codes + (m >> 1) * 32 + (m & 1) * 16;
    uint8_t b = cb[byte_of_vec(v)];
    return v < 16 ? (b & 15) : (b >> 4);
}

float orc_one_over_sqrtD(size_t D) {
    /* HASearcher.hpp:56 -- 1.0 / std::sqrt((float)D), double quotient stored as float */
    if (g_fast) { /* F3: a 12-bit estimate (rsqrtps' accuracy), one Newton-Raphson step */
        const float x = (float)D;
        float r0 = 1.0f / sqrtf(x);
        uint32_t b;
        memcpy(&b, &r0, 4);
        b &= ~((1u << 11) - 1u); /* keep 12 significant bits */
        memcpy(&r0, &b, 4);
        return r0 * (1.5f - 0.5f * x * r0 * r0);
    }
    return (float)(1.0 / (double)sqrtf((float)D));
}

/* accumulate_one_block_high_acc, FastScan.hpp:214-358.
 * Integer part: the reference adds low and high LUT bytes in u16 lanes; codebooks
 * m = 0,2 (mod 4) share one 128-bit half and m = 1,3 the other (:237-257), each half
 * wraps mod 2^16 before the two halves are widened and added in int32 (:264-290):
 *   S = (lo_even mod 2^16) + (lo_odd mod 2^16) + 256*((hi_even mod 2^16)+(hi_odd mod 2^16)).
 * Float part (:295-323), literally:
 *   ipq = float(S + shift) * delta
 *   lower = ((ipq - c1) * c2) * onorm + qnorm^2 + onorm*onorm,  bit set iff lower < distk
 * with c1 = float(0.5*double(sumq) - double(0.58f)), c2 = (-5*qnorm)*one_over_sqrtD. */
uint32_t orc_fastscan_block(const uint8_t* codes, const uint16_t* lut, const float* onorm,
                            float sumq, float qnorm, float delta, int32_t shift,
                            float one_over_sqrtD, float distk, size_t D,
                            float* ipq_out, int32_t* sum_out, float* lower_out) {
    size_t M = D >> 2;
    const float const_bound = 0.58f;
    float c1 = (float)(0.5 * (double)sumq - (double)const_bound);
    float c2 = g_fast ? -5.0f * (qnorm * one_over_sqrtD) /* F4 */
                      : (-5.0f * qnorm) * one_over_sqrtD;
    float qs = qnorm * qnorm;
    uint32_t mask = 0;
    for (int v = 0; v < 32; v++) {
        uint32_t lo[2] = {0, 0}, hi[2] = {0, 0};
        for (size_t m = 0; m < M; m++) {
            uint16_t e = lut[m * 16 + block_nibble(codes, m, v)];
            lo[m & 1] += e & 0xFFu;
            hi[m & 1] += e >> 8;
        }
        int32_t s = (int32_t)((lo[0] & 0xFFFFu) + (lo[1] & 0xFFFFu)) +
                    (int32_t)(((hi[0] & 0xFFFFu) + (hi[1] & 0xFFFFu)) << 8);
        int32_t I = s + shift;
        float ipq = (float)I * delta;
        float t = ipq - c1;
        if (g_fast) { /* F5 */
            t = fmaf(onorm[v], fmaf(t, c2, onorm[v]), qs);
        } else {
            t = t * c2;
            t = t * onorm[v];
            t = t + qs;
            t = t + onorm[v] * onorm[v];
        }
        if (t < distk) mask |= 1u << v;
        if (ipq_out) ipq_out[v] = ipq;
        if (sum_out) sum_out[v] = s;
        if (lower_out) lower_out[v] = t;
    }
    return mask;
}

/* Ex-code layouts written by DataQuantizer::store_compacted_code
 * (Quantizer.hpp:491-631) and read by IP16_fxu8 / IP32_fxu4 / IP64_fxu{2,3,6,7}
 * (space.hpp:109-131, 152-173, 247-497).  Decodes to one byte per dimension. */
int orc_ex_decode(const uint8_t* code, size_t D, int ex_bits, uint8_t* out) {
    size_t d;
    switch (ex_bits) {
    case 8:
        memcpy(out, code, D);
        return 0;
    case 4:
        for (d = 0; d < D; d += 32, code += 16)
            for (int j = 0; j < 16; j++) {
                out[d + j] = code[j] & 15;
                out[d + 16 + j] = code[j] >> 4;
            }
        return 0;
    case 2:
    case 3:
        for (d = 0; d < D; d += 64) {
            for (int j = 0; j < 16; j++)
                for (int s = 0; s < 4; s++) out[d + 16 * s + j] = (code[j] >> (2 * s)) & 3;
            code += 16;
            if (ex_bits == 3) {
                uint64_t T;
                memcpy(&T, code, 8);
                for (int t = 0; t < 64; t++)
                    out[d + t] |= (uint8_t)(((T >> (8 * (t & 7) + (t >> 3))) & 1u) << 2);
                code += 8;
            }
        }
        return 0;
    case 6:
    case 7:
        for (d = 0; d < D; d += 64) {
            const uint8_t *b0 = code, *b1 = code + 16, *b2 = code + 32;
            for (int j = 0; j < 16; j++) {
                out[d + j] = b0[j] & 63;
                out[d + 16 + j] = b1[j] & 63;
                out[d + 32 + j] = (uint8_t)((b2[j] & 15) | ((b0[j] >> 6) << 4));
                out[d + 48 + j] = (uint8_t)((b2[j] >> 4) | ((b1[j] >> 6) << 4));
            }
            code += 48;
            if (ex_bits == 7) {
                uint64_t T;
                memcpy(&T, code, 8);
                for (int t = 0; t < 64; t++)
                    out[d + t] |= (uint8_t)(((T >> (8 * (t & 7) + (t >> 3))) & 1u) << 6);
                code += 8;
            }
        }
        return 0;
    default:
        return -1;
    }
}

/* Inverse of orc_ex_decode: the layout store_compacted_code writes. */
int orc_ex_encode(const uint8_t* raw, size_t D, int ex_bits, uint8_t* out) {
    size_t d;
    switch (ex_bits) {
    case 8:
        memcpy(out, raw, D);
        return 0;
    case 4:
        for (d = 0; d < D; d += 32, out += 16)
            for (int j = 0; j < 16; j++) out[j] = (uint8_t)(raw[d + j] | (raw[d + 16 + j] << 4));
        return 0;
    case 2:
    case 3:
        for (d = 0; d < D; d += 64) {
            for (int j = 0; j < 16; j++) {
                uint8_t b = 0;
                for (int s = 0; s < 4; s++) b |= (uint8_t)((raw[d + 16 * s + j] & 3) << (2 * s));
                out[j] = b;
            }
            out += 16;
            if (ex_bits == 3) {
                uint64_t T = 0;
                for (int t = 0; t < 64; t++)
                    T |= (uint64_t)((raw[d + t] >> 2) & 1) << (8 * (t & 7) + (t >> 3));
                memcpy(out, &T, 8);
                out += 8;
            }
        }
        return 0;
    case 6:
    case 7:
        for (d = 0; d < D; d += 64) {
            for (int j = 0; j < 16; j++) {
                out[j] = (uint8_t)((raw[d + j] & 63) | (((raw[d + 32 + j] >> 4) & 3) << 6));
                out[16 + j] = (uint8_t)((raw[d + 16 + j] & 63) | (((raw[d + 48 + j] >> 4) & 3) << 6));
                out[32 + j] = (uint8_t)((raw[d + 32 + j] & 15) | ((raw[d + 48 + j] & 15) << 4));
            }
            out += 48;
            if (ex_bits == 7) {
                uint64_t T = 0;
                for (int t = 0; t < 64; t++)
                    T |= (uint64_t)((raw[d + t] >> 6) & 1) << (8 * (t & 7) + (t >> 3));
                memcpy(out, &T, 8);
                out += 8;
            }
        }
        return 0;
    default:
        return -1;
    }
}

/* IP*_fxu*: sum over d of u[d]*e[d], lane l = d mod 16 accumulates by FMA in
 * increasing d (every variant walks 16-dim groups in order), then reduce16. */
float orc_ip_ex(const float* u, const uint8_t* code, size_t D, int ex_bits) {
    uint8_t e[8192];
    if (D > sizeof(e) || orc_ex_decode(code, D, ex_bits, e) != 0) return NAN;
    float p[16] = {0};
    for (size_t i = 0; i < D; i += 16)
        for (int l = 0; l < 16; l++) p[l] = fmaf(u[i + l], (float)e[i + l], p[l]);
    return orc_reduce16(p);
}

/* ResultPool, Pool.hpp:10-50. */
void orc_pool_init(orc_pool* p, uint32_t* ids_buf, float* dist_buf, size_t cap) {
    p->ids = ids_buf;
    p->dist = dist_buf;
    p->size = 0;
    p->cap = cap;
}

void orc_pool_insert(orc_pool* p, uint32_t id, float dist) {
    if (p->size == p->cap && dist > p->dist[p->size - 1]) return;
    /* find_bsearch, Pool.hpp:40-49: number of stored distances strictly below dist */
    size_t lo = 0, len = p->size;
    while (len > 1) {
        size_t half = len >> 1;
        len -= half;
        lo += (p->dist[lo + half - 1] < dist) * half;
    }
    if (lo < p->size && p->dist[lo] < dist) lo++;
    size_t n = p->size - lo; /* entries [lo, size) move up; buffers hold cap+1 */
    memmove(&p->ids[lo + 1], &p->ids[lo], n * sizeof(uint32_t));
    memmove(&p->dist[lo + 1], &p->dist[lo], n * sizeof(float));
    p->ids[lo] = id;
    p->dist[lo] = dist;
    p->size += (p->size < p->cap);
}

float orc_pool_distk(const orc_pool* p) {
    return p->size == p->cap ? p->dist[p->size - 1] : FLT_MAX;
}

typedef struct {
    uint32_t id;
    float d;
} cand;

static int cand_cmp(const void* a, const void* b) {
    const cand *x = (const cand*)a, *y = (const cand*)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->id > y->id) - (x->id < y->id);
}

/* FlatInitializer::centroids_distances, Initializer.hpp:55-68 (partial_sort; equal
 * distances are ordered by id here -- the reference leaves their order unspecified). */
void orc_centroid_topn(const float* q, const float* centroids, size_t K, size_t D, size_t nprobe,
                       uint32_t* out_ids, float* out_dist) {
    cand* c = (cand*)malloc(K * sizeof(cand));
    for (size_t i = 0; i < K; i++) {
        c[i].id = (uint32_t)i;
        c[i].d = orc_l2sqr16(q, centroids + i * D, D);
    }
    qsort(c, K, sizeof(cand), cand_cmp);
    for (size_t i = 0; i < nprobe && i < K; i++) {
        out_ids[i] = c[i].id;
        out_dist[i] = c[i].d;
    }
    free(c);
}

/* The re-estimated distance, HASearcher.hpp:225-232, with C's float -> double promotions:
 * the bracket becomes double at "- (FAC_RESCALE - 0.5) * sumq", the product at "* (...)".
 * F6: its multiply-adds fused. */
static float ex_distance(float onorm, float sqr_y, float xip, float y, int R, float ipq, float ip,
                         float sumq) {
    if (g_fast) {
        const float cc = fmaf((float)R, ipq, ip);
        const double g = fma(-((double)R - 0.5), (double)sumq, (double)cc);
        return (float)fma(-(double)(xip * y), g, (double)fmaf(onorm, onorm, sqr_y));
    }
    float a = onorm * onorm + sqr_y;
    float bb = xip * y;
    float cc = (float)R * ipq + ip;
    double g = (double)cc - ((double)R - 0.5) * (double)sumq;
    double h = (double)bb * g;
    return (float)((double)a - h);
}

/* HASearcher::search_cluster + scan_one_block, HASearcher.hpp:82-112, :187-237. */
static void search_cluster(const orc_index* idx, size_t cid, const float* q, float sqr_y,
                           float one_over_sqrtD, float* unit_q, int16_t* qq, uint16_t* lut,
                           orc_pool* pool, orc_stats* st) {
    size_t D = idx->D;
    size_t n = (size_t)idx->cluster_sizes[cid];
    const float* c = idx->centroids + cid * D;
    float y = sqrtf(sqr_y);
    float sumq = orc_normalize_query(unit_q, q, c, y, D);
    float delta;
    orc_quantize_query(qq, unit_q, D, &delta);
    int32_t shift = orc_pack_lut(lut, qq, D);
    int R = 1 << idx->EX;
    size_t block_bytes = 4 * D + 128;
    size_t code_len = D * idx->EX / 8;
    const uint8_t* block = idx->short_data + idx->block_start[cid] * block_bytes;
    size_t v0 = idx->vec_start[cid];
    float distk = orc_pool_distk(pool);
    size_t nblocks = (n + 31) / 32;
    float ipq[32];
    for (size_t b = 0; b < nblocks; b++, block += block_bytes) {
        size_t np = n - b * 32 < 32 ? n - b * 32 : 32;
        const float* onorm = (const float*)(block + 4 * D);
        uint32_t mask = orc_fastscan_block(block, lut, onorm, sumq, y, delta, shift,
                                           one_over_sqrtD, distk, D, ipq, NULL, NULL);
        /* HASearcher.hpp:216 -- intended semantics for a full block: keep all 32 */
        if (np < 32) mask &= (1u << np) - 1u;
        if (st) {
            st->blocks++;
            st->vectors += np;
        }
        while (mask) {
            int j = __builtin_ctz(mask);
            mask &= mask - 1;
            size_t vi = v0 + b * 32 + (size_t)j;
            float ip = orc_ip_ex(unit_q, idx->long_code + vi * code_len, D, (int)idx->EX);
            float ex_dist = ex_distance(onorm[j], sqr_y, idx->ex_factor[vi], y, R, ipq[j], ip,
                                        sumq);
            orc_pool_insert(pool, idx->ids[vi], ex_dist);
            distk = orc_pool_distk(pool);
            if (st) st->survivors++;
        }
    }
}

/* Every vector of cluster cid against one query: the lower bound (FastScan.hpp:295-323) and
 * the re-estimated distance (HASearcher.hpp:225-232), both under the current float mode. */
void orc_pair_estimates(const orc_index* idx, const float* q_rot, size_t cid, float sqr_y,
                        float* lower, float* exd) {
    size_t D = idx->D;
    size_t n = (size_t)idx->cluster_sizes[cid];
    float* unit_q = (float*)malloc(D * sizeof(float));
    int16_t* qq = (int16_t*)malloc(D * sizeof(int16_t));
    uint16_t* lut = (uint16_t*)malloc((D / 4) * 16 * sizeof(uint16_t));
    const float* c = idx->centroids + cid * D;
    float y = sqrtf(sqr_y);
    float sumq = orc_normalize_query(unit_q, q_rot, c, y, D);
    float delta;
    orc_quantize_query(qq, unit_q, D, &delta);
    int32_t shift = orc_pack_lut(lut, qq, D);
    int R = 1 << idx->EX;
    size_t block_bytes = 4 * D + 128, code_len = D * idx->EX / 8;
    const uint8_t* block = idx->short_data + idx->block_start[cid] * block_bytes;
    size_t v0 = idx->vec_start[cid];
    float ipq[32], lw[32];
    for (size_t b = 0; b < (n + 31) / 32; b++, block += block_bytes) {
        size_t np = n - b * 32 < 32 ? n - b * 32 : 32;
        const float* onorm = (const float*)(block + 4 * D);
        orc_fastscan_block(block, lut, onorm, sumq, y, delta, shift, orc_one_over_sqrtD(D),
                           0.0f, D, ipq, NULL, lw);
        for (size_t j = 0; j < np; j++) {
            size_t vi = v0 + b * 32 + j;
            float ip = orc_ip_ex(unit_q, idx->long_code + vi * code_len, D, (int)idx->EX);
            lower[b * 32 + j] = lw[j];
            exd[b * 32 + j] = ex_distance(onorm[j], sqr_y, idx->ex_factor[vi], y, R, ipq[j], ip,
                                          sumq);
        }
    }
    free(unit_q);
    free(qq);
    free(lut);
}

void orc_ivf_search_probes(const orc_index* idx, const float* q_rot, size_t k, size_t nprobe,
                           const uint32_t* probe_ids, const float* probe_sqr_y,
                           uint32_t* ids, float* dists, orc_stats* st) {
    size_t D = idx->D;
    float* unit_q = (float*)malloc(D * sizeof(float));
    int16_t* qq = (int16_t*)malloc(D * sizeof(int16_t));
    uint16_t* lut = (uint16_t*)malloc((D / 4) * 16 * sizeof(uint16_t));
    uint32_t* pid = (uint32_t*)malloc((k + 1) * sizeof(uint32_t));
    float* pd = (float*)malloc((k + 1) * sizeof(float));
    orc_pool pool;
    orc_pool_init(&pool, pid, pd, k);
    float oos = orc_one_over_sqrtD(D);
    for (size_t i = 0; i < nprobe; i++)
        search_cluster(idx, probe_ids[i], q_rot, probe_sqr_y[i], oos, unit_q, qq, lut, &pool, st);
    /* ResultPool::copy_results, Pool.hpp:33: k ids; unused slots are zero here */
    for (size_t i = 0; i < k; i++) {
        ids[i] = i < pool.size ? pid[i] : 0u;
        if (dists) dists[i] = i < pool.size ? pd[i] : FLT_MAX;
    }
    free(unit_q);
    free(qq);
    free(lut);
    free(pid);
    free(pd);
}

/* Round-synchronous batched search (this project's cluster-major schedule, checked
 * against liberq's erq_scan_batched_dev): probe ranks [0,4) exactly as IVF::search; ranks
 * [4,16), [16,64), [64,192), then rounds 128 ranks wide; each round's pruning threshold is
 * the pool's k-th distance at round start; a vector is a candidate if its lower bound <
 * threshold and its estimate <= threshold; candidates are merged into the pool by
 * (distance, id). */
static int dist_id_cmp(const void* a, const void* b) {
    const cand *x = (const cand*)a, *y = (const cand*)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->id > y->id) - (x->id < y->id);
}

void orc_ivf_search_rounds(const orc_index* idx, const float* q_rot, size_t k, size_t nprobe,
                           const uint32_t* probe_ids, const float* probe_sqr_y, uint32_t* ids,
                           float* dists, orc_stats* st) {
    size_t D = idx->D;
    float* unit_q = (float*)malloc(D * sizeof(float));
    int16_t* qq = (int16_t*)malloc(D * sizeof(int16_t));
    uint16_t* lut = (uint16_t*)malloc((D / 4) * 16 * sizeof(uint16_t));
    uint32_t* pid = (uint32_t*)malloc((k + 1) * sizeof(uint32_t));
    float* pd = (float*)malloc((k + 1) * sizeof(float));
    orc_pool pool;
    orc_pool_init(&pool, pid, pd, k);
    float oos = orc_one_over_sqrtD(D);
    /* ranks [0, 4): the exact replay (IVF::search) */
    size_t exact = nprobe < 4 ? nprobe : 4;
    for (size_t r = 0; r < exact; r++)
        search_cluster(idx, probe_ids[r], q_rot, probe_sqr_y[r], oos, unit_q, qq, lut, &pool, st);
    size_t ncap = 1024, nc = 0;
    cand* cs = (cand*)malloc(ncap * sizeof(cand));
    int R = 1 << idx->EX;
    size_t block_bytes = 4 * D + 128, code_len = D * idx->EX / 8;
    for (size_t r0 = 4, r1; r0 < nprobe; r0 = r1) {
        r1 = 4 * r0 < r0 + 128 ? 4 * r0 : r0 + 128; /* [4,16) [16,64) [64,192) then 128 wide */
        float thr = orc_pool_distk(&pool);
        nc = 0;
        for (size_t i = 0; i < pool.size; i++) {
            if (nc == ncap) cs = (cand*)realloc(cs, (ncap *= 2) * sizeof(cand));
            cs[nc].id = pid[i];
            cs[nc++].d = pd[i];
        }
        for (size_t r = r0; r < r1 && r < nprobe; r++) {
            size_t cid = probe_ids[r];
            float sqr_y = probe_sqr_y[r];
            size_t n = (size_t)idx->cluster_sizes[cid];
            const float* c = idx->centroids + cid * D;
            float y = sqrtf(sqr_y);
            float sumq = orc_normalize_query(unit_q, q_rot, c, y, D);
            float delta;
            orc_quantize_query(qq, unit_q, D, &delta);
            int32_t shift = orc_pack_lut(lut, qq, D);
            const uint8_t* block = idx->short_data + idx->block_start[cid] * block_bytes;
            size_t v0 = idx->vec_start[cid];
            float ipq[32];
            for (size_t b = 0; b < (n + 31) / 32; b++, block += block_bytes) {
                size_t np = n - b * 32 < 32 ? n - b * 32 : 32;
                const float* onorm = (const float*)(block + 4 * D);
                uint32_t mask = orc_fastscan_block(block, lut, onorm, sumq, y, delta, shift, oos,
                                                   thr, D, ipq, NULL, NULL);
                if (np < 32) mask &= (1u << np) - 1u;
                if (st) st->survivors += (uint64_t)__builtin_popcount(mask);
                while (mask) {
                    int j = __builtin_ctz(mask);
                    mask &= mask - 1;
                    size_t vi = v0 + b * 32 + (size_t)j;
                    float ip = orc_ip_ex(unit_q, idx->long_code + vi * code_len, D, (int)idx->EX);
                    float ex = ex_distance(onorm[j], sqr_y, idx->ex_factor[vi], y, R, ipq[j], ip,
                                           sumq);
                    if (ex <= thr) {
                        if (st) st->inserts++;
                        if (nc == ncap) cs = (cand*)realloc(cs, (ncap *= 2) * sizeof(cand));
                        cs[nc].id = idx->ids[vi];
                        cs[nc++].d = ex;
                    }
                }
            }
        }
        qsort(cs, nc, sizeof(cand), dist_id_cmp);
        pool.size = nc < k ? nc : k;
        for (size_t i = 0; i < pool.size; i++) {
            pid[i] = cs[i].id;
            pd[i] = cs[i].d;
        }
    }
    for (size_t i = 0; i < k; i++) {
        ids[i] = i < pool.size ? pid[i] : 0u;
        if (dists) dists[i] = i < pool.size ? pd[i] : FLT_MAX;
    }
    free(cs);
    free(unit_q);
    free(qq);
    free(lut);
    free(pid);
    free(pd);
}

void orc_ivf_search(const orc_index* idx, const float* q_rot, size_t k, size_t nprobe,
                    uint32_t* ids, float* dists, orc_stats* st) {
    if (nprobe > idx->K) nprobe = idx->K;
    uint32_t* pids = (uint32_t*)malloc(nprobe * sizeof(uint32_t) + 1);
    float* pdist = (float*)malloc(nprobe * sizeof(float) + 1);
    orc_centroid_topn(q_rot, idx->centroids, idx->K, idx->D, nprobe, pids, pdist);
    orc_ivf_search_probes(idx, q_rot, k, nprobe, pids, pdist, ids, dists, st);
    free(pids);
    free(pdist);
}

void orc_ivf_search_batch(const orc_index* idx, const float* q_rot, size_t nq, size_t k,
                          const uint32_t* nprobe_per_q, uint32_t* ids, float* dists, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (long i = 0; i < (long)nq; i++)
        orc_ivf_search(idx, q_rot + (size_t)i * idx->D, k, nprobe_per_q[i], ids + (size_t)i * k,
                       dists ? dists + (size_t)i * k : NULL, NULL);
}

/* pack_codes (pack_codes.hpp:33-69) on the per-vector nibbles that pack_binary and the
 * byte/nibble swaps produce (Quantizer.hpp:371-384, pack_codes.hpp:75-98): nibble m of
 * a vector holds dims 4m..4m+3 with dim 4m at bit 3.  bits is n x D of 0/1, n <= 32. */
void orc_pack_block_codes(const uint8_t* bits, size_t n, size_t D, uint8_t* block_codes) {
    size_t M = D >> 2;
    memset(block_codes, 0, 4 * D);
    for (size_t m = 0; m < M; m++) {
        uint8_t* cb = block_codes + (m >> 1) * 32 + (m & 1) * 16;
        for (size_t v = 0; v < n; v++) {
            const uint8_t* x = bits + v * D + 4 * m;
            int nib = (x[0] << 3) | (x[1] << 2) | (x[2] << 1) | x[3];
            int j = byte_of_vec((int)v);
            cb[j] |= (uint8_t)(v < 16 ? nib : nib << 4);
        }
    }
}

/* ---------------------------------------------------------------- dataset preparation
 * The reference trains its IVF with faiss (python/ivf.py:26-35: IVF{K},Flat, then each
 * vector's nearest centroid) and its ground truth with numpy (python/compute_gt.py:12-15).
 * This project's trainer, restated: Lloyd k-means whose assignment is the L2Sqr16-nearest
 * centroid (ties by id, the ranking FlatInitializer applies at search time) and whose
 * update is each cluster's mean, summed in double in increasing vector id order and
 * rounded to float; an empty cluster keeps its centroid.  Rows shorter than a multiple of
 * 16 are taken as zero-padded (a zero pair adds nothing to a lane's chain). */
static float l2sqr16_any(const float* x, const float* y, size_t D) {
    float p[16] = {0};
    for (size_t i = 0; i < D; i += 16)
        for (int l = 0; l < 16 && i + l < D; l++) {
            float t = x[i + l] - y[i + l];
            p[l] = fmaf(t, t, p[l]);
        }
    return orc_reduce16(p);
}

void orc_kmeans_assign(const float* x, size_t n, size_t D, const float* c, size_t K,
                       uint32_t* assign, float* dist) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) {
        uint32_t best = 0;
        float bd = l2sqr16_any(x + (size_t)i * D, c, D);
        for (size_t j = 1; j < K; j++) {
            float d = l2sqr16_any(x + (size_t)i * D, c + j * D, D);
            if (d < bd) {
                bd = d;
                best = (uint32_t)j;
            }
        }
        assign[i] = best;
        dist[i] = bd;
    }
}

void orc_kmeans_update(const float* x, size_t n, size_t D, const uint32_t* assign, size_t K,
                       float* c, uint32_t* counts) {
    double* acc = (double*)calloc(K * D, sizeof(double));
    memset(counts, 0, K * sizeof(uint32_t));
    for (size_t i = 0; i < n; i++) {
        double* a = acc + (size_t)assign[i] * D;
        for (size_t d = 0; d < D; d++) a[d] = a[d] + (double)x[i * D + d];
        counts[assign[i]]++;
    }
    for (size_t j = 0; j < K; j++)
        if (counts[j])
            for (size_t d = 0; d < D; d++) c[j * D + d] = (float)(acc[j * D + d] / (double)counts[j]);
    free(acc);
}

void orc_kmeans(const float* x, size_t n, size_t D, size_t K, size_t iters,
                const uint32_t* init_rows, float* c, uint32_t* assign, float* dist) {
    uint32_t* counts = (uint32_t*)malloc(K * sizeof(uint32_t));
    for (size_t j = 0; j < K; j++) memcpy(c + j * D, x + (size_t)init_rows[j] * D, D * sizeof(float));
    for (size_t it = 0; it < iters; it++) {
        orc_kmeans_assign(x, n, D, c, K, assign, dist);
        orc_kmeans_update(x, n, D, assign, K, c, counts);
    }
    orc_kmeans_assign(x, n, D, c, K, assign, dist);
    free(counts);
}

/* The exact squared L2 distance of the ground truth: d_i = (double)x_i - (double)y_i (exact),
 * 16 lane-strided chains of fma(d, d, p) in double, reduced in _mm512_reduce_add_ps order. */
double orc_l2sqr16_f64(const float* x, const float* y, size_t D) {
    double p[16] = {0};
    for (size_t i = 0; i < D; i += 16)
        for (int l = 0; l < 16 && i + l < D; l++) {
            double t = (double)x[i + l] - (double)y[i + l];
            p[l] = fma(t, t, p[l]);
        }
    double t3[8], t6[4];
    for (int i = 0; i < 8; i++) t3[i] = p[i + 8] + p[i];
    for (int j = 0; j < 4; j++) t6[j] = t3[j + 4] + t3[j];
    return (t6[0] + t6[2]) + (t6[1] + t6[3]);
}

typedef struct {
    uint32_t id;
    double d;
} cand64;

static int cand64_cmp(const void* a, const void* b) {
    const cand64 *x = (const cand64*)a, *y = (const cand64*)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->id > y->id) - (x->id < y->id);
}

/* python/compute_gt.py:12-15: per query the k nearest base vectors by orc_l2sqr16_f64,
 * ordered by (distance, id); dists (nullable) the distances rounded to float. */
void orc_knn(const float* base, size_t n, const float* q, size_t nq, size_t D, size_t k,
             uint32_t* ids, float* dists) {
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < (long)nq; i++) {
        cand64* c = (cand64*)malloc(n * sizeof(cand64));
        for (size_t j = 0; j < n; j++) {
            c[j].id = (uint32_t)j;
            c[j].d = orc_l2sqr16_f64(q + (size_t)i * D, base + j * D, D);
        }
        qsort(c, n, sizeof(cand64), cand64_cmp);
        for (size_t j = 0; j < k; j++) {
            ids[(size_t)i * k + j] = j < n ? c[j].id : 0xFFFFFFFFu;
            if (dists) dists[(size_t)i * k + j] = j < n ? (float)c[j].d : FLT_MAX;
        }
        free(c);
    }
}

/* ---------------------------------------------------------------- the non-HA searcher
 * Searcher (Searcher.hpp:14-236), the searcher IVF::search uses when HIGH_ACC_FAST_SCAN is
 * not defined (IVF.hpp:307-311).  Its blocks hold four factors per vector (Quantizer.hpp:
 * 107-121): 4D code bytes, then ||o-c||^2, fac_ip, sumxb and fac_err, 32 floats each. */

/* minus_sum16, space.hpp:714-736: r = q - c, and the sum of r in 16 lane-strided chains
 * reduced by _mm512_reduce_add_ps. */
float orc_minus_sum16(const float* q, const float* c, float* r, size_t D) {
    float p[16] = {0};
    for (size_t i = 0; i < D; i += 16)
        for (int l = 0; l < 16; l++) {
            r[i + l] = q[i + l] - c[i + l];
            p[l] = p[l] + r[i + l];
        }
    return orc_reduce16(p);
}

/* scalar_quantize16<uint16_t, 16>, space.hpp:644-686 (AVX-512 branch): (r - vl) * f32(1 /
 * width) + 0.5, truncated to int32, then to 16 bits. */
void orc_scalar_quantize16(uint16_t* out, const float* r, float vl, float width, size_t D) {
    const float o = (float)(1.0 / (double)width);
    for (size_t i = 0; i < D; i++) {
        float t = (r[i] - vl) * o;
        t = t + 0.5f;
        out[i] = (uint16_t)(int32_t)t;
    }
}

/* Searcher::pack_LUT (Searcher.hpp:129-172): LUT_total[m][j] = sum of the digits of the dims
 * whose bits j holds (bit 3 = dim 4m), in u16. */
void orc_pack_lut_u16(uint16_t* lut, const uint16_t* q, size_t D) {
    for (size_t m = 0; m < D / 4; m++)
        for (int j = 0; j < 16; j++) {
            uint16_t s = 0;
            for (int b = 0; b < 4; b++)
                if (j & (8 >> b)) s = (uint16_t)(s + q[4 * m + b]);
            lut[m * 16 + j] = s;
        }
}

/* accumulate_robust (FastScan.hpp:136-212) over LUT_lower and LUT_upper, combined as
 * Searcher::scan_one_block does (Searcher.hpp:196-199): per byte plane, the even and the odd
 * codebooks are summed in u16 lanes (mod 2^16) and the two sums added in 32 bits; then
 * (upper << 8) + lower. */
void orc_nonha_block_ints(const uint8_t* codes, const uint16_t* lut, size_t D, int32_t* out) {
    for (int v = 0; v < 32; v++) {
        uint32_t lo[2] = {0, 0}, hi[2] = {0, 0};
        for (size_t m = 0; m < D / 4; m++) {
            const uint16_t e = lut[m * 16 + block_nibble(codes, m, v)];
            lo[m & 1] += e & 0xFFu;
            hi[m & 1] += e >> 8;
        }
        const uint32_t lower = (lo[0] & 0xFFFFu) + (lo[1] & 0xFFFFu);
        const uint32_t upper = (hi[0] & 0xFFFFu) + (hi[1] & 0xFFFFu);
        out[v] = (int32_t)((upper << 8) + lower);
    }
}

/* Searcher::search_cluster + scan_one_block (Searcher.hpp:85-236). */
static void nonha_search_cluster(const orc_index* idx, size_t cid, const float* q, float sqr_y,
                                 float* r, uint16_t* qq, uint16_t* lut, orc_pool* pool,
                                 orc_stats* st) {
    const size_t D = idx->D, n = (size_t)idx->cluster_sizes[cid];
    const float* c = idx->centroids + cid * D;
    /* preparing(): the residual, half its sum, the 14-bit unsigned digits, the LUT */
    const float half_sum = (float)(0.5 * (double)orc_minus_sum16(q, c, r, D));
    float vl, vr;
    orc_data_range(r, D, &vl, &vr);
    const float width = (vr - vl) / (float)((1 << 14) - 1);
    orc_scalar_quantize16(qq, r, vl, width, D);
    orc_pack_lut_u16(lut, qq, D);
    const int R = 1 << idx->EX;
    const size_t bb = 4 * D + 512, code_len = D * idx->EX / 8;
    const uint8_t* block = idx->short_data + idx->block_start[cid] * bb;
    const size_t v0 = idx->vec_start[cid];
    const float y = sqrtf(sqr_y);
    float distk = orc_pool_distk(pool);
    int32_t ints[32];
    for (size_t b = 0; b * 32 < n; b++, block += bb) {
        const size_t np = n - b * 32 < 32 ? n - b * 32 : 32;
        const float* x2 = (const float*)(block + 4 * D);
        const float* fip = x2 + 32;
        const float* sumxb = x2 + 64;
        const float* ferr = x2 + 96;
        orc_nonha_block_ints(block, lut, D, ints);
        float lower[32], rip[32];
        for (int j = 0; j < 32; j++) {
            const float sum_sqr = x2[j] + sqr_y;
            const float xbvl = sumxb[j] * vl;
            const float resf = (float)ints[j];
            float ip = resf * width + xbvl;
            ip = ip - half_sum;
            rip[j] = ip;
            ip = ip * fip[j];
            const float err = ferr[j] * y;
            lower[j] = sum_sqr - (ip + err);
        }
        if (st) {
            st->blocks++;
            st->vectors += np;
        }
        for (size_t j = 0; j < np; j++) {
            if (lower[j] > distk) continue;
            const size_t vi = v0 + b * 32 + j;
            const float ipx = orc_ip_ex(r, idx->long_code + vi * code_len, D, (int)idx->EX);
            const float ex_dist = x2[j] + sqr_y -
                                  idx->ex_factor[vi] *
                                      ((float)R * rip[j] + ipx - (float)(R - 1) * half_sum);
            orc_pool_insert(pool, idx->ids[vi], ex_dist);
            distk = orc_pool_distk(pool);
            if (st) st->survivors++;
        }
    }
}

void orc_nonha_search_probes(const orc_index* idx, const float* q_rot, size_t k, size_t nprobe,
                             const uint32_t* probe_ids, const float* probe_sqr_y, uint32_t* ids,
                             float* dists, orc_stats* st) {
    const size_t D = idx->D;
    float* r = (float*)malloc(D * sizeof(float));
    uint16_t* qq = (uint16_t*)malloc(D * sizeof(uint16_t));
    uint16_t* lut = (uint16_t*)malloc((D / 4) * 16 * sizeof(uint16_t));
    uint32_t* pid = (uint32_t*)malloc((k + 1) * sizeof(uint32_t));
    float* pd = (float*)malloc((k + 1) * sizeof(float));
    orc_pool pool;
    orc_pool_init(&pool, pid, pd, k);
    for (size_t i = 0; i < nprobe; i++)
        nonha_search_cluster(idx, probe_ids[i], q_rot, probe_sqr_y[i], r, qq, lut, &pool, st);
    for (size_t i = 0; i < k; i++) {
        ids[i] = i < pool.size ? pid[i] : 0u;
        if (dists) dists[i] = i < pool.size ? pd[i] : FLT_MAX;
    }
    free(r);
    free(qq);
    free(lut);
    free(pid);
    free(pd);
}

void orc_nonha_search_batch(const orc_index* idx, const float* q_rot, size_t nq, size_t k,
                            size_t nprobe, uint32_t* ids, float* dists, int threads) {
    if (nprobe > idx->K) nprobe = idx->K;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (long i = 0; i < (long)nq; i++) {
        uint32_t* pids = (uint32_t*)malloc(nprobe * sizeof(uint32_t) + 1);
        float* pdist = (float*)malloc(nprobe * sizeof(float) + 1);
        const float* q = q_rot + (size_t)i * idx->D;
        orc_centroid_topn(q, idx->centroids, idx->K, idx->D, nprobe, pids, pdist);
        orc_nonha_search_probes(idx, q, k, nprobe, pids, pdist, ids + (size_t)i * k,
                                dists ? dists + (size_t)i * k : NULL, NULL);
        free(pids);
        free(pdist);
    }
}
