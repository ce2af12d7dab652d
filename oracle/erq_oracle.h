/*
 * erq_oracle.h -- CPU restatement of the Extended-RaBitQ "high-accuracy FastScan"
 * IVF search path (HASearcher), used ONLY as the checker of the HIP implementation.
 *
 * TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library.  The product path never calls it.
 *
 * Floating point follows the reference source literally (IEEE single/double ops in
 * source order, FMA only where the source writes an FMA intrinsic, C promotion
 * rules for mixed float/double expressions).  Reductions reproduce
 * _mm512_reduce_add_ps's xor-8/4/2/1 butterfly over 16 lane-strided partial sums.
 *
 * Pinning status of each function is stated in oracle/README.md.
 */
#ifndef ERQ_ORACLE_H
#define ERQ_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* In-memory view of an HA-format IVF index (inc/index/IVF.hpp:28-40, :254-293). */
typedef struct {
    size_t N, DIM, D, K, EX;
    const uint64_t* cluster_sizes; /* K */
    const float* centroids;        /* K x D rotated centroids */
    const uint8_t* short_data;     /* sum_k ceil(n_k/32) blocks of (4D + 128) bytes */
    const uint8_t* long_code;      /* N x D*EX/8 */
    const float* ex_factor;        /* N xipnorm */
    const uint32_t* ids;           /* N */
    const uint64_t* block_start;   /* K+1: first block of cluster k (prefix sums of ceil(n/32)) */
    const uint64_t* vec_start;     /* K+1: first vector of cluster k (prefix sums of n) */
} orc_index;

float orc_reduce16(const float p[16]);
float orc_l2sqr16(const float* x, const float* y, size_t D);
float orc_normalize_query(float* unit_q, const float* q, const float* c, float norm, size_t D);
void orc_data_range(const float* q, size_t D, float* vl, float* vr);
void orc_quantize_query(int16_t* out, const float* u, size_t D, float* delta);
int32_t orc_pack_lut(uint16_t* lut, const int16_t* q, size_t D);
uint32_t orc_fastscan_block(const uint8_t* codes, const uint16_t* lut, const float* onorm,
                            float sumq, float qnorm, float delta, int32_t shift,
                            float one_over_sqrtD, float distk, size_t D,
                            float* ipq_out, int32_t* sum_out, float* lower_out);
int orc_ex_decode(const uint8_t* code, size_t D, int ex_bits, uint8_t* out);
float orc_ip_ex(const float* u, const uint8_t* code, size_t D, int ex_bits);
float orc_one_over_sqrtD(size_t D);

/* ResultPool (inc/index/Pool.hpp:10-50). */
typedef struct {
    uint32_t* ids;
    float* dist;
    size_t size, cap;
} orc_pool;
void orc_pool_init(orc_pool* p, uint32_t* ids_buf, float* dist_buf, size_t cap);
void orc_pool_insert(orc_pool* p, uint32_t id, float dist);
float orc_pool_distk(const orc_pool* p);

/* Centroid ranking (inc/index/Initializer.hpp:55-68). Ties broken by id. */
void orc_centroid_topn(const float* q, const float* centroids, size_t K, size_t D, size_t nprobe,
                       uint32_t* out_ids, float* out_dist);

/* Statistics of one search, for tests and the CPU baseline. */
typedef struct {
    uint64_t blocks, vectors, survivors, inserts;
} orc_stats;

/* IVF::search (inc/index/IVF.hpp:295-324) on a rotated query.  Writes k ids
 * (0-filled past the pool size) and, if dists != NULL, k distances (FLT_MAX-filled). */
void orc_ivf_search(const orc_index* idx, const float* q_rot, size_t k, size_t nprobe,
                    uint32_t* ids, float* dists, orc_stats* st);
/* Same, with an explicit probe list (cluster ids + squared distances) already ranked. */
void orc_ivf_search_probes(const orc_index* idx, const float* q_rot, size_t k, size_t nprobe,
                           const uint32_t* probe_ids, const float* probe_sqr_y,
                           uint32_t* ids, float* dists, orc_stats* st);
/* Round-synchronous batched schedule (see erq_oracle.c).  st (nullable): blocks, vectors
 * and survivors of rank 0 plus the survivors of the later rounds; `inserts` counts the
 * later rounds' candidates (estimated distance <= the round threshold). */
void orc_ivf_search_rounds(const orc_index* idx, const float* q_rot, size_t k, size_t nprobe,
                           const uint32_t* probe_ids, const float* probe_sqr_y, uint32_t* ids,
                           float* dists, orc_stats* st);
/* Batch over nq queries with OpenMP (threads<=0: all). */
void orc_ivf_search_batch(const orc_index* idx, const float* q_rot, size_t nq, size_t k,
                          const uint32_t* nprobe_per_q, uint32_t* ids, float* dists, int threads);

/* Float mode: 0 (default) the source's literal IEEE semantics; 1 the -Ofast rewrites
 * F1..F6 listed in erq_oracle.c, for measuring their effect.  Process-wide. */
void orc_set_fast_math(int on);
int orc_fast_math(void);
/* Lower bounds and re-estimated distances of every vector of cluster cid for one query
 * (arrays of cluster_sizes[cid] floats), under the current float mode. */
void orc_pair_estimates(const orc_index* idx, const float* q_rot, size_t cid, float sqr_y,
                        float* lower, float* exd);

/* Format helpers the reference's builder defines (pack_codes.hpp:33-98,
 * Quantizer.hpp:371-384, :491-631). */
void orc_pack_block_codes(const uint8_t* bits, size_t n, size_t D, uint8_t* block_codes);
int orc_ex_encode(const uint8_t* raw, size_t D, int ex_bits, uint8_t* out);

/* Dataset preparation (python/ivf.py:26-35, python/compute_gt.py:12-15) as this project
 * does it; see erq_oracle.c.  Row-major n x D floats (D any: zero padding implied). */
void orc_kmeans_assign(const float* x, size_t n, size_t D, const float* c, size_t K,
                       uint32_t* assign, float* dist);
void orc_kmeans_update(const float* x, size_t n, size_t D, const uint32_t* assign, size_t K,
                       float* c, uint32_t* counts);
void orc_kmeans(const float* x, size_t n, size_t D, size_t K, size_t iters,
                const uint32_t* init_rows, float* c, uint32_t* assign, float* dist);
double orc_l2sqr16_f64(const float* x, const float* y, size_t D);
void orc_knn(const float* base, size_t n, const float* q, size_t nq, size_t D, size_t k,
             uint32_t* ids, float* dists);

/* The non-HA searcher (Searcher.hpp; IVF::search without HIGH_ACC_FAST_SCAN): blocks of
 * 4D code bytes + 4 x 32 float factors (||o-c||^2, fac_ip, sumxb, fac_err). */
float orc_minus_sum16(const float* q, const float* c, float* r, size_t D);
void orc_scalar_quantize16(uint16_t* out, const float* r, float vl, float width, size_t D);
void orc_pack_lut_u16(uint16_t* lut, const uint16_t* q, size_t D);
void orc_nonha_block_ints(const uint8_t* codes, const uint16_t* lut, size_t D, int32_t* out);
void orc_nonha_search_probes(const orc_index* idx, const float* q_rot, size_t k, size_t nprobe,
                             const uint32_t* probe_ids, const float* probe_sqr_y, uint32_t* ids,
                             float* dists, orc_stats* st);
void orc_nonha_search_batch(const orc_index* idx, const float* q_rot, size_t nq, size_t k,
                            size_t nprobe, uint32_t* ids, float* dists, int threads);

#ifdef __cplusplus
}
#endif
#endif
