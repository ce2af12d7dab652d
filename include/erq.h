/*
 * erq.h -- C-ABI of liberq.so, the MI355X (gfx950) Extended-RaBitQ IVF search library.
 *
 * The reference exposes a C++ class, IVF (inc/index/IVF.hpp:26-91); these entry points
 * are what a binding of that class's search path needs: load an HA index, rotate
 * queries, search.  Plain pointers and sizes only; every call returns ERQ_OK or an
 * error code, with a message in erq_last_error().  No exceptions cross the ABI.
 *
 * Pointer conventions: functions without a _dev suffix take host pointers and run on the
 * handle's own stream, synchronously; _dev functions take device pointers, enqueue on
 * `stream` exactly as given (a hipStream_t; 0 is the null stream) and return without
 * synchronising.  With ERQ_OPT_BATCH_OVERLAP set, the batched searches fork part of their
 * work onto a second stream of the handle and join it back into `stream` before returning
 * (events, so a HIP graph capture of `stream` takes both).
 * Thread safety: calls on one handle are serialised by a per-handle mutex, and a call on one
 * stream is ordered after the device work of the previous call on the handle when that ran
 * on another stream (an event each call records when it returns).  Calls captured into a HIP
 * graph are not ordered that way: the graph's launch orders them, and a replay must not
 * overlap another call on the same handle.
 */
#ifndef ERQ_H
#define ERQ_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ERQ_OK = 0,
    ERQ_EINVAL = 1,      /* bad argument (k == 0, nprobe > K, null pointer, ...) */
    ERQ_EIO = 2,         /* file missing / unreadable / truncated */
    ERQ_EFORMAT = 3,     /* header or sizes inconsistent */
    ERQ_EHIP = 4,        /* HIP runtime error */
    ERQ_ENOMEM = 5,      /* device allocation failed */
    ERQ_EUNSUPPORTED = 6 /* shape outside what the kernels handle */
};

typedef struct erq_index erq_index;

/* Options (erq_set_option). */
enum {
    /* Integer kernel of the per-query scan: 0 auto (= 2), 1 LUT, 2 int8 MFMA over the
     * bit-plane copy of the codes (above D = 2048 with the reference's u16 wraparound
     * reproduced for every vector the lower bound passes).  Results are identical; only
     * speed and the extra bit-plane copy (one more 1-bit code array in HBM) differ. */
    ERQ_OPT_PLAIN_KERNEL = 1,
    /* Queries (waves) per workgroup of the per-query scan, 1..4 (LDS permitting). */
    ERQ_OPT_PLAIN_WAVES = 2,
    /* Workgroups per CU of the persistent batched kernel, 1..8 (default 2: its LDS holds
     * two per CU at D <= 1536). */
    ERQ_OPT_BATCH_WG_PER_CU = 3,
    /* Records (16 bytes each) of the batched search's survivor list per round, 0..2^30
     * (default 2^26).  Survivors beyond it are re-estimated by an inline pass of the scan
     * instead of by the separate re-estimation kernel; results are identical. */
    ERQ_OPT_BATCH_SURVIVOR_CAP = 4,
    /* Order in which the per-query scan takes queries: 1 (default) by scan cost (blocks of
     * the probed clusters), heaviest first, so that light queries fill the last, partial
     * generation of waves; 0 as given.  Results are identical. */
    ERQ_OPT_PLAIN_ORDER = 5,
    /* (query, cluster) pairs the batched search prepares ahead, 1024..2^26 (default 2^23;
     * 3 KB of digits each at D = 1536).  Every round of a chunk of queries is prepared at
     * once, so a chunk's pairs over all ranks past the exact ones must fit it.  Results
     * are identical; only workspace size and speed differ. */
    ERQ_OPT_BATCH_PAIR_BUDGET = 6,
    /* 0 (default): the batched search prepares its rounds on the caller's stream before the
     * exact replay of ranks [0, 4); 1: on a second stream beside it.  Results are identical;
     * both are VALU-bound, so side by side they take about their sum (measured: 23.1 ms per
     * 20k-query learned search without the overlap, 23.5 ms with it). */
    ERQ_OPT_BATCH_OVERLAP = 7,
    /* Batches of at most this many queries (default 16; 0 off) through erq_search,
     * erq_search_batch, erq_search_batch_nprobe (per-query scan) and erq_search_dev take the
     * low-latency scan when D <= 2048: every vector of every probed cluster is estimated in
     * parallel across the GPU, then one wave per query replays its probes in order with the
     * reference's pruning threshold.  Results are identical to the per-query scan; it reads
     * every probed vector's EX-bit code (the per-query scan reads the survivors' only), so it
     * trades throughput for latency.  Its slot buffer (12 bytes per probed vector slot) is
     * allocated on first use, outside HIP graph capture. */
    ERQ_OPT_LATENCY_NQ = 8,
    /* 16-vector tiles per work item of the batched rounds: 16 (default; 8 blocks of 32
     * vectors) or 8.  Smaller items spread a large cluster over more workgroups; each item
     * reads its entries' digits once.  Results are identical. */
    ERQ_OPT_BATCH_ITEM_TILES = 9,
    /* Workgroups per CU of the batched search's preparation kernel, 1..6 (default 6, what its
     * LDS allows).  Results are identical. */
    ERQ_OPT_PREP_WG_PER_CU = 10
};
int erq_set_option(erq_index* h, int option, long value);
long erq_get_option(const erq_index* h, int option);

/* Autotuner: times the result-invariant launch shapes (ERQ_OPT_PLAIN_WAVES, then for the
 * batched search ERQ_OPT_BATCH_WG_PER_CU, ERQ_OPT_BATCH_ITEM_TILES and ERQ_OPT_PREP_WG_PER_CU,
 * each kept at its fastest before the next is swept) on these queries with HIP events and
 * keeps the fastest; best_ms_out (nullable, 2 floats) receives the best per-query and batched
 * times.  It holds the handle for the whole sweep. */
int erq_autotune(erq_index* h, const float* q_rot_dev, size_t nq, size_t k, size_t nprobe,
                 float* best_ms_out, void* stream);

/* Message of the last error on the calling thread ("" if none). */
const char* erq_last_error(void);

/* Replaces IVF::load (IVF.hpp:254-293): reads an index file and uploads it to the current
 * HIP device.  The file does not record its block format; the HA one (HIGH_ACC_FAST_SCAN,
 * 1 factor per vector) or the non-HA one (4 factors) is the one whose size the file has. */
int erq_index_load(const char* path, erq_index** out);

/* The same arrays IVF::load reads (IVF.hpp:261-286), from host memory. */
int erq_index_create(size_t N, size_t DIM, size_t K, size_t EX_BITS, const uint64_t* cluster_sizes,
                     const float* P, const float* centroids, const uint8_t* short_data,
                     const uint8_t* long_code, const float* ex_factor, const uint32_t* ids,
                     erq_index** out);

/* erq_index_create for the non-HA format: the index IVF::construct writes when built without
 * HIGH_ACC_FAST_SCAN (Quantizer.hpp:33-37,107-121), whose blocks hold 4D code bytes and four
 * factors per vector (||o-c||^2, fac_ip, sumxb, fac_err; 4D + 512 bytes per 32 vectors), and
 * which IVF::search scans with Searcher (Searcher.hpp:14-236).  The per-query searches
 * (erq_search, erq_search_batch, erq_search_batch_nprobe without ERQ_SEARCH_BATCHED,
 * erq_search_dev, erq_scan_dev, erq_search_dev_noalloc) run its scan; the batched and learned
 * searches return ERQ_EUNSUPPORTED.  erq_index_load infers the format from the file size. */
int erq_index_create_nonha(size_t N, size_t DIM, size_t K, size_t EX_BITS,
                           const uint64_t* cluster_sizes, const float* P, const float* centroids,
                           const uint8_t* short_data, const uint8_t* long_code,
                           const float* ex_factor, const uint32_t* ids, erq_index** out);

/* Replaces IVF::~IVF (IVF.hpp:107). */
void erq_index_free(erq_index* h);

/* IVF::padded_dim (IVF.hpp:86), IVF::k (IVF.hpp:90) and the other header fields. */
size_t erq_padded_dim(const erq_index* h);
size_t erq_num_clusters(const erq_index* h);
size_t erq_dim(const erq_index* h);
size_t erq_num_vectors(const erq_index* h);
size_t erq_ex_bits(const erq_index* h);
/* Short factors per vector: 1 (HA format) or 4 (non-HA); Quantizer.hpp:33-37. */
size_t erq_short_factors(const erq_index* h);

/* Replaces Rotator::rotate on zero-padded queries (Rotator.hpp:60-63 as used by
 * test_search.cpp:83-92): q is nq x DIM, q_rot is nq x D (host pointers). */
int erq_rotate(erq_index* h, const float* q, size_t nq, float* q_rot);

/* Replaces IVF::search (IVF.hpp:84, :295-324): q_rot is one rotated, padded query of D
 * floats; writes k ids in ascending estimated distance (0 past the pool size), and k
 * distances if dists != NULL (FLT_MAX past the pool size).  Host pointers. */
int erq_search(erq_index* h, const float* q_rot, size_t k, size_t nprobe, uint32_t* results,
               float* dists);

/* IVF::search over nq queries (host pointers), fixed nprobe.  erq_search and this call
 * keep their device buffers in the handle (grown on demand, never per call). */
int erq_search_batch(erq_index* h, const float* q_rot, size_t nq, size_t k, size_t nprobe,
                     uint32_t* results, float* dists);

/* IVF::search over nq queries (host pointers) with each query's own nprobe
 * (nprobe_per_q[i] in [1, min(K, 2048)]): flags 0 runs the exact per-query replay (each
 * query's result equals IVF::search with its nprobe), ERQ_SEARCH_BATCHED the cluster-major
 * rounds of erq_scan_batched_dev. */
enum { ERQ_SEARCH_BATCHED = 1 };
int erq_search_batch_nprobe(erq_index* h, const float* q_rot, size_t nq, size_t k,
                            const uint32_t* nprobe_per_q, uint32_t* results, float* dists,
                            unsigned flags);

/* ---- device-pointer stages of the same path (enqueue on `stream`) ---- */

/* Rotator::rotate: q_dev nq x DIM -> q_rot_dev nq x D (FP32 MFMA GEMM). */
int erq_rotate_dev(erq_index* h, const float* q_dev, size_t nq, float* q_rot_dev, void* stream);

/* The query-to-centroid contraction on MFMA: out[i][c] = ||q_i||^2 + ||c||^2 - 2 q_i.c
 * (nq x K), the ranking key used before the exact distances of Initializer.hpp:59-62. */
int erq_centroid_dist_dev(erq_index* h, const float* q_rot_dev, size_t nq, float* out_dev,
                          void* stream);

/* FlatInitializer::centroids_distances (Initializer.hpp:55-68): for each query the
 * nprobe nearest clusters in ascending squared distance, distances computed exactly as
 * the reference's L2Sqr16 (space.hpp:18-39).  Outputs nq x nprobe. */
int erq_probe_dev(erq_index* h, const float* q_rot_dev, size_t nq, size_t nprobe,
                  uint32_t* probe_ids_dev, float* probe_sqr_y_dev, void* stream);

/* The per-query cluster loop of IVF::search (IVF.hpp:313-323) with HASearcher's scan
 * (HASearcher.hpp:82-237): query i scans clusters probe_ids[off[i] .. off[i+1]) in that
 * order.  Results nq x k.  stats_dev (nullable) accumulates 4 uint64 counters:
 * blocks scanned, vectors scanned, survivors re-ranked, probes processed. */
int erq_scan_dev(erq_index* h, const float* q_rot_dev, size_t nq, size_t k,
                 const uint32_t* probe_off_dev, const uint32_t* probe_ids_dev,
                 const float* probe_sqr_y_dev, uint32_t* ids_dev, float* dists_dev,
                 uint64_t* stats_dev, void* stream);

/* erq_probe_dev followed by erq_scan_dev with a fixed nprobe. */
int erq_search_dev(erq_index* h, const float* q_rot_dev, size_t nq, size_t k, size_t nprobe,
                   uint32_t* ids_dev, float* dists_dev, uint64_t* stats_dev, void* stream);

/* erq_search_dev without any device allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Every intermediate (approximate centroid distances, probe lists,
 * query-order bins) lives in a fixed per-device scratch array of
 * erq_search_noalloc_capacity() bytes (64 MiB), held from the library's first index on that
 * device until the process ends; a call whose need (erq_search_noalloc_bytes: about
 * nq * (4 K + 8 nprobe) bytes) exceeds it fails with ERQ_EUNSUPPORTED.  All calls on one
 * device share that array: two such calls that overlap in time on one GPU (different
 * handles, or streams not ordered by the caller) race.  The int8-MFMA scan runs if the
 * handle already holds its copies of the codes (built by an earlier allocating search),
 * else the LUT scan; results are identical to erq_search_dev's either way. */
int erq_search_dev_noalloc(erq_index* h, const float* q_rot_dev, size_t nq, size_t k,
                           size_t nprobe, uint32_t* ids_dev, float* dists_dev,
                           uint64_t* stats_dev, void* stream);
size_t erq_search_noalloc_bytes(const erq_index* h, size_t nq, size_t nprobe);
size_t erq_search_noalloc_capacity(void);

/* Cluster-major batched form of erq_scan_dev for many queries with uneven probe lists:
 * ranks [0,4) of every query are scanned exactly as IVF::search does; ranks [4,16),
 * [16,64), [64,192) and then 128 ranks at a time are processed in rounds, each round
 * reading every probed cluster's codes once for all queries that probe it (FastScan sums
 * on int8 MFMA) and pruning against each query's k-th distance at round start.
 * max_nprobe bounds every query's probe count.  Results are deterministic; with at most 4
 * probes per query they equal erq_scan_dev's.  No candidate is ever dropped: a query whose
 * round produced more candidates than the per-query buffer holds has that round rescanned
 * alone with an exact per-query pool.  dists_dev is required.  stats_dev (nullable) takes 8
 * counters: [0] blocks, [1] vectors, [3] probes of round 0; [2] survivors re-ranked (all
 * rounds); [4] candidates kept, [5] (query, round) rescans after a candidate overflow, [6]
 * 16-vector tiles scanned in rounds >= 1. */
int erq_scan_batched_dev(erq_index* h, const float* q_rot_dev, size_t nq, size_t k,
                         const uint32_t* probe_off_dev, const uint32_t* probe_ids_dev,
                         const float* probe_sqr_y_dev, size_t max_nprobe, uint32_t* ids_dev,
                         float* dists_dev, uint64_t* stats_dev, void* stream);

/* erq_probe_dev followed by erq_scan_batched_dev with a fixed nprobe. */
int erq_search_batched_dev(erq_index* h, const float* q_rot_dev, size_t nq, size_t k,
                           size_t nprobe, uint32_t* ids_dev, float* dists_dev,
                           uint64_t* stats_dev, void* stream);

/* ---- learned probe predictor (no counterpart in the reference) ----
 * MLP  D -> H -> H -> K with GELU, BF16 weights/activations, FP32 accumulation and FP32
 * master weights; output p[c] = predicted probability that cluster c holds one of the
 * query's top-k neighbours.  Parameters pack [W1 (H x D), b1, W2 (H x H), b2, W3 (K x H), b3]. */
typedef struct erq_mlp erq_mlp;
int erq_mlp_create(size_t in_dim, size_t hidden, size_t out_dim, size_t max_batch, uint64_t seed,
                   erq_mlp** out);
void erq_mlp_free(erq_mlp* m);
size_t erq_mlp_num_params(const erq_mlp* m);
size_t erq_mlp_in_dim(const erq_mlp* m);
size_t erq_mlp_out_dim(const erq_mlp* m);
int erq_mlp_set_params(erq_mlp* m, const float* host_params);
/* Layer 1 reads bf16(scale * x) (default scale 1): with scale = sqrt(D), rotated unit-norm
 * queries enter with unit variance per coordinate. */
int erq_mlp_set_input_scale(erq_mlp* m, float scale);
float erq_mlp_input_scale(const erq_mlp* m);
/* Either output may be NULL; grads are those of the last erq_mlp_loss_grad_dev. */
int erq_mlp_get_params(erq_mlp* m, float* host_params, float* host_grads);
/* prob_dev = p for batch rows of x_dev (batch x in_dim), any batch size.  x_dev and prob_dev
 * are read and written 16 bytes at a time: both must be 16-byte aligned (hipMalloc and torch
 * allocations are), else ERQ_EINVAL. */
int erq_mlp_forward_dev(erq_mlp* m, const float* x_dev, size_t batch, float* prob_dev,
                        void* stream);
/* Forward + backward of the pos-weighted BCE (labels y_dev: batch x out_dim of 0/1, loss =
 * batch mean of the per-query sum); gradients stay in the model; batch % 32 == 0. */
int erq_mlp_loss_grad_dev(erq_mlp* m, const float* x_dev, const float* y_dev, size_t batch,
                          float pos_weight, float* loss_dev, void* stream);
int erq_mlp_adam_step_dev(erq_mlp* m, float lr, float beta1, float beta2, float eps,
                          void* stream);

/* Learned probe lists (CSR: probe_off nq+1, ids/sqr_y up to nq*max_nprobe): each query keeps
 * its most probable clusters until their predicted probability mass reaches `mass` of the
 * mass of its max_nprobe most probable clusters (at least min_nprobe of them), ordered by
 * exact centroid distance; a query whose total predicted mass over all clusters is below
 * min_conf takes the fixed list of its fallback_nprobe nearest clusters.  n_fallback_dev
 * (one u32) is incremented per fallback query. */
int erq_probe_learned_dev(erq_index* h, erq_mlp* m, const float* q_rot_dev, size_t nq, float mass,
                          size_t min_nprobe, size_t max_nprobe, size_t fallback_nprobe,
                          float min_conf, uint32_t* probe_off_dev, uint32_t* probe_ids_dev,
                          float* probe_sqr_y_dev, uint32_t* n_fallback_dev, void* stream);

/* erq_probe_learned_dev followed by erq_scan_batched_dev; stats counter [7] counts fallbacks. */
int erq_search_learned_dev(erq_index* h, erq_mlp* m, const float* q_rot_dev, size_t nq, size_t k,
                           float mass, size_t min_nprobe, size_t max_nprobe,
                           size_t fallback_nprobe, float min_conf, uint32_t* ids_dev,
                           float* dists_dev, uint64_t* stats_dev, void* stream);

/* ---- stage timing (no counterpart in the reference; for measurement) ----
 * Between erq_timing_begin and erq_timing_end, every search call on the handle records a
 * pair of HIP events around each of its stages, on the stream that stage runs on (up to
 * max_records stages in all; later ones go unrecorded; nothing is recorded under HIP graph
 * capture).  erq_timing_end waits for the recorded events and returns, per stage, the summed
 * milliseconds and the number of recorded stages. */
enum {
    ERQ_STAGE_SELECT = 0, /* probe selection: centroid contraction + top-nprobe, or MLP +
                             learned lists */
    ERQ_STAGE_REPLAY = 1, /* the per-query exact scan (all of it, or ranks [0, 4)) */
    ERQ_STAGE_PREP = 2,   /* batched rounds: work lists and per-pair digits */
    ERQ_STAGE_SCAN = 3,   /* batched rounds: the FastScan pass over work items */
    ERQ_STAGE_REEST = 4,  /* batched rounds: ex-bit re-estimate of the survivors */
    ERQ_STAGE_MERGE = 5,  /* batched rounds: candidates into the top-k pools */
    ERQ_STAGE_REDO = 6,   /* batched rounds: rescan of overflowed (query, round)s */
    ERQ_NUM_STAGES = 7
};
int erq_timing_begin(erq_index* h, size_t max_records);
int erq_timing_end(erq_index* h, double* ms_per_stage, uint64_t* count_per_stage);

/* Device pointers of the resident index arrays (for tools that read the HBM layout).
 * which: 0 short blocks, 1 long codes, 2 ex factors, 3 ids, 4 rotated centroids,
 * 5 rotation P, 6 cluster sizes (u32), 7 first block per cluster (u32),
 * 8 first vector per cluster (u32). */
const void* erq_device_array(const erq_index* h, int which);

/* ---- dataset preparation (device pointers; row-major float rows of `dim` values; each
 * call allocates its temporaries and returns after `stream` has drained, except
 * erq_kmeans_update_dev, which only enqueues).  Rows are taken as zero-padded to a multiple
 * of 16 dims, dim <= 8192, row counts < 2^31. */

/* Ground truth, python/compute_gt.py:12-15 (numpy: argsort of the L2 norms, first k): per
 * query, the k base rows with the smallest squared L2 distance -- evaluated in double as 16
 * lane-strided fma chains over the exact float differences, reduced in
 * _mm512_reduce_add_ps order -- ordered by (distance, id).  ids: nq x k; dists (nullable):
 * nq x k, the distances rounded to float.  k <= min(n, 4096).  ERQ_EUNSUPPORTED if more
 * than 65536 base rows lie within the FP32 GEMM's error bound of one query's k-th
 * distance. */
int erq_knn_dev(const float* base, uint64_t n, const float* queries, uint64_t nq, uint32_t dim,
                uint32_t k, uint32_t* ids, float* dists, void* stream);

/* Nearest centroid of each row, the index.quantizer.search(X, 1) of python/ivf.py:31: the
 * centroid with the smallest L2Sqr16 (space.hpp:18-39, the search-time ranking of
 * Initializer.hpp:55-68), ties by id; dist = that squared distance (ivf.py takes its square
 * root for dist_to_centroid). */
int erq_kmeans_assign_dev(const float* x, uint64_t n, uint32_t dim, const float* centroids,
                          uint32_t K, uint32_t* assign, float* dist, void* stream);

/* One Lloyd update: each cluster with members gets their mean (summed in double in
 * increasing row order, rounded to float); a cluster without members keeps its centroid.
 * counts (nullable): K member counts. */
int erq_kmeans_update_dev(const float* x, uint64_t n, uint32_t dim, const uint32_t* assign,
                          uint32_t K, float* centroids, uint32_t* counts, void* stream);

/* IVF training, python/ivf.py:26-35 (faiss IVF{K},Flat train, then quantizer.search):
 * centroids start as the rows init_rows[0..K), then `iters` rounds of assignment + update,
 * then a final assignment into assign / dist. */
int erq_kmeans_dev(const float* x, uint64_t n, uint32_t dim, uint32_t K, uint32_t iters,
                   const uint32_t* init_rows, float* centroids, uint32_t* assign, float* dist,
                   void* stream);

/* ---- index building (device pointers; each call returns after `stream` has drained) */

/* DataQuantizer::exrabitq_codes (Quantizer.hpp:457-489) with fast_quantize (:138-219) and
 * store_compacted_code (:491-631) for n rotated residuals resid[n x D] (D a multiple of 64,
 * <= 8192) and their norms ||o - c|| onorm[n]: the compacted EX-bit codes long_code[n x
 * D*EX/8] and xipnorm ex_factor[n], byte for byte what the host builder (builder.cpp)
 * writes from the same inputs.  ERQ_EINVAL on a non-finite input. */
int erq_build_quantize_dev(const float* resid, const float* onorm, uint64_t n, uint32_t D,
                           uint32_t ex_bits, uint8_t* long_code, float* ex_factor, void* stream);

/* The short blocks (Quantizer.hpp:277-290, pack_codes.hpp:33-98): per cluster (sizes in
 * host memory; resid/onorm in cluster order) ceil(n_k / 32) blocks of 4D code bytes -- the
 * sign bits of resid -- and 32 float norms; unused slots of a tail block are zero. */
int erq_build_pack_dev(const float* resid, const float* onorm, const uint64_t* cluster_sizes,
                       uint32_t K, uint32_t D, uint8_t* short_data, void* stream);

/* IVF::construct (IVF.hpp:116-153) on the GPU from raw vectors data[N x DIM], their
 * centroids[K x DIM] (un-rotated), cluster ids assign[N] and the rotation P[D x D]
 * (D = DIM rounded up to 64): cluster_sizes[K] (host), rotated centroids [K x D], ids[N]
 * (cluster order, ascending within a cluster), short_data (sum ceil(n_k/32) (4D + 128)
 * bytes, sized by the caller from the cluster sizes, e.g. a bincount of assign), long_code
 * [N x D*EX/8] and ex_factor[N].  Rotations are one FP32 GEMM each; residual = rotated
 * vector - rotated centroid; ||o - c|| = f32(sqrt((double)L2Sqr(o, c, DIM))) as
 * Quantizer.hpp:399-401 computes it; then erq_build_quantize_dev and erq_build_pack_dev. */
int erq_build_index_dev(const float* data, uint64_t N, uint32_t DIM, const float* centroids,
                        uint32_t K, const uint32_t* assign, const float* P, uint32_t ex_bits,
                        uint64_t* cluster_sizes, float* rotated_centroids, uint8_t* short_data,
                        uint8_t* long_code, float* ex_factor, uint32_t* ids, void* stream);

/* erq_build_index_dev for the non-HA format (IVF::construct without HIGH_ACC_FAST_SCAN;
 * DataQuantizer::rabitq_factor, Quantizer.hpp:386-437): short_data holds sum ceil(n_k/32)
 * (4D + 512) bytes, four factors per vector after each block's codes -- ||o - c||^2 (L2Sqr),
 * fac_ip = 2 ||o - c|| / sum |u / 2|, sumxb = the positive dims, fac_err = 2 ||o - c||
 * FAC_ERR sqrt(1 - <u, o_bar>^2) / <u, o_bar> -- and ex_factor = xipnorm over
 * f32(sqrt((double)||o - c||^2)), as the host builder writes them. */
int erq_build_index_dev_nonha(const float* data, uint64_t N, uint32_t DIM,
                              const float* centroids, uint32_t K, const uint32_t* assign,
                              const float* P, uint32_t ex_bits, uint64_t* cluster_sizes,
                              float* rotated_centroids, uint8_t* short_data, uint8_t* long_code,
                              float* ex_factor, uint32_t* ids, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* ERQ_H */
