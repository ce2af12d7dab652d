// erq_internal.h -- handle layout and kernel launchers shared by the liberq sources.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <exception>
#include <mutex>
#include <new>

#include "erq.h"
#include <string>
#include <vector>

struct erq_index {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t N = 0, DIM = 0, D = 0, K = 0, EX = 0;
    size_t nblocks = 0;
    // short factors per vector: 1 (the HA format, HIGH_ACC_FAST_SCAN) or 4 (the non-HA
    // format that Searcher reads, Quantizer.hpp:107-121); a block is 4D + 128 * factors bytes
    int factors = 1;
    std::vector<uint64_t> sizes;
    // resident arrays (HA layout, reference inc/index/IVF.hpp:28-40)
    uint8_t* short_data = nullptr;  // nblocks * (4D + 128 * factors)
    uint8_t* long_code = nullptr;   // N * D*EX/8
    float* ex_factor = nullptr;     // N
    uint32_t* ids = nullptr;        // N
    float* centroids = nullptr;     // K x D
    float* cnorm = nullptr;         // K, ||c||^2
    float* P = nullptr;             // D x D
    uint32_t* csize = nullptr;      // K
    uint32_t* bstart = nullptr;     // K + 1
    uint32_t* vstart = nullptr;     // K + 1
    // the low-latency scan's slot buffers (grow-only) and its batch-size limit
    void* lat = nullptr;
    size_t lat_cap = 0;
    int latency_nq = 16;
    std::vector<uint64_t> top_pad;  // sums of the i largest cluster sizes rounded up to 32
    // grow-only workspace
    void* ws = nullptr;
    size_t ws_bytes = 0;
    uint8_t* bp = nullptr;  // bit-plane copy of the 1-bit codes (batched path, built on first use)
    uint32_t* exl = nullptr;    // lane-major EX-bit codes (with bp, built on first use)
    uint32_t* order = nullptr;  // per-query scan order scratch (counts, cursor, order)
    size_t order_cap = 0;       // u32 entries of `order`
    int plain_order = 1;        // ERQ_OPT_PLAIN_ORDER
    int num_cu = 0;
    int plain_kernel = 0;   // ERQ_OPT_PLAIN_KERNEL
    int plain_waves = 1;    // ERQ_OPT_PLAIN_WAVES
    int batch_wg_per_cu = 2;  // ERQ_OPT_BATCH_WG_PER_CU
    int batch_item_tiles = 16;  // ERQ_OPT_BATCH_ITEM_TILES
    int prep_wg_per_cu = 6;     // ERQ_OPT_PREP_WG_PER_CU
    size_t batch_surv_cap = size_t(1) << 26;  // ERQ_OPT_BATCH_SURVIVOR_CAP
    size_t batch_pair_budget = size_t(1) << 23;  // ERQ_OPT_BATCH_PAIR_BUDGET
    int batch_overlap = 0;                       // ERQ_OPT_BATCH_OVERLAP
    // second stream of the batched search: the rounds' work lists and preparation run on it
    // while the exact ranks run on the caller's stream (fork and join by events)
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // ordering between calls on different streams: every call records ev_last on its stream
    // when it returns, and the next call on another stream waits for it first (the buffers
    // above are shared by all calls on the handle)
    hipEvent_t ev_last = nullptr;
    hipStream_t last_stream = nullptr;
    bool last_valid = false;
    // device buffers of the host-pointer calls (erq_search, erq_search_batch*), grow-only
    float* hq = nullptr;         // hq_cap x D rotated queries
    uint32_t* hids = nullptr;    // hq_cap x hk
    float* hdists = nullptr;     // hq_cap x hk
    uint32_t* hnp = nullptr;     // hq_cap + 1 (per-query probe counts, then CSR offsets)
    size_t hq_cap = 0, hk_cap = 0;
    // pinned host staging of those calls' small copies (queries in, ids and distances out),
    // grow-only up to kPinMax bytes; larger copies go to and from the caller's memory
    void* pin = nullptr;
    size_t pin_cap = 0;
    int depth = 0;  // nesting of entry points on this thread (only the outermost orders)
    // stage timing (erq_timing_begin / erq_timing_end): event pairs recorded around stages
    std::vector<hipEvent_t> tim_ev;   // 2 per record
    std::vector<int> tim_stage;       // stage of each record in use
    bool tim_on = false;
    // recursive: the host-pointer calls hold it while they call the device-pointer ones
    std::recursive_mutex mu;
};

namespace erq {

void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int workspace(erq_index* h, size_t bytes, void** out);

// Runs an entry point's body; a C++ exception becomes an error code (a failed host
// allocation is ERQ_ENOMEM).
template <typename F>
int guarded(F&& f) noexcept {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return ERQ_ENOMEM;
    } catch (const std::exception& ex) {
        set_error(std::string("internal error: ") + ex.what());
        return ERQ_EINVAL;
    } catch (...) {
        set_error("internal error");
        return ERQ_EINVAL;
    }
}

struct ScanArgs {
    const float* q_rot;
    const uint32_t* probe_off;
    const uint32_t* probe_ids;
    const float* probe_sqr_y;
    const float* centroids;
    const uint8_t* short_data;
    const uint8_t* long_code;
    const float* ex_factor;
    const uint32_t* ids;
    const uint32_t* csize;
    const uint32_t* bstart;
    const uint32_t* vstart;
    uint32_t* out_ids;
    float* out_dists;
    unsigned long long* stats;
    int D, EX, k, nq, K;
    int max_probes;  // > 0: scan at most this many leading probes per query
    const uint8_t* bp;  // bit-plane codes (nullptr: LUT kernel)
    const uint32_t* exl;  // lane-major EX-bit codes (with bp)
    int exl_w;
    int nchunk;         // 512-dim chunks of bp
    int waves_per_block;  // queries (waves) per workgroup
    float one_over_sqrtD;
    const uint32_t* qorder;  // nullable: the order queries are taken in (a permutation of nq)
    // Round mode (the batched search's redo of a round whose candidates overflowed): the
    // queries qlist[1 .. qlist[0]] rescan probe ranks [r0, r1) against the k-th distance of
    // the pool in out_ids / out_dists (fixed for the round), and merge every candidate whose
    // estimate is at most that distance into the pool by (distance, id).
    const uint32_t* qlist;
    int r0, r1;
    // Estimate mode (the low-latency search): waves take the (query, probe) pairs p <
    // probe_off[nq] and write, for every vector of the probed cluster, its lower bound,
    // re-estimated distance and id at slot est_off[p] + 32 b + j (block b, vector j; a tail
    // block's unused slots get an infinite lower bound); the replay then walks each query's
    // slots in order with the reference's pruning threshold.
    float* est_lower;
    float* est_exd;
    uint32_t* est_id;
    uint32_t* est_off;
    float* est_bmin;  // per 32-slot block (slot / 32): the smallest lower bound of its slots
    // Non-HA int8 MFMA scan (a.bp set, nchunk <= 4): [0] count, then the queries it handed
    // back to the LUT kernel (a digit above 16383); nullptr: the LUT kernel alone.
    uint32_t* nonha_redo;
};
hipError_t launch_scan(const ScanArgs& a, hipStream_t s);
// The per-query scan of a non-HA index (Searcher.hpp:85-236): the LUT kernel of erq_nonha.hip
// with a.short_data in the 4-factor block layout (a.bp unused).
hipError_t launch_scan_nonha(const ScanArgs& a, hipStream_t s);
// The round-mode rescan (a.qlist set): a fixed grid of waves that take the listed queries.
hipError_t launch_scan_round(const ScanArgs& a, hipStream_t s);
// The low-latency scan (a.est_* set, a.bp set, nchunk <= 4): slot offsets, the per-vector
// estimates of every (query, probe) pair in parallel, then one wave per query replays the
// probes in order -- the same results as launch_scan.
hipError_t launch_scan_latency(const ScanArgs& a, hipStream_t s);
// Bytes of LDS the round-mode kernel needs per wave (D, k), and whether it fits.
size_t scan_round_wave_bytes(int D, int k);
// qorder[nq] = the queries by scan cost, heaviest first (cost: the 32-vector blocks of the
// probed clusters, the first max_probes when > 0, shifted right by `shift`, K bins);
// scratch holds K + 1 u32.
hipError_t launch_query_order(const uint32_t* probe_off, const uint32_t* probe_ids,
                              const uint32_t* csize, int nq, int K, int max_probes, int shift,
                              uint32_t* scratch, uint32_t* qorder, hipStream_t s);

// C[m][n] = sum_k A[m][k] * B(k, n); B row-major [K x N] (b_trans = 0) or stored as
// B^T row-major [N x K] (b_trans = 1).  epi = 0: store; epi = 1: rn[m] + cn[n] - 2*acc.
hipError_t launch_gemm(const float* A, int lda, const float* B, int ldb, int b_trans, float* C,
                       int ldc, int M, int N, int K, int epi, const float* rn, const float* cn,
                       hipStream_t s);
// approx[q][c] = ||Q_q||^2 + cnorm[c] - 2 Q_q . centroid_c and qnorm[q] = ||Q_q||^2 (Q: nq rows
// of D floats); one kernel for up to 8 queries, else launch_row_sqnorm and launch_gemm.
hipError_t launch_query_centroid_dist(const float* Q, int nq, int D, const float* centroids,
                                      int K, const float* cnorm, float* approx, float* qnorm,
                                      hipStream_t s);
hipError_t launch_row_sqnorm(const float* X, int rows, int cols, int ld, float* out, hipStream_t s);
hipError_t launch_pad_rows(const float* src, int rows, int cols, float* dst, int ldd,
                           hipStream_t s);

// Ranked probe lists from approximate distances (see erq_probe.hip).  With cand (nq x
// kProbeCandWords words) and nq <= kProbeSplitNq the selection runs as three kernels that
// spread the candidates' exact distances over many workgroups, overwriting approx; it can
// then also write the lists' CSR offsets (fixed_off, nq + 1 entries: q * nprobe).
constexpr int kProbeCap = 2048;  // candidates kept per query
constexpr size_t kProbeSplitNq = 16;
constexpr size_t kProbeCandWords = kProbeCap + 1;
hipError_t launch_probe_select(float* approx, const float* qnorm, const float* cnorm,
                               const float* q_rot, const float* centroids, int nq, int K, int D,
                               int nprobe, uint32_t* probe_ids, float* probe_d, hipStream_t s,
                               uint32_t* cand = nullptr, uint32_t* fixed_off = nullptr);

// CSR probe lists (ids and distances) from padded per-query rows of `stride` entries: query q
// keeps the first off[q + 1] - off[q] of its row.
hipError_t launch_compact_probes(const uint32_t* pad_ids, const float* pad_d, int stride,
                                 const uint32_t* off, int nq, uint32_t* csr_ids, float* csr_d,
                                 hipStream_t s);

// prob (nq x K) is overwritten for the queries that take the fixed list (their exact
// distances); fb_list (nq) and fb_count (1) are scratch.
hipError_t launch_adaptive_probe(float* prob, const float* q_rot, const float* centroids,
                                 int nq, int K, int D, float mass, int min_np, int max_np,
                                 float min_conf, int fb_np, uint32_t* pad_ids, float* pad_d,
                                 uint32_t* len, uint32_t* off, uint32_t* csr_ids, float* csr_d,
                                 uint32_t* n_fallback, uint32_t* fb_list, uint32_t* fb_count,
                                 hipStream_t s);

}  // namespace erq
