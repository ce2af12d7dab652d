// erq_batch.h -- arguments of the cluster-major batched search (erq_batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace erq {

struct BatchArgs {
    const float* q_rot;        // nq x D
    const uint32_t* probe_off; // nq + 1 (CSR over probe slots)
    const float* probe_sqr_y;  // per slot: squared distance to the cluster centroid
    const float* centroids;
    const uint8_t* short_data; // HA blocks (factors are read from here)
    const uint8_t* bp;         // bit-plane copy of the 1-bit codes
    const uint8_t* long_code;
    const uint32_t* exl;       // lane-major EX-bit codes, exl_w dwords per vector lane
    int exl_w;
    const float* ex_factor;
    const uint32_t* ids;
    const uint32_t* csize;
    const uint32_t* bstart;
    const uint32_t* vstart;
    const uint32_t* items;     // 3 words per item: cluster, (first entry << 4) | count, first tile
    const uint32_t* entries;   // (probe slot, query) per entry, grouped by cluster
    uint32_t* n_items;         // [0] items of this round, [1] its first item, [2] entries, [3] queued
    uint32_t* ovf;             // queued item halves (2 x item + half) for the inline pass
    const uint32_t* probe_ids; // per slot: cluster id
    int8_t* adig;              // per entry: 2 * Dp int8 digits in MFMA operand order
    void* qparams;             // per entry: query parameters
    float* pool_d;             // nq x k, ascending (FLT_MAX padded)
    uint32_t* cand_cnt;        // nq
    float* cand_d;             // nq x cand_cap
    uint32_t* cand_i;          // nq x cand_cap
    uint4* surv;               // survivor records: entry, vector, FastScan product scaled by
                               // delta (above 4 chunks: the exact integer), |o - c|,
                               // one slice of surv_cap / grid records per scan workgroup
    uint32_t* surv_count;      // per scan workgroup: records in its slice
    size_t surv_cap;
    unsigned long long* stats; // 8 counters (see erq.h)
    uint32_t* redo;            // the round's queries whose candidates overflowed: [0] count,
                               // then the queries (rescanned by the round-mode scan)
    uint32_t* slow;            // the round's queries for the general merge: [0] count, then
                               // the queries (zero before the round)
    int cand_cap;
    int item_tiles;            // 16-vector tiles per work item: 8 or 16 (4 or 8 blocks)
    int num_cu;                // compute units (sizes the re-estimate's grid)
    int D, EX, k, nq, nchunk;
    float one_over_sqrtD;
};

struct RoundBufs {
    const uint32_t* probe_ids;
    uint32_t* counts;  // bins ((round, cluster): R x K)
    uint32_t* cursor;  // bins
    uint32_t* local;   // 2 per bin: entries and items before it in its block of 256 bins
    uint32_t* btot;    // 2 per block of 256 bins
    uint32_t* wgbase;  // kCountGroups x bins: a counting workgroup's entries before its own in a bin
    uint32_t* pool_i;  // nq x k
    int K;
};

// 512-dim chunks of the bit-plane codes: ceil(D / 512), rounded up to 6 or 8 above 4 (the
// kernels are built for 1, 2, 3, 4, 6 and 8 chunks; the padding chunks' codes and digits are 0).
int bp_nchunk(int D);
size_t batch_lds_bytes(int D);         // main scan pass
size_t batch_inline_lds_bytes(int D);  // inline pass
size_t prep_wave_bytes(int D);
// Upper bound on the work items of one round of at most `pairs` (query, cluster) pairs.
size_t batch_items_bound(size_t pairs, size_t K, size_t N, size_t max_csize, int item_tiles);
hipError_t launch_relayout(const uint8_t* short_data, size_t nblocks, int D, int factors, uint8_t* bp,
                           hipStream_t s);
hipError_t launch_zero(uint32_t* p, size_t n, hipStream_t s);
// exl (N x 16 x exl_words(D, EX) u32, erq_device.h) from long_code.
hipError_t launch_exl_relayout(const uint8_t* long_code, size_t N, int D, int EX, uint32_t* exl,
                               hipStream_t s);
constexpr int kFirstRoundRank = 4;  // ranks below it are the exact per-query replay's
constexpr int kMaxRounds = 32;
// Workgroups of the LDS-privatized bin counting and scatter (each a contiguous range of
// queries); used while R x K bins of 4 bytes fit kCountLds.
constexpr int kCountGroups = 16;
constexpr size_t kCountLds = 128 * 1024;
// The work lists of every round past the exact ranks (ranks [4, r_end), R rounds:
// [4,16), [16,64), [64,192), then 128 wide) and the per-entry preparation (digits and
// parameters, no threshold), which need only the probe lists.  ba.items / entries / adig /
// qparams cover all rounds; ba.n_items = rinfo, 4 (R + 1) words (round r at 4 r: items,
// first item, entries, queued items; totals at 4 R); rb.counts / cursor hold R x K bins.
// prep_grid: workgroups (4 waves each, a multiple of 8) of the preparation kernel.
hipError_t launch_prep_all(const BatchArgs& ba, const RoundBufs& rb, int R, int r_end,
                           int prep_grid, hipStream_t s);
// One round's scan, re-estimate and merge (ba.n_items = rinfo + 4 r), with each query's
// threshold read from pool_d (its k-th distance when the round starts).
hipError_t launch_round_scan(const BatchArgs& ba, const RoundBufs& rb, int grid, hipStream_t s);
// One stage of it: 0 the scan (main and inline passes), 1 the re-estimate, 2 the merge.
hipError_t launch_round_stage(const BatchArgs& ba, const RoundBufs& rb, int grid, int stage,
                              hipStream_t s);

}  // namespace erq
