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
    const uint32_t* items;     // 3 x n_items: cluster, (first entry << 4) | count, first tile
    const uint32_t* entries;   // (probe slot, query) per entry, grouped by cluster
    uint32_t* n_items;         // [0] items of this round, [2] entries, [3] queued items
    uint32_t* ovf;             // items queued for the inline pass (slice full)
    const uint32_t* probe_ids; // per slot: cluster id
    int8_t* adig;              // per entry: 2 * Dp int8 digits in MFMA operand order
    void* qparams;             // per entry: query parameters
    float* pool_d;             // nq x k, ascending (FLT_MAX padded)
    uint32_t* cand_cnt;        // nq
    float* cand_d;             // nq x cand_cap
    uint32_t* cand_i;          // nq x cand_cap
    uint4* surv;               // survivor records: entry, cluster, vector, FastScan product,
                               // one slice of surv_cap / grid records per scan workgroup
    uint32_t* surv_count;      // per scan workgroup: records in its slice
    size_t surv_cap;
    unsigned long long* stats; // 8 counters (see erq.h)
    int cand_cap;
    int D, EX, k, nq, nchunk;
    float one_over_sqrtD;
};

struct RoundBufs {
    const uint32_t* probe_ids;
    uint32_t* counts;  // K
    uint32_t* cursor;  // K
    uint32_t* pool_i;  // nq x k
    int K;
};

int bp_nchunk(int D);
size_t batch_lds_bytes(int D);
size_t prep_wave_bytes(int D);
// Upper bound on the work items of one round of at most `pairs` (query, cluster) pairs.
size_t batch_items_bound(size_t pairs, size_t K, size_t N, size_t max_csize);
hipError_t launch_relayout(const uint8_t* short_data, size_t nblocks, int D, uint8_t* bp,
                           hipStream_t s);
hipError_t launch_zero(uint32_t* p, size_t n, hipStream_t s);
// exl (N x 16 x exl_words(D, EX) u32, erq_device.h) from long_code.
hipError_t launch_exl_relayout(const uint8_t* long_code, size_t N, int D, int EX, uint32_t* exl,
                               hipStream_t s);
hipError_t launch_round(const BatchArgs& ba, const RoundBufs& rb, int r0, int r1, int grid,
                        hipStream_t s);

}  // namespace erq
