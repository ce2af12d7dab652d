// erq_probe.hip -- top-nprobe cluster selection with reference-exact distances.
//
// The MFMA contraction gives a[c] = ||q||^2 + ||c||^2 - 2 q.c, which ranks clusters but
// is not the value the reference uses: FlatInitializer::centroids_distances
// (Initializer.hpp:55-68) ranks and returns L2Sqr16 (space.hpp:18-39), 16 lane-strided
// FMA chains of (q-c)^2 reduced by _mm512_reduce_add_ps.  Each query therefore:
//   1. bounds |a[c] - L2Sqr16(c)| <= E[c] = (4D + 64) * 2^-24 * (||q||^2 + ||c||^2),
//   2. finds U = the nprobe-th smallest a[c] + E[c] (radix select on float keys),
//   3. keeps every c with a[c] - E[c] <= U (a superset of the exact top-nprobe),
//   4. recomputes L2Sqr16 for those in the reference's order, sorts by (distance, id)
//      and returns the first nprobe.
// Equal distances are ordered by id (std::partial_sort leaves their order unspecified).
#include <algorithm>

#include "erq_device.h"
#include "erq_internal.h"

namespace erq {

namespace {

constexpr int kThreads = 256;
constexpr int kCapC = kProbeCap;  // candidate capacity per query

__device__ __forceinline__ bool cand_less(float da, uint32_t ia, float db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

// Block-wide radix select: the key of rank `rank` (1-based) among keys[0..n).  Per 8-bit
// digit: an LDS histogram, then wave 0 finds the digit holding the rank (each lane sums
// four bins, a shuffle prefix scan over the lanes).
__device__ uint32_t radix_select(const uint32_t* keys, int n, uint32_t rank, uint32_t* hist,
                                 uint32_t* sh_sel, uint32_t* sh_rank) {
    const int tid = threadIdx.x;
    uint32_t prefix = 0, pmask = 0;
    if (tid == 0) *sh_rank = rank;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = tid; i < 256; i += kThreads) hist[i] = 0;
        __syncthreads();
        for (int c = tid; c < n; c += kThreads) {
            uint32_t kv = keys[c];
            if ((kv & pmask) == prefix) atomicAdd(&hist[(kv >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid < 64) {
            const uint32_t h0 = hist[4 * tid], h1 = hist[4 * tid + 1], h2 = hist[4 * tid + 2],
                           h3 = hist[4 * tid + 3];
            const uint32_t own = h0 + h1 + h2 + h3;
            uint32_t incl = own;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t v = __shfl_up(incl, o);
                if (tid >= o) incl += v;
            }
            const uint32_t r = *sh_rank, excl = incl - own;
            // the lane whose bins span rank r (the last lane if r exceeds the count)
            const bool mine = (excl < r && r <= incl) || (tid == 63 && r > incl);
            if (mine) {
                uint32_t cum = excl, sel = 4 * tid + 3;
                const uint32_t hs[4] = {h0, h1, h2, h3};
                for (int j = 0; j < 4; j++) {
                    if (cum + hs[j] >= r) {
                        sel = 4 * tid + j;
                        break;
                    }
                    cum += hs[j];
                }
                *sh_sel = sel;
                *sh_rank = r - cum;
            }
        }
        __syncthreads();
        prefix |= *sh_sel << shift;
        pmask |= 255u << shift;
        __syncthreads();
    }
    return prefix;
}

// L2Sqr16 of q (in LDS) against centroid c by the 16 lanes of this thread's group.
__device__ __forceinline__ float l2sqr16_group(const float* qs, const float* c, int D, int l,
                                               bool active) {
    float p = 0.f;
    if (active) {
#pragma unroll 8
        for (int d = l; d < D; d += 16) {
            float t = qs[d] - c[d];
            p = fmaf(t, t, p);
        }
    }
    return butterfly16(p);
}

__global__ void __launch_bounds__(kThreads)
    probe_select_kernel(const float* __restrict__ approx, const float* __restrict__ qnorm,
                        const float* __restrict__ cnorm, const float* __restrict__ q_rot,
                        const float* __restrict__ centroids, int K, int D, int nprobe,
                        uint32_t* __restrict__ probe_ids, float* __restrict__ probe_d) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* keys = reinterpret_cast<uint32_t*>(smem);          // K
    float* cd = reinterpret_cast<float*>(smem + 4 * (size_t)K);  // kCapC
    uint32_t* ci = reinterpret_cast<uint32_t*>(cd + kCapC);      // kCapC
    float* qs = reinterpret_cast<float*>(ci + kCapC);            // D
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sh_sel, sh_rank, sh_cnt;

    const int tid = threadIdx.x;
    const int qi = blockIdx.x;
    const int grp = tid >> 4, l = tid & 15;
    const float* arow = approx + (size_t)qi * K;
    const float qn = qnorm[qi];
    const float coef = (float)(4 * D + 64) * 5.9604644775390625e-8f;  // 2^-24
    const float* q = q_rot + (size_t)qi * D;
    for (int d = tid; d < D; d += kThreads) qs[d] = q[d];
    for (int c = tid; c < K; c += kThreads) keys[c] = float_key(arow[c] + coef * (qn + cnorm[c]));
    if (tid == 0) sh_cnt = 0;
    __syncthreads();

    const float U = key_float(radix_select(keys, K, (uint32_t)nprobe, hist, &sh_sel, &sh_rank));
    for (int c = tid; c < K; c += kThreads) {
        if (arow[c] - coef * (qn + cnorm[c]) <= U) {
            uint32_t slot = atomicAdd(&sh_cnt, 1u);
            if (slot < (uint32_t)kCapC) ci[slot] = (uint32_t)c;
        }
    }
    __syncthreads();
    int cnt = (int)sh_cnt;
    if (cnt <= kCapC) {
        for (int base = 0; base < cnt; base += kThreads / 16) {
            const int i = base + grp;
            const bool act = i < cnt;
            float p = l2sqr16_group(qs, centroids + (size_t)(act ? ci[i] : 0) * D, D, l, act);
            if (act && l == 0) cd[i] = p;
        }
    } else {
        // bound too loose for the candidate buffer: rank all K by the exact distance
        for (int base = 0; base < K; base += kThreads / 16) {
            const int c = base + grp;
            const bool act = c < K;
            float p = l2sqr16_group(qs, centroids + (size_t)(act ? c : 0) * D, D, l, act);
            if (act && l == 0) keys[c] = float_key(p);
        }
        __syncthreads();
        if (tid == 0) sh_cnt = 0;
        const uint32_t U2 = radix_select(keys, K, (uint32_t)nprobe, hist, &sh_sel, &sh_rank);
        for (int c = tid; c < K; c += kThreads) {
            if (keys[c] <= U2) {
                uint32_t slot = atomicAdd(&sh_cnt, 1u);
                if (slot < (uint32_t)kCapC) {
                    ci[slot] = (uint32_t)c;
                    cd[slot] = key_float(keys[c]);
                }
            }
        }
        __syncthreads();
        cnt = min((int)sh_cnt, kCapC);
    }
    int P2 = 1;
    while (P2 < cnt) P2 <<= 1;
    for (int i = cnt + tid; i < P2; i += kThreads) {
        cd[i] = FLT_MAX;
        ci[i] = 0xFFFFFFFFu;
    }
    __syncthreads();

    // bitonic sort by (distance, id)
    for (int size = 2; size <= P2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = tid; t < (P2 >> 1); t += kThreads) {
                int lo = 2 * t - (t & (stride - 1));
                int hi = lo + stride;
                bool up = (lo & size) == 0;
                float dl = cd[lo], dh = cd[hi];
                uint32_t il = ci[lo], ih = ci[hi];
                bool swap = up ? cand_less(dh, ih, dl, il) : cand_less(dl, il, dh, ih);
                if (swap) {
                    cd[lo] = dh;
                    cd[hi] = dl;
                    ci[lo] = ih;
                    ci[hi] = il;
                }
            }
            __syncthreads();
        }
    for (int i = tid; i < nprobe; i += kThreads) {
        probe_ids[(size_t)qi * nprobe + i] = ci[i];
        probe_d[(size_t)qi * nprobe + i] = cd[i];
    }
}

// Learned probe lists.  Per query: rank clusters by predicted probability p (descending,
// ties by id) and keep the shortest prefix whose mass reaches `mass` of the mass of the
// max_np most probable clusters (at least min_np, at most max_np), then order the kept
// clusters by their exact L2Sqr16 distance (the order IVF::search scans in).  A query
// whose total p is below min_conf takes the fixed list instead: the fb_np nearest clusters
// by exact L2Sqr16 (ties by id), computed here over all K -- the list probe_select_kernel
// returns.
__device__ void bitonic_by_key(float* cd, uint32_t* ci, int P) {
    const int tid = threadIdx.x;
    for (int size = 2; size <= P; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = tid; t < (P >> 1); t += kThreads) {
                int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
                bool up = (lo & size) == 0;
                float dl = cd[lo], dh = cd[hi];
                uint32_t il = ci[lo], ih = ci[hi];
                if (up ? cand_less(dh, ih, dl, il) : cand_less(dl, il, dh, ih)) {
                    cd[lo] = dh;
                    cd[hi] = dl;
                    ci[lo] = ih;
                    ci[hi] = il;
                }
            }
            __syncthreads();
        }
}

// keys <= U into (cd, ci) (capped at kCapC), padded to a power of two and sorted by
// (value, id); returns the count kept
__device__ int gather_sorted(const uint32_t* keys, int K, uint32_t U, float* cd, uint32_t* ci,
                             uint32_t* sh_cnt) {
    const int tid = threadIdx.x;
    if (tid == 0) *sh_cnt = 0;
    __syncthreads();
    for (int c = tid; c < K; c += kThreads)
        if (keys[c] <= U) {
            uint32_t slot = atomicAdd(sh_cnt, 1u);
            if (slot < (uint32_t)kCapC) {
                ci[slot] = (uint32_t)c;
                cd[slot] = key_float(keys[c]);
            }
        }
    __syncthreads();
    const int cnt = min((int)*sh_cnt, kCapC);
    int P2 = 1;
    while (P2 < cnt) P2 <<= 1;
    for (int i = cnt + tid; i < P2; i += kThreads) {
        cd[i] = FLT_MAX;
        ci[i] = 0xFFFFFFFFu;
    }
    __syncthreads();
    bitonic_by_key(cd, ci, P2);
    return cnt;
}

__global__ void __launch_bounds__(kThreads)
    adaptive_select_kernel(const float* __restrict__ prob, const float* __restrict__ q_rot,
                           const float* __restrict__ centroids, int K, int D, float mass,
                           int min_np, int max_np, float min_conf, int fb_np,
                           uint32_t* __restrict__ out_ids, float* __restrict__ out_d,
                           uint32_t* __restrict__ out_len, uint32_t* __restrict__ n_fallback,
                           uint32_t* __restrict__ fb_list, uint32_t* __restrict__ fb_count) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* keys = reinterpret_cast<uint32_t*>(smem);
    float* cd = reinterpret_cast<float*>(smem + 4 * (size_t)K);
    uint32_t* ci = reinterpret_cast<uint32_t*>(cd + kCapC);
    float* qs = reinterpret_cast<float*>(ci + kCapC);
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sh_sel, sh_rank, sh_cnt, sh_n;
    __shared__ float sh_sum[kThreads / 64], sh_part[kThreads];
    const int tid = threadIdx.x, qi = blockIdx.x;
    const int grp = tid >> 4, l = tid & 15;
    const float* prow = prob + (size_t)qi * K;
    float part = 0.f;
    // 16 loads per thread in flight, then folded in the same order as one at a time
    for (int c0 = 0; c0 < K; c0 += 16 * kThreads) {
        float pv[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int c = c0 + tid + kThreads * j;
            pv[j] = c < K ? prow[c] : 0.f;
        }
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int c = c0 + tid + kThreads * j;
            if (c < K) {
                part += pv[j];
                keys[c] = float_key(-pv[j]);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    if ((tid & 63) == 0) sh_sum[tid >> 6] = part;
    const float* q = q_rot + (size_t)qi * D;
    for (int d = tid; d < D; d += kThreads) qs[d] = q[d];
    __syncthreads();
    float total = 0.f;
    for (int w = 0; w < kThreads / 64; w++) total += sh_sum[w];

    if (total < min_conf) {  // uniform
        // the fixed list needs every cluster's exact distance: listed here, computed across
        // the GPU by fallback_dist_kernel and selected by fallback_pick_kernel (one workgroup
        // would take the K distances alone, long after the other queries are done)
        if (tid == 0) {
            fb_list[atomicAdd(fb_count, 1u)] = (uint32_t)qi;
            out_len[qi] = (uint32_t)fb_np;
            atomicAdd(n_fallback, 1u);
        }
        return;
    }

    const int T = min(max_np, K);
    const uint32_t U = radix_select(keys, K, (uint32_t)T, hist, &sh_sel, &sh_rank);
    const int cnt = gather_sorted(keys, K, U, cd, ci, &sh_cnt);  // cd = -p, most probable first
    // n = shortest prefix of the sorted list whose mass reaches mass * (mass of the first L);
    // prefix sums blocked by thread (thread t sums entries [t*per, (t+1)*per))
    const int L = min(T, cnt);
    const int per = (L + kThreads - 1) / kThreads;
    float own = 0.f;
    for (int i = tid * per; i < min(L, (tid + 1) * per); i++) own += -cd[i];
    sh_part[tid] = own;
    if (tid == 0) sh_n = (uint32_t)L;
    __syncthreads();
    for (int o = 1; o < kThreads; o <<= 1) {  // inclusive scan of the thread sums
        const float v = tid >= o ? sh_part[tid - o] : 0.f;
        __syncthreads();
        sh_part[tid] += v;
        __syncthreads();
    }
    const float goal = mass * sh_part[kThreads - 1];
    float acc = tid ? sh_part[tid - 1] : 0.f;
    for (int i = tid * per; i < min(L, (tid + 1) * per); i++) {
        acc += -cd[i];
        if (acc >= goal) {
            atomicMin(&sh_n, (uint32_t)(i + 1));
            break;
        }
    }
    __syncthreads();
    const int n = max(min(max((int)sh_n, min_np), T), 1);
    __syncthreads();
    // exact distances of the kept clusters, then order them by (distance, id)
    for (int base = 0; base < n; base += kThreads / 16) {
        const int i = base + grp;
        const bool act = i < n;
        float pdist = l2sqr16_group(qs, centroids + (size_t)(act ? ci[i] : 0) * D, D, l, act);
        if (act && l == 0) cd[i] = pdist;
    }
    __syncthreads();
    int P = 1;
    while (P < n) P <<= 1;
    for (int i = n + tid; i < P; i += kThreads) {
        cd[i] = FLT_MAX;
        ci[i] = 0xFFFFFFFFu;
    }
    __syncthreads();
    bitonic_by_key(cd, ci, P);
    for (int i = tid; i < n; i += kThreads) {
        out_ids[(size_t)qi * max_np + i] = ci[i];
        out_d[(size_t)qi * max_np + i] = cd[i];
    }
    if (tid == 0) out_len[qi] = (uint32_t)n;
}

// Exact L2Sqr16 distances of the listed (fallback) queries to every cluster, written over
// their probability rows (no longer needed): workgroup (x, y) takes clusters
// [kFbChunk y, kFbChunk (y + 1)) of list entries x, x + gridDim.x, ...
constexpr int kFbChunk = 64;
__global__ void __launch_bounds__(kThreads)
    fallback_dist_kernel(float* __restrict__ prob, const float* __restrict__ q_rot,
                         const float* __restrict__ centroids, int K, int D,
                         const uint32_t* __restrict__ fb_list, const uint32_t* __restrict__ fb_count) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* qs = reinterpret_cast<float*>(smem);
    const int tid = threadIdx.x, grp = tid >> 4, l = tid & 15;
    const uint32_t cnt = *fb_count;
    const int c0 = (int)blockIdx.y * kFbChunk;
    for (uint32_t li = blockIdx.x; li < cnt; li += gridDim.x) {  // uniform
        const uint32_t qi = fb_list[li];
        __syncthreads();
        for (int d = tid; d < D; d += kThreads) qs[d] = q_rot[(size_t)qi * D + d];
        __syncthreads();
        for (int base = c0; base < min(K, c0 + kFbChunk); base += kThreads / 16) {
            const int c = base + grp;
            const bool act = c < min(K, c0 + kFbChunk);
            const float dist = l2sqr16_group(qs, centroids + (size_t)(act ? c : 0) * D, D, l, act);
            if (act && l == 0) prob[(size_t)qi * K + c] = dist;
        }
    }
}

// The fixed list of each listed query: its fb_np nearest clusters by exact distance (ties by
// id), from the distances fallback_dist_kernel wrote.
__global__ void __launch_bounds__(kThreads)
    fallback_pick_kernel(const float* __restrict__ dist, int K, int max_np, int fb_np,
                         const uint32_t* __restrict__ fb_list, const uint32_t* __restrict__ fb_count,
                         uint32_t* __restrict__ out_ids, float* __restrict__ out_d) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* keys = reinterpret_cast<uint32_t*>(smem);
    float* cd = reinterpret_cast<float*>(smem + 4 * (size_t)K);
    uint32_t* ci = reinterpret_cast<uint32_t*>(cd + kCapC);
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sh_sel, sh_rank, sh_cnt;
    const int tid = threadIdx.x;
    const uint32_t cnt = *fb_count;
    for (uint32_t li = blockIdx.x; li < cnt; li += gridDim.x) {  // uniform
        const uint32_t qi = fb_list[li];
        __syncthreads();
        for (int c = tid; c < K; c += kThreads) keys[c] = float_key(dist[(size_t)qi * K + c]);
        __syncthreads();
        const uint32_t U = radix_select(keys, K, (uint32_t)fb_np, hist, &sh_sel, &sh_rank);
        gather_sorted(keys, K, U, cd, ci, &sh_cnt);
        for (int i = tid; i < fb_np; i += kThreads) {
            out_ids[(size_t)qi * max_np + i] = ci[i];
            out_d[(size_t)qi * max_np + i] = cd[i];
        }
    }
}

// Small batches: probe_select_kernel's steps over three kernels, so that the candidates'
// exact distances -- most of one query's selection -- spread over many workgroups.  cand
// holds per query kProbeCandWords words: the candidate count (above kCapC: every cluster is
// a candidate), then the candidates.  The same candidates, distances and order as
// probe_select_kernel, so the same lists.
__global__ void __launch_bounds__(kThreads)
    probe_cand_kernel(const float* __restrict__ approx, const float* __restrict__ qnorm,
                      const float* __restrict__ cnorm, int K, int D, int nprobe,
                      uint32_t* __restrict__ cand) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* keys = reinterpret_cast<uint32_t*>(smem);  // K
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sh_sel, sh_rank, sh_cnt;
    const int tid = threadIdx.x, qi = blockIdx.x;
    const float* arow = approx + (size_t)qi * K;
    const float qn = qnorm[qi];
    const float coef = (float)(4 * D + 64) * 5.9604644775390625e-8f;  // 2^-24
    for (int c = tid; c < K; c += kThreads) keys[c] = float_key(arow[c] + coef * (qn + cnorm[c]));
    if (tid == 0) sh_cnt = 0;
    __syncthreads();
    const float U = key_float(radix_select(keys, K, (uint32_t)nprobe, hist, &sh_sel, &sh_rank));
    uint32_t* cq = cand + (size_t)qi * kProbeCandWords;
    for (int c = tid; c < K; c += kThreads) {
        if (arow[c] - coef * (qn + cnorm[c]) <= U) {
            const uint32_t slot = atomicAdd(&sh_cnt, 1u);
            if (slot < (uint32_t)kCapC) cq[1 + slot] = (uint32_t)c;
        }
    }
    __syncthreads();
    if (tid == 0) cq[0] = sh_cnt;
}

// Workgroup (q, y): the exact L2Sqr16 distances of query q's candidates 16 y .. 16 y + 15
// (one per 16-lane group), into dist[q K + i] -- for cluster i when every cluster is one.
__global__ void __launch_bounds__(kThreads)
    cand_dist_kernel(const uint32_t* __restrict__ cand, const float* __restrict__ q_rot,
                     const float* __restrict__ centroids, int K, int D,
                     float* __restrict__ dist) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* qs = reinterpret_cast<float*>(smem);  // D
    const int tid = threadIdx.x, qi = blockIdx.x, grp = tid >> 4, l = tid & 15;
    const uint32_t* cq = cand + (size_t)qi * kProbeCandWords;
    const uint32_t cnt = cq[0];
    const bool all = cnt > (uint32_t)kCapC;
    const int n = all ? K : (int)cnt;
    const int i0 = (int)blockIdx.y * (kThreads / 16);
    if (i0 >= n) return;  // uniform
    for (int d = tid; d < D; d += kThreads) qs[d] = q_rot[(size_t)qi * D + d];
    __syncthreads();
    const int i = i0 + grp;
    const bool act = i < n;
    const uint32_t c = act ? (all ? (uint32_t)i : cq[1 + i]) : 0u;
    const float p = l2sqr16_group(qs, centroids + (size_t)c * D, D, l, act);
    if (act && l == 0) dist[(size_t)qi * K + i] = p;
}

// One workgroup per query: the first nprobe candidates by (distance, id), as
// probe_select_kernel orders them (every cluster a candidate: the radix-selected nprobe
// nearest, as its fallback branch).
__global__ void __launch_bounds__(kThreads)
    cand_sort_kernel(const uint32_t* __restrict__ cand, const float* __restrict__ dist, int K,
                     int nprobe, uint32_t* __restrict__ probe_ids, float* __restrict__ probe_d,
                     uint32_t* __restrict__ fixed_off) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* keys = reinterpret_cast<uint32_t*>(smem);          // K
    float* cd = reinterpret_cast<float*>(smem + 4 * (size_t)K);  // kCapC
    uint32_t* ci = reinterpret_cast<uint32_t*>(cd + kCapC);      // kCapC
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sh_sel, sh_rank, sh_cnt;
    const int tid = threadIdx.x, qi = blockIdx.x;
    const uint32_t* cq = cand + (size_t)qi * kProbeCandWords;
    const float* drow = dist + (size_t)qi * K;
    if (fixed_off && qi == 0)  // the lists' CSR offsets, every query nprobe entries
        for (int i = tid; i <= (int)gridDim.x; i += kThreads) fixed_off[i] = (uint32_t)(i * nprobe);
    int cnt = (int)cq[0];
    if (cnt <= kCapC) {
        for (int i = tid; i < cnt; i += kThreads) {
            ci[i] = cq[1 + i];
            cd[i] = drow[i];
        }
    } else {
        for (int c = tid; c < K; c += kThreads) keys[c] = float_key(drow[c]);
        if (tid == 0) sh_cnt = 0;
        __syncthreads();
        const uint32_t U2 = radix_select(keys, K, (uint32_t)nprobe, hist, &sh_sel, &sh_rank);
        for (int c = tid; c < K; c += kThreads) {
            if (keys[c] <= U2) {
                const uint32_t slot = atomicAdd(&sh_cnt, 1u);
                if (slot < (uint32_t)kCapC) {
                    ci[slot] = (uint32_t)c;
                    cd[slot] = key_float(keys[c]);
                }
            }
        }
        __syncthreads();
        cnt = min((int)sh_cnt, kCapC);
    }
    int P2 = 1;
    while (P2 < cnt) P2 <<= 1;
    for (int i = cnt + tid; i < P2; i += kThreads) {
        cd[i] = FLT_MAX;
        ci[i] = 0xFFFFFFFFu;
    }
    __syncthreads();
    bitonic_by_key(cd, ci, P2);
    for (int i = tid; i < nprobe; i += kThreads) {
        probe_ids[(size_t)qi * nprobe + i] = ci[i];
        probe_d[(size_t)qi * nprobe + i] = cd[i];
    }
}

__global__ void zero_u32_kernel(uint32_t* p) { *p = 0; }

// CSR offsets from per-query lengths by one workgroup, 8192 lengths at a time: loaded into
// LDS with coalesced reads, thread t sums lengths 8 t .. 8 t + 7, the sums are scanned across
// the workgroup, and the offsets leave through LDS again.
__global__ void __launch_bounds__(1024) lens_to_offsets_kernel(const uint32_t* __restrict__ len,
                                                               int nq, uint32_t* __restrict__ off) {
    constexpr int kPer = 8, kTile = 1024 * kPer;
    __shared__ uint32_t tile[kTile];
    __shared__ uint32_t wsum[16];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    uint32_t carry = 0;
    for (int base = 0; base < nq; base += kTile) {
        for (int i = t; i < kTile; i += 1024) tile[i] = base + i < nq ? len[base + i] : 0u;
        __syncthreads();
        uint32_t v[kPer], s = 0;
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            v[j] = tile[kPer * t + j];
            s += v[j];
        }
        uint32_t x = s;  // inclusive scan over the wave, then over the waves
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        uint32_t pre = carry, tot = 0;
        for (int w = 0; w < 16; w++) {
            pre += w < wv ? wsum[w] : 0u;
            tot += wsum[w];
        }
        uint32_t run = pre + x - s;  // lengths before this thread's eight
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            tile[kPer * t + j] = run;
            run += v[j];
        }
        __syncthreads();
        for (int i = t; i < kTile && base + i < nq; i += 1024) off[base + i] = tile[i];
        carry += tot;
        __syncthreads();  // the tile and the wave sums are reused by the next chunk
    }
    if (t == 0) off[nq] = carry;
}

__global__ void compact_probes_kernel(const uint32_t* __restrict__ pid, const float* __restrict__ pd,
                                      int stride, const uint32_t* __restrict__ off, int nq,
                                      uint32_t* __restrict__ out_ids, float* __restrict__ out_d) {
    const int q = blockIdx.x;
    if (q >= nq) return;
    const uint32_t a = off[q], n = off[q + 1] - a;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        out_ids[a + i] = pid[(size_t)q * stride + i];
        out_d[a + i] = pd[(size_t)q * stride + i];
    }
}

}  // namespace

hipError_t launch_adaptive_probe(float* prob, const float* q_rot, const float* centroids,
                                 int nq, int K, int D, float mass, int min_np, int max_np,
                                 float min_conf, int fb_np, uint32_t* pad_ids, float* pad_d,
                                 uint32_t* len, uint32_t* off, uint32_t* csr_ids, float* csr_d,
                                 uint32_t* n_fallback, uint32_t* fb_list, uint32_t* fb_count,
                                 hipStream_t s) {
    if (nq == 0) return hipSuccess;
    const size_t lds = 4 * (size_t)K + 8 * (size_t)kCapC + 4 * (size_t)D;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, fb_count);
    hipLaunchKernelGGL(adaptive_select_kernel, dim3(nq), dim3(kThreads), lds, s, prob, q_rot,
                       centroids, K, D, mass, min_np, max_np, min_conf, fb_np, pad_ids, pad_d, len,
                       n_fallback, fb_list, fb_count);
    const unsigned fb_x = (unsigned)std::min(nq, 16);
    hipLaunchKernelGGL(fallback_dist_kernel, dim3(fb_x, (K + kFbChunk - 1) / kFbChunk),
                       dim3(kThreads), 4 * (size_t)D, s, prob, q_rot, centroids, K, D, fb_list,
                       fb_count);
    hipLaunchKernelGGL(fallback_pick_kernel, dim3((unsigned)std::min(nq, 512)), dim3(kThreads),
                       4 * (size_t)K + 8 * (size_t)kCapC, s, prob, K, max_np, fb_np, fb_list,
                       fb_count, pad_ids, pad_d);
    hipLaunchKernelGGL(lens_to_offsets_kernel, dim3(1), dim3(1024), 0, s, len, nq, off);
    hipLaunchKernelGGL(compact_probes_kernel, dim3(nq), dim3(64), 0, s, pad_ids, pad_d, max_np,
                       off, nq, csr_ids, csr_d);
    return hipGetLastError();
}

hipError_t launch_compact_probes(const uint32_t* pad_ids, const float* pad_d, int stride,
                                 const uint32_t* off, int nq, uint32_t* csr_ids, float* csr_d,
                                 hipStream_t s) {
    if (nq == 0) return hipSuccess;
    hipLaunchKernelGGL(compact_probes_kernel, dim3(nq), dim3(64), 0, s, pad_ids, pad_d, stride,
                       off, nq, csr_ids, csr_d);
    return hipGetLastError();
}

size_t probe_select_lds(int K, int D) { return 4 * (size_t)K + 8 * (size_t)kCapC + 4 * (size_t)D; }

hipError_t launch_probe_select(float* approx, const float* qnorm, const float* cnorm,
                               const float* q_rot, const float* centroids, int nq, int K, int D,
                               int nprobe, uint32_t* probe_ids, float* probe_d, hipStream_t s,
                               uint32_t* cand, uint32_t* fixed_off) {
    if (nq == 0) return hipSuccess;
    if (nprobe < 1 || nprobe > kCapC || nprobe > K) return hipErrorInvalidValue;
    if (fixed_off && !(cand && (size_t)nq <= kProbeSplitNq)) return hipErrorInvalidValue;
    if (cand && (size_t)nq <= kProbeSplitNq) {
        const size_t sort_lds = 4 * (size_t)K + 8 * (size_t)kCapC;
        if (sort_lds > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute(
                reinterpret_cast<const void*>(&cand_sort_kernel),
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sort_lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(probe_cand_kernel, dim3(nq), dim3(kThreads), 4 * (size_t)K, s, approx,
                           qnorm, cnorm, K, D, nprobe, cand);
        hipLaunchKernelGGL(cand_dist_kernel, dim3(nq, (K + kThreads / 16 - 1) / (kThreads / 16)),
                           dim3(kThreads), 4 * (size_t)D, s, cand, q_rot, centroids, K, D, approx);
        hipLaunchKernelGGL(cand_sort_kernel, dim3(nq), dim3(kThreads), sort_lds, s, cand, approx, K,
                           nprobe, probe_ids, probe_d, fixed_off);
        return hipGetLastError();
    }
    size_t lds = probe_select_lds(K, D);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&probe_select_kernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(probe_select_kernel, dim3(nq), dim3(kThreads), lds, s, approx, qnorm, cnorm,
                       q_rot, centroids, K, D, nprobe, probe_ids, probe_d);
    return hipGetLastError();
}

}  // namespace erq
