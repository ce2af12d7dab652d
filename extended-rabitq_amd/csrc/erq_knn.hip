// erq_knn.hip -- dataset preparation on the GPU: the exact k nearest neighbours of the
// ground truth (python/compute_gt.py:12-15, numpy brute force there) and Lloyd k-means IVF
// training with each vector's nearest centroid (python/ivf.py:26-35, faiss IVF{K},Flat
// there).
//
// Nearest neighbours, per block of query rows:
//   1. a[j] = ||q||^2 + ||y_j||^2 - 2 q.y_j by the FP32 MFMA GEMM (erq_gemm.hip);
//      |a[j] - ||q - y_j||^2| <= E_j = (D + D/32 + 32) 2^-24 (||q||^2 + ||y_j||^2): at most
//      D + 1 roundings in the dot product (|q.y| terms bounded by (q^2 + y^2) / 2), D/64 + 6
//      in each FP32 norm, three in the epilogue;
//   2. per query (one workgroup): a threshold U with at least k entries a[j] + E_j <= U,
//      from an order statistic of a strided sample of the row (checked by a count; a higher
//      order statistic if the count falls short), then every j with a[j] - E_j <= U -- a
//      superset of the exact k nearest -- into the query's candidate list;
//   3. each candidate's distance in double (16 lane-strided fma chains over the exact
//      differences, _mm512_reduce_add_ps order), then the k smallest by (distance, id).
// k-means: assignment = the same GEMM + the probe selection's exact L2Sqr16 ranking with
// nprobe 1 (launch_probe_select, ties by id); update = per cluster, one workgroup walks the
// assignment array in order, gathers its members (an ordered compaction) and sums them in
// double in increasing id order, so the centroids do not depend on scheduling.  An empty
// cluster keeps its centroid.
#include <algorithm>
#include <vector>

#include "erq.h"
#include "erq_device.h"
#include "erq_internal.h"

namespace erq {

namespace {

constexpr int kT = 256;
constexpr int kSample = 4096;    // strided sample of a row for the threshold
constexpr int kLdsCand = 4096;   // candidates sorted in LDS
constexpr uint32_t kCandCap = 65536;  // candidates per query in the global list
constexpr int kMaxDim = 8192;

__device__ __forceinline__ uint64_t double_key(double d) {
    const uint64_t u = (uint64_t)__double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double key_double(uint64_t k) {
    const uint64_t u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double((long long)u);
}

// Block-wide sum of one value per thread (every thread gets it).
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* red) {
    const int tid = threadIdx.x;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// Block-wide radix select: the key of rank `rank` (1-based, <= n) among key(0..n), 8 bits
// per pass from the top.
template <typename KT, typename F>
__device__ KT radix_select(uint32_t n, uint32_t rank, F key, uint32_t* hist, uint32_t* sh) {
    const int tid = threadIdx.x;
    KT prefix = 0, pmask = 0;
    if (tid == 0) sh[1] = rank;
    for (int shift = (int)sizeof(KT) * 8 - 8; shift >= 0; shift -= 8) {
        for (int i = tid; i < 256; i += kT) hist[i] = 0;
        __syncthreads();
        for (uint32_t c = tid; c < n; c += kT) {
            const KT kv = key(c);
            if ((kv & pmask) == prefix) atomicAdd(&hist[(uint32_t)(kv >> shift) & 255u], 1u);
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
            const uint32_t r = sh[1], excl = incl - own;
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
                sh[0] = sel;
                sh[1] = r - cum;
            }
        }
        __syncthreads();
        prefix |= (KT)sh[0] << shift;
        pmask |= (KT)255 << shift;
        __syncthreads();
    }
    return prefix;
}

// Ascending bitonic sort of P (a power of two) (key, id) pairs in LDS.
template <typename KT>
__device__ void bitonic_pairs(KT* key, uint32_t* id, int P) {
    const int tid = threadIdx.x;
    for (int size = 2; size <= P; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = tid; t < (P >> 1); t += kT) {
                const int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
                const bool up = (lo & size) == 0;
                const KT kl = key[lo], kh = key[hi];
                const uint32_t il = id[lo], ih = id[hi];
                const bool hl = kh < kl || (kh == kl && ih < il);
                const bool lh = kl < kh || (kl == kh && il < ih);
                if (up ? hl : lh) {
                    key[lo] = kh;
                    key[hi] = kl;
                    id[lo] = ih;
                    id[hi] = il;
                }
            }
            __syncthreads();
        }
}

// Squared L2 in double by one 16-lane group: lane l chains dims l, l + 16, ... with the
// exact float differences, then the xor-8/4/2/1 butterfly (orc_l2sqr16_f64).
__device__ __forceinline__ double l2sqr16_f64_group(const float* qs, const float* y, int D, int l,
                                                    bool act) {
    double p = 0.0;
    if (act)
        for (int d = l; d < D; d += 16) {
            const double t = (double)qs[d] - (double)y[d];
            p = fma(t, t, p);
        }
    p = p + __shfl_xor(p, 8);
    p = p + __shfl_xor(p, 4);
    p = p + __shfl_xor(p, 2);
    p = p + __shfl_xor(p, 1);
    return p;
}

struct KnnArgs {
    const float* approx;  // rows x n, row stride n
    const float* qn;      // rows
    const float* yn;      // n
    const float* Q;       // rows x D
    const float* Y;       // n x D
    uint32_t n;
    int D, k;
    float coef;
    uint32_t* cand_id;   // rows x kCandCap
    uint64_t* cand_key;  // rows x kCandCap
    uint32_t* ids;       // rows x k
    float* dists;        // rows x k (nullable)
    uint32_t* err;       // bit 0: candidate list full; bit 1: more than kLdsCand ties
};

__global__ void __launch_bounds__(kT) knn_select_kernel(KnnArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint64_t* skey = reinterpret_cast<uint64_t*>(smem);                 // kLdsCand
    uint32_t* sid = reinterpret_cast<uint32_t*>(skey + kLdsCand);       // kLdsCand
    float* qs = reinterpret_cast<float*>(sid + kLdsCand);               // D
    float* samp = reinterpret_cast<float*>(skey);                       // kSample (phase 1)
    __shared__ uint32_t hist[256], sh[4], red[4], cnt;
    const int tid = threadIdx.x, l = tid & 15, grp = tid >> 4;
    const uint32_t row = blockIdx.x, n = a.n, k = (uint32_t)a.k;
    const float* arow = a.approx + (size_t)row * n;
    const float qn = a.qn[row], coef = a.coef;
    auto hi_of = [&](uint32_t j) { return arow[j] + coef * (qn + a.yn[j]); };
    auto lo_of = [&](uint32_t j) { return arow[j] - coef * (qn + a.yn[j]); };

    // 1. threshold: an order statistic of a strided sample, raised until >= k entries
    //    have a + E at or below it (the sample is the whole row when n <= kSample)
    const uint32_t s = n < (uint32_t)kSample ? n : (uint32_t)kSample;
    for (uint32_t i = tid; i < (uint32_t)kSample; i += kT)
        samp[i] = i < s ? hi_of((uint32_t)(((uint64_t)i * n) / s)) : FLT_MAX;
    __syncthreads();
    {
        uint32_t* sidx = sid;  // unused ids of the float sort
        for (uint32_t i = tid; i < (uint32_t)kSample; i += kT) sidx[i] = i;
        __syncthreads();
        bitonic_pairs<float>(samp, sidx, kSample);
    }
    uint32_t j = s == n ? k : min(s, (uint32_t)(((uint64_t)k * s + n - 1) / n) + 2u);
    float U = samp[j - 1];
    while (true) {  // block-uniform
        uint32_t c = 0;
        for (uint32_t i = tid; i < n; i += kT) c += hi_of(i) <= U;
        if (block_sum(c, red) >= k) break;
        if (j >= s) {
            U = FLT_MAX;
            break;
        }
        j = min(s, 2 * j);
        U = samp[j - 1];
    }
    __syncthreads();

    // 2. candidates: every j with a - E <= U
    if (tid == 0) cnt = 0;
    __syncthreads();
    uint32_t* cid = a.cand_id + (size_t)row * kCandCap;
    for (uint32_t i = tid; i < n; i += kT)
        if (lo_of(i) <= U) {
            const uint32_t slot = atomicAdd(&cnt, 1u);
            if (slot < kCandCap) cid[slot] = i;
        }
    __syncthreads();
    const uint32_t nc = cnt;
    uint32_t* oid = a.ids + (size_t)row * k;
    float* od = a.dists ? a.dists + (size_t)row * k : nullptr;
    auto fail = [&](uint32_t bit) {
        for (uint32_t i = tid; i < k; i += kT) {
            oid[i] = 0xFFFFFFFFu;
            if (od) od[i] = FLT_MAX;
        }
        if (tid == 0) atomicOr(a.err, bit);
    };
    if (nc > kCandCap) {
        fail(1u);
        return;
    }

    // 3. exact distances in double
    const int D = a.D;
    const float* q = a.Q + (size_t)row * D;
    for (int d = tid; d < D; d += kT) qs[d] = q[d];
    __syncthreads();
    const bool in_lds = nc <= (uint32_t)kLdsCand;
    uint64_t* gkey = a.cand_key + (size_t)row * kCandCap;
    for (uint32_t base = 0; base < nc; base += kT / 16) {
        const uint32_t i = base + grp;
        const bool act = i < nc;
        const uint32_t yi = act ? cid[i] : 0u;
        const double dd = l2sqr16_f64_group(qs, a.Y + (size_t)yi * D, D, l, act);
        if (act && l == 0) {
            if (in_lds) {
                skey[i] = double_key(dd);
                sid[i] = yi;
            } else {
                gkey[i] = double_key(dd);
            }
        }
    }
    __syncthreads();

    // 4. the k smallest by (distance, id)
    uint32_t m = nc;
    if (!in_lds) {
        const uint64_t T = radix_select<uint64_t>(nc, k, [&](uint32_t i) { return gkey[i]; },
                                                  hist, sh);
        if (tid == 0) cnt = 0;
        __syncthreads();
        for (uint32_t i = tid; i < nc; i += kT)
            if (gkey[i] <= T) {
                const uint32_t slot = atomicAdd(&cnt, 1u);
                if (slot < (uint32_t)kLdsCand) {
                    skey[slot] = gkey[i];
                    sid[slot] = cid[i];
                }
            }
        __syncthreads();
        m = cnt;
        if (m > (uint32_t)kLdsCand) {
            fail(2u);
            return;
        }
    }
    int P2 = 1;
    while ((uint32_t)P2 < m) P2 <<= 1;
    for (int i = (int)m + tid; i < P2; i += kT) {
        skey[i] = ~0ull;
        sid[i] = 0xFFFFFFFFu;
    }
    __syncthreads();
    bitonic_pairs<uint64_t>(skey, sid, P2);
    for (uint32_t i = tid; i < k; i += kT) {
        oid[i] = sid[i];
        if (od) od[i] = (float)key_double(skey[i]);
    }
}

size_t knn_lds(int D) { return (size_t)kLdsCand * 12 + 4 * (size_t)D; }

// ---------------------------------------------------------------- k-means update
// One workgroup per cluster: walks assign[0..n) 1024 entries at a time (thread t takes
// entries 4t .. 4t+3 of the step), compacts the cluster's members in order into LDS and
// adds their rows to per-thread double sums (dims t + 256 i) in that order.
template <int DPT>
__global__ void __launch_bounds__(kT)
    kmeans_update_kernel(const float* __restrict__ X, int ldx, uint32_t n, int D,
                         const uint32_t* __restrict__ assign, float* __restrict__ C,
                         uint32_t* __restrict__ counts) {
    __shared__ uint32_t members[4 * kT];
    __shared__ uint32_t wtot[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t cl = blockIdx.x;
    double acc[DPT];
#pragma unroll
    for (int i = 0; i < DPT; i++) acc[i] = 0.0;
    uint32_t total = 0;
    for (uint32_t base = 0; base < n; base += 4 * kT) {  // block-uniform
        const uint32_t e = base + 4 * (uint32_t)tid;
        bool m[4];
        uint32_t own = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            m[i] = e + i < n && assign[e + i] == cl;
            own += m[i];
        }
        uint32_t incl = own;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        if (lane == 63) wtot[wid] = incl;
        __syncthreads();
        uint32_t pos = incl - own;
        for (int w = 0; w < wid; w++) pos += wtot[w];
        const uint32_t nm = wtot[0] + wtot[1] + wtot[2] + wtot[3];
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (m[i]) members[pos++] = e + i;
        __syncthreads();
        for (uint32_t r = 0; r < nm; r++) {
            const float* row = X + (size_t)members[r] * ldx;
#pragma unroll
            for (int i = 0; i < DPT; i++) {
                const int d = tid + kT * i;
                if (d < D) acc[i] = acc[i] + (double)row[d];
            }
        }
        total += nm;
        __syncthreads();
    }
    if (total) {
#pragma unroll
        for (int i = 0; i < DPT; i++) {
            const int d = tid + kT * i;
            if (d < D) C[(size_t)cl * D + d] = (float)(acc[i] / (double)total);
        }
    }
    if (tid == 0 && counts) counts[cl] = total;
}

__global__ void gather_rows_kernel(const float* __restrict__ X, int ldx, int D,
                                   const uint32_t* __restrict__ rows, uint32_t nrows,
                                   float* __restrict__ out) {
    const size_t total = (size_t)nrows * D;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / D, d = i % D;
        out[i] = X[(size_t)rows[r] * ldx + d];
    }
}

hipError_t launch_kmeans_update(const float* X, int ldx, uint32_t n, int D,
                                const uint32_t* assign, uint32_t K, float* C, uint32_t* counts,
                                hipStream_t s) {
    if (K == 0) return hipSuccess;
    const int dpt = (D + kT - 1) / kT;
    dim3 grid(K), block(kT);
    if (dpt <= 1)
        hipLaunchKernelGGL(kmeans_update_kernel<1>, grid, block, 0, s, X, ldx, n, D, assign, C, counts);
    else if (dpt <= 2)
        hipLaunchKernelGGL(kmeans_update_kernel<2>, grid, block, 0, s, X, ldx, n, D, assign, C, counts);
    else if (dpt <= 4)
        hipLaunchKernelGGL(kmeans_update_kernel<4>, grid, block, 0, s, X, ldx, n, D, assign, C, counts);
    else if (dpt <= 8)
        hipLaunchKernelGGL(kmeans_update_kernel<8>, grid, block, 0, s, X, ldx, n, D, assign, C, counts);
    else if (dpt <= 16)
        hipLaunchKernelGGL(kmeans_update_kernel<16>, grid, block, 0, s, X, ldx, n, D, assign, C, counts);
    else if (dpt <= 32)
        hipLaunchKernelGGL(kmeans_update_kernel<32>, grid, block, 0, s, X, ldx, n, D, assign, C, counts);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// ---------------------------------------------------------------- host helpers
// Device buffers of one call, freed after the stream has drained.
struct Temps {
    hipStream_t s;
    std::vector<void*> p;
    explicit Temps(hipStream_t st) : s(st) {}
    template <typename T>
    int get(size_t count, T** out) {
        void* v = nullptr;
        hipError_t e = hipMalloc(&v, std::max<size_t>(count * sizeof(T), 16));
        if (e != hipSuccess) return hip_fail(e, "device allocation");
        p.push_back(v);
        *out = static_cast<T*>(v);
        return ERQ_OK;
    }
    ~Temps() {
        if (!p.empty()) (void)hipStreamSynchronize(s);
        for (void* v : p) (void)hipFree(v);
    }
};

#define ERQ_TRY(x)                   \
    do {                             \
        const int rc_ = (x);         \
        if (rc_ != ERQ_OK) return rc_; \
    } while (0)
#define ERQ_HIP(x, what)                            \
    do {                                            \
        const hipError_t e_ = (x);                  \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

int pad16(int d) { return (d + 15) & ~15; }

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// Rows of x with a stride that is a multiple of 16 floats and 16-byte alignment (the GEMM's
// operand layout): x itself when it already is, else a zero-padded copy.
int padded_rows(Temps& t, const float* x, size_t rows, int D, const float** out) {
    const int Dp = pad16(D);
    if (Dp == D && aligned16(x)) {
        *out = x;
        return ERQ_OK;
    }
    float* p = nullptr;
    ERQ_TRY(t.get(rows * Dp, &p));
    ERQ_HIP(launch_pad_rows(x, (int)rows, D, p, Dp, t.s), "pad rows");
    *out = p;
    return ERQ_OK;
}

int kmeans_assign_impl(Temps& t, const float* Xp, size_t n, int Dp, const float* Cp, uint32_t K,
                       const float* xn, uint32_t* assign, float* dist) {
    hipStream_t s = t.s;
    float *cn = nullptr, *approx = nullptr;
    ERQ_TRY(t.get(K, &cn));
    ERQ_HIP(launch_row_sqnorm(Cp, (int)K, Dp, Dp, cn, s), "centroid norms");
    const size_t chunk = std::max<size_t>(1, std::min<size_t>(n, ((size_t)1 << 28) / K));
    ERQ_TRY(t.get(chunk * K, &approx));
    for (size_t r0 = 0; r0 < n; r0 += chunk) {
        const int rows = (int)std::min(chunk, n - r0);
        ERQ_HIP(launch_gemm(Xp + r0 * Dp, Dp, Cp, Dp, 1, approx, (int)K, rows, (int)K, Dp, 1,
                            xn + r0, cn, s),
                "assignment gemm");
        ERQ_HIP(launch_probe_select(approx, xn + r0, cn, Xp + r0 * Dp, Cp, rows, (int)K, Dp, 1,
                                    assign + r0, dist + r0, s),
                "nearest centroid");
    }
    return ERQ_OK;
}

int check_common(const void* x, uint64_t n, uint32_t dim) {
    if (!x) {
        set_error("null pointer");
        return ERQ_EINVAL;
    }
    if (n == 0 || n >= 0x7FFFFFFFull) {
        set_error("row count must be in [1, 2^31)");
        return ERQ_EINVAL;
    }
    if (dim == 0 || dim > (uint32_t)kMaxDim) {
        set_error("dim must be in [1, 8192]");
        return ERQ_EUNSUPPORTED;
    }
    return ERQ_OK;
}

}  // namespace

}  // namespace erq

using namespace erq;

extern "C" int erq_knn_dev(const float* base, uint64_t n, const float* queries, uint64_t nq,
                           uint32_t dim, uint32_t k, uint32_t* ids, float* dists, void* stream) {
    return guarded([&]() -> int {
        ERQ_TRY(check_common(base, n, dim));
        if (nq == 0) return ERQ_OK;
        if (!queries || !ids || nq >= 0x7FFFFFFFull) {
            set_error("bad query arguments");
            return ERQ_EINVAL;
        }
        if (k == 0 || k > n || k > (uint32_t)kLdsCand) {
            set_error("k must be in [1, min(n, 4096)]");
            return ERQ_EINVAL;
        }
        hipStream_t s = static_cast<hipStream_t>(stream);
        Temps t(s);
        const int Dp = pad16((int)dim);
        const float *Yp = nullptr, *Qp = nullptr;
        ERQ_TRY(padded_rows(t, base, n, (int)dim, &Yp));
        ERQ_TRY(padded_rows(t, queries, nq, (int)dim, &Qp));
        float *yn = nullptr, *qn = nullptr, *approx = nullptr;
        uint32_t *cand_id = nullptr, *err = nullptr;
        uint64_t* cand_key = nullptr;
        ERQ_TRY(t.get(n, &yn));
        ERQ_TRY(t.get(nq, &qn));
        ERQ_TRY(t.get(1, &err));
        ERQ_HIP(launch_row_sqnorm(Yp, (int)n, Dp, Dp, yn, s), "base norms");
        ERQ_HIP(launch_row_sqnorm(Qp, (int)nq, Dp, Dp, qn, s), "query norms");
        ERQ_HIP(hipMemsetAsync(err, 0, 4, s), "clear flags");
        size_t chunk = std::min<size_t>(nq, 4096);
        chunk = std::min(chunk, std::max<size_t>(1, ((size_t)1 << 29) / n));
        chunk = std::min(chunk, std::max<size_t>(1, ((size_t)1 << 30) / (12 * (size_t)kCandCap)));
        ERQ_TRY(t.get(chunk * n, &approx));
        ERQ_TRY(t.get(chunk * kCandCap, &cand_id));
        ERQ_TRY(t.get(chunk * kCandCap, &cand_key));
        const size_t lds = knn_lds(Dp);
        ERQ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&knn_select_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                "knn lds");
        const float coef = (float)(Dp + Dp / 32 + 32) * 5.9604644775390625e-8f;  // 2^-24
        for (size_t r0 = 0; r0 < nq; r0 += chunk) {
            const int rows = (int)std::min(chunk, nq - r0);
            ERQ_HIP(launch_gemm(Qp + r0 * Dp, Dp, Yp, Dp, 1, approx, (int)n, rows, (int)n, Dp, 1,
                                qn + r0, yn, s),
                    "distance gemm");
            KnnArgs a{approx, qn + r0, yn, Qp + r0 * Dp, Yp, (uint32_t)n, Dp, (int)k, coef,
                      cand_id, cand_key, ids + r0 * k, dists ? dists + r0 * k : nullptr, err};
            hipLaunchKernelGGL(knn_select_kernel, dim3(rows), dim3(kT), lds, s, a);
            ERQ_HIP(hipGetLastError(), "knn select");
        }
        uint32_t flags = 0;
        ERQ_HIP(hipMemcpyAsync(&flags, err, 4, hipMemcpyDeviceToHost, s), "flags");
        ERQ_HIP(hipStreamSynchronize(s), "knn");
        if (flags) {
            set_error(flags & 1u ? "more than 65536 candidates within the GEMM error bound of "
                                   "one query's k-th distance"
                                 : "more than 4096 base vectors tie at one query's k-th distance");
            return ERQ_EUNSUPPORTED;
        }
        return ERQ_OK;
    });
}

extern "C" int erq_kmeans_assign_dev(const float* x, uint64_t n, uint32_t dim,
                                     const float* centroids, uint32_t K, uint32_t* assign,
                                     float* dist, void* stream) {
    return guarded([&]() -> int {
        ERQ_TRY(check_common(x, n, dim));
        if (!centroids || !assign || !dist || K == 0) {
            set_error("bad centroid arguments");
            return ERQ_EINVAL;
        }
        hipStream_t s = static_cast<hipStream_t>(stream);
        Temps t(s);
        const int Dp = pad16((int)dim);
        const float *Xp = nullptr, *Cp = nullptr;
        float* xn = nullptr;
        ERQ_TRY(padded_rows(t, x, n, (int)dim, &Xp));
        ERQ_TRY(padded_rows(t, centroids, K, (int)dim, &Cp));
        ERQ_TRY(t.get(n, &xn));
        ERQ_HIP(launch_row_sqnorm(Xp, (int)n, Dp, Dp, xn, s), "row norms");
        return kmeans_assign_impl(t, Xp, n, Dp, Cp, K, xn, assign, dist);
    });
}

extern "C" int erq_kmeans_update_dev(const float* x, uint64_t n, uint32_t dim,
                                     const uint32_t* assign, uint32_t K, float* centroids,
                                     uint32_t* counts, void* stream) {
    return guarded([&]() -> int {
        ERQ_TRY(check_common(x, n, dim));
        if (!assign || !centroids || K == 0) {
            set_error("bad centroid arguments");
            return ERQ_EINVAL;
        }
        ERQ_HIP(launch_kmeans_update(x, (int)dim, (uint32_t)n, (int)dim, assign, K, centroids,
                                     counts, static_cast<hipStream_t>(stream)),
                "k-means update");
        return ERQ_OK;
    });
}

extern "C" int erq_kmeans_dev(const float* x, uint64_t n, uint32_t dim, uint32_t K,
                              uint32_t iters, const uint32_t* init_rows, float* centroids,
                              uint32_t* assign, float* dist, void* stream) {
    return guarded([&]() -> int {
        ERQ_TRY(check_common(x, n, dim));
        if (!init_rows || !centroids || !assign || !dist || K == 0 || K > n) {
            set_error("bad k-means arguments (K must be in [1, n])");
            return ERQ_EINVAL;
        }
        hipStream_t s = static_cast<hipStream_t>(stream);
        Temps t(s);
        const int D = (int)dim, Dp = pad16(D);
        const float* Xp = nullptr;
        float *xn = nullptr, *Cp = nullptr;
        ERQ_TRY(padded_rows(t, x, n, D, &Xp));
        ERQ_TRY(t.get(n, &xn));
        ERQ_TRY(t.get((size_t)K * Dp, &Cp));
        const bool pad_c = Dp != D || !aligned16(centroids);
        ERQ_HIP(launch_row_sqnorm(Xp, (int)n, Dp, Dp, xn, s), "row norms");
        const int gb = (int)std::min<size_t>(((size_t)K * D + 255) / 256, 65536);
        hipLaunchKernelGGL(gather_rows_kernel, dim3(gb), dim3(256), 0, s, x, D, D, init_rows, K,
                           centroids);
        ERQ_HIP(hipGetLastError(), "initial centroids");
        for (uint32_t it = 0; it <= iters; it++) {
            const float* C = centroids;
            if (pad_c) {
                ERQ_HIP(launch_pad_rows(centroids, (int)K, D, Cp, Dp, s), "pad centroids");
                C = Cp;
            }
            ERQ_TRY(kmeans_assign_impl(t, Xp, n, Dp, C, K, xn, assign, dist));
            if (it == iters) break;
            ERQ_HIP(launch_kmeans_update(Xp, Dp, (uint32_t)n, D, assign, K, centroids, nullptr, s),
                    "k-means update");
        }
        ERQ_HIP(hipStreamSynchronize(s), "k-means");
        return ERQ_OK;
    });
}
