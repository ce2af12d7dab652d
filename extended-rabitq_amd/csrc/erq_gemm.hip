// erq_gemm.hip -- FP32 MFMA GEMM for query rotation (q * P, Rotator.hpp:60-63) and the
// query-to-centroid contraction (Initializer.hpp:55-68, as ||q||^2 + ||c||^2 - 2 q.c).
//
// Tiling: 256 threads (4 waves as 2 x 2), 128 x 128 output tile, K step 16 staged in
// LDS (k-major, so each MFMA operand read is 16 consecutive floats per lane group).
// Each wave owns 64 x 64 = 4 x 4 tiles of v_mfma_f32_16x16x4f32; operands for lane i:
// A[row = i%16][k = i/16], B[k = i/16][col = i%16]; result reg j: row 4*(i/16)+j, col i%16.
#include "erq_device.h"
#include "erq_internal.h"

namespace erq {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 16, PAD = 4;

template <int BT, int EPI>
__global__ void __launch_bounds__(256) gemm_kernel(const float* __restrict__ A, int lda,
                                                   const float* __restrict__ B, int ldb,
                                                   float* __restrict__ C, int ldc, int M, int N,
                                                   int K, const float* __restrict__ rn,
                                                   const float* __restrict__ cn) {
    __shared__ __align__(16) float As[BK][BM + PAD];
    __shared__ __align__(16) float Bs[BK][BN + PAD];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;

    floatx4 acc[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    float4 ra[2], rb[2];
    auto load_tile = [&](int k0) {
        // A tile: 128 rows x 16 k as 512 float4; thread -> (row = tid/4 + 64r, quad = tid%4)
        for (int r = 0; r < 2; r++) {
            int row = (tid >> 2) + 64 * r, q4 = tid & 3;
            int gm = m0 + row;
            ra[r] = gm < M ? *reinterpret_cast<const float4*>(A + (size_t)gm * lda + k0 + 4 * q4)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (BT) {  // B^T rows are n
            for (int r = 0; r < 2; r++) {
                int row = (tid >> 2) + 64 * r, q4 = tid & 3;
                int gn = n0 + row;
                rb[r] = gn < N
                            ? *reinterpret_cast<const float4*>(B + (size_t)gn * ldb + k0 + 4 * q4)
                            : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        } else {  // B rows are k: 16 x 128 as 512 float4; thread -> (k = tid/32 + 8r, quad)
            for (int r = 0; r < 2; r++) {
                int kr = (tid >> 5) + 8 * r, q4 = tid & 31;
                int gn = n0 + 4 * q4;
                rb[r] = gn < N ? *reinterpret_cast<const float4*>(B + (size_t)(k0 + kr) * ldb + gn)
                               : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
    };
    auto store_tile = [&]() {
        for (int r = 0; r < 2; r++) {
            int row = (tid >> 2) + 64 * r, q4 = tid & 3;
            As[4 * q4 + 0][row] = ra[r].x;
            As[4 * q4 + 1][row] = ra[r].y;
            As[4 * q4 + 2][row] = ra[r].z;
            As[4 * q4 + 3][row] = ra[r].w;
        }
        if (BT) {
            for (int r = 0; r < 2; r++) {
                int row = (tid >> 2) + 64 * r, q4 = tid & 3;
                Bs[4 * q4 + 0][row] = rb[r].x;
                Bs[4 * q4 + 1][row] = rb[r].y;
                Bs[4 * q4 + 2][row] = rb[r].z;
                Bs[4 * q4 + 3][row] = rb[r].w;
            }
        } else {
            for (int r = 0; r < 2; r++) {
                int kr = (tid >> 5) + 8 * r, q4 = tid & 31;
                *reinterpret_cast<float4*>(&Bs[kr][4 * q4]) = rb[r];
            }
        }
    };

    load_tile(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
        __syncthreads();
        store_tile();
        __syncthreads();
        if (k0 + BK < K) load_tile(k0 + BK);
        const int li = lane & 15, lk = lane >> 4;
        for (int kk = 0; kk < BK; kk += 4) {
            float af[4], bf[4];
            for (int t = 0; t < 4; t++) {
                af[t] = As[kk + lk][wm * 64 + t * 16 + li];
                bf[t] = Bs[kk + lk][wn * 64 + t * 16 + li];
            }
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    }

    const int col_l = lane & 15, rowg = (lane >> 4) * 4;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            const int gn = n0 + wn * 64 + j * 16 + col_l;
            if (gn >= N) continue;
            for (int r = 0; r < 4; r++) {
                const int gm = m0 + wm * 64 + i * 16 + rowg + r;
                if (gm >= M) continue;
                float v = acc[i][j][r];
                if (EPI == 1) v = fmaf(-2.0f, v, rn[gm] + cn[gn]);
                C[(size_t)gm * ldc + gn] = v;
            }
        }
}

__global__ void row_sqnorm_kernel(const float* __restrict__ X, int rows, int cols, int ld,
                                  float* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= rows) return;
    const float* x = X + (size_t)row * ld;
    float s = 0.f;
    for (int c = lane; c < cols; c += 64) s = fmaf(x[c], x[c], s);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) out[row] = s;
}

__global__ void pad_rows_kernel(const float* __restrict__ src, int rows, int cols,
                                float* __restrict__ dst, int ldd) {
    const size_t total = (size_t)rows * ldd;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        size_t r = i / ldd, c = i % ldd;
        dst[i] = c < (size_t)cols ? src[r * cols + c] : 0.f;
    }
}

// C = A B^T for at most kSkinnyRows rows of A (a few queries against the centroid rows): one
// wave per column n, its lanes over k (16 bytes each), the rows' partial sums reduced across
// the wave -- full occupancy where the tiled kernel would launch a handful of workgroups.
// Any order of the FP32 sum keeps the error within the bound the probe selection assumes.
constexpr int kSkinnyRows = 8;

// EPI 0: store; 1: rn[m] + cn[n] - 2 acc; 2: as 1 with rn[m] = ||A_m||^2 summed by every wave
// in the same order (so every column uses the same value), the wave of column 0 storing it to
// rn_out for the caller.
template <int EPI>
__global__ void __launch_bounds__(256) skinny_kernel(const float* __restrict__ A, int lda,
                                                     const float* __restrict__ B, int ldb,
                                                     float* __restrict__ C, int ldc, int M, int N,
                                                     int K, const float* __restrict__ rn,
                                                     const float* __restrict__ cn,
                                                     float* __restrict__ rn_out = nullptr) {
    const int lane = threadIdx.x & 63;
    const int n = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (n >= N) return;  // wave-uniform
    const float* b = B + (size_t)n * ldb;
    float acc[kSkinnyRows], sq[kSkinnyRows];
#pragma unroll
    for (int m = 0; m < kSkinnyRows; m++) acc[m] = sq[m] = 0.f;
    for (int k0 = 4 * lane; k0 < K; k0 += 256) {
        const float4 bv = *reinterpret_cast<const float4*>(b + k0);
#pragma unroll
        for (int m = 0; m < kSkinnyRows; m++)
            if (m < M) {
                const float4 av = *reinterpret_cast<const float4*>(A + (size_t)m * lda + k0);
                acc[m] = fmaf(av.x, bv.x, acc[m]);
                acc[m] = fmaf(av.y, bv.y, acc[m]);
                acc[m] = fmaf(av.z, bv.z, acc[m]);
                acc[m] = fmaf(av.w, bv.w, acc[m]);
                if (EPI == 2) {
                    sq[m] = fmaf(av.x, av.x, sq[m]);
                    sq[m] = fmaf(av.y, av.y, sq[m]);
                    sq[m] = fmaf(av.z, av.z, sq[m]);
                    sq[m] = fmaf(av.w, av.w, sq[m]);
                }
            }
    }
#pragma unroll
    for (int m = 0; m < kSkinnyRows; m++)
        for (int o = 32; o > 0; o >>= 1) {
            acc[m] += __shfl_xor(acc[m], o);
            if (EPI == 2) sq[m] += __shfl_xor(sq[m], o);
        }
    if (lane < kSkinnyRows && lane < M) {
        float v = acc[0], r = sq[0];
#pragma unroll
        for (int m = 1; m < kSkinnyRows; m++)
            if (lane == m) {
                v = acc[m];
                r = sq[m];
            }
        if (EPI == 1) v = fmaf(-2.0f, v, rn[lane] + cn[n]);
        if (EPI == 2) {
            v = fmaf(-2.0f, v, r + cn[n]);
            if (n == 0) rn_out[lane] = r;
        }
        C[(size_t)lane * ldc + n] = v;
    }
}

}  // namespace

hipError_t launch_query_centroid_dist(const float* Q, int nq, int D, const float* centroids,
                                      int K, const float* cnorm, float* approx, float* qnorm,
                                      hipStream_t s) {
    if (nq == 0 || K == 0) return hipSuccess;
    if (nq > kSkinnyRows || D % BK != 0) {
        hipError_t e = launch_row_sqnorm(Q, nq, D, D, qnorm, s);
        if (e != hipSuccess) return e;
        return launch_gemm(Q, D, centroids, D, 1, approx, K, nq, K, D, 1, qnorm, cnorm, s);
    }
    hipLaunchKernelGGL(skinny_kernel<2>, dim3((K + 3) / 4), dim3(256), 0, s, Q, D, centroids, D,
                       approx, K, nq, K, D, nullptr, cnorm, qnorm);
    return hipGetLastError();
}

hipError_t launch_gemm(const float* A, int lda, const float* B, int ldb, int b_trans, float* C,
                       int ldc, int M, int N, int K, int epi, const float* rn, const float* cn,
                       hipStream_t s) {
    if (K % BK != 0 || lda % 4 != 0 || ldb % 4 != 0 || (!b_trans && N % 4 != 0))
        return hipErrorInvalidValue;
    if (M == 0 || N == 0) return hipSuccess;
    if (b_trans && M <= kSkinnyRows) {
        const dim3 g((N + 3) / 4), b(256);
        if (epi)
            hipLaunchKernelGGL(skinny_kernel<1>, g, b, 0, s, A, lda, B, ldb, C, ldc, M, N, K, rn, cn);
        else
            hipLaunchKernelGGL(skinny_kernel<0>, g, b, 0, s, A, lda, B, ldb, C, ldc, M, N, K, rn, cn);
        return hipGetLastError();
    }
    dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM), block(256);
    if (b_trans) {
        if (epi)
            hipLaunchKernelGGL((gemm_kernel<1, 1>), grid, block, 0, s, A, lda, B, ldb, C, ldc, M, N, K, rn, cn);
        else
            hipLaunchKernelGGL((gemm_kernel<1, 0>), grid, block, 0, s, A, lda, B, ldb, C, ldc, M, N, K, rn, cn);
    } else {
        if (epi)
            hipLaunchKernelGGL((gemm_kernel<0, 1>), grid, block, 0, s, A, lda, B, ldb, C, ldc, M, N, K, rn, cn);
        else
            hipLaunchKernelGGL((gemm_kernel<0, 0>), grid, block, 0, s, A, lda, B, ldb, C, ldc, M, N, K, rn, cn);
    }
    return hipGetLastError();
}

hipError_t launch_row_sqnorm(const float* X, int rows, int cols, int ld, float* out, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    dim3 grid((rows + 3) / 4), block(256);
    hipLaunchKernelGGL(row_sqnorm_kernel, grid, block, 0, s, X, rows, cols, ld, out);
    return hipGetLastError();
}

hipError_t launch_pad_rows(const float* src, int rows, int cols, float* dst, int ldd,
                           hipStream_t s) {
    if (rows == 0) return hipSuccess;
    size_t total = (size_t)rows * ldd;
    int blocks = (int)std::min<size_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(pad_rows_kernel, dim3(blocks), dim3(256), 0, s, src, rows, cols, dst, ldd);
    return hipGetLastError();
}

}  // namespace erq
