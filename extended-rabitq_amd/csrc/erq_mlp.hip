// erq_mlp.hip -- learned probe predictor: MLP  D -> H -> H -> K  (GELU), BF16 weights and
// activations on v_mfma_f32_16x16x32_bf16 with FP32 accumulation, trained with a
// pos-weighted BCE loss on "cluster holds one of the query's true top-k" labels, Adam on
// FP32 master weights.  p = sigmoid(W3 gelu(W2 gelu(W1 x + b1) + b2) + b3).
//
// Weights are stored out-major ([out x in], as a Linear layer), so the forward GEMM reads
// both operands k-contiguous; the backward products transpose what enters transposed and
// run the same kernel.  GEMM tile: 128 (or 64) x 128 per 256-thread workgroup (2 x 2 waves
// of 64 x 64 = 4 x 4 MFMA tiles), K step 32 double-buffered in LDS as [row][k] so each
// lane's MFMA operand (8 consecutive k) is one 16-byte LDS read.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "erq.h"
#include "erq_device.h"
#include "erq_internal.h"

struct erq_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    int in = 0, hid = 0, out = 0, max_batch = 0;
    size_t nparams = 0;
    float* master = nullptr;  // FP32 [W1 b1 W2 b2 W3 b3]
    float* grad = nullptr;    // same layout
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    __bf16* wb = nullptr;     // BF16 copies of W1, W2, W3 (no biases)
    long step = 0;
    float xscale = 1.f;       // inputs enter layer 1 as bf16(xscale * x)
    // activations (max_batch rows)
    __bf16* xb = nullptr;     // B x in
    float* z1 = nullptr;      // B x hid (pre-activation)
    __bf16* h1 = nullptr;
    float* z2 = nullptr;
    __bf16* h2 = nullptr;
    float* logit = nullptr;   // B x out
    __bf16* dlog = nullptr;   // B x out
    __bf16* dz2 = nullptr;    // B x hid
    __bf16* dz1 = nullptr;
    float* tmp = nullptr;     // B x hid fp32 scratch
    __bf16* tA = nullptr;     // transposed operands of the backward products:
    __bf16* tB = nullptr;     //   max(out, hid) x B, max(hid, in) x B
    __bf16* wt = nullptr;     //   and hid x max(out, hid) (a transposed weight)
    float* loss = nullptr;    // 1
    float* part = nullptr;    // partial sums of the loss and of the bias gradients
    // inference activations for batches above max_batch (inf_rows rows, made on first use)
    size_t inf_rows = 0;
    __bf16* ih1 = nullptr;
    __bf16* ih2 = nullptr;
    std::mutex mu;
};

namespace erq {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int GBN = 128, GBK = 32, GPAD = 8;
constexpr size_t kInferRows = 32768;  // rows per inference chunk above max_batch

// EPI_BIAS_GELU keeps the pre-activation (training); EPI_BIAS_GELU_H stores only the BF16
// activation and EPI_BIAS_SIGMOID the probability sigmoid(acc + bias) (inference).
enum Epi {
    EPI_F32 = 0, EPI_BIAS_F32 = 1, EPI_BIAS_GELU = 2, EPI_BIAS_GELU_H = 3, EPI_BIAS_SIGMOID = 4
};

__device__ __forceinline__ float gelu(float z) {
    return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
           z * 0.39894228040143268f * expf(-0.5f * z * z);
}
__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

// C(m, n) = sum_k A(m, k) B(n, k): both operands k-contiguous, A [M x K] (lda), B [N x K]
// (ldb).  Workgroup tile BM x 128 (BM = 64 or 128), K step 32, LDS double-buffered: the next
// step's tiles are loaded into registers while the current step's feed the MFMAs, then
// stored to the other buffer, one barrier per step.  4 waves as 2 x 2, wave tile BM/2 x 64
// = TM x 4 tiles of v_mfma_f32_16x16x32_bf16.  AF: A is fp32 and enters as bf16(ascale * a)
// (layer 1's input conversion fused into the load).  Edges in M, N and K are zero-filled;
// 16-byte loads need lda and ldb to be multiples of 8 and 16-byte aligned bases.  Each
// output sums its k in the same order whatever BM, so a row's result does not depend on
// the batch it is computed in.
template <int BM, int EPI, bool AF>
__global__ void __launch_bounds__(256) gemm_nt_kernel(const void* __restrict__ Av, int lda,
                                                      const __bf16* __restrict__ B, int ldb,
                                                      int M, int N, int K, float ascale,
                                                      float* __restrict__ Cf, int ldc,
                                                      __bf16* __restrict__ Cb,
                                                      const float* __restrict__ bias) {
    constexpr int TM = BM / 32;               // 16-row tiles per wave along m
    constexpr int ACH = BM * GBK / 8 / 256;   // 16-byte A chunks per thread per step
    constexpr int BCH = GBN * GBK / 8 / 256;  // same for B
    // one LDS buffer: the double-buffered A and B tiles, then the epilogue's staging
    constexpr int kTileBytes = 2 * (BM + GBN) * (GBK + GPAD) * 2, kStageBytes = 4 * 16 * 68 * 4;
    __shared__ __align__(16) unsigned char lds[kTileBytes > kStageBytes ? kTileBytes : kStageBytes];
    auto As = reinterpret_cast<__bf16 (*)[BM][GBK + GPAD]>(lds);
    auto Bs = reinterpret_cast<__bf16 (*)[GBN][GBK + GPAD]>(lds + 2 * BM * (GBK + GPAD) * 2);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * GBN;
    const __bf16 zero = (__bf16)0.0f;
    f4 acc[TM][4];  // every loop over it is unrolled, so it stays in registers
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    bf16x8 ra[ACH], rb[BCH];

    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < ACH; i++) {
            const int c = tid + i * 256, gm = m0 + (c >> 2), k = k0 + (c & 3) * 8;
            if (AF) {
                const float* src = reinterpret_cast<const float*>(Av) + (size_t)gm * lda + k;
                if (gm < M && k + 8 <= K) {
                    const f4 a0 = *reinterpret_cast<const f4*>(src);
                    const f4 a1 = *reinterpret_cast<const f4*>(src + 4);
                    for (int e = 0; e < 4; e++) {
                        ra[i][e] = (__bf16)(ascale * a0[e]);
                        ra[i][4 + e] = (__bf16)(ascale * a1[e]);
                    }
                } else {
                    for (int e = 0; e < 8; e++)
                        ra[i][e] = (gm < M && k + e < K) ? (__bf16)(ascale * src[e]) : zero;
                }
            } else {
                const __bf16* src = reinterpret_cast<const __bf16*>(Av) + (size_t)gm * lda + k;
                if (gm < M && k + 8 <= K) ra[i] = *reinterpret_cast<const bf16x8*>(src);
                else
                    for (int e = 0; e < 8; e++) ra[i][e] = (gm < M && k + e < K) ? src[e] : zero;
            }
        }
#pragma unroll
        for (int i = 0; i < BCH; i++) {
            const int c = tid + i * 256, gn = n0 + (c >> 2), k = k0 + (c & 3) * 8;
            const __bf16* src = B + (size_t)gn * ldb + k;
            if (gn < N && k + 8 <= K) rb[i] = *reinterpret_cast<const bf16x8*>(src);
            else
                for (int e = 0; e < 8; e++) rb[i][e] = (gn < N && k + e < K) ? src[e] : zero;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < ACH; i++) {
            const int c = tid + i * 256;
            *reinterpret_cast<bf16x8*>(&As[buf][c >> 2][(c & 3) * 8]) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < BCH; i++) {
            const int c = tid + i * 256;
            *reinterpret_cast<bf16x8*>(&Bs[buf][c >> 2][(c & 3) * 8]) = rb[i];
        }
    };

    const int nk = (K + GBK - 1) / GBK;
    load(0);
    store(0);
    __syncthreads();
    const int li = lane & 15, lg = lane >> 4;
    for (int kt = 0; kt < nk; kt++) {
        const int cur = kt & 1;
        if (kt + 1 < nk) load((kt + 1) * GBK);  // in flight during this step's MFMAs
        bf16x8 af[TM], bfr[4];
#pragma unroll
        for (int i = 0; i < TM; i++)
            af[i] = *reinterpret_cast<const bf16x8*>(&As[cur][wm * (BM / 2) + i * 16 + li][lg * 8]);
#pragma unroll
        for (int j = 0; j < 4; j++)
            bfr[j] = *reinterpret_cast<const bf16x8*>(&Bs[cur][wn * 64 + j * 16 + li][lg * 8]);
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0,
                                                                    0, 0);
        // the other buffer was last read in the previous step, before its closing barrier
        if (kt + 1 < nk) store(cur ^ 1);
        __syncthreads();
    }
    // Epilogue through LDS (free after the last barrier): per 16-row slice of the wave tile,
    // the accumulators are staged as 16 x 64 floats, then each lane takes 4 consecutive
    // columns of a row, so a row's 64 outputs leave as one contiguous segment.
    float (*T)[68] = reinterpret_cast<float (*)[68]>(lds) + wid * 16;
    const int col = lane & 15, rg = (lane >> 4) * 4;
    const int sr = lane >> 4, sc = (lane & 15) * 4, gn = n0 + wn * 64 + sc;
    f4 bv = f4{0.f, 0.f, 0.f, 0.f};
    if (EPI != EPI_F32) {
#pragma unroll
        for (int e = 0; e < 4; e++) bv[e] = gn + e < N ? bias[gn + e] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < TM; i++) {
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) T[rg + r][j * 16 + col] = acc[i][j][r];
        __syncthreads();
#pragma unroll
        for (int pass = 0; pass < 4; pass++) {
            const int row = pass * 4 + sr, gm = m0 + wm * (BM / 2) + i * 16 + row;
            const f4 v = *reinterpret_cast<const f4*>(&T[row][sc]);
            if (gm < M) {
                f4 out;
                bf16x4 hb;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const float z = EPI == EPI_F32 ? v[e] : v[e] + bv[e];
                    out[e] = EPI == EPI_BIAS_SIGMOID ? sigmoid(z) : z;
                    if (EPI == EPI_BIAS_GELU || EPI == EPI_BIAS_GELU_H) hb[e] = (__bf16)gelu(z);
                }
                const size_t o = (size_t)gm * ldc + gn;
                if (gn + 4 <= N) {
                    if (EPI != EPI_BIAS_GELU_H) *reinterpret_cast<f4*>(Cf + o) = out;
                    if (EPI == EPI_BIAS_GELU || EPI == EPI_BIAS_GELU_H)
                        *reinterpret_cast<bf16x4*>(Cb + o) = hb;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        if (gn + e < N) {
                            if (EPI != EPI_BIAS_GELU_H) Cf[o + e] = out[e];
                            if (EPI == EPI_BIAS_GELU || EPI == EPI_BIAS_GELU_H) Cb[o + e] = hb[e];
                        }
                }
            }
        }
        __syncthreads();  // the slice is read before the next one is staged
    }
}

// C = A B^T with the operand layouts of gemm_nt_kernel; A is bf16 unless AF (fp32, scaled).
// ldc is the row stride of whichever output the epilogue writes (Cf, Cb or both).
template <int EPI, bool AF = false>
hipError_t gemm(const void* A, int lda, const __bf16* B, int ldb, int M, int N, int K,
                float* Cf, int ldc, __bf16* Cb, const float* bias, hipStream_t s,
                float ascale = 1.f) {
    if ((lda | ldb | ldc) % 8 || ((uintptr_t)A | (uintptr_t)B | (uintptr_t)Cf | (uintptr_t)Cb) % 16)
        return hipErrorInvalidValue;
    if (M <= 0 || N <= 0) return hipSuccess;
    const unsigned gx = (unsigned)((N + GBN - 1) / GBN);
    // 128-row tiles when they still give the chip several workgroups per CU, else 64
    if ((size_t)gx * ((M + 127) / 128) >= 1024) {
        hipLaunchKernelGGL((gemm_nt_kernel<128, EPI, AF>), dim3(gx, (M + 127) / 128), dim3(256), 0,
                           s, A, lda, B, ldb, M, N, K, ascale, Cf, ldc, Cb, bias);
    } else {
        hipLaunchKernelGGL((gemm_nt_kernel<64, EPI, AF>), dim3(gx, (M + 63) / 64), dim3(256), 0,
                           s, A, lda, B, ldb, M, N, K, ascale, Cf, ldc, Cb, bias);
    }
    return hipGetLastError();
}

// out [C x R] = in [R x C]^T, bf16, 64 x 64 tiles through LDS
__global__ void __launch_bounds__(256) transpose_bf16_kernel(const __bf16* __restrict__ in, int R,
                                                             int C, __bf16* __restrict__ out) {
    __shared__ __bf16 t[64][66];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int i = ty; i < 64; i += 4) {
        const int r = r0 + i, c = c0 + tx;
        t[i][tx] = (r < R && c < C) ? in[(size_t)r * C + c] : (__bf16)0.0f;
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
        const int c = c0 + i, r = r0 + tx;
        if (c < C && r < R) out[(size_t)c * R + r] = t[tx][i];
    }
}

hipError_t transpose(const __bf16* in, int R, int C, __bf16* out, hipStream_t s) {
    hipLaunchKernelGGL(transpose_bf16_kernel, dim3((C + 63) / 64, (R + 63) / 64), dim3(256), 0, s,
                       in, R, C, out);
    return hipGetLastError();
}

__global__ void f32_to_bf16_kernel(const float* __restrict__ x, __bf16* __restrict__ y, size_t n,
                                   float scale) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] = (__bf16)(scale * x[i]);
}

// BCE with logits, positive weight w: l = w*y*softplus(-z) + (1-y)*softplus(z);
// dl/dz = w*y*(p-1) + (1-y)*p.  Loss summed over clusters, averaged over the batch.
__global__ void bce_kernel(const float* __restrict__ z, const float* __restrict__ y,
                           __bf16* __restrict__ dz, size_t n, float w, float inv_batch,
                           float* __restrict__ loss) {
    float part = 0.f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const float zi = z[i], yi = y[i];
        const float p = 1.0f / (1.0f + expf(-zi));
        const float sp_pos = fmaxf(zi, 0.f) + log1pf(expf(-fabsf(zi)));  // softplus(z)
        const float sp_neg = sp_pos - zi;                                   // softplus(-z)
        part += w * yi * sp_neg + (1.f - yi) * sp_pos;
        dz[i] = (__bf16)((w * yi * (p - 1.f) + (1.f - yi) * p) * inv_batch);
    }
    // one partial per workgroup (summed in a fixed order by sum_partials_kernel)
    __shared__ float wsum[4];
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) loss[blockIdx.x] = ((wsum[0] + wsum[1]) + (wsum[2] + wsum[3])) * inv_batch;
}

// out[0] = sum of in[0..n) in a fixed order (single workgroup of 256)
__global__ void __launch_bounds__(256) sum_partials_kernel(const float* __restrict__ in, int n,
                                                           float* __restrict__ out) {
    __shared__ float sh[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += in[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// dz = dh * gelu'(z), dh fp32 [B x H]
__global__ void gelu_back_kernel(const float* __restrict__ dh, const float* __restrict__ z,
                                 __bf16* __restrict__ dz, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dz[i] = (__bf16)(dh[i] * gelu_grad(z[i]));
}

// db[n] = sum_b dz[b][n]
// Column sums of dz [B x N] in two passes with a fixed order: workgroup (column tile of 64,
// row slice rs of kColSlices) -> part[rs][n]; then db[n] = sum over rs.
constexpr int kColSlices = 16;

__global__ void __launch_bounds__(256) colsum_partial_kernel(const __bf16* __restrict__ dz, int B,
                                                             int N, float* __restrict__ part) {
    __shared__ float sh[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + lane, rs = blockIdx.y;
    const int r0 = (int)((long)B * rs / kColSlices), r1 = (int)((long)B * (rs + 1) / kColSlices);
    float s = 0.f;
    if (n < N)
        for (int b = r0 + wv; b < r1; b += 4) s += (float)dz[(size_t)b * N + n];
    sh[wv][lane] = s;
    __syncthreads();
    if (wv == 0 && n < N) part[(size_t)rs * N + n] = (sh[0][lane] + sh[1][lane]) + (sh[2][lane] + sh[3][lane]);
}

__global__ void colsum_final_kernel(const float* __restrict__ part, int N, float* __restrict__ db) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int rs = 0; rs < kColSlices; rs++) s += part[(size_t)rs * N + n];
    db[n] = s;
}

hipError_t colsum(const __bf16* dz, int B, int N, float* part, float* db, hipStream_t s) {
    hipLaunchKernelGGL(colsum_partial_kernel, dim3((N + 63) / 64, kColSlices), dim3(256), 0, s, dz,
                       B, N, part);
    hipLaunchKernelGGL(colsum_final_kernel, dim3((N + 255) / 256), dim3(256), 0, s, part, N, db);
    return hipGetLastError();
}

__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, size_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        float gi = g[i];
        float mi = b1 * m[i] + (1.f - b1) * gi;
        float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

__global__ void zero_f32_kernel(float* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.f;
}

int grid_for(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 8192); }

struct Layout {
    size_t w1, b1, w2, b2, w3, b3;
};
Layout layout(const erq_mlp* m) {
    Layout L;
    L.w1 = 0;
    L.b1 = L.w1 + (size_t)m->hid * m->in;
    L.w2 = L.b1 + m->hid;
    L.b2 = L.w2 + (size_t)m->hid * m->hid;
    L.w3 = L.b2 + m->hid;
    L.b3 = L.w3 + (size_t)m->out * m->hid;
    return L;
}

hipError_t refresh_bf16(erq_mlp* m, hipStream_t s) {
    Layout L = layout(m);
    size_t n1 = (size_t)m->hid * m->in, n2 = (size_t)m->hid * m->hid, n3 = (size_t)m->out * m->hid;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(grid_for(n1)), dim3(256), 0, s, m->master + L.w1,
                       m->wb, n1, 1.f);
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(grid_for(n2)), dim3(256), 0, s, m->master + L.w2,
                       m->wb + n1, n2, 1.f);
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(grid_for(n3)), dim3(256), 0, s, m->master + L.w3,
                       m->wb + n1 + n2, n3, 1.f);
    return hipGetLastError();
}

// training forward on B <= max_batch rows: xb = bf16(xscale x), z1, h1, z2, h2 and the
// logits (fp32) kept for the backward pass
hipError_t forward(erq_mlp* m, const float* x, int B, hipStream_t s) {
    Layout L = layout(m);
    const size_t n1 = (size_t)m->hid * m->in, n2 = (size_t)m->hid * m->hid;
    const __bf16 *W1 = m->wb, *W2 = m->wb + n1, *W3 = m->wb + n1 + n2;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(grid_for((size_t)B * m->in)), dim3(256), 0, s, x,
                       m->xb, (size_t)B * m->in, m->xscale);
    hipError_t e = gemm<EPI_BIAS_GELU>(m->xb, m->in, W1, m->in, B, m->hid, m->in, m->z1, m->hid,
                                       m->h1, m->master + L.b1, s);
    if (e == hipSuccess)
        e = gemm<EPI_BIAS_GELU>(m->h1, m->hid, W2, m->hid, B, m->hid, m->hid, m->z2, m->hid,
                                m->h2, m->master + L.b2, s);
    if (e == hipSuccess)
        e = gemm<EPI_BIAS_F32>(m->h2, m->hid, W3, m->hid, B, m->out, m->hid, m->logit, m->out,
                               nullptr, m->master + L.b3, s);
    return e;
}

// inference on B rows: probabilities straight into prob (B x out), the input conversion
// fused into layer 1 and the sigmoid into layer 3 (the same values as the training
// forward's bf16 input, activations and logits); h1, h2: B x hid scratch
hipError_t infer(erq_mlp* m, const float* x, int B, __bf16* h1, __bf16* h2, float* prob,
                 hipStream_t s) {
    Layout L = layout(m);
    const size_t n1 = (size_t)m->hid * m->in, n2 = (size_t)m->hid * m->hid;
    const __bf16 *W1 = m->wb, *W2 = m->wb + n1, *W3 = m->wb + n1 + n2;
    hipError_t e = gemm<EPI_BIAS_GELU_H, true>(x, m->in, W1, m->in, B, m->hid, m->in, nullptr,
                                               m->hid, h1, m->master + L.b1, s, m->xscale);
    if (e == hipSuccess)
        e = gemm<EPI_BIAS_GELU_H>(h1, m->hid, W2, m->hid, B, m->hid, m->hid, nullptr, m->hid, h2,
                                  m->master + L.b2, s);
    if (e == hipSuccess)
        e = gemm<EPI_BIAS_SIGMOID>(h2, m->hid, W3, m->hid, B, m->out, m->hid, prob, m->out,
                                   nullptr, m->master + L.b3, s);
    return e;
}

// backward after forward with labels y (fp32 0/1): grads into m->grad, loss into m->loss.
// Every product runs as A B^T with k-contiguous operands: the activations and the weights
// that enter transposed are transposed first (tA, tB, wt).
hipError_t backward(erq_mlp* m, const float* y, int B, float pos_weight, hipStream_t s) {
    Layout L = layout(m);
    const int in = m->in, hid = m->hid, out = m->out;
    const size_t n1 = (size_t)hid * in, n2 = (size_t)hid * hid;
    const __bf16 *W2 = m->wb + n1, *W3 = m->wb + n1 + n2;
    const size_t no = (size_t)B * out, nh = (size_t)B * hid;
    const int gb = grid_for(no);
    hipLaunchKernelGGL(bce_kernel, dim3(gb), dim3(256), 0, s, m->logit, y, m->dlog, no,
                       pos_weight, 1.0f / B, m->part);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, m->part, gb, m->loss);
    // layer 3: dW3 [out x hid] = dlog^T h2; db3; dh2 = dlog W3
    hipError_t e = transpose(m->dlog, B, out, m->tA, s);
    if (!e) e = transpose(m->h2, B, hid, m->tB, s);
    if (!e) e = gemm<EPI_F32>(m->tA, B, m->tB, B, out, hid, B, m->grad + L.w3, hid, nullptr,
                              nullptr, s);
    if (!e) e = colsum(m->dlog, B, out, m->part, m->grad + L.b3, s);
    if (!e) e = transpose(W3, out, hid, m->wt, s);
    if (!e) e = gemm<EPI_F32>(m->dlog, out, m->wt, out, B, hid, out, m->tmp, hid, nullptr,
                              nullptr, s);
    if (e) return e;
    hipLaunchKernelGGL(gelu_back_kernel, dim3(grid_for(nh)), dim3(256), 0, s, m->tmp, m->z2,
                       m->dz2, nh);
    // layer 2: dW2 = dz2^T h1; db2; dh1 = dz2 W2
    e = transpose(m->dz2, B, hid, m->tA, s);
    if (!e) e = transpose(m->h1, B, hid, m->tB, s);
    if (!e) e = gemm<EPI_F32>(m->tA, B, m->tB, B, hid, hid, B, m->grad + L.w2, hid, nullptr,
                              nullptr, s);
    if (!e) e = colsum(m->dz2, B, hid, m->part, m->grad + L.b2, s);
    if (!e) e = transpose(W2, hid, hid, m->wt, s);
    if (!e) e = gemm<EPI_F32>(m->dz2, hid, m->wt, hid, B, hid, hid, m->tmp, hid, nullptr,
                              nullptr, s);
    if (e) return e;
    hipLaunchKernelGGL(gelu_back_kernel, dim3(grid_for(nh)), dim3(256), 0, s, m->tmp, m->z1,
                       m->dz1, nh);
    // layer 1: dW1 = dz1^T xb; db1
    e = transpose(m->dz1, B, hid, m->tA, s);
    if (!e) e = transpose(m->xb, B, in, m->tB, s);
    if (!e) e = gemm<EPI_F32>(m->tA, B, m->tB, B, hid, in, B, m->grad + L.w1, in, nullptr,
                              nullptr, s);
    if (!e) e = colsum(m->dz1, B, hid, m->part, m->grad + L.b1, s);
    if (e) return e;
    return hipGetLastError();
}

}  // namespace
}  // namespace erq

using namespace erq;

namespace {

void mlp_release(erq_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void* ptrs[] = {m->master, m->grad, m->adam_m, m->adam_v, m->wb, m->xb, m->z1, m->h1,
                    m->z2, m->h2, m->logit, m->dlog, m->dz2, m->dz1, m->tmp, m->loss,
                    m->part, m->tA, m->tB, m->wt, m->ih1, m->ih2};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

template <typename T>
hipError_t alloc(T** p, size_t n) {
    return hipMalloc(p, std::max<size_t>(n, 1) * sizeof(T));
}

hipStream_t mlp_stream(erq_mlp*, void* s) { return (hipStream_t)s; }

}  // namespace

extern "C" {

int erq_mlp_create(size_t in_dim, size_t hidden, size_t out_dim, size_t max_batch, uint64_t seed,
                   erq_mlp** out) {
    if (!out) {
        set_error("erq_mlp_create: null argument");
        return ERQ_EINVAL;
    }
    *out = nullptr;
    if (in_dim == 0 || hidden == 0 || out_dim == 0 || max_batch == 0 || in_dim % 32 ||
        hidden % 32 || out_dim % 8 || max_batch % 32) {
        set_error("erq_mlp_create: in/hidden must be multiples of 32, out of 8, max_batch of 32");
        return ERQ_EINVAL;
    }
    erq_mlp* m = new erq_mlp();
    m->in = (int)in_dim;
    m->hid = (int)hidden;
    m->out = (int)out_dim;
    m->max_batch = (int)max_batch;
    Layout L = layout(m);
    m->nparams = L.b3 + out_dim;
    hipError_t e = hipGetDevice(&m->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    const size_t B = max_batch, P = m->nparams;
    const size_t nw = (size_t)hidden * in_dim + (size_t)hidden * hidden + (size_t)out_dim * hidden;
    if (e == hipSuccess) e = alloc(&m->master, P);
    if (e == hipSuccess) e = alloc(&m->grad, P);
    if (e == hipSuccess) e = alloc(&m->adam_m, P);
    if (e == hipSuccess) e = alloc(&m->adam_v, P);
    if (e == hipSuccess) e = alloc(&m->wb, nw);
    if (e == hipSuccess) e = alloc(&m->xb, B * in_dim);
    if (e == hipSuccess) e = alloc(&m->z1, B * hidden);
    if (e == hipSuccess) e = alloc(&m->h1, B * hidden);
    if (e == hipSuccess) e = alloc(&m->z2, B * hidden);
    if (e == hipSuccess) e = alloc(&m->h2, B * hidden);
    if (e == hipSuccess) e = alloc(&m->logit, B * out_dim);
    if (e == hipSuccess) e = alloc(&m->dlog, B * out_dim);
    if (e == hipSuccess) e = alloc(&m->dz2, B * hidden);
    if (e == hipSuccess) e = alloc(&m->dz1, B * hidden);
    if (e == hipSuccess) e = alloc(&m->tmp, B * hidden);
    if (e == hipSuccess) e = alloc(&m->tA, B * std::max(out_dim, hidden));
    if (e == hipSuccess) e = alloc(&m->tB, B * std::max(hidden, in_dim));
    if (e == hipSuccess) e = alloc(&m->wt, hidden * std::max(out_dim, hidden));
    if (e == hipSuccess) e = alloc(&m->loss, 1);
    if (e == hipSuccess)
        e = alloc(&m->part, std::max<size_t>(8192, (size_t)kColSlices * std::max(out_dim, hidden)));
    if (e != hipSuccess) {
        mlp_release(m);
        return hip_fail(e, "erq_mlp_create");
    }
    // He-uniform weights, zero biases (host-side init from a seeded generator)
    std::vector<float> h(P, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](size_t off, size_t n, size_t fan_in) {
        std::uniform_real_distribution<float> U(-1.f, 1.f);
        float a = std::sqrt(6.0f / (float)fan_in);
        for (size_t i = 0; i < n; i++) h[off + i] = a * U(rng);
    };
    fill(L.w1, (size_t)hidden * in_dim, in_dim);
    fill(L.w2, (size_t)hidden * hidden, hidden);
    fill(L.w3, (size_t)out_dim * hidden, hidden);
    e = hipMemcpy(m->master, h.data(), P * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->adam_m, 0, P * 4);
    if (e == hipSuccess) e = hipMemset(m->adam_v, 0, P * 4);
    if (e == hipSuccess) e = refresh_bf16(m, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) {
        mlp_release(m);
        return hip_fail(e, "erq_mlp_create init");
    }
    *out = m;
    return ERQ_OK;
}

void erq_mlp_free(erq_mlp* m) { mlp_release(m); }

int erq_mlp_set_input_scale(erq_mlp* m, float scale) {
    if (!m || !(scale > 0.f) || !std::isfinite(scale)) {
        set_error("erq_mlp_set_input_scale: null model or scale not positive and finite");
        return ERQ_EINVAL;
    }
    std::lock_guard<std::mutex> lk(m->mu);
    m->xscale = scale;
    return ERQ_OK;
}

float erq_mlp_input_scale(const erq_mlp* m) { return m ? m->xscale : 0.f; }

size_t erq_mlp_num_params(const erq_mlp* m) { return m ? m->nparams : 0; }
size_t erq_mlp_in_dim(const erq_mlp* m) { return m ? (size_t)m->in : 0; }
size_t erq_mlp_out_dim(const erq_mlp* m) { return m ? (size_t)m->out : 0; }

int erq_mlp_set_params(erq_mlp* m, const float* host_params) {
    if (!m || !host_params) {
        set_error("erq_mlp_set_params: null argument");
        return ERQ_EINVAL;
    }
    std::lock_guard<std::mutex> lk(m->mu);
    hipError_t e = hipMemcpy(m->master, host_params, m->nparams * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = refresh_bf16(m, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    return e == hipSuccess ? ERQ_OK : hip_fail(e, "erq_mlp_set_params");
}

int erq_mlp_get_params(erq_mlp* m, float* host_params, float* host_grads) {
    if (!m || (!host_params && !host_grads)) {
        set_error("erq_mlp_get_params: null argument");
        return ERQ_EINVAL;
    }
    std::lock_guard<std::mutex> lk(m->mu);
    hipError_t e = hipStreamSynchronize(m->stream);
    if (e == hipSuccess && host_params)
        e = hipMemcpy(host_params, m->master, m->nparams * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && host_grads)
        e = hipMemcpy(host_grads, m->grad, m->nparams * 4, hipMemcpyDeviceToHost);
    return e == hipSuccess ? ERQ_OK : hip_fail(e, "erq_mlp_get_params");
}

int erq_mlp_forward_dev(erq_mlp* m, const float* x_dev, size_t batch, float* prob_dev,
                        void* stream) {
    if (!m || (batch && (!x_dev || !prob_dev))) {
        set_error("erq_mlp_forward_dev: null argument");
        return ERQ_EINVAL;
    }
    if (((uintptr_t)x_dev | (uintptr_t)prob_dev) % 16) {
        set_error("erq_mlp_forward_dev: x_dev and prob_dev must be 16-byte aligned");
        return ERQ_EINVAL;
    }
    std::lock_guard<std::mutex> lk(m->mu);
    hipStream_t s = mlp_stream(m, stream);
    if (batch == 0) return ERQ_OK;
    if (batch <= (size_t)m->max_batch) {
        hipError_t e = infer(m, x_dev, (int)batch, m->h1, m->h2, prob_dev, s);
        return e == hipSuccess ? ERQ_OK : hip_fail(e, "mlp forward");
    }
    // large batches: chunks of up to kInferRows rows through the inference buffers (the
    // same per-row results: a row's sums do not depend on the rows beside it)
    const size_t rows = std::min<size_t>(batch, kInferRows);
    if (m->inf_rows < rows) {
        void* old[] = {m->ih1, m->ih2};
        for (void* p : old)
            if (p) (void)hipFree(p);
        m->ih1 = m->ih2 = nullptr;
        m->inf_rows = 0;
        hipError_t e = alloc(&m->ih1, rows * m->hid);
        if (e == hipSuccess) e = alloc(&m->ih2, rows * m->hid);
        if (e != hipSuccess) return hip_fail(e, "mlp inference buffers");
        m->inf_rows = rows;
    }
    for (size_t b0 = 0; b0 < batch; b0 += rows) {
        const int B = (int)std::min<size_t>(rows, batch - b0);
        hipError_t e = infer(m, x_dev + b0 * m->in, B, m->ih1, m->ih2, prob_dev + b0 * m->out, s);
        if (e != hipSuccess) return hip_fail(e, "mlp forward");
    }
    return ERQ_OK;
}

int erq_mlp_loss_grad_dev(erq_mlp* m, const float* x_dev, const float* y_dev, size_t batch,
                          float pos_weight, float* loss_dev, void* stream) {
    if (!m || !x_dev || !y_dev) {
        set_error("erq_mlp_loss_grad_dev: null argument");
        return ERQ_EINVAL;
    }
    if (batch == 0 || batch > (size_t)m->max_batch || batch % 32) {
        set_error("erq_mlp_loss_grad_dev: batch must be a multiple of 32 up to max_batch");
        return ERQ_EINVAL;
    }
    std::lock_guard<std::mutex> lk(m->mu);
    hipStream_t s = mlp_stream(m, stream);
    hipError_t e = forward(m, x_dev, (int)batch, s);
    if (e == hipSuccess) e = backward(m, y_dev, (int)batch, pos_weight, s);
    if (e == hipSuccess && loss_dev)
        e = hipMemcpyAsync(loss_dev, m->loss, 4, hipMemcpyDeviceToDevice, s);
    return e == hipSuccess ? ERQ_OK : hip_fail(e, "mlp loss/grad");
}

int erq_mlp_adam_step_dev(erq_mlp* m, float lr, float beta1, float beta2, float eps,
                          void* stream) {
    if (!m) {
        set_error("erq_mlp_adam_step_dev: null argument");
        return ERQ_EINVAL;
    }
    std::lock_guard<std::mutex> lk(m->mu);
    hipStream_t s = mlp_stream(m, stream);
    m->step++;
    float c1 = 1.f - std::pow(beta1, (float)m->step), c2 = 1.f - std::pow(beta2, (float)m->step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(m->nparams)), dim3(256), 0, s, m->master,
                       m->grad, m->adam_m, m->adam_v, m->nparams, lr, beta1, beta2, eps, c1, c2);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = refresh_bf16(m, s);
    return e == hipSuccess ? ERQ_OK : hip_fail(e, "mlp adam");
}

}  // extern "C"
