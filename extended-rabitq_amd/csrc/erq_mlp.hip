// erq_mlp.hip -- learned probe predictor: MLP  D -> H -> H -> K  (GELU), BF16 weights and
// activations on v_mfma_f32_16x16x32_bf16 with FP32 accumulation, trained with a
// pos-weighted BCE loss on "cluster holds one of the query's true top-k" labels, Adam on
// FP32 master weights.  p = sigmoid(W3 gelu(W2 gelu(W1 x + b1) + b2) + b3).
//
// Weights are stored out-major ([out x in], as a Linear layer), so the forward GEMM reads
// both operands k-contiguous.  GEMM tile: 64 x 64 per 256-thread workgroup (2 x 2 waves of
// 32 x 32 = 2 x 2 MFMA tiles), K step 32 staged in LDS as [row][k] so each lane's MFMA
// operand (8 consecutive k) is one 16-byte LDS read.
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
    float* loss = nullptr;    // 1
    float* part = nullptr;    // partial sums of the loss and of the bias gradients
    // inference activations for batches above max_batch (inf_rows rows, made on first use)
    size_t inf_rows = 0;
    __bf16* ixb = nullptr;
    float* iz = nullptr;      // pre-activations of both hidden layers (not kept)
    __bf16* ih1 = nullptr;
    __bf16* ih2 = nullptr;
    std::mutex mu;
};

namespace erq {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int GBM = 64, GBN = 64, GBK = 32, GPAD = 8;
constexpr size_t kInferRows = 32768;  // rows per inference chunk above max_batch

enum Epi { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_BIAS_F32 = 2, EPI_ACC_F32 = 3 };

__device__ __forceinline__ float gelu(float z) {
    return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
           z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// C(m, n) = sum_k A(m, k) B(k, n).  TA: A stored [K x M] (else [M x K]); TB: B stored
// [N x K] (else [K x N]).  Edges in M, N and K are zero-filled; 16-byte loads need the
// leading dimensions to be multiples of 8 and, along m (TA) or n (!TB), M or N too.
template <int TA, int TB, int EPI>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(const __bf16* __restrict__ A, int lda,
                                                        const __bf16* __restrict__ B, int ldb,
                                                        int M, int N, int K,
                                                        float* __restrict__ Cf, int ldc,
                                                        __bf16* __restrict__ Cb,
                                                        const float* __restrict__ bias) {
    __shared__ __align__(16) __bf16 As[GBM][GBK + GPAD];
    __shared__ __align__(16) __bf16 Bs[GBN][GBK + GPAD];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
    const __bf16 zero = (__bf16)0.0f;
    f4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < K; k0 += GBK) {
        // ---- stage A (64 x 32) ----
        if (!TA) {
            int row = tid >> 2, kq = (tid & 3) * 8, gm = m0 + row;
            bf16x8 v;
            const __bf16* src = A + (size_t)gm * lda + k0 + kq;
            if (gm < M && k0 + kq + 8 <= K) v = *reinterpret_cast<const bf16x8*>(src);
            else
                for (int e = 0; e < 8; e++) v[e] = (gm < M && k0 + kq + e < K) ? src[e] : zero;
            *reinterpret_cast<bf16x8*>(&As[row][kq]) = v;
        } else {
            int kk = tid >> 3, mq = (tid & 7) * 8, gm = m0 + mq;
            bf16x8 v;
            if (gm < M && k0 + kk < K)
                v = *reinterpret_cast<const bf16x8*>(A + (size_t)(k0 + kk) * lda + gm);
            else for (int e = 0; e < 8; e++) v[e] = zero;
            for (int e = 0; e < 8; e++) As[mq + e][kk] = v[e];
        }
        // ---- stage B (32 x 64) as [n][k] ----
        if (TB) {
            int row = tid >> 2, kq = (tid & 3) * 8, gn = n0 + row;
            bf16x8 v;
            const __bf16* src = B + (size_t)gn * ldb + k0 + kq;
            if (gn < N && k0 + kq + 8 <= K) v = *reinterpret_cast<const bf16x8*>(src);
            else
                for (int e = 0; e < 8; e++) v[e] = (gn < N && k0 + kq + e < K) ? src[e] : zero;
            *reinterpret_cast<bf16x8*>(&Bs[row][kq]) = v;
        } else {
            int kk = tid >> 3, nq = (tid & 7) * 8, gn = n0 + nq;
            bf16x8 v;
            if (gn < N && k0 + kk < K)
                v = *reinterpret_cast<const bf16x8*>(B + (size_t)(k0 + kk) * ldb + gn);
            else for (int e = 0; e < 8; e++) v[e] = zero;
            for (int e = 0; e < 8; e++) Bs[nq + e][kk] = v[e];
        }
        __syncthreads();
        const int li = lane & 15, lg = lane >> 4;
        bf16x8 af[2], bfr[2];
        for (int i = 0; i < 2; i++)
            af[i] = *reinterpret_cast<const bf16x8*>(&As[wm * 32 + i * 16 + li][lg * 8]);
        for (int j = 0; j < 2; j++)
            bfr[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn * 32 + j * 16 + li][lg * 8]);
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0,
                                                                    0, 0);
        __syncthreads();
    }
    const int col = lane & 15, rg = (lane >> 4) * 4;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            const int gn = n0 + wn * 32 + j * 16 + col;
            if (gn >= N) continue;
            for (int r = 0; r < 4; r++) {
                const int gm = m0 + wm * 32 + i * 16 + rg + r;
                if (gm >= M) continue;
                float v = acc[i][j][r];
                const size_t o = (size_t)gm * ldc + gn;
                if (EPI == EPI_F32) {
                    Cf[o] = v;
                } else if (EPI == EPI_ACC_F32) {
                    Cf[o] += v;
                } else if (EPI == EPI_BIAS_F32) {
                    Cf[o] = v + bias[gn];
                } else {  // bias + GELU: keep the pre-activation, store the BF16 activation
                    float z = v + bias[gn];
                    Cf[o] = z;
                    Cb[o] = (__bf16)gelu(z);
                }
            }
        }
}

template <int TA, int TB, int EPI>
hipError_t gemm(const __bf16* A, int lda, const __bf16* B, int ldb, int M, int N, int K,
                float* Cf, int ldc, __bf16* Cb, const float* bias, hipStream_t s) {
    if ((lda | ldb) % 8) return hipErrorInvalidValue;
    dim3 grid((N + GBN - 1) / GBN, (M + GBM - 1) / GBM);
    hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, EPI>), grid, dim3(256), 0, s, A, lda, B, ldb,
                       M, N, K, Cf, ldc, Cb, bias);
    return hipGetLastError();
}

__global__ void f32_to_bf16_kernel(const float* __restrict__ x, __bf16* __restrict__ y, size_t n,
                                   float scale) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] = (__bf16)(scale * x[i]);
}

__global__ void sigmoid_kernel(const float* __restrict__ z, float* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 1.0f / (1.0f + expf(-z[i]));
}

__global__ void sigmoid_inplace_kernel(float* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 1.0f / (1.0f + expf(-p[i]));
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

// forward on B rows from fp32 inputs into the given activation buffers (B rows each);
// logits (B x out, fp32) into logit
struct Acts {
    __bf16* xb;
    float* z1;
    __bf16* h1;
    float* z2;
    __bf16* h2;
    float* logit;
};

hipError_t forward_into(erq_mlp* m, const Acts& A, const float* x, int B, hipStream_t s) {
    Layout L = layout(m);
    const size_t n1 = (size_t)m->hid * m->in, n2 = (size_t)m->hid * m->hid;
    const __bf16 *W1 = m->wb, *W2 = m->wb + n1, *W3 = m->wb + n1 + n2;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(grid_for((size_t)B * m->in)), dim3(256), 0, s, x,
                       A.xb, (size_t)B * m->in, m->xscale);
    hipError_t e = gemm<0, 1, EPI_BIAS_GELU>(A.xb, m->in, W1, m->in, B, m->hid, m->in, A.z1,
                                             m->hid, A.h1, m->master + L.b1, s);
    if (e == hipSuccess)
        e = gemm<0, 1, EPI_BIAS_GELU>(A.h1, m->hid, W2, m->hid, B, m->hid, m->hid, A.z2, m->hid,
                                      A.h2, m->master + L.b2, s);
    if (e == hipSuccess)
        e = gemm<0, 1, EPI_BIAS_F32>(A.h2, m->hid, W3, m->hid, B, m->out, m->hid, A.logit,
                                     m->out, nullptr, m->master + L.b3, s);
    return e;
}

// forward on batch B (B <= max_batch) from fp32 inputs; leaves z1,h1,z2,h2,logit
hipError_t forward(erq_mlp* m, const float* x, int B, hipStream_t s) {
    return forward_into(m, Acts{m->xb, m->z1, m->h1, m->z2, m->h2, m->logit}, x, B, s);
}

// backward after forward with labels y (fp32 0/1): grads into m->grad, loss into m->loss
hipError_t backward(erq_mlp* m, const float* y, int B, float pos_weight, hipStream_t s) {
    Layout L = layout(m);
    const size_t n1 = (size_t)m->hid * m->in, n2 = (size_t)m->hid * m->hid;
    const __bf16 *W2 = m->wb + n1, *W3 = m->wb + n1 + n2;
    const size_t no = (size_t)B * m->out, nh = (size_t)B * m->hid;
    hipError_t e = hipSuccess;
    const int gb = grid_for(no);
    hipLaunchKernelGGL(bce_kernel, dim3(gb), dim3(256), 0, s, m->logit, y, m->dlog, no,
                       pos_weight, 1.0f / B, m->part);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, m->part, gb, m->loss);
    // layer 3: dW3[out x hid] = dlog^T h2 ; db3 ; dh2 = dlog W3
    e = gemm<1, 0, EPI_F32>(m->dlog, m->out, m->h2, m->hid, m->out, m->hid, B,
                            m->grad + L.w3, m->hid, nullptr, nullptr, s);
    if (e) return e;
    e = colsum(m->dlog, B, m->out, m->part, m->grad + L.b3, s);
    if (e) return e;
    e = gemm<0, 0, EPI_F32>(m->dlog, m->out, W3, m->hid, B, m->hid, m->out, m->tmp, m->hid,
                            nullptr, nullptr, s);
    if (e) return e;
    hipLaunchKernelGGL(gelu_back_kernel, dim3(grid_for(nh)), dim3(256), 0, s, m->tmp, m->z2,
                       m->dz2, nh);
    // layer 2
    e = gemm<1, 0, EPI_F32>(m->dz2, m->hid, m->h1, m->hid, m->hid, m->hid, B, m->grad + L.w2,
                            m->hid, nullptr, nullptr, s);
    if (e) return e;
    e = colsum(m->dz2, B, m->hid, m->part, m->grad + L.b2, s);
    if (e) return e;
    e = gemm<0, 0, EPI_F32>(m->dz2, m->hid, W2, m->hid, B, m->hid, m->hid, m->tmp, m->hid,
                            nullptr, nullptr, s);
    if (e) return e;
    hipLaunchKernelGGL(gelu_back_kernel, dim3(grid_for(nh)), dim3(256), 0, s, m->tmp, m->z1,
                       m->dz1, nh);
    // layer 1
    e = gemm<1, 0, EPI_F32>(m->dz1, m->hid, m->xb, m->in, m->hid, m->in, B, m->grad + L.w1,
                            m->in, nullptr, nullptr, s);
    if (e) return e;
    e = colsum(m->dz1, B, m->hid, m->part, m->grad + L.b1, s);
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
                    m->part, m->ixb, m->iz, m->ih1, m->ih2};
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
    std::lock_guard<std::mutex> lk(m->mu);
    hipStream_t s = mlp_stream(m, stream);
    if (batch <= (size_t)m->max_batch) {
        if (batch == 0) return ERQ_OK;
        hipError_t e = forward(m, x_dev, (int)batch, s);
        if (e != hipSuccess) return hip_fail(e, "mlp forward");
        const size_t n = batch * m->out;
        hipLaunchKernelGGL(sigmoid_kernel, dim3(grid_for(n)), dim3(256), 0, s, m->logit, prob_dev,
                           n);
        e = hipGetLastError();
        return e == hipSuccess ? ERQ_OK : hip_fail(e, "mlp sigmoid");
    }
    // large batches: chunks of up to kInferRows rows through the inference buffers, the
    // logits written straight into prob_dev and turned into probabilities there
    const size_t rows = std::min<size_t>(batch, kInferRows);
    if (m->inf_rows < rows) {
        void* old[] = {m->ixb, m->iz, m->ih1, m->ih2};
        for (void* p : old)
            if (p) (void)hipFree(p);
        m->ixb = nullptr;
        m->iz = nullptr;
        m->ih1 = m->ih2 = nullptr;
        m->inf_rows = 0;
        hipError_t e = alloc(&m->ixb, rows * m->in);
        if (e == hipSuccess) e = alloc(&m->iz, rows * m->hid);
        if (e == hipSuccess) e = alloc(&m->ih1, rows * m->hid);
        if (e == hipSuccess) e = alloc(&m->ih2, rows * m->hid);
        if (e != hipSuccess) return hip_fail(e, "mlp inference buffers");
        m->inf_rows = rows;
    }
    for (size_t b0 = 0; b0 < batch; b0 += rows) {
        const int B = (int)std::min<size_t>(rows, batch - b0);
        float* pr = prob_dev + b0 * m->out;
        hipError_t e = forward_into(m, Acts{m->ixb, m->iz, m->ih1, m->iz, m->ih2, pr},
                                    x_dev + b0 * m->in, B, s);
        if (e != hipSuccess) return hip_fail(e, "mlp forward");
        const size_t n = (size_t)B * m->out;
        hipLaunchKernelGGL(sigmoid_inplace_kernel, dim3(grid_for(n)), dim3(256), 0, s, pr, n);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "mlp sigmoid");
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
