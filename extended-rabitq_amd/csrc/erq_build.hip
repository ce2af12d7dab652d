// erq_build.hip -- the HA index builder on the GPU (IVF::construct, IVF.hpp:116-153, with
// DataQuantizer, Quantizer.hpp:138-631): the same bytes as the host builder (builder.cpp)
// from the same rotated residuals.
//
// Per vector, one wave (quantize_kernel):
//   * |u| = |r * f32(1 / sqrt(sum r^2))| with the sum in double in dim order;
//   * the EX-bit code by the critical-point sweep of fast_quantize (Quantizer.hpp:138-219):
//     every breakpoint t = b / |u_i| of every dim is visited in ascending (t, i) order -- the
//     order the reference's min-heap pops them -- updating the numerator (a double sum, kept
//     sequential: lane 0 adds one term per event) and the integer denominator, and the first
//     t with the largest numerator / sqrt(denominator) wins.  The events are produced window
//     by window: each window is a lexicographic interval of (t, i) holding at most kCap events
//     (shrunk in t, or split by dim index when all of them share one t), generated from each
//     dim's next breakpoint, sorted in LDS (bitonic), then swept.  Only the numerator chain is
//     sequential; denominators (exact integers), ratios and the first-maximum search run on
//     all lanes;
//   * codes flipped where r <= 0, stored in the compacted EX-bit layout (Quantizer.hpp:
//     491-631); xipnorm = f32(1 / sum (c + 1/2) |u|) * 2 * ||o - c||.
// Per 32-vector block, pack_kernel writes the 1-bit FastScan codes (sign of r, nibble per
// four dims, pack_codes.hpp:33-98) and the ||o - c|| factors; unused slots are zero.
// Doubles follow the host code operation for operation (IEEE add, multiply, divide and
// square root, no contraction); float-to-int conversions reproduce x86 cvttsd2si, whose
// out-of-range result is INT_MIN.
#include <algorithm>
#include <vector>

#include "erq.h"
#include "erq_device.h"
#include "erq_internal.h"

namespace erq {

namespace {

constexpr int kCap = 512;  // events per window

__device__ __forceinline__ int cvtt(double x) {  // (int)x as x86-64 computes it
    return (x < 2147483648.0 && x > -2147483649.0) ? (int)x : (int)0x80000000;
}

struct QuantArgs {
    const float* r;      // n x D residuals
    const float* onorm;  // n: ||o - c|| (HA), or ||o - c||^2 when fac is set (non-HA)
    uint64_t n;
    int D, EX;
    uint8_t* long_code;  // n x D*EX/8
    float* xipnorm;      // n
    uint32_t* flags;     // bit 0: a non-finite input
    float* fac;          // non-HA (nullable): n x 3 -- fac_ip, sumxb, fac_err
};

// LDS of one wave: o[D] float, cur[D] and bl[D] bytes, then the window's events.
__host__ __device__ inline size_t quant_lds(int D) {
    return 4 * (size_t)D + 2 * (size_t)D + (size_t)kCap * (8 + 4 + 8);
}

// Byte `pos` of the compacted EX-bit code of raw codes c[0..D) (store_compacted_code).
__device__ __forceinline__ uint8_t compact_byte(const uint8_t* c, int EX, int pos) {
    switch (EX) {
    case 8:
        return c[pos];
    case 4: {
        const int d = (pos >> 4) * 32, j = pos & 15;
        return (uint8_t)(c[d + j] | (c[d + 16 + j] << 4));
    }
    case 2:
    case 3: {
        const int per = EX == 2 ? 16 : 24;
        const int d = (pos / per) * 64, r = pos % per;
        if (r < 16)
            return (uint8_t)((c[d + r] & 3) | ((c[d + 16 + r] & 3) << 2) |
                             ((c[d + 32 + r] & 3) << 4) | ((c[d + 48 + r] & 3) << 6));
        const int tb = r - 16;
        uint32_t v = 0;
        for (int k = 0; k < 8; k++) v |= (uint32_t)((c[d + 8 * k + tb] >> 2) & 1) << k;
        return (uint8_t)v;
    }
    default: {  // 6, 7
        const int per = EX == 6 ? 48 : 56;
        const int d = (pos / per) * 64, r = pos % per;
        if (r < 16) return (uint8_t)((c[d + r] & 63) | (((c[d + 32 + r] >> 4) & 3) << 6));
        if (r < 32) {
            const int j = r - 16;
            return (uint8_t)((c[d + 16 + j] & 63) | (((c[d + 48 + j] >> 4) & 3) << 6));
        }
        if (r < 48) {
            const int j = r - 32;
            return (uint8_t)((c[d + 32 + j] & 15) | ((c[d + 48 + j] & 15) << 4));
        }
        const int tb = r - 48;
        uint32_t v = 0;
        for (int k = 0; k < 8; k++) v |= (uint32_t)((c[d + 8 * k + tb] >> 6) & 1) << k;
        return (uint8_t)v;
    }
    }
}

__device__ __forceinline__ double wave_min_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ double wave_max_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// Ascending bitonic sort of P (a power of two, <= kCap) events by (t, packed id) in LDS.
__device__ void sort_events(double* key, uint32_t* id, int P, int lane) {
    for (int size = 2; size <= P; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < (P >> 1); t += kWave) {
                const int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
                const bool up = (lo & size) == 0;
                const double kl = key[lo], kh = key[hi];
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
            wave_sync();
        }
}

__global__ void __launch_bounds__(64) quantize_kernel(QuantArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int D = a.D, EX = a.EX, lane = threadIdx.x;
    float* o = reinterpret_cast<float*>(smem);
    uint8_t* cur = smem + 4 * (size_t)D;
    uint8_t* bl = cur + D;
    double* ekey = reinterpret_cast<double*>(smem + 6 * (size_t)D);  // 6 D is a multiple of 8
    double* enm = ekey + kCap;
    uint32_t* eid = reinterpret_cast<uint32_t*>(enm + kCap);
    const int top = (1 << EX) - 1;
    const double eps = 1e-5;

    for (uint64_t v = blockIdx.x; v < a.n; v += gridDim.x) {
        const float* x = a.r + v * (size_t)D;
        // |u|: the squared norm in double, dim by dim (lane 0), from a copy in LDS
        bool finite = true;
        for (int d = lane; d < D; d += kWave) {
            const float xv = x[d];
            finite = finite && isfinite(xv);
            o[d] = xv;
        }
        const float on = a.onorm[v];
        if (__ballot(!finite) != 0 || !isfinite(on)) {
            if (lane == 0) atomicOr(a.flags, 1u);
            continue;  // wave-uniform
        }
        wave_sync();
        double ss = 0.0;
        if (lane == 0)
            for (int d = 0; d < D; d++) ss = ss + (double)o[d] * (double)o[d];
        ss = __shfl(ss, 0);
        const float inv = ss > 0 ? (float)(1.0 / sqrt(ss)) : 0.f;
        float mx = -1.f;
        int npos = 0;
        for (int d = lane; d < D; d += kWave) {
            const float u = fabsf(o[d] * inv);
            npos += o[d] > 0 ? 1 : 0;
            o[d] = u;
            mx = fmaxf(mx, u);
        }
        mx = wave_max(mx);
        if (a.fac) {
            // the non-HA factors (DataQuantizer::rabitq_factor, Quantizer.hpp:386-437) as the
            // host builder evaluates them: <u, o_bar> = sum |u| f32(1/sqrt(D)) in float and
            // sum |u / 2| in double, both in dim order (lane 0); sumxb = the positive dims
            npos = wave_sum_i(npos);
            wave_sync();
            if (lane == 0) {
                const float fn = (float)(1.0 / sqrt((double)D));
                float oob = 0.f;
                double ip = 0.0;
                for (int d = 0; d < D; d++) {
                    oob = oob + o[d] * fn;
                    ip = ip + 0.5 * (double)o[d];
                }
                if (!isfinite(oob)) oob = 0.8f;
                const float dist2c = (float)sqrt((double)on);
                const double fac_err = 2.0 / sqrt((double)(D - 1));
                const float oo = oob * oob;
                a.fac[3 * v + 0] = (float)(1.0 / ip * 2 * (double)dist2c);
                a.fac[3 * v + 1] = (float)npos;
                a.fac[3 * v + 2] =
                    (float)(sqrt((1.0 - (double)oo) / (double)oo) * fac_err * 2 * (double)dist2c);
            }
        }
        const double max_o = (double)mx;
        double best_t = 0.0;
        float ipn = 1.f;
        wave_sync();
        if (max_o > 0) {
            const double t_start = (double)(top / 3) / max_o;
            const double t_end = (double)(top + 10) / max_o;
            // initial codes; each dim's events are breakpoints b = c+1 .. bl (b = c+1 always,
            // later ones while b <= top and b / o < t_end)
            long long sq = 0;
            for (int i = lane; i < D; i += kWave) {
                const double oi = (double)o[i];
                const int c = cvtt(t_start * oi + eps);
                sq += (long long)(c * c + c);
                int b = min(top, max(c + 1, cvtt(t_end * oi)));
                while (b < top && (double)(b + 1) / oi < t_end) b++;
                while (b > c + 1 && !((double)b / oi < t_end)) b--;
                cur[i] = (uint8_t)c;
                bl[i] = (uint8_t)b;
            }
            for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
            double sqr_den = (double)(D / 4) + (double)sq;  // D * 0.25 + sum (c^2 + c), exact
            wave_sync();
            double num = 0.0;
            if (lane == 0)
                for (int i = 0; i < D; i++) num = num + ((double)cur[i] + 0.5) * (double)o[i];
            num = __shfl(num, 0);
            double best_ip = 0.0;
            double w = 0.0;  // window width in t (0: not yet sized)
            while (true) {  // wave-uniform: one window per iteration
                // the smallest pending breakpoint
                double tmin = INFINITY;
                bool any = false;
                for (int i = lane; i < D; i += kWave)
                    if (cur[i] < bl[i]) {
                        any = true;
                        tmin = fmin(tmin, (double)(cur[i] + 1) / (double)o[i]);
                    }
                if (__ballot(any) == 0) break;
                tmin = wave_min_d(tmin);
                // events of the window (t, i) <= (T1, I1) (lexicographic, I1 exclusive of i)
                auto count = [&](double T1, int I1) {
                    int cnt = 0;
                    for (int i = lane; i < D; i += kWave) {
                        const double oi = (double)o[i];
                        for (int c = cur[i]; c < bl[i]; c++) {
                            const double t = (double)(c + 1) / oi;
                            if (t < T1 || (t == T1 && i < I1)) cnt++;
                            else break;
                        }
                    }
                    return wave_sum_i(cnt);
                };
                if (!(w > 0)) w = (double)(kCap / 2) / (max_o * (double)D) * 8.0;
                double T1 = tmin + w;
                int I1 = D;
                int n = count(T1, I1);
                if (n > kCap) {
                    // the events at exactly tmin (at least one)
                    const int ntie = count(tmin, D);
                    if (ntie > kCap) {
                        // more than a window share t = tmin: split them by dim index, the
                        // largest I1 in [1, D] with count(tmin, I1) <= kCap (count(tmin, i + 1)
                        // = 1 for the smallest pending dim i, so the window is never empty)
                        T1 = tmin;
                        int lo = 0, hi = D + 1;  // count(lo) <= kCap < count(hi)
                        while (hi - lo > 1) {
                            const int mid = (lo + hi) >> 1;
                            if (count(T1, mid) <= kCap) lo = mid;
                            else hi = mid;
                        }
                        I1 = lo;
                        n = count(T1, I1);
                    } else {
                        // halve in t; T1 strictly decreases and stays above tmin, else the
                        // window is the tie group itself (which fits)
                        while (n > kCap) {
                            const double T1n = tmin + (T1 - tmin) * 0.5;
                            if (!(T1n > tmin) || !(T1n < T1)) {
                                T1 = tmin;
                                n = ntie;
                                break;
                            }
                            T1 = T1n;
                            n = count(T1, I1);
                        }
                    }
                }
                if (I1 == D && T1 > tmin) {  // next window's width (finite)
                    const double wi = T1 - tmin;
                    w = n < kCap / 4 ? wi * 2.0 : (n > (3 * kCap) / 4 ? wi * 0.75 : wi);
                    w = fmin(w, 1e300);
                }
                // emit: each lane's events at its prefix offset
                int mine = 0;
                for (int i = lane; i < D; i += kWave) {
                    const double oi = (double)o[i];
                    for (int c = cur[i]; c < bl[i]; c++) {
                        const double t = (double)(c + 1) / oi;
                        if (t < T1 || (t == T1 && i < I1)) mine++;
                        else break;
                    }
                }
                int pos = mine;
                for (int off = 1; off < 64; off <<= 1) {
                    const int t = __shfl_up(pos, off);
                    if (lane >= off) pos += t;
                }
                pos -= mine;
                for (int i = lane; i < D; i += kWave) {
                    const double oi = (double)o[i];
                    int c = cur[i];
                    for (; c < bl[i]; c++) {
                        const double t = (double)(c + 1) / oi;
                        if (!(t < T1 || (t == T1 && i < I1))) break;
                        ekey[pos] = t;
                        eid[pos] = ((uint32_t)i << 9) | (uint32_t)(c + 1);
                        pos++;
                    }
                    cur[i] = (uint8_t)c;
                }
                int P = 1;
                while (P < n) P <<= 1;
                for (int j = n + lane; j < P; j += kWave) {
                    ekey[j] = INFINITY;
                    eid[j] = 0xFFFFFFFFu;
                }
                wave_sync();
                sort_events(ekey, eid, P, lane);
                // the numerator after each event, in event order (lane 0)
                if (lane == 0) {
                    double s = num;
                    for (int j = 0; j < n; j++) {
                        s = s + (double)o[eid[j] >> 9];
                        enm[j] = s;
                    }
                    num = s;
                }
                num = __shfl(num, 0);
                wave_sync();
                // denominators (exact integers), ratios, the first strict improvement
                for (int j0 = 0; j0 < n; j0 += kWave) {
                    const int j = j0 + lane;
                    const bool ok = j < n;
                    const int b2 = ok ? 2 * (int)(eid[j] & 511u) : 0;
                    int pre = b2;
                    for (int off = 1; off < 64; off <<= 1) {
                        const int t = __shfl_up(pre, off);
                        if (lane >= off) pre += t;
                    }
                    const double den = sqr_den + (double)pre;
                    const double ip = ok ? enm[j] / sqrt(den) : -1.0;
                    sqr_den = sqr_den + (double)__shfl(pre, 63);
                    const double m = wave_max_d(ip);
                    if (m > best_ip) {
                        const uint64_t at = __ballot(ok && ip == m);
                        const int first = __builtin_ctzll(at);
                        best_ip = m;
                        best_t = ekey[j0 + first];
                    }
                }
                wave_sync();
            }
            // final codes
            // (bl marks an overflowed conversion: the host's int code is then INT_MIN, which
            // enters the numerator as such and is stored as byte 0)
            for (int i = lane; i < D; i += kWave) {
                int c = cvtt(best_t * (double)o[i] + eps);
                if (c > top) c = top;
                cur[i] = (uint8_t)c;
                bl[i] = c < 0;
            }
            wave_sync();
            double numer = 0.0;
            if (lane == 0)
                for (int i = 0; i < D; i++) {
                    const double c = bl[i] ? -2147483648.0 : (double)cur[i];
                    numer = numer + (c + 0.5) * (double)o[i];
                }
            numer = __shfl(numer, 0);
            ipn = (float)(1.0 / numer);
            if (!isfinite(ipn)) ipn = 1.f;
        } else {
            for (int i = lane; i < D; i += kWave) cur[i] = 0;
        }
        wave_sync();
        // sign flip (code of r <= 0 becomes ~code & mask) and the compacted layout
        const int mask = top;
        for (int d = lane; d < D; d += kWave)
            if (!(x[d] > 0)) cur[d] = (uint8_t)((~cur[d]) & mask);
        wave_sync();
        uint8_t* out = a.long_code + v * ((size_t)D * EX / 8);
        for (int p = lane; p < D * EX / 8; p += kWave) out[p] = compact_byte(cur, EX, p);
        if (lane == 0)
            a.xipnorm[v] = a.fac ? (float)((double)(ipn * 2) * sqrt((double)on)) : ipn * 2 * on;
        wave_sync();
    }
}

// One thread per byte of the packed blocks: 4D code bytes then the factors, 32 floats each
// (as 128 bytes): ||o - c|| (HA), or ||o - c||^2, fac_ip, sumxb and fac_err when fac is set
// (non-HA, Quantizer.hpp:107-121).
__global__ void pack_kernel(const float* __restrict__ r, const float* __restrict__ onorm, int D,
                            uint32_t v_base, const uint32_t* __restrict__ blk_first,
                            const uint8_t* __restrict__ blk_count, size_t nblocks,
                            const float* __restrict__ fac, uint8_t* __restrict__ out) {
    const size_t BB = 4 * (size_t)D + (fac ? 512 : 128);
    const size_t total = nblocks * BB;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total;
         q += (size_t)gridDim.x * blockDim.x) {
        const size_t blk = q / BB;
        const int byte = (int)(q % BB);
        const uint32_t first = blk_first[blk] - v_base;  // r and onorm start at v_base
        const int cnt = blk_count[blk];
        if (byte < 4 * D) {
            const int m = (byte >> 5) * 2 + ((byte >> 4) & 1);  // codebook: dims 4m..4m+3
            const int s0 = vec_of_byte(byte & 15), s1 = s0 + 16;
            uint32_t val = 0;
            for (int h = 0; h < 2; h++) {
                const int s = h ? s1 : s0;
                if (s < cnt) {
                    const float* x = r + ((size_t)first + s) * D + 4 * m;
                    const uint32_t nib = ((x[0] > 0) << 3) | ((x[1] > 0) << 2) |
                                         ((x[2] > 0) << 1) | (uint32_t)(x[3] > 0);
                    val |= nib << (4 * h);
                }
            }
            out[q] = (uint8_t)val;
        } else {
            const int fb = byte - 4 * D, row = fb >> 7;
            const int s = (fb & 127) >> 2, bi = fb & 3;
            float f = 0.f;
            if (s < cnt) f = row == 0 ? onorm[first + s] : fac[3 * ((size_t)first + s) + row - 1];
            out[q] = (uint8_t)(__float_as_uint(f) >> (8 * bi));
        }
    }
}

}  // namespace

hipError_t launch_quantize(const float* r, const float* onorm, uint64_t n, int D, int EX,
                           uint8_t* long_code, float* xipnorm, uint32_t* flags, hipStream_t s,
                           float* fac = nullptr) {
    if (n == 0) return hipSuccess;
    const size_t lds = quant_lds(D);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&quantize_kernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return e;
    }
    QuantArgs a{r, onorm, n, D, EX, long_code, xipnorm, flags, fac};
    const unsigned grid = (unsigned)std::min<uint64_t>(n, 1u << 20);
    hipLaunchKernelGGL(quantize_kernel, dim3(grid), dim3(64), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_pack(const float* r, const float* onorm, int D, uint32_t v_base,
                       const uint32_t* blk_first, const uint8_t* blk_count, size_t nblocks,
                       uint8_t* out, hipStream_t s, const float* fac = nullptr) {
    if (nblocks == 0) return hipSuccess;
    const size_t total = nblocks * (4 * (size_t)D + (fac ? 512 : 128));
    const int grid = (int)std::min<size_t>((total + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(pack_kernel, dim3(grid), dim3(256), 0, s, r, onorm, D, v_base, blk_first,
                       blk_count, nblocks, fac, out);
    return hipGetLastError();
}

namespace {

// ---------------------------------------------------------------- IVF::construct inputs
__global__ void cluster_count_kernel(const uint32_t* __restrict__ assign, uint64_t n, uint32_t K,
                                     uint32_t* __restrict__ counts, uint32_t* __restrict__ flags) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = assign[i];
        if (k < K) atomicAdd(&counts[k], 1u);
        else atomicOr(flags, 2u);
    }
}

// ids in cluster order, ascending within a cluster (IVF.hpp:126-140 collects each cluster's
// ids in increasing order): one workgroup per cluster walks assign[] 1024 entries at a time
// and appends its members in order.
__global__ void __launch_bounds__(256)
    cluster_order_kernel(const uint32_t* __restrict__ assign, uint64_t n,
                         const uint64_t* __restrict__ vstart, uint32_t* __restrict__ ids) {
    __shared__ uint32_t wtot[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t cl = blockIdx.x;
    uint64_t at = vstart[cl];
    for (uint64_t base = 0; base < n; base += 1024) {  // block-uniform
        const uint64_t e = base + 4 * (uint64_t)tid;
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
        uint64_t pos = at + incl - own;
        for (int w = 0; w < wid; w++) pos += wtot[w];
        const uint32_t nm = wtot[0] + wtot[1] + wtot[2] + wtot[3];
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (m[i]) ids[pos++] = (uint32_t)(e + i);
        at += nm;
        __syncthreads();
    }
}

// rows of ids[0..rows) from data (DIM floats each), zero-padded to D
__global__ void gather_pad_kernel(const float* __restrict__ data, int DIM,
                                  const uint32_t* __restrict__ ids, uint32_t rows, int D,
                                  float* __restrict__ out) {
    const size_t total = (size_t)rows * D;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total;
         q += (size_t)gridDim.x * blockDim.x) {
        const size_t r = q / D;
        const int d = (int)(q % D);
        out[q] = d < DIM ? data[(size_t)ids[r] * DIM + d] : 0.f;
    }
}

// XP.row(i) - CP (Quantizer.hpp:333-335), in place on the rotated rows
__global__ void residual_kernel(float* __restrict__ xr, const uint32_t* __restrict__ ids,
                                const uint32_t* __restrict__ assign, uint32_t rows, int D,
                                const float* __restrict__ crot) {
    const size_t total = (size_t)rows * D;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total;
         q += (size_t)gridDim.x * blockDim.x) {
        const size_t r = q / D;
        const int d = (int)(q % D);
        xr[q] = xr[q] - crot[(size_t)assign[ids[r]] * D + d];
    }
}

// ||o - c|| = f32(sqrt((double)L2Sqr(o, c, DIM))) (Quantizer.hpp:399-401; space.hpp:41-51:
// L2Sqr16 over the largest multiple of 16 dims -- 16 lane-strided fma chains, butterfly --
// then the remaining dims one by one), one 16-lane group per vector; squared: L2Sqr itself
// (the non-HA factor ||o - c||^2).
__global__ void onorm_kernel(const float* __restrict__ data, int DIM,
                             const float* __restrict__ cent, const uint32_t* __restrict__ ids,
                             const uint32_t* __restrict__ assign, uint32_t rows,
                             float* __restrict__ onorm, int squared) {
    const int l = threadIdx.x & 15;
    const uint32_t g = (blockIdx.x * (uint32_t)blockDim.x + threadIdx.x) >> 4;
    const bool act = g < rows;
    const uint32_t id = act ? ids[g] : 0u;
    const float* x = data + (size_t)id * DIM;
    const float* c = cent + (size_t)(act ? assign[id] : 0u) * DIM;
    const int n16 = DIM & ~15;
    float p = 0.f;
    if (act)
        for (int d = l; d < n16; d += 16) {
            const float t = x[d] - c[d];
            p = fmaf(t, t, p);
        }
    p = butterfly16(p);
    if (act && l == 0) {
        for (int d = n16; d < DIM; d++) {
            const float t = x[d] - c[d];
            p = fmaf(t, t, p);
        }
        onorm[g] = squared ? p : (float)sqrt((double)p);
    }
}

int grid_for(size_t total, int per) {
    return (int)std::max<size_t>(1, std::min<size_t>((total + per - 1) / per, 1 << 20));
}

}  // namespace

}  // namespace erq

using namespace erq;

namespace {

bool ex_supported(uint32_t ex) { return ex == 2 || ex == 3 || ex == 4 || ex == 6 || ex == 7 || ex == 8; }

}  // namespace

extern "C" int erq_build_quantize_dev(const float* resid, const float* onorm, uint64_t n,
                                      uint32_t D, uint32_t ex_bits, uint8_t* long_code,
                                      float* xipnorm, void* stream) {
    return guarded([&]() -> int {
        if (n == 0) return ERQ_OK;
        if (!resid || !onorm || !long_code || !xipnorm) {
            set_error("null pointer");
            return ERQ_EINVAL;
        }
        if (D == 0 || D % 64 != 0 || D > 8192 || !ex_supported(ex_bits)) {
            set_error("D must be a multiple of 64 up to 8192 and EX one of 2, 3, 4, 6, 7, 8");
            return ERQ_EUNSUPPORTED;
        }
        hipStream_t s = static_cast<hipStream_t>(stream);
        uint32_t* flags = nullptr;
        hipError_t e = hipMalloc(&flags, 4);
        if (e != hipSuccess) return hip_fail(e, "flags");
        uint32_t h = 0;
        e = hipMemsetAsync(flags, 0, 4, s);
        if (e == hipSuccess)
            e = launch_quantize(resid, onorm, n, (int)D, (int)ex_bits, long_code, xipnorm, flags, s);
        if (e == hipSuccess) e = hipMemcpyAsync(&h, flags, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(flags);
        if (e != hipSuccess) return hip_fail(e, "quantize");
        if (h) {
            set_error("non-finite residual or norm");
            return ERQ_EINVAL;
        }
        return ERQ_OK;
    });
}

extern "C" int erq_build_pack_dev(const float* resid, const float* onorm,
                                  const uint64_t* cluster_sizes, uint32_t K, uint32_t D,
                                  uint8_t* short_data, void* stream) {
    return guarded([&]() -> int {
        if (!resid || !onorm || !cluster_sizes || !short_data) {
            set_error("null pointer");
            return ERQ_EINVAL;
        }
        if (D == 0 || D % 64 != 0 || D > 8192) {
            set_error("D must be a multiple of 64 up to 8192");
            return ERQ_EUNSUPPORTED;
        }
        std::vector<uint32_t> first;
        std::vector<uint8_t> count;
        uint64_t v0 = 0;
        for (uint32_t k = 0; k < K; k++) {
            const uint64_t nk = cluster_sizes[k];
            for (uint64_t b = 0; b < nk; b += 32) {
                first.push_back((uint32_t)(v0 + b));
                count.push_back((uint8_t)std::min<uint64_t>(32, nk - b));
            }
            v0 += nk;
        }
        if (v0 >= 0xFFFFFFFFull) {
            set_error("too many vectors");
            return ERQ_EUNSUPPORTED;
        }
        const size_t nb = first.size();
        if (nb == 0) return ERQ_OK;
        hipStream_t s = static_cast<hipStream_t>(stream);
        uint32_t* d_first = nullptr;
        uint8_t* d_count = nullptr;
        hipError_t e = hipMalloc(&d_first, nb * 4);
        if (e == hipSuccess) e = hipMalloc(&d_count, nb);
        if (e == hipSuccess) e = hipMemcpyAsync(d_first, first.data(), nb * 4, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(d_count, count.data(), nb, hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = launch_pack(resid, onorm, (int)D, 0u, d_first, d_count, nb, short_data, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(d_first);
        (void)hipFree(d_count);
        if (e != hipSuccess) return hip_fail(e, "pack blocks");
        return ERQ_OK;
    });
}

namespace {

// IVF::construct on the GPU for either block format (nonha: four factors per vector).
int build_index(const float* data, uint64_t N, uint32_t DIM, const float* centroids, uint32_t K,
                const uint32_t* assign, const float* P, uint32_t ex_bits, uint64_t* cluster_sizes,
                float* rotated_centroids, uint8_t* short_data, uint8_t* long_code,
                float* ex_factor, uint32_t* ids, void* stream, bool nonha) {
    return guarded([&]() -> int {
        if (!data || !centroids || !assign || !P || !cluster_sizes || !rotated_centroids ||
            !short_data || !long_code || !ex_factor || !ids) {
            set_error("null pointer");
            return ERQ_EINVAL;
        }
        if (N == 0 || N >= 0xFFFFFFFFull || K == 0 || DIM == 0 || DIM > 8192 ||
            !ex_supported(ex_bits)) {
            set_error("bad shape: N in [1, 2^32), K >= 1, DIM in [1, 8192], EX in {2,3,4,6,7,8}");
            return ERQ_EINVAL;
        }
        const int D = (int)((DIM + 63) / 64 * 64);
        const int EX = (int)ex_bits;
        hipStream_t s = static_cast<hipStream_t>(stream);
        std::vector<void*> temps;
        struct Free {
            std::vector<void*>& t;
            hipStream_t s;
            ~Free() {
                (void)hipStreamSynchronize(s);
                for (void* p : t) (void)hipFree(p);
            }
        } fr{temps, s};
        auto alloc = [&](size_t bytes, auto** out) -> int {
            void* p = nullptr;
            const hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 16));
            if (e != hipSuccess) return hip_fail(e, "device allocation");
            temps.push_back(p);
            *out = static_cast<std::remove_reference_t<decltype(*out)>>(p);
            return ERQ_OK;
        };
#define ERQ_B(x)                          \
    do {                                  \
        const int rc_ = (x);              \
        if (rc_ != ERQ_OK) return rc_;    \
    } while (0)
#define ERQ_BH(x, what)                                 \
    do {                                                \
        const hipError_t e_ = (x);                      \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)
        // 1. cluster sizes and ids in cluster order
        uint32_t *counts = nullptr, *flags = nullptr;
        ERQ_B(alloc(4 * (size_t)K, &counts));
        ERQ_B(alloc(4, &flags));
        ERQ_BH(hipMemsetAsync(counts, 0, 4 * (size_t)K, s), "clear counts");
        ERQ_BH(hipMemsetAsync(flags, 0, 4, s), "clear flags");
        hipLaunchKernelGGL(cluster_count_kernel, dim3(grid_for(N, 256)), dim3(256), 0, s, assign,
                           N, K, counts, flags);
        ERQ_BH(hipGetLastError(), "cluster count");
        std::vector<uint32_t> hc(K);
        uint32_t hf = 0;
        ERQ_BH(hipMemcpyAsync(hc.data(), counts, 4 * (size_t)K, hipMemcpyDeviceToHost, s), "counts");
        ERQ_BH(hipMemcpyAsync(&hf, flags, 4, hipMemcpyDeviceToHost, s), "flags");
        ERQ_BH(hipStreamSynchronize(s), "counts");
        if (hf & 2u) {
            set_error("cluster id out of range");
            return ERQ_EINVAL;
        }
        std::vector<uint64_t> vstart(K + 1, 0), bstart(K + 1, 0);
        for (uint32_t k = 0; k < K; k++) {
            cluster_sizes[k] = hc[k];
            vstart[k + 1] = vstart[k] + hc[k];
            bstart[k + 1] = bstart[k] + (hc[k] + 31) / 32;
        }
        uint64_t* d_vstart = nullptr;
        ERQ_B(alloc(8 * (size_t)(K + 1), &d_vstart));
        ERQ_BH(hipMemcpyAsync(d_vstart, vstart.data(), 8 * (size_t)(K + 1), hipMemcpyHostToDevice, s),
               "vstart");
        hipLaunchKernelGGL(cluster_order_kernel, dim3(K), dim3(256), 0, s, assign, N, d_vstart, ids);
        ERQ_BH(hipGetLastError(), "cluster order");
        // 2. rotated centroids: C_pad * P (Rotator.hpp:60-63, Initializer.hpp:70-76)
        float* cpad = nullptr;
        ERQ_B(alloc(4 * (size_t)K * D, &cpad));
        ERQ_BH(launch_pad_rows(centroids, (int)K, (int)DIM, cpad, D, s), "pad centroids");
        ERQ_BH(launch_gemm(cpad, D, P, D, 0, rotated_centroids, D, (int)K, D, D, 0, nullptr,
                           nullptr, s),
               "rotate centroids");
        // 3. chunks of whole clusters: rotate, residual, norms, quantize, pack
        uint64_t maxc = 0;
        for (uint32_t k = 0; k < K; k++) maxc = std::max<uint64_t>(maxc, hc[k]);
        const uint64_t chunk = std::max<uint64_t>(maxc, std::max<uint64_t>(
                                   1024, ((uint64_t)1 << 30) / (8 * (uint64_t)D)));
        float *xp = nullptr, *xr = nullptr, *on = nullptr, *fac = nullptr;
        ERQ_B(alloc(4 * chunk * D, &xp));
        ERQ_B(alloc(4 * chunk * D, &xr));
        ERQ_B(alloc(4 * chunk, &on));
        if (nonha) ERQ_B(alloc(12 * chunk, &fac));
        std::vector<uint32_t> bfirst;
        std::vector<uint8_t> bcount;
        for (uint32_t k = 0; k < K; k++)
            for (uint64_t b = 0; b < hc[k]; b += 32) {
                bfirst.push_back((uint32_t)(vstart[k] + b));
                bcount.push_back((uint8_t)std::min<uint64_t>(32, hc[k] - b));
            }
        const size_t nb = bfirst.size();
        uint32_t* d_bfirst = nullptr;
        uint8_t* d_bcount = nullptr;
        ERQ_B(alloc(4 * nb, &d_bfirst));
        ERQ_B(alloc(nb, &d_bcount));
        if (nb) {
            ERQ_BH(hipMemcpyAsync(d_bfirst, bfirst.data(), 4 * nb, hipMemcpyHostToDevice, s), "blocks");
            ERQ_BH(hipMemcpyAsync(d_bcount, bcount.data(), nb, hipMemcpyHostToDevice, s), "blocks");
        }
        const size_t BB = 4 * (size_t)D + (nonha ? 512 : 128), CL = (size_t)D * EX / 8;
        for (uint32_t k0 = 0; k0 < K;) {
            uint32_t k1 = k0 + 1;
            while (k1 < K && vstart[k1 + 1] - vstart[k0] <= chunk) k1++;
            const uint64_t v0 = vstart[k0], rows = vstart[k1] - v0;
            if (rows) {
                const size_t tot = (size_t)rows * D;
                hipLaunchKernelGGL(gather_pad_kernel, dim3(grid_for(tot, 256)), dim3(256), 0, s,
                                   data, (int)DIM, ids + v0, (uint32_t)rows, D, xp);
                ERQ_BH(hipGetLastError(), "gather");
                ERQ_BH(launch_gemm(xp, D, P, D, 0, xr, D, (int)rows, D, D, 0, nullptr, nullptr, s),
                       "rotate vectors");
                hipLaunchKernelGGL(residual_kernel, dim3(grid_for(tot, 256)), dim3(256), 0, s, xr,
                                   ids + v0, assign, (uint32_t)rows, D, rotated_centroids);
                ERQ_BH(hipGetLastError(), "residual");
                hipLaunchKernelGGL(onorm_kernel, dim3(grid_for(rows * 16, 256)), dim3(256), 0, s,
                                   data, (int)DIM, centroids, ids + v0, assign, (uint32_t)rows, on,
                                   nonha ? 1 : 0);
                ERQ_BH(hipGetLastError(), "norms");
                ERQ_BH(launch_quantize(xr, on, rows, D, EX, long_code + v0 * CL, ex_factor + v0,
                                       flags, s, fac),
                       "quantize");
                const size_t b0 = bstart[k0], b1 = bstart[k1];
                ERQ_BH(launch_pack(xr, on, D, (uint32_t)v0, d_bfirst + b0, d_bcount + b0, b1 - b0,
                                   short_data + b0 * BB, s, fac),
                       "pack");
            }
            k0 = k1;
        }
        ERQ_BH(hipMemcpyAsync(&hf, flags, 4, hipMemcpyDeviceToHost, s), "flags");
        ERQ_BH(hipStreamSynchronize(s), "build");
        if (hf & 1u) {
            set_error("non-finite vector, centroid or norm");
            return ERQ_EINVAL;
        }
        return ERQ_OK;
#undef ERQ_B
#undef ERQ_BH
    });
}

}  // namespace

extern "C" int erq_build_index_dev(const float* data, uint64_t N, uint32_t DIM,
                                   const float* centroids, uint32_t K, const uint32_t* assign,
                                   const float* P, uint32_t ex_bits, uint64_t* cluster_sizes,
                                   float* rotated_centroids, uint8_t* short_data,
                                   uint8_t* long_code, float* ex_factor, uint32_t* ids,
                                   void* stream) {
    return build_index(data, N, DIM, centroids, K, assign, P, ex_bits, cluster_sizes,
                       rotated_centroids, short_data, long_code, ex_factor, ids, stream, false);
}

extern "C" int erq_build_index_dev_nonha(const float* data, uint64_t N, uint32_t DIM,
                                         const float* centroids, uint32_t K,
                                         const uint32_t* assign, const float* P, uint32_t ex_bits,
                                         uint64_t* cluster_sizes, float* rotated_centroids,
                                         uint8_t* short_data, uint8_t* long_code,
                                         float* ex_factor, uint32_t* ids, void* stream) {
    return build_index(data, N, DIM, centroids, K, assign, P, ex_bits, cluster_sizes,
                       rotated_centroids, short_data, long_code, ex_factor, ids, stream, true);
}
