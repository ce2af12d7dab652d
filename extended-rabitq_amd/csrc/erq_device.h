// erq_device.h -- device helpers shared by the liberq kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>

namespace erq {

constexpr int kWave = 64;

// Make this wave's LDS writes visible to its other lanes (no workgroup barrier, so
// waves of one workgroup may run independent queries and exit independently).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Sum over each 16-lane group in the order of _mm512_reduce_add_ps: lanes i and i^8,
// then ^4, ^2, ^1 (IEEE addition is commutative, so every lane of the group ends with
// the same, reference-ordered, value).
__device__ __forceinline__ float butterfly16(float v) {
    v = v + __shfl_xor(v, 8);
    v = v + __shfl_xor(v, 4);
    v = v + __shfl_xor(v, 2);
    v = v + __shfl_xor(v, 1);
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)v, o), hi = __shfl_xor((unsigned)(v >> 32), o);
        v += ((unsigned long long)hi << 32) | lo;
    }
    return v;
}

// Adds this wave's counts (summed over its lanes) to a stats counter; lane 0 only, and
// only when nonzero, so that idle workgroups do not contend on the counters.
__device__ __forceinline__ void wave_stat_add(unsigned long long* stat, unsigned long long v) {
    v = wave_sum_u64(v);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(stat, v);
}

// Byte position (0..15) inside a codebook's 16 bytes that holds vector v%16 in the
// FastScan block (reference pack_codes.hpp:44-65, perm0 = {0,8,1,9,...,7,15}).
__device__ __forceinline__ int byte_of_vec(int v) {
    int w = v & 15;
    return w < 8 ? 2 * w : 2 * (w - 8) + 1;
}
// Inverse: vector (0..15) whose low nibble is at byte j.
__device__ __forceinline__ int vec_of_byte(int j) { return (j & 1) ? (j >> 1) + 8 : (j >> 1); }

// Ex-code values of dims 16s + l (s = 0..3) of one 64-dim chunk (8*EX bytes) for lane
// l = 0..15 of a 16-lane group: layouts of DataQuantizer::store_compacted_code
// (reference Quantizer.hpp:491-631), read as IP16_fxu8/IP32_fxu4/IP64_fxu{2,3,6,7} do.
template <int EX>
__device__ __forceinline__ void decode4(const uint8_t* __restrict__ c, int l, float e[4]) {
    if constexpr (EX == 8) {
        for (int s = 0; s < 4; s++) e[s] = (float)c[16 * s + l];
    } else if constexpr (EX == 4) {
        uint32_t b0 = c[l], b1 = c[16 + l];
        e[0] = (float)(b0 & 15u);
        e[1] = (float)(b0 >> 4);
        e[2] = (float)(b1 & 15u);
        e[3] = (float)(b1 >> 4);
    } else if constexpr (EX == 2 || EX == 3) {
        uint32_t b = c[l];
        uint32_t tb = 0;
        if constexpr (EX == 3) tb = c[16 + (l & 7)] >> (l >> 3);
        for (int s = 0; s < 4; s++) {
            uint32_t v = (b >> (2 * s)) & 3u;
            if constexpr (EX == 3) v |= ((tb >> (2 * s)) & 1u) << 2;
            e[s] = (float)v;
        }
    } else {  // 6, 7
        uint32_t b0 = c[l], b1 = c[16 + l], b2 = c[32 + l];
        uint32_t v[4];
        v[0] = b0 & 63u;
        v[1] = b1 & 63u;
        v[2] = (b2 & 15u) | ((b0 >> 6) << 4);
        v[3] = (b2 >> 4) | ((b1 >> 6) << 4);
        if constexpr (EX == 7) {
            uint32_t tb = c[48 + (l & 7)] >> (l >> 3);
            for (int s = 0; s < 4; s++) v[s] |= ((tb >> (2 * s)) & 1u) << 6;
        }
        for (int s = 0; s < 4; s++) e[s] = (float)v[s];
    }
}

// Lane-major copy of the EX-bit codes ("exl"), for the 16-lane re-estimate: lane l of a
// vector's group handles dims d = 16 i + l (i = 0 .. D/16 - 1), in increasing i, which is
// the order of the reference's lane-strided FMA chains.  Per vector 16 x W dwords, lane l's
// W dwords contiguous (16-byte aligned).  Values go four to a dword's bytes: dword w holds
// values w vpd .. w vpd + vpd - 1, value w vpd + 4 h + t in byte t at bit 8 t + sp h, so the
// four values 4 h .. 4 h + 3 of a dword are the bytes of (dword >> sp h) & mask and convert
// with v_cvt_f32_ubyte0..3.  sp = 2 for EX = 2, 4 for EX = 3, 4 and 8 for EX = 6..8.
__host__ __device__ constexpr int exl_sp(int ex) { return ex <= 2 ? 2 : (ex <= 4 ? 4 : 8); }
__host__ __device__ constexpr int exl_npb(int ex) { return 8 / exl_sp(ex); }  // values per byte
__host__ __device__ constexpr int exl_vpd(int ex) { return 4 * exl_npb(ex); }
__host__ __device__ inline int exl_words(int D, int ex) {
    const int vpd = exl_vpd(ex), w = (D / 16 + vpd - 1) / vpd;
    return (w + 3) & ~3;
}

// The residual u in LDS, lane-major like exl: u[16 i + l] at ul[l * ul_stride(D) + i].  The
// stride (a multiple of 4, = 4 m with m odd) spreads the 16 lanes' rows over distinct banks.
__host__ __device__ inline int ul_stride(int D) { return D / 16 + ((D / 64) % 2 == 0 ? 4 : 8); }
__device__ __forceinline__ int ul_index(int d, int S) { return (d & 15) * S + (d >> 4); }

// high_acc_quantize16 (space.hpp:695-711) of one value: the IEEE float quotient
// tmp = u / delta, rounded half away from zero in double as the reference writes it.
// rd = 1.0 / (double)delta.  The double product u * rd is within 2^-52 (relative) of
// u / delta, and a quotient of two floats never lies closer than 2^-49 (relative) to a
// float rounding boundary without being the float itself, so rounding the product to
// float gives exactly the IEEE float division.
// The reference's (int)(tmp + 0.5 - [tmp < 0]), evaluated in double, rounds half away from
// zero.  Here: (double)tmp + copysign(0.5, tmp), truncated toward zero by the conversion.
// The sum is exact whenever |tmp| >= 2^-29 (tmp has 24 significant bits and |tmp| < 2^14, so
// the sum needs at most 53); below that it rounds to a value of magnitude < 1, which
// truncates to 0 as it should.  For |tmp| < 0.5 the exact sum has magnitude < 1 (0 after
// truncation); otherwise trunc(|tmp| + 0.5) = floor(|tmp| + 0.5).  Four instructions after
// tmp instead of roundf's seven.
__device__ __forceinline__ int quantize_digit(float u, double rd) {
    const float tmp = (float)((double)u * rd);
    const double t = (double)tmp;
    return (int)(t + copysign(0.5, t));
}

// Byte 0 of each of a, b, c, d as the bytes 0..3 of one word.
__device__ __forceinline__ uint32_t pack_low_bytes(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    const uint32_t ab = __builtin_amdgcn_perm(b, a, 0x0c0c0400u);
    const uint32_t cd = __builtin_amdgcn_perm(d, c, 0x0c0c0400u);
    return __builtin_amdgcn_perm(cd, ab, 0x05040100u);
}

// The reference's per-(query, cluster) preparation (normalize_query16 and the digit scale
// of high_acc_quantize16, space.hpp:695-778) by one wave, in two parts.
// residual_delta: u = (q - c) * (1 / y) stored lane-major in uq (stride S), or u = 1/sqrt(D)
// when y <= 1e-5 (then returns false), and delta = max |u| / 8191.  D is a multiple of 64;
// q and c are 16-byte aligned.  NL: 256-dim slices of q and c loaded before any is used
// (loads past D are clamped into the row and dropped; max and min do not depend on the
// order the slices are taken in).
template <int NL = 1>
__device__ __forceinline__ bool residual_delta(const float* __restrict__ q,
                                               const float* __restrict__ c, float y, int D, int S,
                                               float* uq, int lane, float& fac, float& delta) {
    float mx = 0.f, mn = 0.f;
    const bool res = y > 1e-5f;
    if (res) {
        fac = (float)(1.0 / (double)y);
        for (int d0 = 4 * lane; d0 < D; d0 += 256 * NL) {
            float4 qa[NL], ca[NL];
#pragma unroll
            for (int j = 0; j < NL; j++) {
                const int d = min(d0 + 256 * j, D - 4);
                qa[j] = *reinterpret_cast<const float4*>(q + d);
                ca[j] = *reinterpret_cast<const float4*>(c + d);
            }
#pragma unroll
            for (int j = 0; j < NL; j++) {
                const int d = d0 + 256 * j;
                if (d < D) {
                    const float u0 = (qa[j].x - ca[j].x) * fac, u1 = (qa[j].y - ca[j].y) * fac;
                    const float u2 = (qa[j].z - ca[j].z) * fac, u3 = (qa[j].w - ca[j].w) * fac;
                    mx = fmaxf(fmaxf(mx, u0), fmaxf(u1, fmaxf(u2, u3)));
                    mn = fminf(fminf(mn, u0), fminf(u1, fminf(u2, u3)));
                    uq[ul_index(d, S)] = u0;
                    uq[ul_index(d + 1, S)] = u1;
                    uq[ul_index(d + 2, S)] = u2;
                    uq[ul_index(d + 3, S)] = u3;
                }
            }
        }
    } else {
        fac = 0.f;
        const float value = (float)(1.0 / (double)sqrtf((float)D));
        for (int d = lane; d < D; d += kWave) uq[ul_index(d, S)] = value;
        mx = value;
    }
    mx = wave_max(mx);
    mn = wave_min(mn);
    delta = fmaxf(fabsf(mn), fabsf(mx)) / 8191.0f;
    return res;
}

// u = (q - c) * fac stored lane-major in uq, as residual_delta stores it, with NL 256-dim
// slices of q and c loaded before any is stored (one memory wait per NL slices).
template <int NL>
__device__ __forceinline__ void stage_residual(const float* __restrict__ q,
                                               const float* __restrict__ c, float fac, int D,
                                               int S, float* uq, int lane) {
    for (int d0 = 4 * lane; d0 < D; d0 += 256 * NL) {
        float4 qa[NL], ca[NL];
#pragma unroll
        for (int j = 0; j < NL; j++) {
            const int d = min(d0 + 256 * j, D - 4);
            qa[j] = *reinterpret_cast<const float4*>(q + d);
            ca[j] = *reinterpret_cast<const float4*>(c + d);
        }
#pragma unroll
        for (int j = 0; j < NL; j++) {
            const int d = d0 + 256 * j;
            if (d < D) {
                uq[ul_index(d, S)] = (qa[j].x - ca[j].x) * fac;
                uq[ul_index(d + 1, S)] = (qa[j].y - ca[j].y) * fac;
                uq[ul_index(d + 2, S)] = (qa[j].z - ca[j].z) * fac;
                uq[ul_index(d + 3, S)] = (qa[j].w - ca[j].w) * fac;
            }
        }
    }
}

// sumq: the sum of u in _mm512 lane order (16 lane-strided chains over the lane-major rows,
// then butterfly16), by the 16 lanes of a group (l = lane & 15); every lane of the group
// gets it.  When y <= 1e-5 the reference takes D / sqrt(D) instead (prepare_residual).
__device__ __forceinline__ float residual_sum16(const float* uq, int S, int D, int l) {
    const float4* row = reinterpret_cast<const float4*>(uq + l * S);
    float part = 0.f;
    for (int i = 0; i < (D >> 6); i++) {
        const float4 v = row[i];
        part = part + v.x;
        part = part + v.y;
        part = part + v.z;
        part = part + v.w;
    }
    return butterfly16(part);
}

// Both parts by one wave (lanes 0..15 run the sum).
template <int NL = 1>
__device__ __forceinline__ void prepare_residual(const float* __restrict__ q,
                                                 const float* __restrict__ c, float y, int D,
                                                 int S, float* uq, int lane, float& fac,
                                                 float& sumq, float& delta) {
    const bool res = residual_delta<NL>(q, c, y, D, S, uq, lane, fac, delta);
    wave_sync();
    if (res) {
        const float part = residual_sum16(uq, S, D, lane & 15);
        sumq = __shfl(part, 0);
    } else {
        sumq = (float)D * (float)(1.0 / (double)sqrtf((float)D));
    }
}

// The int8 MFMA digits of the quantized residual (digit 0: q mod 128, digit 1: floor(q /
// 128)) in operand order: 16-byte units (chunk c, plane p, lane group g) at unit index
// 32 c + 4 p + g, low-digit row then high-digit row; byte 4 jw + b of a row is dim
// 512 c + 128 g + 32 jw + 8 b + p.  One task = (c, g, jw, plane pair 2 pp, 2 pp + 1): the
// 8 dims 512 c + 128 g + 32 jw + 8 b + 2 pp + {0, 1}; task t of a wave has jw = t & 3,
// pp = (t >> 2) & 3, g = (t >> 4) & 3, c = t >> 6 (lane-major u: conflict-free LDS reads
// within 16 lanes).  Dims >= D (a whole 32-dim block, D being a multiple of 64) get 0.
// dst is the pair's 2 * nchunk * 512 bytes as dwords.  qv (nullable) also receives the
// digits q themselves as int16 in the bit-plane order of the codes (dim 512 c + 8 B + p at
// qv[(64 c + B) * 8 + p]), which the wrap correction reads.
__device__ __forceinline__ void quantize_digits(const float* uq, int S, int D, int nchunk,
                                                double rd, uint32_t* dst, int lane,
                                                int16_t* qv = nullptr) {
    for (int t = lane; t < nchunk * 64; t += kWave) {
        const int jw = t & 3, pp = (t >> 2) & 3, g = (t >> 4) & 3, c = t >> 6;
        const int base = 512 * c + 128 * g + 32 * jw + 2 * pp;
        uint32_t lo0 = 0, hi0 = 0, lo1 = 0, hi1 = 0;
        int q0[4] = {0, 0, 0, 0}, q1[4] = {0, 0, 0, 0};
        if (base < D) {
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int d = base + 8 * b;
                q0[b] = quantize_digit(uq[ul_index(d, S)], rd);
                q1[b] = quantize_digit(uq[ul_index(d + 1, S)], rd);
            }
            // digit 0 = q & 127 (byte b of the word), digit 1 = floor(q / 128) (its low byte:
            // bits 7..14 of q, as |q| < 2^13).  The low halves of two q share a word, so one
            // shift serves both high digits.
            const uint32_t a01 = __builtin_amdgcn_perm((uint32_t)q0[1], (uint32_t)q0[0], 0x05040100u);
            const uint32_t a23 = __builtin_amdgcn_perm((uint32_t)q0[3], (uint32_t)q0[2], 0x05040100u);
            const uint32_t b01 = __builtin_amdgcn_perm((uint32_t)q1[1], (uint32_t)q1[0], 0x05040100u);
            const uint32_t b23 = __builtin_amdgcn_perm((uint32_t)q1[3], (uint32_t)q1[2], 0x05040100u);
            lo0 = __builtin_amdgcn_perm(a23, a01, 0x06040200u) & 0x7F7F7F7Fu;
            lo1 = __builtin_amdgcn_perm(b23, b01, 0x06040200u) & 0x7F7F7F7Fu;
            hi0 = __builtin_amdgcn_perm(a23 >> 7, a01 >> 7, 0x06040200u);
            hi1 = __builtin_amdgcn_perm(b23 >> 7, b01 >> 7, 0x06040200u);
        }
        const int un = 32 * c + 8 * pp + g;  // plane 2 pp; plane 2 pp + 1 is un + 4
        dst[8 * un + jw] = lo0;
        dst[8 * un + 4 + jw] = hi0;
        dst[8 * un + 32 + jw] = lo1;
        dst[8 * un + 36 + jw] = hi1;
        if (qv) {
            // byte B = 16 g + 4 jw + b, planes 2 pp and 2 pp + 1: one dword per b
            uint32_t* qd = reinterpret_cast<uint32_t*>(qv);
#pragma unroll
            for (int b = 0; b < 4; b++)
                qd[(64 * c + 16 * g + 4 * jw + b) * 4 + pp] =
                    ((uint32_t)q0[b] & 0xFFFFu) | ((uint32_t)q1[b] << 16);
        }
    }
}

// ---- the reference's u16 wraparound above D = 2048 (FastScan.hpp:237-290) ----
// The reference adds the low and the high bytes of the shifted LUT entries
// L'_m[code] = sum_{k<4} bit(4m+k) q_{4m+k} - sum_{k<4} min(0, q_{4m+k})
//            = sum_{k<4} (bit(4m+k) ? max(q_{4m+k}, 0) : max(-q_{4m+k}, 0))
// in u16 lanes, separately for even and odd codebooks, and takes each of the four half-sums
// mod 2^16.  A high byte is at most 127 (L' <= 4 * 8191), so the high half-sums stay below
// 2^16 for D <= 4096; a low half-sum reaches 2^16 only above D = 2048 and stays below 2^17.
// So the reference's integer is the exact inner product minus 65536 per low half-sum that
// reached 2^16, and that integer is never larger than the exact one: a vector the lower
// bound rejects on the exact integer is rejected by the reference too.
// lo_wraps returns the number of wrapped low half-sums (0, 1 or 2) of one vector for the
// 16 lanes of a group (l = lane & 15; every lane of the group gets the result).  code(c)
// returns the lane's dword of chunk c of the vector's bit-plane codes (bytes 4l .. 4l+3,
// byte B bit p = dim 512 c + 8 B + p), qv the digits as quantize_digits stores them.  In a
// byte, bits 0..3 belong to an even codebook, bits 4..7 to the next (odd) one.
template <typename CF>
__device__ __forceinline__ int lo_wraps(const int16_t* __restrict__ qv, int nchunk, int l,
                                        const CF& code) {
    int se = 0, so = 0;
    for (int c = 0; c < nchunk; c++) {
        const uint32_t dw = code(c);
        const uint4* qq = reinterpret_cast<const uint4*>(qv + (size_t)(64 * c + 4 * l) * 8);
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const uint4 w = qq[b];  // dims 8 B .. 8 B + 7 of byte B = 4 l + b, two per dword
            const uint32_t x = dw >> (8 * b);
            const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
            int e = 0, o = 0;
#pragma unroll
            for (int h = 0; h < 4; h++) {
                const int qa = (int)(int16_t)(ws[h] & 0xFFFFu), qb = (int)(int16_t)(ws[h] >> 16);
                const int va = max(((x >> (2 * h)) & 1u) ? qa : -qa, 0);
                const int vb = max(((x >> (2 * h + 1)) & 1u) ? qb : -qb, 0);
                if (h < 2) e += va + vb; else o += va + vb;
            }
            se += e & 255;
            so += o & 255;
        }
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) {
        se += __shfl_xor(se, off);
        so += __shfl_xor(so, off);
    }
    return (se >> 16) + (so >> 16);
}

// The digits q alone (no MFMA operands), in quantize_digits' qv order, from the lane-major
// residual: dims 512 c + 8 B + 2 h + {0, 1} as one dword per task.
__device__ __forceinline__ void store_qv(const float* uq, int S, int D, int nchunk, double rd,
                                         int16_t* qv, int lane) {
    uint32_t* qd = reinterpret_cast<uint32_t*>(qv);
    for (int t = lane; t < nchunk * 256; t += kWave) {
        const int d = 2 * t;  // = 512 c + 8 B + 2 h
        uint32_t v = 0;
        if (d < D) {
            const int q0 = quantize_digit(uq[ul_index(d, S)], rd);
            const int q1 = quantize_digit(uq[ul_index(d + 1, S)], rd);
            v = ((uint32_t)q0 & 0xFFFFu) | ((uint32_t)q1 << 16);
        }
        qd[t] = v;
    }
}

// The values 4 h .. 4 h + 3 of an exl dword x as floats (h < exl_npb(EX)).  The byte mask
// is applied by an opaque v_and so that each byte then converts with one v_cvt_f32_ubyteN
// (the compiler would otherwise fold mask and shift into a v_bfe_u32 per value).
template <int EX>
__device__ __forceinline__ float4 exl_group(uint32_t x, int h) {
    constexpr uint32_t m = ((1u << EX) - 1u) * 0x01010101u;
    uint32_t y = x >> (exl_sp(EX) * h);
    if constexpr (m != 0xFFFFFFFFu) asm("v_and_b32 %0, %1, %2" : "=v"(y) : "s"(m), "v"(y));
    return make_float4((float)(y & 0xFFu), (float)((y >> 8) & 0xFFu), (float)((y >> 16) & 0xFFu),
                       (float)(y >> 24));
}

// fma(u_i, e_i, part) for the four values of group h of dword x, in order.
template <int EX>
__device__ __forceinline__ void ex_group_fma(float4 u, uint32_t x, int h, float& part) {
    const float4 e = exl_group<EX>(x, h);
    part = fmaf(u.x, e.x, part);
    part = fmaf(u.y, e.y, part);
    part = fmaf(u.z, e.z, part);
    part = fmaf(u.w, e.w, part);
}

// The chain over 16-byte code words b4 .. b4 + 3 (those below nq) of a lane's exl codes,
// held in c4: part = fma(u_i, e_i, part) for their values in order.  nfd = full dwords,
// rem = 4-value groups of the last, partly used dword (ex_chain).
template <int EX>
__device__ __forceinline__ void ex_chain_words(const float4* __restrict__ u4, const uint4 (&c4)[4],
                                               int b4, int nq, int nfd, int rem, float& part) {
    constexpr int npb = exl_npb(EX);
#pragma unroll
    for (int j4 = 0; j4 < 4; j4++) {
        if (b4 + j4 >= nq) break;
        const uint32_t cs[4] = {c4[j4].x, c4[j4].y, c4[j4].z, c4[j4].w};
        const int dw0 = 4 * (b4 + j4);
        if (dw0 + 4 <= nfd) {
            // every value of the word: no branch; the next dword's u is read while this
            // one's FMAs run (scheduling barriers keep the compiler from hoisting every
            // read of the word, which costs registers)
            float4 ua[npb], ub[npb];
#pragma unroll
            for (int h = 0; h < npb; h++) ua[h] = u4[dw0 * npb + h];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (j < 3) {
#pragma unroll
                    for (int h = 0; h < npb; h++) ub[h] = u4[(dw0 + j + 1) * npb + h];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int h = 0; h < npb; h++) ex_group_fma<EX>(ua[h], cs[j], h, part);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int h = 0; h < npb; h++) ua[h] = ub[h];
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int dw = dw0 + j;
                const int nh = dw < nfd ? npb : (dw == nfd ? rem : 0);
#pragma unroll
                for (int h = 0; h < npb; h++)
                    if (h < nh) ex_group_fma<EX>(u4[dw * npb + h], cs[j], h, part);
            }
        }
    }
}

// The shape of a lane's exl codes for nvals values: full dwords, groups of the last dword,
// 16-byte words holding values.
template <int EX>
__device__ __forceinline__ void ex_chain_shape(int nvals, int& nfd, int& rem, int& nq) {
    constexpr int npb = exl_npb(EX);
    const int ng = nvals >> 2;
    nfd = ng / npb;
    rem = ng - nfd * npb;
    nq = (nfd + (rem ? 1 : 0) + 3) >> 2;
}

// One lane's share of <u, e> for the re-estimate: part = fma(u_i, e_i, part) over
// i = 0 .. nvals-1 in order (nvals a multiple of 4); ur = the lane's u row (u_i = ur[i], for
// dim 16 i + l, 16-byte aligned), code = the lane's W exl dwords.  u is read four values at
// a time; only the last 16-byte code word can be partly used.
template <int EX>
__device__ __forceinline__ float ex_chain(const float* __restrict__ ur,
                                          const uint32_t* __restrict__ code, int nvals, int W) {
    const float4* u4 = reinterpret_cast<const float4*>(ur);
    const uint4* cw = reinterpret_cast<const uint4*>(code);
    int nfd, rem, nq;
    ex_chain_shape<EX>(nvals, nfd, rem, nq);
    float part = 0.f;
    for (int b4 = 0; b4 < nq; b4 += 4) {
        uint4 c4[4];  // up to four 16-byte loads in flight before the chain uses them
#pragma unroll
        for (int j4 = 0; j4 < 4; j4++)
            if (b4 + j4 < nq) c4[j4] = cw[b4 + j4];
        ex_chain_words<EX>(u4, c4, b4, nq, nfd, rem, part);
    }
    (void)W;
    return part;
}

// ex_chain for two vectors at once (codes ca, cb; cb unused when !has_b): their loads are
// issued together and each u value feeds both chains, each chain in the order of ex_chain.
template <int EX>
__device__ __forceinline__ void ex_chain2(const float* __restrict__ ur,
                                          const uint32_t* __restrict__ ca,
                                          const uint32_t* __restrict__ cb, bool has_b,
                                          int nvals, int W, float& pa, float& pb) {
    constexpr int npb = exl_npb(EX);
    const float4* u4 = reinterpret_cast<const float4*>(ur);
    const uint4* wa = reinterpret_cast<const uint4*>(ca);
    const uint4* wb = reinterpret_cast<const uint4*>(cb);
    const int ng = nvals >> 2;
    const int nfd = ng / npb, rem = ng - nfd * npb;
    const int nq = (nfd + (rem ? 1 : 0) + 3) >> 2;
    pa = 0.f;
    pb = 0.f;
    for (int b4 = 0; b4 < nq; b4 += 4) {
        uint4 xa[4], xb[4];
#pragma unroll
        for (int j4 = 0; j4 < 4; j4++) {
            if (b4 + j4 < nq) {
                xa[j4] = wa[b4 + j4];
                xb[j4] = has_b ? wb[b4 + j4] : make_uint4(0, 0, 0, 0);
            }
        }
#pragma unroll
        for (int j4 = 0; j4 < 4; j4++) {
            if (b4 + j4 >= nq) break;
            const uint32_t sa[4] = {xa[j4].x, xa[j4].y, xa[j4].z, xa[j4].w};
            const uint32_t sb[4] = {xb[j4].x, xb[j4].y, xb[j4].z, xb[j4].w};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int dw = 4 * (b4 + j4) + j;
                const int nh = dw < nfd ? npb : (dw == nfd ? rem : 0);
#pragma unroll
                for (int h = 0; h < npb; h++)
                    if (h < nh) {
                        const float4 u = u4[dw * npb + h];
                        ex_group_fma<EX>(u, sa[j], h, pa);
                        ex_group_fma<EX>(u, sb[j], h, pb);
                    }
            }
        }
    }
    (void)W;
}

// Float key with the same order as the float (for radix selection / sorting).
__device__ __forceinline__ uint32_t float_key(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

}  // namespace erq
