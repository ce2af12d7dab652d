// builder.cpp -- host-side ExRaBitQ HA index quantizer (OpenMP).
//
// Produces the on-disk/in-HBM HA layout that the search path reads:
//   short blocks : per 32 vectors, 4D bytes of packed 1-bit FastScan codes followed by
//                  32 float factors ||o - c||          (Quantizer.hpp:86-125, :277-290)
//   long codes   : per vector D*EX/8 bytes of the compacted EX-bit code (:491-631)
//   ex factors   : per vector xipnorm = 2*||o - c|| / <o_bar + 1/2, |u|> (:457-489)
// The caller supplies, per vector, the rotated residual r = o*P - c*P and ||o - c||.
// The EX-bit code search is the critical-point enumeration of DataQuantizer::fast_quantize
// (Quantizer.hpp:138-219): it has to choose the same codes, so it follows the same steps.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Position of vector v%16 inside a codebook's 16 bytes (pack_codes.hpp:44: perm0 places
// vector w < 8 at byte 2w and vector w+8 at byte 2w+1).
inline int byte_of_vec(int v) {
    int w = v & 15;
    return w < 8 ? 2 * w : 2 * (w - 8) + 1;
}

// Per-thread buffers of the rescale search.
struct Scratch {
    struct Ev {
        double t;
        uint32_t i;
    };
    std::vector<int> cur;
    std::vector<Ev> ev, sorted, late;
    std::vector<uint32_t> cnt;
    explicit Scratch(size_t D) : cur(D) {}
};

inline bool ev_less(const Scratch::Ev& a, const Scratch::Ev& b) {
    return a.t < b.t || (a.t == b.t && a.i < b.i);
}

// Critical-point search for the rescale factor t maximising <o_bar(t)+1/2, o>/||o_bar(t)+1/2||
// over o_bar(t)_i = floor(t*o_i + eps) clipped to 2^EX - 1; o holds |u| (all >= 0).
// Writes the codes and returns 1/<o_bar + 1/2, o> (1.0 if not finite).
//
// DataQuantizer::fast_quantize pops its events -- t = (v+1)/o_i, the next breakpoint of
// dimension i -- from a min-heap of (t, i), pushing a dimension's next breakpoint when it
// pops the previous one (while below t_end).  A dimension's breakpoints increase with v, so
// the pops are every event it ever holds in ascending (t, i) order.  Here those events are
// listed up front and put in that order directly: a counting sort on a bucket of t that is
// monotone in t, then insertion sort by (t, i) (each bucket holds a few), with the first
// events at or past t_end (small magnitudes only) sorted on their own and taken last.  The
// sweep then makes the same updates in the same order as the heap.
float quantize_magnitudes(const float* o, size_t D, int ex_bits, uint8_t* code, Scratch& s) {
    const double eps = 1e-5;
    const int n_enum = 10;
    const int top = (1 << ex_bits) - 1;
    double max_o = -1;
    for (size_t i = 0; i < D; i++)
        if (o[i] > max_o) max_o = o[i];
    if (!(max_o > 0)) {
        // zero residual (a vector equal to its centroid, e.g. a singleton cluster): the
        // reference divides by max_o = 0 here; every magnitude code is 0 and the factor
        // takes the reference's non-finite fallback (Quantizer.hpp:210-213).
        std::memset(code, 0, D);
        return 1.f;
    }
    double t_start = (double)(top / 3) / max_o;
    double t_end = (double)(top + n_enum) / max_o;

    double sqr_den = D * 0.25;
    double num = 0;
    s.ev.clear();
    s.late.clear();
    for (size_t i = 0; i < D; i++) {
        const int c = int(t_start * o[i] + eps);
        s.cur[i] = c;
        sqr_den += c * c + c;
        num += (c + 0.5) * o[i];
        const double oi = o[i];
        const double t0 = (double)(c + 1) / oi;  // pushed whatever its value
        if (t0 < t_end)
            s.ev.push_back({t0, (uint32_t)i});
        else
            s.late.push_back({t0, (uint32_t)i});
        for (int v = c + 1; v < top; v++) {  // pushed while below t_end
            const double tn = (double)(v + 1) / oi;
            if (!(tn < t_end)) break;
            s.ev.push_back({tn, (uint32_t)i});
        }
    }
    const size_t n = s.ev.size();
    s.sorted.resize(n);
    if (n > 0) {
        double t_lo = t_end;
        for (const auto& e : s.ev) t_lo = std::min(t_lo, e.t);
        const size_t B = 4 * n + 1;
        const double scale = (double)B / (t_end - t_lo);
        auto bucket = [&](double t) {
            const size_t b = (size_t)((t - t_lo) * scale);
            return b < B ? b : B - 1;
        };
        s.cnt.assign(B + 1, 0);
        for (const auto& e : s.ev) s.cnt[bucket(e.t) + 1]++;
        for (size_t b = 0; b < B; b++) s.cnt[b + 1] += s.cnt[b];
        for (const auto& e : s.ev) s.sorted[s.cnt[bucket(e.t)]++] = e;
        for (size_t k = 1; k < n; k++) {
            const Scratch::Ev x = s.sorted[k];
            size_t j = k;
            while (j > 0 && ev_less(x, s.sorted[j - 1])) {
                s.sorted[j] = s.sorted[j - 1];
                j--;
            }
            s.sorted[j] = x;
        }
    }
    std::sort(s.late.begin(), s.late.end(), ev_less);

    double best_ip = 0, best_t = 0;
    auto pop = [&](const Scratch::Ev& e) {
        const int v = ++s.cur[e.i];
        sqr_den += 2 * v;
        num += o[e.i];
        const double ip = num / std::sqrt(sqr_den);
        if (ip > best_ip) {
            best_ip = ip;
            best_t = e.t;
        }
    };
    for (const auto& e : s.sorted) pop(e);
    for (const auto& e : s.late) pop(e);

    double numer = 0;
    for (size_t i = 0; i < D; i++) {
        int c = int(best_t * o[i] + eps);
        if (c > top) c = top;
        code[i] = (uint8_t)c;
        numer += (c + 0.5) * o[i];
    }
    float ipn = (float)(1.0 / numer);
    if (!std::isfinite(ipn)) ipn = 1.f;
    return ipn;
}

// Compacted EX-bit layouts (Quantizer.hpp:491-631).
void compact_code(const uint8_t* raw, size_t D, int ex_bits, uint8_t* out) {
    switch (ex_bits) {
    case 8:
        std::memcpy(out, raw, D);
        break;
    case 4:
        for (size_t d = 0; d < D; d += 32, out += 16)
            for (int j = 0; j < 16; j++) out[j] = (uint8_t)(raw[d + j] | (raw[d + 16 + j] << 4));
        break;
    case 2:
    case 3:
        for (size_t d = 0; d < D; d += 64) {
            for (int j = 0; j < 16; j++)
                out[j] = (uint8_t)((raw[d + j] & 3) | ((raw[d + 16 + j] & 3) << 2) |
                                   ((raw[d + 32 + j] & 3) << 4) | ((raw[d + 48 + j] & 3) << 6));
            out += 16;
            if (ex_bits == 3) {
                uint64_t T = 0;
                for (int t = 0; t < 64; t++)
                    T |= (uint64_t)((raw[d + t] >> 2) & 1) << (8 * (t & 7) + (t >> 3));
                std::memcpy(out, &T, 8);
                out += 8;
            }
        }
        break;
    case 6:
    case 7:
        for (size_t d = 0; d < D; d += 64) {
            for (int j = 0; j < 16; j++) {
                out[j] = (uint8_t)((raw[d + j] & 63) | (((raw[d + 32 + j] >> 4) & 3) << 6));
                out[16 + j] =
                    (uint8_t)((raw[d + 16 + j] & 63) | (((raw[d + 48 + j] >> 4) & 3) << 6));
                out[32 + j] = (uint8_t)((raw[d + 32 + j] & 15) | ((raw[d + 48 + j] & 15) << 4));
            }
            out += 48;
            if (ex_bits == 7) {
                uint64_t T = 0;
                for (int t = 0; t < 64; t++)
                    T |= (uint64_t)((raw[d + t] >> 6) & 1) << (8 * (t & 7) + (t >> 3));
                std::memcpy(out, &T, 8);
                out += 8;
            }
        }
        break;
    }
}

}  // namespace

extern "C" {

// The rescale search of one vector: o = |u| (D values) -> codes in [0, 2^EX - 1]; returns
// 1/<o_bar + 1/2, o> (1.0 if not finite).  The per-vector step of erqb_quantize_cluster.
float erqb_quantize_magnitudes(const float* o, size_t D, int ex_bits, uint8_t* code) {
    Scratch s(D);
    return quantize_magnitudes(o, D, ex_bits, code, s);
}

// Quantize one cluster: r[n x D] rotated residuals (cluster-local order), onorm[n].
// blocks: ceil(n/32) * (4D + 128) bytes; long_code: n * D*EX/8; xipnorm: n.
int erqb_quantize_cluster(const float* r, const float* onorm, size_t n, size_t D, int ex_bits,
                          uint8_t* blocks, uint8_t* long_code, float* xipnorm) {
    if (D % 64 != 0) return -1;
    if (!(ex_bits == 2 || ex_bits == 3 || ex_bits == 4 || ex_bits == 6 || ex_bits == 7 ||
          ex_bits == 8))
        return -1;
    const size_t M = D / 4;
    const size_t block_bytes = 4 * D + 128;
    const size_t code_len = D * (size_t)ex_bits / 8;
    const int mask = (1 << ex_bits) - 1;
    const size_t nblocks = (n + 31) / 32;
    std::memset(blocks, 0, nblocks * block_bytes);

    std::vector<float> absu(D);
    std::vector<uint8_t> raw(D);
    Scratch scratch(D);
    for (size_t v = 0; v < n * D; v++)
        if (!std::isfinite(r[v])) return -2;
    for (size_t v = 0; v < n; v++)
        if (!std::isfinite(onorm[v])) return -2;
    for (size_t v = 0; v < n; v++) {
        const float* x = r + v * D;
        double ss = 0;
        for (size_t d = 0; d < D; d++) ss += (double)x[d] * x[d];
        float inv = ss > 0 ? (float)(1.0 / std::sqrt(ss)) : 0.f;
        for (size_t d = 0; d < D; d++) absu[d] = std::fabs(x[d] * inv);

        // 1-bit code: sign of the rotated residual (Quantizer.hpp:356-361), packed into
        // the FastScan block (nibble m = dims 4m..4m+3, dim 4m at bit 3).
        uint8_t* blk = blocks + (v / 32) * block_bytes;
        int vb = (int)(v % 32);
        int j = byte_of_vec(vb);
        for (size_t m = 0; m < M; m++) {
            const float* xm = x + 4 * m;
            int nib = ((xm[0] > 0) << 3) | ((xm[1] > 0) << 2) | ((xm[2] > 0) << 1) | (xm[3] > 0);
            uint8_t* cb = blk + (m >> 1) * 32 + (m & 1) * 16;
            cb[j] |= (uint8_t)(vb < 16 ? nib : nib << 4);
        }
        reinterpret_cast<float*>(blk + 4 * D)[vb] = onorm[v];

        // EX-bit code on |u|, flipped where the sign bit is 0 (Quantizer.hpp:470-488).
        float ipn = quantize_magnitudes(absu.data(), D, ex_bits, raw.data(), scratch);
        xipnorm[v] = ipn * 2 * onorm[v];
        for (size_t d = 0; d < D; d++)
            if (!(x[d] > 0)) raw[d] = (uint8_t)((~raw[d]) & mask);
        compact_code(raw.data(), D, ex_bits, long_code + v * code_len);
    }
    return 0;
}

// The non-HA format (Searcher.hpp; DataQuantizer::rabitq_factor, Quantizer.hpp:386-437): the
// same codes, four factors per vector after a block's 4D code bytes -- ||o - c||^2 (x2[v],
// given), fac_ip = 2 ||o - c|| / sum |u| / 2, the number of positive dims, and the error
// factor 2 ||o - c|| FAC_ERR sqrt(1 - <u, o_bar>^2) / <u, o_bar> -- and xipnorm over
// ||o - c|| = sqrt((double) x2).  Sums over dims run in dim order (float for <u, o_bar>,
// double for the others).  blocks: ceil(n/32) * (4D + 512) bytes.
int erqb_quantize_cluster_nonha(const float* r, const float* x2, size_t n, size_t D, int ex_bits,
                                uint8_t* blocks, uint8_t* long_code, float* xipnorm) {
    if (D % 64 != 0) return -1;
    if (!(ex_bits == 2 || ex_bits == 3 || ex_bits == 4 || ex_bits == 6 || ex_bits == 7 ||
          ex_bits == 8))
        return -1;
    const size_t M = D / 4;
    const size_t block_bytes = 4 * D + 512;
    const size_t code_len = D * (size_t)ex_bits / 8;
    const int mask = (1 << ex_bits) - 1;
    const size_t nblocks = (n + 31) / 32;
    const double fac_norm = 1 / std::sqrt((double)D);
    const double fac_err = 2.0 / std::sqrt((double)(D - 1));
    std::memset(blocks, 0, nblocks * block_bytes);
    std::vector<float> u(D), absu(D);
    std::vector<uint8_t> raw(D);
    Scratch scratch(D);
    for (size_t v = 0; v < n * D; v++)
        if (!std::isfinite(r[v])) return -2;
    for (size_t v = 0; v < n; v++)
        if (!std::isfinite(x2[v])) return -2;
    for (size_t v = 0; v < n; v++) {
        const float* x = r + v * D;
        double ss = 0;
        for (size_t d = 0; d < D; d++) ss += (double)x[d] * x[d];
        const float inv = ss > 0 ? (float)(1.0 / std::sqrt(ss)) : 0.f;
        for (size_t d = 0; d < D; d++) {
            u[d] = x[d] * inv;
            absu[d] = std::fabs(u[d]);
        }
        uint8_t* blk = blocks + (v / 32) * block_bytes;
        const int vb = (int)(v % 32), j = byte_of_vec(vb);
        for (size_t m = 0; m < M; m++) {
            const float* xm = x + 4 * m;
            int nib = ((xm[0] > 0) << 3) | ((xm[1] > 0) << 2) | ((xm[2] > 0) << 1) | (xm[3] > 0);
            uint8_t* cb = blk + (m >> 1) * 32 + (m & 1) * 16;
            cb[j] |= (uint8_t)(vb < 16 ? nib : nib << 4);
        }
        const float dist2c = (float)std::sqrt((double)x2[v]);
        float o_obar = 0.f;
        double ip = 0;
        int sumxb = 0;
        const float fn = (float)fac_norm;
        for (size_t d = 0; d < D; d++) {
            const bool pos = x[d] > 0;
            o_obar += u[d] * (pos ? fn : -fn);
            ip += std::abs(0.5 * u[d]);
            sumxb += pos;
        }
        if (!std::isfinite(o_obar)) o_obar = 0.8f;
        float* fac = reinterpret_cast<float*>(blk + 4 * D);
        fac[vb] = x2[v];
        fac[32 + vb] = (float)(1.0 / ip * 2 * dist2c);
        fac[64 + vb] = (float)sumxb;
        fac[96 + vb] = (float)(std::sqrt((1.0 - o_obar * o_obar) / (o_obar * o_obar)) * fac_err *
                               2 * dist2c);
        const float ipn = quantize_magnitudes(absu.data(), D, ex_bits, raw.data(), scratch);
        xipnorm[v] = (float)(ipn * 2 * std::sqrt((double)x2[v]));
        for (size_t d = 0; d < D; d++)
            if (!(x[d] > 0)) raw[d] = (uint8_t)((~raw[d]) & mask);
        compact_code(raw.data(), D, ex_bits, long_code + v * code_len);
    }
    return 0;
}

// All clusters, OpenMP over clusters.  r/onorm are in cluster order (cluster 0's
// vectors first); outputs are the contiguous SHORT_DATA / LONG_CODE / EX_FACTOR arrays.
int erqb_quantize_all(const float* r, const float* onorm, const uint64_t* cluster_sizes, size_t K,
                      size_t D, int ex_bits, uint8_t* short_data, uint8_t* long_code,
                      float* xipnorm, int threads) {
    std::vector<size_t> vstart(K + 1, 0), bstart(K + 1, 0);
    for (size_t k = 0; k < K; k++) {
        vstart[k + 1] = vstart[k] + cluster_sizes[k];
        bstart[k + 1] = bstart[k] + (cluster_sizes[k] + 31) / 32;
    }
    const size_t block_bytes = 4 * D + 128;
    const size_t code_len = D * (size_t)ex_bits / 8;
    int err = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (long k = 0; k < (long)K; k++) {
        size_t v0 = vstart[k];
        err |= erqb_quantize_cluster(r + v0 * D, onorm + v0, cluster_sizes[k], D, ex_bits,
                                     short_data + bstart[k] * block_bytes,
                                     long_code + v0 * code_len, xipnorm + v0) != 0;
    }
    return err ? -1 : 0;
}

// erqb_quantize_all for the non-HA format (x2 = ||o - c||^2 per vector).
int erqb_quantize_all_nonha(const float* r, const float* x2, const uint64_t* cluster_sizes,
                            size_t K, size_t D, int ex_bits, uint8_t* short_data,
                            uint8_t* long_code, float* xipnorm, int threads) {
    std::vector<size_t> vstart(K + 1, 0), bstart(K + 1, 0);
    for (size_t k = 0; k < K; k++) {
        vstart[k + 1] = vstart[k] + cluster_sizes[k];
        bstart[k + 1] = bstart[k] + (cluster_sizes[k] + 31) / 32;
    }
    const size_t block_bytes = 4 * D + 512;
    const size_t code_len = D * (size_t)ex_bits / 8;
    int err = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (long k = 0; k < (long)K; k++) {
        size_t v0 = vstart[k];
        err |= erqb_quantize_cluster_nonha(r + v0 * D, x2 + v0, cluster_sizes[k], D, ex_bits,
                                           short_data + bstart[k] * block_bytes,
                                           long_code + v0 * code_len, xipnorm + v0) != 0;
    }
    return err ? -1 : 0;
}

}  // extern "C"
