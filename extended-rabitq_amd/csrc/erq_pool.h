// erq_pool.h -- the per-query result pool (ResultPool, Pool.hpp:10-50) kept in a wave's LDS
// slice, shared by the per-query scan kernels.
#pragma once
#include "erq_device.h"

namespace erq {

struct WaveLds {
    float* uq;
    uint16_t* lut;
    int* sums;
    float* pd;
    uint32_t* pid;
    float* nd;      // a block's re-estimated survivors (MFMA kernel), up to 32
    uint32_t* nid;
};

__device__ __forceinline__ size_t wave_lds_bytes(int D, int k) {
    int kcap = (k + 63) & ~63;
    return (size_t)12 * D + 128 + (size_t)8 * kcap;
}

// ResultPool::insert + distk (Pool.hpp:15-31) with the pool in LDS, wave-parallel.
// dist/id are wave-uniform.  New entries go before equal ones; a full pool rejects
// only strictly larger distances.
__device__ __forceinline__ void pool_insert(const WaveLds& w, int& psize, int k, int lane,
                                            float dist, uint32_t id) {
    if (psize == k && dist > w.pd[k - 1]) return;
    int pos = 0;
    for (int base = 0; base < psize; base += kWave) {
        int i = base + lane;
        bool lt = (i < psize) && (w.pd[i] < dist);
        pos += __popcll(ballot(lt));
    }
    const int lim = psize < k - 1 ? psize : k - 1;
    for (int base = lim & ~(kWave - 1); base >= 0; base -= kWave) {
        int i = base + lane;
        bool mv = (i > pos) && (i <= lim);
        float vd = 0.f;
        uint32_t vid = 0;
        if (mv) {
            vd = w.pd[i - 1];
            vid = w.pid[i - 1];
        }
        wave_sync();
        if (mv) {
            w.pd[i] = vd;
            w.pid[i] = vid;
        }
        wave_sync();
    }
    if (lane == 0) {
        w.pd[pos] = dist;
        w.pid[pos] = id;
    }
    wave_sync();
    psize += (psize < k);
}

// ns successive pool_insert calls (entries nd[t], nid[t], t = 0 .. ns-1 in insertion
// order, ns <= 32) as one merge.  pool_insert keeps the pool as the first min(k, size)
// entries of all inserted ones in the order (distance ascending; of equal distances the
// later-inserted first): a new entry lands after the pool entries strictly below it, and a
// full pool drops its last entry.  So new entry t ends at (pool entries with a distance
// below its own) + (new entries before it in that order), and pool entry i at
// i + (new entries with a distance <= its own).  nd is overwritten.
__device__ __forceinline__ void pool_insert_bulk(const WaveLds& w, int& psize, int k, int lane,
                                                 float* nd, const uint32_t* nid, int ns) {
    float dt = 0.f;
    uint32_t it = 0;
    int rt = 0, at = k;
    if (lane < ns) {
        dt = nd[lane];
        it = nid[lane];
    }
    for (int u = 0; u < ns; u++) {  // wave-uniform; the other lanes' values by shuffle
        const float du = __shfl(dt, u);
        rt += (du < dt || (du == dt && u > lane)) ? 1 : 0;
    }
    if (lane < ns) {
        int lo = 0, hi = psize;  // pool entries with pd < dt
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (w.pd[mid] < dt) lo = mid + 1; else hi = mid;
        }
        at = lo + rt;
    }
    wave_sync();
    if (lane < ns) nd[rt] = dt;  // the new distances, ascending
    wave_sync();
    // pool entries move up, highest 64 first; each lane reads before any lane writes
    for (int base = (psize - 1) & ~(kWave - 1); psize > 0 && base >= 0; base -= kWave) {
        const int i = base + lane;
        float vd = 0.f;
        uint32_t vid = 0;
        int dst = k;
        if (i < psize) {
            vd = w.pd[i];
            vid = w.pid[i];
            int lo = 0, hi = ns;  // new entries with a distance <= vd
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (nd[mid] <= vd) lo = mid + 1; else hi = mid;
            }
            dst = i + lo;
        }
        wave_sync();
        if (dst < k) {
            w.pd[dst] = vd;
            w.pid[dst] = vid;
        }
        wave_sync();
    }
    if (at < k) {
        w.pd[at] = dt;
        w.pid[at] = it;
    }
    wave_sync();
    psize = min(k, psize + ns);
}

// pool_insert_bulk for k <= 128 without chains of dependent LDS reads: the pool's two
// 64-entry halves (entry lane and lane + 64) and the new entries sit in registers, and every
// count is one pass over the new entries read with v_readlane (a wave-uniform index): new
// entry t's place = (pool entries below it) + (new entries before it), pool entry i's = i +
// (new entries with a distance <= its own), as in pool_insert_bulk.  All entries are read
// before any is written.  nd and nid are not changed.
__device__ __forceinline__ void pool_insert_bulk128(const WaveLds& w, int& psize, int k, int lane,
                                                    const float* nd, const uint32_t* nid, int ns) {
    float dt = FLT_MAX;
    uint32_t it = 0;
    if (lane < ns) {
        dt = nd[lane];
        it = nid[lane];
    }
    const bool h0 = lane < psize, h1 = lane + kWave < psize;
    const float v0 = h0 ? w.pd[lane] : 0.f, v1 = h1 ? w.pd[lane + kWave] : 0.f;
    const uint32_t i0 = h0 ? w.pid[lane] : 0u, i1 = h1 ? w.pid[lane + kWave] : 0u;
    int rt = 0, lo = 0, m0 = 0, m1 = 0;
    for (int u = 0; u < ns; u++) {  // wave-uniform
        const float du = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dt), u));
        rt += (du < dt || (du == dt && u > lane)) ? 1 : 0;  // new entries placed before t
        m0 += du <= v0 ? 1 : 0;                            // new entries placed before v0
        m1 += du <= v1 ? 1 : 0;
        const int below = __popcll(ballot(h0 && v0 < du)) + __popcll(ballot(h1 && v1 < du));
        if (lane == u) lo = below;  // pool entries below new entry u
    }
    wave_sync();  // every lane has read its pool entries
    const int d0 = lane + m0, d1 = lane + kWave + m1, at = lo + rt;
    if (h0 && d0 < k) {
        w.pd[d0] = v0;
        w.pid[d0] = i0;
    }
    if (h1 && d1 < k) {
        w.pd[d1] = v1;
        w.pid[d1] = i1;
    }
    if (lane < ns && at < k) {
        w.pd[at] = dt;
        w.pid[at] = it;
    }
    wave_sync();
    psize = min(k, psize + ns);
}

}  // namespace erq
