// erq_capi.hip -- the extern "C" boundary of liberq.so (declared in include/erq.h).
//
// Owns the resident index (the HA arrays of IVF::load, IVF.hpp:254-293, copied to HBM
// unchanged), a per-handle stream and a grow-only workspace, and sequences the search
// stages: rotation GEMM -> centroid GEMM -> probe selection -> per-query scan (or the
// cluster-major rounds).
//
// Every call on a handle holds its mutex, and is ordered after the previous call's device
// work when the two run on different streams (an event recorded when each call returns), so
// the workspace and the other shared buffers are never in use by two calls at once.  No
// exception crosses the ABI: bodies run through guarded().
#include <algorithm>
#include <cmath>
#include <functional>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "erq.h"
#include "erq_batch.h"
#include "erq_device.h"
#include "erq_internal.h"

namespace erq {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? ERQ_ENOMEM : ERQ_EHIP;
}

int workspace(erq_index* h, size_t bytes, void** out) {
    if (bytes > h->ws_bytes) {
        if (h->ws) {
            // kernels of earlier calls may still read it on a caller's stream
            hipError_t e = hipDeviceSynchronize();
            if (e != hipSuccess) return hip_fail(e, "workspace sync");
            (void)hipFree(h->ws);
            h->ws = nullptr;
            h->ws_bytes = 0;
        }
        size_t want = bytes + bytes / 4;
        hipError_t e = hipMalloc(&h->ws, want);
        if (e != hipSuccess) return hip_fail(e, "workspace alloc");
        h->ws_bytes = want;
    }
    *out = h->ws;
    return ERQ_OK;
}

}  // namespace erq

using namespace erq;

namespace {

constexpr size_t kMaxK = 8192;   // probe selection keeps all centroid keys in LDS
constexpr size_t kMaxD = 4096;   // bit-plane kernels: at most 8 chunks of 512 dims
constexpr int kCandCap = 8192;   // batched rounds: candidates kept per query per round (more:
                                 // the query's round is redone by the round-mode scan)
// Digits prepared ahead by the batched search (2 * Dp bytes per (query, cluster) pair): the
// pair budget is capped so that they stay within 24 GiB.
constexpr size_t kDigitBytes = size_t(3) << 33;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Scratch of erq_search_dev_noalloc: one array per device, held until the process ends.  In
// host code `scratch` only names the symbol; its device address comes from
// hipGetSymbolAddress, resolved once per device when the first index is created there.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// With `device` current: the device address of its scratch (resolved on the first call).
void resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void* ptr = nullptr;
    g_scratch_err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = g_scratch_err[device] == hipSuccess ? ptr : nullptr;
    g_scratch_done[device] = true;
}

bool ex_ok(size_t ex) { return ex == 2 || ex == 3 || ex == 4 || ex == 6 || ex == 7 || ex == 8; }

bool capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

// One call on a handle: holds the handle's mutex; the outermost call on a thread waits for
// the previous call's work when it ran on another stream, and records its own end.  Under
// HIP graph capture neither is done: the graph's launch orders the captured work.
struct Call {
    erq_index* h;
    hipStream_t s;
    std::unique_lock<std::recursive_mutex> lk;
    bool outer;
    Call(erq_index* h_, hipStream_t s_) : h(h_), s(s_), lk(h_->mu), outer(h_->depth++ == 0) {
        if (outer && h->last_valid && h->last_stream != s && !capturing(s))
            (void)hipStreamWaitEvent(s, h->ev_last, 0);
    }
    ~Call() {
        if (outer && !capturing(s) && hipEventRecord(h->ev_last, s) == hipSuccess) {
            h->last_stream = s;
            h->last_valid = true;
        }
        h->depth--;
    }
};

// Records HIP events around one stage on the stream it runs on while timing is on (not under
// graph capture, and not beyond the records reserved by erq_timing_begin).
struct StageTimer {
    erq_index* h;
    hipStream_t s;
    long slot = -1;
    StageTimer(erq_index* h_, int stage, hipStream_t s_) : h(h_), s(s_) {
        if (!h->tim_on || capturing(s)) return;
        const size_t n = h->tim_stage.size();
        if (2 * (n + 1) > h->tim_ev.size()) return;
        if (hipEventRecord(h->tim_ev[2 * n], s) != hipSuccess) return;
        h->tim_stage.push_back(stage);  // capacity reserved: no allocation
        slot = (long)n;
    }
    ~StageTimer() {
        if (slot >= 0) (void)hipEventRecord(h->tim_ev[2 * slot + 1], s);
    }
};

void timing_release(erq_index* h) {
    for (hipEvent_t e : h->tim_ev)
        if (e) (void)hipEventDestroy(e);
    h->tim_ev.clear();
    h->tim_stage.clear();
    h->tim_on = false;
}

template <typename T>
int upload(T** dst, const T* src, size_t count, const char* what) {
    if (count == 0) {
        *dst = nullptr;
        return ERQ_OK;
    }
    hipError_t e = hipMalloc(dst, count * sizeof(T));
    if (e != hipSuccess) return hip_fail(e, what);
    e = hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, what);
    return ERQ_OK;
}

void release(erq_index* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->last_valid) (void)hipEventSynchronize(h->ev_last);
    void* ptrs[] = {h->short_data, h->long_code, h->ex_factor, h->ids,    h->centroids,
                    h->cnorm,      h->P,         h->csize,     h->bstart, h->vstart,
                    h->ws,         h->bp,        h->order,     h->exl,    h->hq,
                    h->hids,       h->hdists,    h->hnp,       h->lat};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (h->pin) (void)hipHostFree(h->pin);
    if (h->aux) {
        (void)hipStreamSynchronize(h->aux);
        (void)hipStreamDestroy(h->aux);
    }
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_join) (void)hipEventDestroy(h->ev_join);
    if (h->ev_last) (void)hipEventDestroy(h->ev_last);
    timing_release(h);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

struct Ws {
    char* base;
    size_t off = 0;
    template <typename T>
    T* take(size_t n) {
        off = align_up(off, 256);
        T* p = reinterpret_cast<T*>(base + off);
        off += n * sizeof(T);
        return p;
    }
};

// _dev entry points run on the caller's stream exactly as given (0 is the null stream).
hipStream_t pick(erq_index*, void* s) { return static_cast<hipStream_t>(s); }

// Device buffers of the host-pointer calls: nq rotated queries, nq x k results, nq + 1
// probe offsets; grown (never shrunk) on demand, so repeated calls allocate nothing.
int host_bufs(erq_index* h, size_t nq, size_t k) {
    if (nq <= h->hq_cap && k <= h->hk_cap) return ERQ_OK;
    const size_t cq = std::max(nq, h->hq_cap), ck = std::max(k, h->hk_cap);
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(e, "host buffers sync");
    void* old[] = {h->hq, h->hids, h->hdists, h->hnp};
    for (void* p : old)
        if (p) (void)hipFree(p);
    h->hq = nullptr;
    h->hids = nullptr;
    h->hdists = nullptr;
    h->hnp = nullptr;
    h->hq_cap = h->hk_cap = 0;
    e = hipMalloc(&h->hq, cq * h->D * 4);
    if (e == hipSuccess) e = hipMalloc(&h->hids, cq * ck * 4);
    if (e == hipSuccess) e = hipMalloc(&h->hdists, cq * ck * 4);
    if (e == hipSuccess) e = hipMalloc(&h->hnp, (cq + 1) * 4);
    if (e != hipSuccess) return hip_fail(e, "host-call buffers");
    h->hq_cap = cq;
    h->hk_cap = ck;
    return ERQ_OK;
}

// Pinned staging of at least `bytes` for a host-pointer call's copies, or null when that is
// above kPinMax or cannot be allocated (the copies then use the caller's memory).  The
// previous call's copies have completed: every host-pointer call synchronises before it
// returns, under the handle's lock.
constexpr size_t kPinMax = 1u << 20;
void* pinned(erq_index* h, size_t bytes) {
    if (bytes > kPinMax) return nullptr;
    if (bytes <= h->pin_cap) return h->pin;
    if (h->pin) (void)hipHostFree(h->pin);
    h->pin = nullptr;
    h->pin_cap = 0;
    const size_t want = std::max<size_t>(bytes, 64 * 1024);
    if (hipHostMalloc(&h->pin, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();  // the pageable path still works
        h->pin = nullptr;
        return nullptr;
    }
    h->pin_cap = want;
    return h->pin;
}

}  // namespace

extern "C" {

const char* erq_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace {

// erq_index_create for either block format (factors per vector: 1 HA, 4 non-HA).
int create_index(size_t N, size_t DIM, size_t K, size_t EX_BITS, const uint64_t* cluster_sizes,
                 const float* P, const float* centroids, const uint8_t* short_data,
                 const uint8_t* long_code, const float* ex_factor, const uint32_t* ids,
                 int factors, erq_index** out) {
    return guarded([&]() -> int {
        if (!out || !cluster_sizes || !centroids || !short_data || !long_code || !ex_factor ||
            !ids) {
            set_error("erq_index_create: null argument");
            return ERQ_EINVAL;
        }
        *out = nullptr;
        if (!ex_ok(EX_BITS)) {
            set_error("erq_index_create: EX_BITS must be one of 2,3,4,6,7,8");
            return ERQ_EFORMAT;
        }
        if (K == 0 || DIM == 0) {
            set_error("erq_index_create: empty index");
            return ERQ_EFORMAT;
        }
        const size_t D = align_up(DIM, 64);
        if (K > kMaxK || D > kMaxD) {
            set_error("erq_index_create: K > 8192 or padded dim > 4096 is not supported");
            return ERQ_EUNSUPPORTED;
        }
        if (N > 0xFFFFFFFFull) {
            set_error("erq_index_create: N exceeds 32-bit ids");
            return ERQ_EUNSUPPORTED;
        }
        size_t total = 0, nblocks = 0;
        std::vector<uint32_t> cs(K), bs(K + 1), vs(K + 1);
        for (size_t k = 0; k < K; k++) {
            if (cluster_sizes[k] > N) {
                set_error("erq_index_create: a cluster is larger than N");
                return ERQ_EFORMAT;
            }
            cs[k] = (uint32_t)cluster_sizes[k];
            bs[k] = (uint32_t)nblocks;
            vs[k] = (uint32_t)total;
            total += cluster_sizes[k];
            nblocks += (cluster_sizes[k] + 31) / 32;
        }
        bs[K] = (uint32_t)nblocks;
        vs[K] = (uint32_t)total;
        if (total != N) {
            set_error("erq_index_create: cluster sizes do not sum to N");
            return ERQ_EFORMAT;
        }
        erq_index* h = new erq_index();
        hipError_t e = hipGetDevice(&h->device);
        if (e != hipSuccess) {
            delete h;
            return hip_fail(e, "hipGetDevice");
        }
        resolve_scratch(h->device);
        e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_last, hipEventDisableTiming);
        if (e != hipSuccess) {
            release(h);
            return hip_fail(e, "hipStreamCreate");
        }
        h->N = N;
        h->DIM = DIM;
        h->D = D;
        h->K = K;
        h->EX = EX_BITS;
        h->nblocks = nblocks;
        h->factors = factors;
        h->sizes.assign(cluster_sizes, cluster_sizes + K);
        const size_t BB = 4 * D + 128 * (size_t)factors;
        int rc;
        if ((rc = upload(&h->short_data, short_data, nblocks * BB, "upload short data")) ||
            (rc = upload(&h->long_code, long_code, N * D * EX_BITS / 8, "upload long codes")) ||
            (rc = upload(&h->ex_factor, ex_factor, N, "upload ex factors")) ||
            (rc = upload(&h->ids, ids, N, "upload ids")) ||
            (rc = upload(&h->centroids, centroids, K * D, "upload centroids")) ||
            (rc = upload(&h->csize, cs.data(), K, "upload sizes")) ||
            (rc = upload(&h->bstart, bs.data(), K + 1, "upload block starts")) ||
            (rc = upload(&h->vstart, vs.data(), K + 1, "upload vector starts"))) {
            release(h);
            return rc;
        }
        if (P && (rc = upload(&h->P, P, D * D, "upload rotation"))) {
            release(h);
            return rc;
        }
        e = hipMalloc(&h->cnorm, K * sizeof(float));
        if (e != hipSuccess) {
            release(h);
            return hip_fail(e, "alloc centroid norms");
        }
        e = launch_row_sqnorm(h->centroids, (int)K, (int)D, (int)D, h->cnorm, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) {
            release(h);
            return hip_fail(e, "centroid norms");
        }
        *out = h;
        return ERQ_OK;
    });
}

}  // namespace

extern "C" {

int erq_index_create(size_t N, size_t DIM, size_t K, size_t EX_BITS, const uint64_t* cluster_sizes,
                     const float* P, const float* centroids, const uint8_t* short_data,
                     const uint8_t* long_code, const float* ex_factor, const uint32_t* ids,
                     erq_index** out) {
    return create_index(N, DIM, K, EX_BITS, cluster_sizes, P, centroids, short_data, long_code,
                        ex_factor, ids, 1, out);
}

int erq_index_create_nonha(size_t N, size_t DIM, size_t K, size_t EX_BITS,
                           const uint64_t* cluster_sizes, const float* P, const float* centroids,
                           const uint8_t* short_data, const uint8_t* long_code,
                           const float* ex_factor, const uint32_t* ids, erq_index** out) {
    return create_index(N, DIM, K, EX_BITS, cluster_sizes, P, centroids, short_data, long_code,
                        ex_factor, ids, 4, out);
}

int erq_index_load(const char* path, erq_index** out) {
    return guarded([&]() -> int {
        if (!path || !out) {
            set_error("erq_index_load: null argument");
            return ERQ_EINVAL;
        }
        *out = nullptr;
        std::ifstream in(path, std::ios::binary);
        if (!in) {
            set_error(std::string("erq_index_load: cannot open ") + path);
            return ERQ_EIO;
        }
        in.seekg(0, std::ios::end);
        const std::streamoff fend = in.tellg();
        in.seekg(0, std::ios::beg);
        if (fend < 0) {
            set_error("erq_index_load: cannot size the file");
            return ERQ_EIO;
        }
        const uint64_t fsize = (uint64_t)fend;
        uint64_t hdr[4];
        if (!in.read(reinterpret_cast<char*>(hdr), sizeof(hdr))) {
            set_error("erq_index_load: truncated header");
            return ERQ_EIO;
        }
        const uint64_t N = hdr[0], DIM = hdr[1], K = hdr[2], EX = hdr[3];
        if (!ex_ok(EX) || K == 0 || DIM == 0 || K > (1u << 24) || DIM > (1u << 16) ||
            N > 0xFFFFFFFFull) {
            set_error("erq_index_load: bad header");
            return ERQ_EFORMAT;
        }
        // the header fixes every section's size: check the file holds them all before any
        // buffer is sized from it
        if (fsize < 32 + 8 * K) {
            set_error("erq_index_load: truncated cluster sizes");
            return ERQ_EIO;
        }
        const uint64_t D = align_up(DIM, 64);
        std::vector<uint64_t> sizes(K);
        if (!in.read(reinterpret_cast<char*>(sizes.data()), K * 8)) {
            set_error("erq_index_load: truncated cluster sizes");
            return ERQ_EIO;
        }
        uint64_t total = 0, nblocks = 0;
        for (auto s : sizes) {
            if (s > N) {
                set_error("erq_index_load: cluster sizes do not sum to N");
                return ERQ_EFORMAT;
            }
            total += s;
            nblocks += (s + 31) / 32;
        }
        if (total != N) {
            set_error("erq_index_load: cluster sizes do not sum to N");
            return ERQ_EFORMAT;
        }
        // The file does not record the block format: the HA one (1 factor per vector) or the
        // non-HA one (4, Quantizer.hpp:107-121) is the one whose size the file has.
        const uint64_t rest = 4 * D * D + 4 * K * D + N * D * EX / 8 + 4 * N + 4 * N;
        const uint64_t body_ha = rest + nblocks * (4 * D + 128);
        const uint64_t body_nonha = rest + nblocks * (4 * D + 512);
        const uint64_t head = 32 + 8 * K;
        int factors = 1;
        if (fsize < head + body_ha) {
            set_error("erq_index_load: truncated index body");
            return ERQ_EIO;
        }
        if (fsize == head + body_nonha && body_nonha != body_ha) {
            factors = 4;
        } else if (fsize != head + body_ha) {
            set_error(fsize < head + body_nonha
                          ? "erq_index_load: bytes after the index body, or a truncated non-HA "
                            "index (the size matches neither block format)"
                          : "erq_index_load: bytes after the index body");
            return ERQ_EFORMAT;
        }
        if (K > kMaxK || D > kMaxD) {
            set_error("erq_index_load: K > 8192 or padded dim > 4096 is not supported");
            return ERQ_EUNSUPPORTED;
        }
        std::vector<float> P(D * D), C(K * D), ef(N);
        std::vector<uint8_t> sd(nblocks * (4 * D + 128 * (uint64_t)factors)), lc(N * D * EX / 8);
        std::vector<uint32_t> ids(N);
        bool ok = in.read(reinterpret_cast<char*>(P.data()), P.size() * 4) &&
                  in.read(reinterpret_cast<char*>(C.data()), C.size() * 4) &&
                  in.read(reinterpret_cast<char*>(sd.data()), sd.size()) &&
                  in.read(reinterpret_cast<char*>(lc.data()), lc.size()) &&
                  in.read(reinterpret_cast<char*>(ef.data()), ef.size() * 4) &&
                  in.read(reinterpret_cast<char*>(ids.data()), ids.size() * 4);
        if (!ok) {
            set_error("erq_index_load: truncated index body");
            return ERQ_EIO;
        }
        return create_index(N, DIM, K, EX, sizes.data(), P.data(), C.data(), sd.data(),
                            lc.data(), ef.data(), ids.data(), factors, out);
    });
}

void erq_index_free(erq_index* h) { release(h); }

size_t erq_padded_dim(const erq_index* h) { return h ? h->D : 0; }
size_t erq_num_clusters(const erq_index* h) { return h ? h->K : 0; }
size_t erq_dim(const erq_index* h) { return h ? h->DIM : 0; }
size_t erq_num_vectors(const erq_index* h) { return h ? h->N : 0; }
size_t erq_ex_bits(const erq_index* h) { return h ? h->EX : 0; }
size_t erq_short_factors(const erq_index* h) { return h ? (size_t)h->factors : 0; }

const void* erq_device_array(const erq_index* h, int which) {
    if (!h) return nullptr;
    switch (which) {
    case 0: return h->short_data;
    case 1: return h->long_code;
    case 2: return h->ex_factor;
    case 3: return h->ids;
    case 4: return h->centroids;
    case 5: return h->P;
    case 6: return h->csize;
    case 7: return h->bstart;
    case 8: return h->vstart;
    default: return nullptr;
    }
}

int erq_rotate_dev(erq_index* h, const float* q_dev, size_t nq, float* q_rot_dev, void* stream) {
    return guarded([&]() -> int {
        if (!h || (!q_dev && nq) || (!q_rot_dev && nq)) {
            set_error("erq_rotate_dev: null argument");
            return ERQ_EINVAL;
        }
        if (!h->P) {
            set_error("erq_rotate_dev: index has no rotation matrix");
            return ERQ_EINVAL;
        }
        if (nq == 0) return ERQ_OK;
        hipStream_t s = pick(h, stream);
        Call call(h, s);
        const float* A = q_dev;
        if (h->DIM != h->D) {
            void* ws;
            int rc = workspace(h, nq * h->D * sizeof(float), &ws);
            if (rc) return rc;
            hipError_t e =
                launch_pad_rows(q_dev, (int)nq, (int)h->DIM, (float*)ws, (int)h->D, s);
            if (e != hipSuccess) return hip_fail(e, "pad queries");
            A = (const float*)ws;
        }
        hipError_t e = launch_gemm(A, (int)h->D, h->P, (int)h->D, 0, q_rot_dev, (int)h->D,
                                   (int)nq, (int)h->D, (int)h->D, 0, nullptr, nullptr, s);
        if (e != hipSuccess) return hip_fail(e, "rotation gemm");
        return ERQ_OK;
    });
}

int erq_centroid_dist_dev(erq_index* h, const float* q_rot_dev, size_t nq, float* out_dev,
                          void* stream) {
    return guarded([&]() -> int {
        if (!h || (nq && (!q_rot_dev || !out_dev))) {
            set_error("erq_centroid_dist_dev: null argument");
            return ERQ_EINVAL;
        }
        if (nq == 0) return ERQ_OK;
        hipStream_t s = pick(h, stream);
        Call call(h, s);
        void* wsp;
        int rc = workspace(h, align_up(nq * 4, 256), &wsp);
        if (rc) return rc;
        float* qn = (float*)wsp;
        hipError_t e = launch_row_sqnorm(q_rot_dev, (int)nq, (int)h->D, (int)h->D, qn, s);
        if (e == hipSuccess)
            e = launch_gemm(q_rot_dev, (int)h->D, h->centroids, (int)h->D, 1, out_dev,
                            (int)h->K, (int)nq, (int)h->K, (int)h->D, 1, qn, h->cnorm, s);
        if (e != hipSuccess) return hip_fail(e, "centroid gemm");
        return ERQ_OK;
    });
}

// Workspace bytes of probe_locked (approximate distances and query norms), at offset 0.
static size_t probe_ws_bytes(size_t nq, size_t K) {
    const size_t cand = nq <= kProbeSplitNq ? align_up(nq * kProbeCandWords * 4, 256) : 0;
    return align_up(nq * K * 4, 256) + align_up(nq * 4, 256) + cand + 512;
}

// Probe selection with its scratch (probe_ws_bytes) at `base`.
static int probe_into(erq_index* h, const float* q_rot_dev, size_t nq, size_t nprobe,
                      uint32_t* pid, float* pd, hipStream_t s, char* base,
                      uint32_t* fixed_off = nullptr) {
    StageTimer tim(h, ERQ_STAGE_SELECT, s);
    const size_t K = h->K, D = h->D;
    Ws ws{base};
    float* approx = ws.take<float>(nq * K);
    float* qn = ws.take<float>(nq);
    uint32_t* cand = nq <= kProbeSplitNq ? ws.take<uint32_t>(nq * kProbeCandWords) : nullptr;
    hipError_t e = launch_query_centroid_dist(q_rot_dev, (int)nq, (int)D, h->centroids, (int)K,
                                              h->cnorm, approx, qn, s);
    if (e != hipSuccess) return hip_fail(e, "centroid gemm");
    e = launch_probe_select(approx, qn, h->cnorm, q_rot_dev, h->centroids, (int)nq, (int)K,
                            (int)D, (int)nprobe, pid, pd, s, cand, cand ? fixed_off : nullptr);
    if (e != hipSuccess) return hip_fail(e, "probe selection");
    return ERQ_OK;
}

static int probe_locked(erq_index* h, const float* q_rot_dev, size_t nq, size_t nprobe,
                        uint32_t* pid, float* pd, hipStream_t s, uint32_t* fixed_off = nullptr) {
    void* wsp;
    int rc = workspace(h, probe_ws_bytes(nq, h->K), &wsp);
    if (rc) return rc;
    return probe_into(h, q_rot_dev, nq, nprobe, pid, pd, s, (char*)wsp, fixed_off);
}

int erq_probe_dev(erq_index* h, const float* q_rot_dev, size_t nq, size_t nprobe,
                  uint32_t* probe_ids_dev, float* probe_sqr_y_dev, void* stream) {
    return guarded([&]() -> int {
        if (!h || (nq && (!q_rot_dev || !probe_ids_dev || !probe_sqr_y_dev))) {
            set_error("erq_probe_dev: null argument");
            return ERQ_EINVAL;
        }
        if (nprobe == 0 || nprobe > h->K || nprobe > 2048) {
            set_error("erq_probe_dev: nprobe must be in [1, min(K, 2048)]");
            return ERQ_EINVAL;
        }
        if (nq == 0) return ERQ_OK;
        hipStream_t s = pick(h, stream);
        Call call(h, s);
        return probe_locked(h, q_rot_dev, nq, nprobe, probe_ids_dev, probe_sqr_y_dev, s);
    });
}

static int ensure_bp(erq_index* h, hipStream_t s);

static size_t max_cluster_size(const erq_index* h) {
    uint64_t m = 0;
    for (uint64_t v : h->sizes) m = std::max(m, v);
    return (size_t)m;
}

// The per-query scan's arguments (bp / exl set when the MFMA kernel runs).
static ScanArgs scan_args(const erq_index* h, const float* q_rot_dev, size_t nq, size_t k,
                          const uint32_t* off, const uint32_t* pid, const float* pd,
                          uint32_t* ids, float* dists, uint64_t* stats, int max_probes) {
    ScanArgs a{};
    a.q_rot = q_rot_dev;
    a.probe_off = off;
    a.probe_ids = pid;
    a.probe_sqr_y = pd;
    a.centroids = h->centroids;
    a.short_data = h->short_data;
    a.long_code = h->long_code;
    a.ex_factor = h->ex_factor;
    a.ids = h->ids;
    a.csize = h->csize;
    a.bstart = h->bstart;
    a.vstart = h->vstart;
    a.out_ids = ids;
    a.out_dists = dists;
    a.stats = reinterpret_cast<unsigned long long*>(stats);
    a.D = (int)h->D;
    a.EX = (int)h->EX;
    a.k = (int)k;
    a.nq = (int)nq;
    a.K = (int)h->K;
    a.max_probes = max_probes;
    a.bp = nullptr;
    a.exl = h->exl;
    a.exl_w = exl_words((int)h->D, (int)h->EX);
    a.nchunk = bp_nchunk((int)h->D);
    a.waves_per_block = h->plain_waves;
    // HASearcher.hpp:56
    a.one_over_sqrtD = (float)(1.0 / (double)std::sqrt((float)h->D));
    a.qorder = nullptr;
    a.qlist = nullptr;
    a.r0 = a.r1 = 0;
    return a;
}

// The per-query scan.  order_buf (nullable, nq + K + 1 u32): scratch of the query order
// given by the caller instead of the handle's; with it nothing is allocated, and the
// bit-plane kernel runs only if its copies of the codes already exist (else the LUT kernel,
// with identical results).
// The low-latency scan of a small batch (launch_scan_latency): every vector of every probed
// cluster estimated in parallel, then one wave per query replays the probes in order.  Only
// for probe lists without repeated clusters (those of the library's own probe selection):
// a query then has at most nblocks * 32 slots.  The slot buffers are per handle, grow-only.
// Slot and pair bounds of a batch whose queries probe at most max_np clusters each.
static size_t latency_bytes(erq_index* h, size_t nq, size_t max_np, size_t* slots_out) {
    if (h->top_pad.empty()) {  // sums of the i largest cluster sizes rounded up to 32
        std::vector<uint64_t> v(h->sizes.size());
        for (size_t i = 0; i < v.size(); i++) v[i] = (h->sizes[i] + 31) / 32 * 32;
        std::sort(v.begin(), v.end(), std::greater<uint64_t>());
        h->top_pad.assign(v.size() + 1, 0);
        for (size_t i = 0; i < v.size(); i++) h->top_pad[i + 1] = h->top_pad[i] + v[i];
    }
    const size_t slots = nq * h->top_pad[std::min(max_np, h->top_pad.size() - 1)];
    if (slots_out) *slots_out = slots;
    // slots; pair offsets and first items; block minima
    return align_up(slots * 12 + (nq * max_np + 1) * 8, 256) + slots / 32 * 4 + 256;
}

static int scan_latency_locked(erq_index* h, ScanArgs& a, size_t nq, size_t max_np,
                               hipStream_t s) {
    size_t slots = 0;
    const size_t need = latency_bytes(h, nq, max_np, &slots);
    if (need > h->lat_cap) {
        if (h->lat) {
            hipError_t e = hipDeviceSynchronize();  // earlier calls may still read it
            if (e != hipSuccess) return hip_fail(e, "latency buffer sync");
            (void)hipFree(h->lat);
            h->lat = nullptr;
            h->lat_cap = 0;
        }
        hipError_t e = hipMalloc(&h->lat, need);
        if (e != hipSuccess) return hip_fail(e, "latency buffer");
        h->lat_cap = need;
    }
    char* p = static_cast<char*>(h->lat);
    a.est_lower = reinterpret_cast<float*>(p);
    a.est_exd = a.est_lower + slots;
    a.est_id = reinterpret_cast<uint32_t*>(a.est_exd + slots);
    a.est_off = a.est_id + slots;
    a.est_bmin = reinterpret_cast<float*>(p + align_up(slots * 12 + (nq * max_np + 1) * 8, 256));
    StageTimer tim(h, ERQ_STAGE_REPLAY, s);
    hipError_t e = launch_scan_latency(a, s);
    if (e != hipSuccess) return hip_fail(e, "latency scan");
    return ERQ_OK;
}

static int scan_locked(erq_index* h, const float* q_rot_dev, size_t nq, size_t k,
                       const uint32_t* off, const uint32_t* pid, const float* pd, uint32_t* ids,
                       float* dists, uint64_t* stats, hipStream_t s, int max_probes = 0,
                       uint32_t* order_buf = nullptr, size_t own_lists = 0) {
    ScanArgs a = scan_args(h, q_rot_dev, nq, k, off, pid, pd, ids, dists, stats, max_probes);
    const bool nonha = h->factors != 1;  // Searcher's format: the kernels of erq_nonha.hip
    // auto or MFMA: bit-plane integer path (the non-HA format's up to 4 chunks, where its u16
    // half-sums cannot wrap)
    if (h->plain_kernel != 1 && (!nonha || a.nchunk <= 4)) {
        if (!order_buf) {
            int rc = ensure_bp(h, s);
            if (rc) return rc;
        }
        if (h->bp && h->exl) {
            a.bp = h->bp;
            a.exl = h->exl;
        }
    }
    // own_lists > 0: the probe lists come from the library's selection (no repeated
    // cluster), at most own_lists probes per query
    if (own_lists > 0 && !order_buf && !nonha && a.bp && a.nchunk <= 4 && h->latency_nq > 0 &&
        nq <= (size_t)h->latency_nq &&
        (!capturing(s) || latency_bytes(h, nq, own_lists, nullptr) <= h->lat_cap))
        return scan_latency_locked(h, a, nq, own_lists, s);
    const bool order = h->plain_order && nq > 1;
    const bool redo = nonha && a.bp;  // the non-HA MFMA scan's hand-back list
    if (order || redo) {
        // queries by scan cost, heaviest first: the result of each query is unchanged;
        // after them the hand-back list (nq + 1)
        const size_t need = nq + h->K + 1 + (redo ? nq + 1 : 0);
        if (!order_buf && need > h->order_cap) {
            if (h->order) {
                hipError_t e = hipDeviceSynchronize();  // earlier calls may still read it
                if (e != hipSuccess) return hip_fail(e, "order sync");
                (void)hipFree(h->order);
                h->order = nullptr;
                h->order_cap = 0;
            }
            const size_t want = need + need / 4;
            hipError_t e = hipMalloc(&h->order, want * 4);
            if (e != hipSuccess) return hip_fail(e, "order alloc");
            h->order_cap = want;
        }
        // cost bins: a query scans at most every block (max_probes > 0: at most that many
        // of the largest clusters); shift so that the bound lands below K
        size_t bound = h->nblocks;
        if (max_probes > 0)
            bound = std::min(bound, (size_t)max_probes * ((max_cluster_size(h) + 31) / 32));
        int shift = 0;
        while ((bound >> shift) >= h->K) shift++;
        uint32_t* ob = order_buf ? order_buf : h->order;
        if (order) {
            hipError_t e = launch_query_order(off, pid, h->csize, (int)nq, (int)h->K,
                                              max_probes, shift, ob, ob + h->K + 1, s);
            if (e != hipSuccess) return hip_fail(e, "query order");
            a.qorder = ob + h->K + 1;
        }
        if (redo) a.nonha_redo = ob + h->K + 1 + nq;
    }
    StageTimer tim(h, ERQ_STAGE_REPLAY, s);
    hipError_t e = nonha ? launch_scan_nonha(a, s) : launch_scan(a, s);
    if (e != hipSuccess) return hip_fail(e, "scan");
    return ERQ_OK;
}

int erq_set_option(erq_index* h, int option, long value) {
    return guarded([&]() -> int {
        if (!h) {
            set_error("erq_set_option: null handle");
            return ERQ_EINVAL;
        }
        std::lock_guard<std::recursive_mutex> lk(h->mu);
        if (option == ERQ_OPT_PLAIN_KERNEL && value >= 0 && value <= 2) {
            h->plain_kernel = (int)value;
            return ERQ_OK;
        }
        if (option == ERQ_OPT_PLAIN_WAVES && value >= 1 && value <= 4) {
            h->plain_waves = (int)value;
            return ERQ_OK;
        }
        if (option == ERQ_OPT_BATCH_WG_PER_CU && value >= 1 && value <= 8) {
            h->batch_wg_per_cu = (int)value;
            return ERQ_OK;
        }
        if (option == ERQ_OPT_PLAIN_ORDER && (value == 0 || value == 1)) {
            h->plain_order = (int)value;
            return ERQ_OK;
        }
        if (option == ERQ_OPT_BATCH_SURVIVOR_CAP && value >= 0 && value <= (1L << 30)) {
            h->batch_surv_cap = (size_t)value;
            return ERQ_OK;
        }
        if (option == ERQ_OPT_BATCH_PAIR_BUDGET && value >= 1024 && value <= (1L << 26)) {
            h->batch_pair_budget = (size_t)value;
            return ERQ_OK;
        }
        if (option == ERQ_OPT_BATCH_OVERLAP && (value == 0 || value == 1)) {
            h->batch_overlap = (int)value;
            return ERQ_OK;
        }
        if (option == ERQ_OPT_BATCH_ITEM_TILES && (value == 8 || value == 16)) {
            h->batch_item_tiles = (int)value;
            return ERQ_OK;
        }
        if (option == ERQ_OPT_PREP_WG_PER_CU && value >= 1 && value <= 6) {
            h->prep_wg_per_cu = (int)value;
            return ERQ_OK;
        }
        if (option == ERQ_OPT_LATENCY_NQ && value >= 0 && value <= 4096) {
            h->latency_nq = (int)value;
            return ERQ_OK;
        }
        set_error("erq_set_option: unknown option or value");
        return ERQ_EINVAL;
    });
}

long erq_get_option(const erq_index* h, int option) {
    if (!h) return -1;
    switch (option) {
    case ERQ_OPT_PLAIN_KERNEL: return h->plain_kernel;
    case ERQ_OPT_PLAIN_WAVES: return h->plain_waves;
    case ERQ_OPT_BATCH_WG_PER_CU: return h->batch_wg_per_cu;
    case ERQ_OPT_BATCH_SURVIVOR_CAP: return (long)h->batch_surv_cap;
    case ERQ_OPT_PLAIN_ORDER: return h->plain_order;
    case ERQ_OPT_BATCH_PAIR_BUDGET: return (long)h->batch_pair_budget;
    case ERQ_OPT_BATCH_OVERLAP: return h->batch_overlap;
    case ERQ_OPT_LATENCY_NQ: return h->latency_nq;
    case ERQ_OPT_BATCH_ITEM_TILES: return h->batch_item_tiles;
    case ERQ_OPT_PREP_WG_PER_CU: return h->prep_wg_per_cu;
    default: return -1;
    }
}

int erq_scan_dev(erq_index* h, const float* q_rot_dev, size_t nq, size_t k,
                 const uint32_t* probe_off_dev, const uint32_t* probe_ids_dev,
                 const float* probe_sqr_y_dev, uint32_t* ids_dev, float* dists_dev,
                 uint64_t* stats_dev, void* stream) {
    return guarded([&]() -> int {
        if (!h || (nq && (!q_rot_dev || !probe_off_dev || !probe_ids_dev || !probe_sqr_y_dev ||
                          !ids_dev))) {
            set_error("erq_scan_dev: null argument");
            return ERQ_EINVAL;
        }
        if (k == 0 || k > 1024) {
            set_error("erq_scan_dev: k must be in [1, 1024]");
            return ERQ_EINVAL;
        }
        if (nq == 0) return ERQ_OK;
        hipStream_t s = pick(h, stream);
        Call call(h, s);
        return scan_locked(h, q_rot_dev, nq, k, probe_off_dev, probe_ids_dev, probe_sqr_y_dev,
                           ids_dev, dists_dev, stats_dev, s);
    });
}

__global__ static void add_u32_to_u64(const uint32_t* src, uint64_t* dst) { *dst += *src; }

__global__ static void fill_offsets(uint32_t* off, size_t nq, uint32_t nprobe) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i <= nq) off[i] = (uint32_t)(i * nprobe);
}

// Workspace bytes of a probe list of nq queries x nprobe (ids, distances, offsets).
static size_t list_ws_bytes(size_t nq, size_t nprobe) {
    return align_up(nq * nprobe * 4, 256) * 2 + align_up((nq + 1) * 4, 256) + 512;
}

int erq_search_dev(erq_index* h, const float* q_rot_dev, size_t nq, size_t k, size_t nprobe,
                   uint32_t* ids_dev, float* dists_dev, uint64_t* stats_dev, void* stream) {
    return guarded([&]() -> int {
        if (!h || (nq && (!q_rot_dev || !ids_dev))) {
            set_error("erq_search_dev: null argument");
            return ERQ_EINVAL;
        }
        if (k == 0 || k > 1024) {
            set_error("erq_search_dev: k must be in [1, 1024]");
            return ERQ_EINVAL;
        }
        if (nprobe == 0 || nprobe > h->K || nprobe > 2048) {
            set_error("erq_search_dev: nprobe must be in [1, min(K, 2048)]");
            return ERQ_EINVAL;
        }
        if (nq == 0) return ERQ_OK;
        if (nq * nprobe >= 0xFFFFFFFFull) {
            set_error("erq_search_dev: nq * nprobe exceeds 32-bit probe offsets");
            return ERQ_EUNSUPPORTED;
        }
        hipStream_t s = pick(h, stream);
        Call call(h, s);
        // probe lists live in the second part of the workspace; size the whole thing first
        const size_t probe_part = probe_ws_bytes(nq, h->K);
        void* wsp;
        int rc = workspace(h, probe_part + list_ws_bytes(nq, nprobe), &wsp);
        if (rc) return rc;
        Ws ws{(char*)wsp};
        ws.off = probe_part;
        uint32_t* pid = ws.take<uint32_t>(nq * nprobe);
        float* pd = ws.take<float>(nq * nprobe);
        uint32_t* off = ws.take<uint32_t>(nq + 1);
        // up to kProbeSplitNq queries: the selection's last kernel writes the offsets too
        const bool split = nq <= kProbeSplitNq;
        rc = probe_locked(h, q_rot_dev, nq, nprobe, pid, pd, s, split ? off : nullptr);
        if (rc) return rc;
        if (!split) {
            hipLaunchKernelGGL(fill_offsets, dim3((unsigned)((nq + 256) / 256)), dim3(256), 0, s,
                               off, nq, (uint32_t)nprobe);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return hip_fail(e, "offsets");
        }
        return scan_locked(h, q_rot_dev, nq, k, off, pid, pd, ids_dev, dists_dev, stats_dev, s, 0,
                           nullptr, nprobe);
    });
}

// erq_search_dev with every intermediate in the device's scratch array: no allocation.
int erq_search_dev_noalloc(erq_index* h, const float* q_rot_dev, size_t nq, size_t k,
                           size_t nprobe, uint32_t* ids_dev, float* dists_dev,
                           uint64_t* stats_dev, void* stream) {
    return guarded([&]() -> int {
        if (!h || (nq && (!q_rot_dev || !ids_dev))) {
            set_error("erq_search_dev_noalloc: null argument");
            return ERQ_EINVAL;
        }
        if (k == 0 || k > 1024) {
            set_error("erq_search_dev_noalloc: k must be in [1, 1024]");
            return ERQ_EINVAL;
        }
        if (nprobe == 0 || nprobe > h->K || nprobe > 2048) {
            set_error("erq_search_dev_noalloc: nprobe must be in [1, min(K, 2048)]");
            return ERQ_EINVAL;
        }
        if (nq == 0) return ERQ_OK;
        const size_t need = erq_search_noalloc_bytes(h, nq, nprobe);
        if (need > kScratchBytes) {
            set_error("erq_search_dev_noalloc: " + std::to_string(nq) + " queries need " +
                      std::to_string(need) + " bytes of scratch, more than the " +
                      std::to_string(kScratchBytes) + " it has; search fewer at a time");
            return ERQ_EUNSUPPORTED;
        }
        if (h->device < 0 || h->device >= kMaxDevices || !g_scratch_ptr[h->device]) {
            set_error(std::string("erq_search_dev_noalloc: no scratch on this device: ") +
                      (h->device >= 0 && h->device < kMaxDevices
                           ? hipGetErrorString(g_scratch_err[h->device])
                           : "device index out of range"));
            return ERQ_EHIP;
        }
        hipStream_t s = pick(h, stream);
        Call call(h, s);
        // layout: probe selection's approximate distances and norms, the probe lists and
        // offsets, the query-order bins
        char* base = static_cast<char*>(g_scratch_ptr[h->device]);
        const size_t probe_part = probe_ws_bytes(nq, h->K);
        Ws ws{base, probe_part};
        uint32_t* pid = ws.take<uint32_t>(nq * nprobe);
        float* pd = ws.take<float>(nq * nprobe);
        uint32_t* off = ws.take<uint32_t>(nq + 1);
        uint32_t* order = ws.take<uint32_t>(2 * nq + h->K + 2);  // order bins, hand-back list
        int rc = probe_into(h, q_rot_dev, nq, nprobe, pid, pd, s, base);
        if (rc) return rc;
        hipLaunchKernelGGL(fill_offsets, dim3((unsigned)((nq + 256) / 256)), dim3(256), 0, s, off,
                           nq, (uint32_t)nprobe);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "offsets");
        return scan_locked(h, q_rot_dev, nq, k, off, pid, pd, ids_dev, dists_dev, stats_dev, s, 0,
                           order);
    });
}

size_t erq_search_noalloc_bytes(const erq_index* h, size_t nq, size_t nprobe) {
    if (!h) return 0;
    Ws ws{nullptr, probe_ws_bytes(nq, h->K)};
    ws.take<uint32_t>(nq * nprobe);
    ws.take<float>(nq * nprobe);
    ws.take<uint32_t>(nq + 1);
    ws.take<uint32_t>(2 * nq + h->K + 2);
    return ws.off;
}

size_t erq_search_noalloc_capacity(void) { return kScratchBytes; }

// The bit-plane copy of the 1-bit codes and the lane-major copy of the EX-bit codes,
// built on the stream on first use of a kernel that reads them.
static int ensure_bp(erq_index* h, hipStream_t s) {
    if (h->bp && h->exl) return ERQ_OK;
    if (h->bp) {
        (void)hipFree(h->bp);
        h->bp = nullptr;
    }
    const size_t bytes = h->nblocks * (size_t)bp_nchunk((int)h->D) * 32 * 64;
    hipError_t e = hipMalloc(&h->bp, bytes ? bytes : 64);
    if (e != hipSuccess) {
        h->bp = nullptr;
        return hip_fail(e, "bit-plane codes alloc");
    }
    e = launch_relayout(h->short_data, h->nblocks, (int)h->D, h->factors, h->bp, s);
    if (e != hipSuccess) return hip_fail(e, "bit-plane relayout");
    const size_t exl_bytes = h->N * 16 * (size_t)exl_words((int)h->D, (int)h->EX) * 4;
    e = hipMalloc(&h->exl, exl_bytes ? exl_bytes : 64);
    if (e != hipSuccess) {
        h->exl = nullptr;
        return hip_fail(e, "lane-major ex-code alloc");
    }
    e = launch_exl_relayout(h->long_code, h->N, (int)h->D, (int)h->EX, h->exl, s);
    if (e != hipSuccess) return hip_fail(e, "ex-code relayout");
    if (!h->num_cu) {
        e = hipDeviceGetAttribute(&h->num_cu, hipDeviceAttributeMultiprocessorCount, h->device);
        if (e != hipSuccess) return hip_fail(e, "device attribute");
    }
    return ERQ_OK;
}

// Round schedule of the batched search: ranks [0, kExactRanks) by the exact per-query
// replay, then [4,16), [16,64), [64,192) and from there rounds of at most 128 ranks.
constexpr size_t kExactRanks = kFirstRoundRank;
static size_t next_round(size_t r0) { return std::min(4 * r0, r0 + 128); }

constexpr size_t kMaxBatchGrid = 4096;    // workgroups of the persistent batched kernel

// (query, cluster) pairs prepared ahead: the option, capped so the digits fit kDigitBytes.
static size_t pair_budget(const erq_index* h) {
    const size_t Dp = (size_t)bp_nchunk((int)h->D) * 512;
    return std::max<size_t>(1024, std::min(h->batch_pair_budget, kDigitBytes / (2 * Dp)));
}

// Queries per chunk: every round of a chunk is prepared ahead, so the chunk's pairs over all
// ranks past the exact ones must fit the pair budget.
static size_t batch_chunk_nq(size_t nq, size_t max_nprobe, size_t budget) {
    const size_t width = std::max<size_t>(1, max_nprobe - kExactRanks);
    return std::max<size_t>(1, std::min(nq, budget / width));
}

// The batched search's workspace, in take order (sizing with a null base gives the bytes).
struct BatchBufs {
    size_t cq = 0, R = 0, r_end = 0;
    uint32_t *counts, *cursor, *local, *btot, *wgbase, *rinfo, *items, *ovf, *entries, *cand_cnt,
        *cand_i, *surv_count, *redo, *slow;
    float* cand_d;
    int8_t* adig;
    char* qparams;
    uint4* surv;
};

static void batch_layout(const erq_index* h, size_t nq, size_t max_nprobe, int cap, Ws& ws,
                         BatchBufs& b) {
    const size_t K = h->K, Dp = (size_t)bp_nchunk((int)h->D) * 512;
    const size_t maxc = max_cluster_size(h);
    b.cq = batch_chunk_nq(nq, max_nprobe, pair_budget(h));
    b.R = 0;
    size_t pairs = 0, items = 0, max_items = 0;
    for (size_t r0 = kExactRanks; r0 < max_nprobe; r0 = next_round(r0)) {
        const size_t r1 = std::min(next_round(r0), max_nprobe), pr = b.cq * (r1 - r0) + 8;
        const size_t it = batch_items_bound(pr, K, h->N, maxc, h->batch_item_tiles);
        pairs += pr;
        items += it;
        max_items = std::max(max_items, it);
        b.R++;
        b.r_end = r1;
    }
    b.counts = ws.take<uint32_t>(b.R * K);
    b.cursor = ws.take<uint32_t>(b.R * K);
    b.local = ws.take<uint32_t>(2 * b.R * K);
    b.btot = ws.take<uint32_t>(2 * ((b.R * K + 255) / 256));
    b.wgbase = ws.take<uint32_t>((size_t)kCountGroups * b.R * K);
    b.rinfo = ws.take<uint32_t>(4 * (b.R + 1));
    b.items = ws.take<uint32_t>(3 * items);
    b.ovf = ws.take<uint32_t>(2 * max_items);
    b.entries = ws.take<uint32_t>(2 * pairs);
    b.cand_cnt = ws.take<uint32_t>(b.cq);
    b.cand_d = ws.take<float>(b.cq * (size_t)cap);
    b.cand_i = ws.take<uint32_t>(b.cq * (size_t)cap);
    b.redo = ws.take<uint32_t>(b.R * (b.cq + 1));
    b.slow = ws.take<uint32_t>(b.R * (b.cq + 1));
    b.adig = ws.take<int8_t>(pairs * 2 * Dp);
    b.qparams = ws.take<char>(pairs * 48);
    b.surv = ws.take<uint4>(h->batch_surv_cap);
    b.surv_count = ws.take<uint32_t>(kMaxBatchGrid);
}

static int ensure_aux(erq_index* h) {
    hipError_t e = hipSuccess;
    if (!h->aux) e = hipStreamCreateWithFlags(&h->aux, hipStreamNonBlocking);
    if (e == hipSuccess && !h->ev_fork)
        e = hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess && !h->ev_join)
        e = hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming);
    return e == hipSuccess ? ERQ_OK : hip_fail(e, "batched search stream");
}

static size_t batch_ws_bytes(const erq_index* h, size_t nq, size_t max_nprobe, int cap) {
    if (max_nprobe <= kExactRanks) return 256;  // the exact replay alone
    Ws ws{nullptr};
    BatchBufs b;
    batch_layout(h, nq, max_nprobe, cap, ws, b);
    return ws.off + 256;
}

// The batched rounds read the HA block layout (and its bit-plane copy): a non-HA index is
// searched by the per-query scan only.
static int require_ha(const erq_index* h, const char* what) {
    if (h->factors == 1) return ERQ_OK;
    set_error(std::string(what) +
              ": the batched search reads the HA block format; this index is non-HA (4 factors "
              "per vector): use the per-query search");
    return ERQ_EUNSUPPORTED;
}

static int batched_locked(erq_index* h, const float* q_rot_dev, size_t nq, size_t k,
                          const uint32_t* off, const uint32_t* pid, const float* pd,
                          size_t max_nprobe, uint32_t* ids, float* dists, uint64_t* stats,
                          hipStream_t s, char* ws_base, size_t ws_off) {
    if (int rc = require_ha(h, "batched search")) return rc;
    int rc = ensure_bp(h, s);
    if (rc) return rc;
    // ranks [0, kExactRanks): the exact per-query scan of each query's nearest clusters
    if (max_nprobe <= kExactRanks)
        return scan_locked(h, q_rot_dev, nq, k, off, pid, pd, ids, dists, stats, s,
                           (int)kExactRanks);
    const int cap = kCandCap;
    Ws ws{ws_base, ws_off};
    BatchBufs b;
    batch_layout(h, nq, max_nprobe, cap, ws, b);
    if (b.R > (size_t)kMaxRounds) {
        set_error("batched search: too many rounds");
        return ERQ_EUNSUPPORTED;
    }
    rc = ensure_aux(h);
    if (rc) return rc;
    const size_t cq = b.cq;
    const int grid = (int)std::min<size_t>(
        kMaxBatchGrid, (size_t)std::max(1, h->num_cu) * std::max(1, h->batch_wg_per_cu));
    // a multiple of 8 workgroups (the preparation's XCD-aware split); 6 fit per CU (LDS)
    const int prep_grid = h->prep_wg_per_cu * ((std::max(1, h->num_cu) + 7) / 8 * 8);
    std::vector<size_t> rlo, rhi;  // each round's probe ranks
    for (size_t r0 = kExactRanks; r0 < max_nprobe; r0 = next_round(r0)) {
        rlo.push_back(r0);
        rhi.push_back(std::min(next_round(r0), max_nprobe));
    }
    BatchArgs ba;
    RoundBufs rb;
    size_t cur_q0 = 0, cur_n = 0;
    auto setup = [&](size_t q0, size_t n) {
        cur_q0 = q0;
        cur_n = n;
        ba.q_rot = q_rot_dev + q0 * h->D;
        ba.probe_off = off + q0;
        ba.probe_sqr_y = pd;
        ba.probe_ids = pid;
        ba.centroids = h->centroids;
        ba.short_data = h->short_data;
        ba.bp = h->bp;
        ba.long_code = h->long_code;
        ba.exl = h->exl;
        ba.exl_w = exl_words((int)h->D, (int)h->EX);
        ba.ex_factor = h->ex_factor;
        ba.ids = h->ids;
        ba.csize = h->csize;
        ba.bstart = h->bstart;
        ba.vstart = h->vstart;
        ba.items = b.items;
        ba.ovf = b.ovf;
        ba.entries = b.entries;
        ba.n_items = b.rinfo;
        ba.adig = b.adig;
        ba.qparams = b.qparams;
        ba.pool_d = dists + q0 * k;
        ba.cand_cnt = b.cand_cnt;
        ba.cand_d = b.cand_d;
        ba.cand_i = b.cand_i;
        ba.surv = b.surv;
        ba.surv_count = b.surv_count;
        ba.surv_cap = h->batch_surv_cap;
        ba.stats = reinterpret_cast<unsigned long long*>(stats);
        ba.redo = b.redo;
        ba.slow = b.slow;
        ba.cand_cap = cap;
        ba.item_tiles = h->batch_item_tiles;
        ba.num_cu = h->num_cu;
        ba.D = (int)h->D;
        ba.EX = (int)h->EX;
        ba.k = (int)k;
        ba.nq = (int)n;
        ba.nchunk = bp_nchunk((int)h->D);
        ba.one_over_sqrtD = (float)(1.0 / (double)std::sqrt((float)h->D));
        rb = RoundBufs{pid, b.counts, b.cursor, b.local, b.btot, b.wgbase, ids + q0 * k, (int)h->K};
    };
    auto rounds = [&]() -> hipError_t {
        // the round-mode rescan of the queries whose candidates overflowed (none, normally:
        // its waves read an empty list and exit)
        ScanArgs sa = scan_args(h, q_rot_dev + cur_q0 * h->D, cur_n, k, off + cur_q0, pid, pd,
                                ids + cur_q0 * k, dists + cur_q0 * k, nullptr, 0);
        sa.bp = h->bp;
        sa.exl = h->exl;
        static const int kStages[3] = {ERQ_STAGE_SCAN, ERQ_STAGE_REEST, ERQ_STAGE_MERGE};
        for (size_t r = 0; r < b.R; r++) {
            BatchArgs br = ba;
            br.n_items = b.rinfo + 4 * r;
            br.redo = b.redo + r * (cq + 1);
            br.slow = b.slow + r * (cq + 1);
            for (int st = 0; st < 3; st++) {
                StageTimer tim(h, kStages[st], s);
                hipError_t e = launch_round_stage(br, rb, grid, st, s);
                if (e != hipSuccess) return e;
            }
            sa.qlist = br.redo;
            sa.r0 = (int)rlo[r];
            sa.r1 = (int)rhi[r];
            StageTimer tim(h, ERQ_STAGE_REDO, s);
            hipError_t e = launch_scan_round(sa, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    // the first chunk's rounds are all prepared on the second stream while the exact ranks
    // of every query run on the caller's; joined before the first round's scan (or, without
    // the overlap, prepared on the caller's stream first)
    setup(0, std::min(cq, nq));
    if (!h->batch_overlap) {
        hipError_t e;
        {
            StageTimer tim(h, ERQ_STAGE_PREP, s);
            e = launch_prep_all(ba, rb, (int)b.R, (int)b.r_end, prep_grid, s);
        }
        if (e != hipSuccess) return hip_fail(e, "batched round preparation");
        rc = scan_locked(h, q_rot_dev, nq, k, off, pid, pd, ids, dists, stats, s,
                         (int)kExactRanks);
        if (rc) return rc;
        for (size_t q0 = 0; q0 < nq; q0 += cq) {
            const size_t n = std::min(cq, nq - q0);
            if (q0) {
                setup(q0, n);
                StageTimer tim(h, ERQ_STAGE_PREP, s);
                e = launch_prep_all(ba, rb, (int)b.R, (int)b.r_end, prep_grid, s);
                if (e != hipSuccess) return hip_fail(e, "batched round preparation");
            }
            e = launch_zero(b.cand_cnt, n, s);
            if (e == hipSuccess) e = launch_zero(b.redo, b.R * (cq + 1), s);
            if (e == hipSuccess) e = launch_zero(b.slow, b.R * (cq + 1), s);
            if (e == hipSuccess) e = rounds();
            if (e != hipSuccess) return hip_fail(e, "batched round");
        }
        return ERQ_OK;
    }
    hipError_t e = hipEventRecord(h->ev_fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(h->aux, h->ev_fork, 0);
    if (e == hipSuccess) {
        StageTimer tim(h, ERQ_STAGE_PREP, h->aux);
        e = launch_prep_all(ba, rb, (int)b.R, (int)b.r_end, prep_grid, h->aux);
    }
    hipError_t ej = hipEventRecord(h->ev_join, h->aux);
    if (e != hipSuccess || ej != hipSuccess) {
        (void)hipStreamWaitEvent(s, h->ev_join, 0);
        return hip_fail(e != hipSuccess ? e : ej, "batched round preparation");
    }
    rc = scan_locked(h, q_rot_dev, nq, k, off, pid, pd, ids, dists, stats, s, (int)kExactRanks);
    e = hipStreamWaitEvent(s, h->ev_join, 0);  // joined even if the exact ranks failed
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "batched search join");
    for (size_t q0 = 0; q0 < nq; q0 += cq) {
        const size_t n = std::min(cq, nq - q0);
        if (q0) {
            setup(q0, n);
            StageTimer tim(h, ERQ_STAGE_PREP, s);
            e = launch_prep_all(ba, rb, (int)b.R, (int)b.r_end, prep_grid, s);
            if (e != hipSuccess) return hip_fail(e, "batched round preparation");
        }
        e = launch_zero(b.cand_cnt, n, s);
        if (e == hipSuccess) e = launch_zero(b.redo, b.R * (cq + 1), s);
        if (e == hipSuccess) e = launch_zero(b.slow, b.R * (cq + 1), s);
        if (e == hipSuccess) e = rounds();
        if (e != hipSuccess) return hip_fail(e, "batched round");
    }
    return ERQ_OK;
}

int erq_scan_batched_dev(erq_index* h, const float* q_rot_dev, size_t nq, size_t k,
                         const uint32_t* probe_off_dev, const uint32_t* probe_ids_dev,
                         const float* probe_sqr_y_dev, size_t max_nprobe, uint32_t* ids_dev,
                         float* dists_dev, uint64_t* stats_dev, void* stream) {
    return guarded([&]() -> int {
        if (!h || (nq && (!q_rot_dev || !probe_off_dev || !probe_ids_dev || !probe_sqr_y_dev ||
                          !ids_dev || !dists_dev))) {
            set_error("erq_scan_batched_dev: null argument (dists_dev is required)");
            return ERQ_EINVAL;
        }
        if (k == 0 || k > 1024) {
            set_error("erq_scan_batched_dev: k must be in [1, 1024]");
            return ERQ_EINVAL;
        }
        if (int rc = require_ha(h, "erq_scan_batched_dev")) return rc;
        if (nq == 0) return ERQ_OK;
        if (nq * std::max<size_t>(max_nprobe, 1) >= (1u << 27)) {
            set_error("erq_scan_batched_dev: nq * max_nprobe must stay below 2^27");
            return ERQ_EUNSUPPORTED;
        }
        hipStream_t s = pick(h, stream);
        Call call(h, s);
        void* wsp;
        int rc = workspace(h, batch_ws_bytes(h, nq, max_nprobe, kCandCap), &wsp);
        if (rc) return rc;
        return batched_locked(h, q_rot_dev, nq, k, probe_off_dev, probe_ids_dev, probe_sqr_y_dev,
                              max_nprobe, ids_dev, dists_dev, stats_dev, s, (char*)wsp, 0);
    });
}

int erq_search_batched_dev(erq_index* h, const float* q_rot_dev, size_t nq, size_t k,
                           size_t nprobe, uint32_t* ids_dev, float* dists_dev,
                           uint64_t* stats_dev, void* stream) {
    return guarded([&]() -> int {
        if (!h || (nq && (!q_rot_dev || !ids_dev || !dists_dev))) {
            set_error("erq_search_batched_dev: null argument (dists_dev is required)");
            return ERQ_EINVAL;
        }
        if (k == 0 || k > 1024) {
            set_error("erq_search_batched_dev: k must be in [1, 1024]");
            return ERQ_EINVAL;
        }
        if (int rc = require_ha(h, "erq_search_batched_dev")) return rc;
        if (nprobe == 0 || nprobe > h->K || nprobe > 2048) {
            set_error("erq_search_batched_dev: nprobe must be in [1, min(K, 2048)]");
            return ERQ_EINVAL;
        }
        if (nq == 0) return ERQ_OK;
        if (nq * nprobe >= (1u << 27)) {
            set_error("erq_search_batched_dev: nq * nprobe must stay below 2^27");
            return ERQ_EUNSUPPORTED;
        }
        hipStream_t s = pick(h, stream);
        Call call(h, s);
        const size_t probe_part = probe_ws_bytes(nq, h->K);
        const size_t list_part = list_ws_bytes(nq, nprobe);
        const size_t batch_part = batch_ws_bytes(h, nq, nprobe, kCandCap);
        void* wsp;
        int rc = workspace(h, probe_part + list_part + batch_part, &wsp);
        if (rc) return rc;
        Ws ws{(char*)wsp};
        ws.off = probe_part;
        uint32_t* pid = ws.take<uint32_t>(nq * nprobe);
        float* pd = ws.take<float>(nq * nprobe);
        uint32_t* off = ws.take<uint32_t>(nq + 1);
        rc = probe_locked(h, q_rot_dev, nq, nprobe, pid, pd, s);
        if (rc) return rc;
        hipLaunchKernelGGL(fill_offsets, dim3((unsigned)((nq + 256) / 256)), dim3(256), 0, s, off,
                           nq, (uint32_t)nprobe);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "offsets");
        return batched_locked(h, q_rot_dev, nq, k, off, pid, pd, nprobe, ids_dev, dists_dev,
                              stats_dev, s, (char*)wsp, probe_part + list_part);
    });
}

static size_t learned_ws_bytes(size_t nq, size_t K, size_t max_np) {
    return 8 * 256 + align_up(nq * K * 4, 256) + 2 * align_up(nq * max_np * 4, 256) +
           2 * align_up(nq * 4, 256) + 256;
}

static int learned_locked(erq_index* h, erq_mlp* m, const float* q_rot_dev, size_t nq,
                          float mass, size_t min_np, size_t max_np, size_t fb_np, float min_conf,
                          uint32_t* off, uint32_t* ids, float* sqr_y, uint32_t* n_fallback,
                          hipStream_t s, char* base, size_t learned_off) {
    StageTimer tim(h, ERQ_STAGE_SELECT, s);
    Ws ws{base, learned_off};
    float* prob = ws.take<float>(nq * h->K);
    uint32_t* pad_ids = ws.take<uint32_t>(nq * max_np);
    float* pad_d = ws.take<float>(nq * max_np);
    uint32_t* len = ws.take<uint32_t>(nq);
    uint32_t* fb_list = ws.take<uint32_t>(nq);
    uint32_t* fb_count = ws.take<uint32_t>(1);
    int rc = erq_mlp_forward_dev(m, q_rot_dev, nq, prob, s);
    if (rc) return rc;
    hipError_t e = launch_adaptive_probe(prob, q_rot_dev, h->centroids, (int)nq, (int)h->K,
                                         (int)h->D, mass, (int)min_np, (int)max_np, min_conf,
                                         (int)fb_np, pad_ids, pad_d, len, off, ids, sqr_y,
                                         n_fallback, fb_list, fb_count, s);
    if (e != hipSuccess) return hip_fail(e, "learned probe selection");
    return ERQ_OK;
}

static int learned_args_ok(erq_index* h, erq_mlp* m, size_t min_np, size_t max_np, size_t fb_np,
                           const char* who) {
    if (erq_mlp_in_dim(m) != h->D || erq_mlp_out_dim(m) != h->K) {
        set_error(std::string(who) + ": the model must map D inputs to K clusters");
        return ERQ_EINVAL;
    }
    if (max_np == 0 || max_np > h->K || max_np > 2048 || fb_np == 0 || fb_np > max_np ||
        min_np > max_np) {
        set_error(std::string(who) + ": need 1 <= fallback <= max_nprobe <= min(K, 2048), "
                  "min_nprobe <= max_nprobe");
        return ERQ_EINVAL;
    }
    return ERQ_OK;
}

int erq_probe_learned_dev(erq_index* h, erq_mlp* m, const float* q_rot_dev, size_t nq, float mass,
                          size_t min_nprobe, size_t max_nprobe, size_t fallback_nprobe,
                          float min_conf, uint32_t* probe_off_dev, uint32_t* probe_ids_dev,
                          float* probe_sqr_y_dev, uint32_t* n_fallback_dev, void* stream) {
    return guarded([&]() -> int {
        if (!h || !m || (nq && (!q_rot_dev || !probe_off_dev || !probe_ids_dev ||
                                !probe_sqr_y_dev || !n_fallback_dev))) {
            set_error("erq_probe_learned_dev: null argument");
            return ERQ_EINVAL;
        }
        int rc = learned_args_ok(h, m, min_nprobe, max_nprobe, fallback_nprobe,
                                 "erq_probe_learned_dev");
        if (rc || nq == 0) return rc;
        hipStream_t s = pick(h, stream);
        Call call(h, s);
        void* wsp;
        rc = workspace(h, learned_ws_bytes(nq, h->K, max_nprobe), &wsp);
        if (rc) return rc;
        return learned_locked(h, m, q_rot_dev, nq, mass, min_nprobe, max_nprobe,
                              fallback_nprobe, min_conf, probe_off_dev, probe_ids_dev,
                              probe_sqr_y_dev, n_fallback_dev, s, (char*)wsp, 0);
    });
}

int erq_search_learned_dev(erq_index* h, erq_mlp* m, const float* q_rot_dev, size_t nq, size_t k,
                           float mass, size_t min_nprobe, size_t max_nprobe,
                           size_t fallback_nprobe, float min_conf, uint32_t* ids_dev,
                           float* dists_dev, uint64_t* stats_dev, void* stream) {
    return guarded([&]() -> int {
        if (!h || !m || (nq && (!q_rot_dev || !ids_dev || !dists_dev))) {
            set_error("erq_search_learned_dev: null argument (dists_dev is required)");
            return ERQ_EINVAL;
        }
        if (k == 0 || k > 1024) {
            set_error("erq_search_learned_dev: k must be in [1, 1024]");
            return ERQ_EINVAL;
        }
        int rc = learned_args_ok(h, m, min_nprobe, max_nprobe, fallback_nprobe,
                                 "erq_search_learned_dev");
        if (rc || nq == 0) return rc;
        if ((rc = require_ha(h, "erq_search_learned_dev"))) return rc;
        if (nq * max_nprobe >= (1u << 27)) {
            set_error("erq_search_learned_dev: nq * max_nprobe must stay below 2^27");
            return ERQ_EUNSUPPORTED;
        }
        hipStream_t s = pick(h, stream);
        Call call(h, s);
        const size_t learned_part = learned_ws_bytes(nq, h->K, max_nprobe);
        const size_t list_part = list_ws_bytes(nq, max_nprobe) + align_up(8, 256);
        const size_t batch_part = batch_ws_bytes(h, nq, max_nprobe, kCandCap);
        void* wsp;
        rc = workspace(h, learned_part + list_part + batch_part, &wsp);
        if (rc) return rc;
        Ws ws{(char*)wsp, learned_part};
        uint32_t* off = ws.take<uint32_t>(nq + 1);
        uint32_t* pid = ws.take<uint32_t>(nq * max_nprobe);
        float* pd = ws.take<float>(nq * max_nprobe);
        uint32_t* nfb = ws.take<uint32_t>(2);
        hipError_t e = launch_zero(nfb, 2, s);
        if (e != hipSuccess) return hip_fail(e, "zero");
        rc = learned_locked(h, m, q_rot_dev, nq, mass, min_nprobe, max_nprobe, fallback_nprobe,
                            min_conf, off, pid, pd, nfb, s, (char*)wsp, 0);
        if (rc) return rc;
        if (stats_dev) {
            // counter 7: queries that took the fixed-nprobe fallback
            hipLaunchKernelGGL(add_u32_to_u64, dim3(1), dim3(1), 0, s, nfb, stats_dev + 7);
            e = hipGetLastError();
            if (e != hipSuccess) return hip_fail(e, "stats");
        }
        return batched_locked(h, q_rot_dev, nq, k, off, pid, pd, max_nprobe, ids_dev, dists_dev,
                              stats_dev, s, (char*)wsp, learned_part + list_part);
    });
}

// Host-pointer search of nq rotated queries through the per-handle device buffers; probe
// counts per query (np_host, nullable: every query takes `nprobe`).
static int host_search(erq_index* h, const float* q_rot, size_t nq, size_t k, size_t nprobe,
                       const uint32_t* np_host, unsigned flags, uint32_t* results, float* dists) {
    const bool batched = (flags & ERQ_SEARCH_BATCHED) != 0;
    if (batched)
        if (int rc = require_ha(h, "erq_search_batch (ERQ_SEARCH_BATCHED)")) return rc;
    hipStream_t s = h->stream;
    Call call(h, s);
    int rc = host_bufs(h, nq, k);
    if (rc) return rc;
    const size_t in_bytes = nq * h->D * 4, out_bytes = nq * k * 4;
    char* pin = static_cast<char*>(pinned(h, std::max(in_bytes, 2 * out_bytes)));
    if (pin) std::memcpy(pin, q_rot, in_bytes);
    hipError_t e = hipMemcpyAsync(h->hq, pin ? (const void*)pin : (const void*)q_rot, in_bytes,
                                  hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "copy queries");
    if (!np_host) {
        rc = batched ? erq_search_batched_dev(h, h->hq, nq, k, nprobe, h->hids, h->hdists, nullptr, s)
                     : erq_search_dev(h, h->hq, nq, k, nprobe, h->hids, h->hdists, nullptr, s);
    } else {
        // CSR offsets of the per-query probe counts on the host; the probe lists are the
        // max-count lists of every query, cut to its own count
        std::vector<uint32_t> offs(nq + 1);
        size_t max_np = 0, tot = 0;
        for (size_t i = 0; i < nq; i++) {
            offs[i] = (uint32_t)tot;
            tot += np_host[i];
            max_np = std::max<size_t>(max_np, np_host[i]);
        }
        offs[nq] = (uint32_t)tot;
        e = hipMemcpyAsync(h->hnp, offs.data(), (nq + 1) * 4, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "copy probe offsets");
        const size_t probe_part = probe_ws_bytes(nq, h->K);
        const size_t pad_part = list_ws_bytes(nq, max_np);
        const size_t csr_part = list_ws_bytes(tot, 1);
        const size_t batch_part = batched ? batch_ws_bytes(h, nq, max_np, kCandCap) : 0;
        void* wsp;
        rc = workspace(h, probe_part + pad_part + csr_part + batch_part, &wsp);
        if (rc) return rc;
        Ws ws{(char*)wsp, probe_part};
        uint32_t* pad_id = ws.take<uint32_t>(nq * max_np);
        float* pad_d = ws.take<float>(nq * max_np);
        ws.off = probe_part + pad_part;
        uint32_t* cid = ws.take<uint32_t>(tot);
        float* cd = ws.take<float>(tot);
        rc = probe_locked(h, h->hq, nq, max_np, pad_id, pad_d, s);
        if (rc) return rc;
        e = launch_compact_probes(pad_id, pad_d, (int)max_np, h->hnp, (int)nq, cid, cd, s);
        if (e != hipSuccess) return hip_fail(e, "probe lists");
        rc = batched ? batched_locked(h, h->hq, nq, k, h->hnp, cid, cd, max_np, h->hids,
                                      h->hdists, nullptr, s, (char*)wsp,
                                      probe_part + pad_part + csr_part)
                     : scan_locked(h, h->hq, nq, k, h->hnp, cid, cd, h->hids, h->hdists, nullptr,
                                   s, 0, nullptr, max_np);
    }
    if (rc) return rc;
    // the staging buffer is free again: the query copy precedes the kernels on the stream
    e = hipMemcpyAsync(pin ? (void*)pin : (void*)results, h->hids, out_bytes,
                       hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && dists)
        e = hipMemcpyAsync(pin ? (void*)(pin + out_bytes) : (void*)dists, h->hdists, out_bytes,
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "copy results");
    if (pin) {
        std::memcpy(results, pin, out_bytes);
        if (dists) std::memcpy(dists, pin + out_bytes, out_bytes);
    }
    return ERQ_OK;
}

int erq_search_batch(erq_index* h, const float* q_rot, size_t nq, size_t k, size_t nprobe,
                     uint32_t* results, float* dists) {
    return guarded([&]() -> int {
        if (!h || (nq && (!q_rot || !results))) {
            set_error("erq_search_batch: null argument");
            return ERQ_EINVAL;
        }
        if (k == 0 || k > 1024 || nprobe == 0 || nprobe > h->K || nprobe > 2048) {
            set_error("erq_search_batch: need k in [1, 1024], nprobe in [1, min(K, 2048)]");
            return ERQ_EINVAL;
        }
        if (nq == 0) return ERQ_OK;
        return host_search(h, q_rot, nq, k, nprobe, nullptr, 0, results, dists);
    });
}

int erq_search_batch_nprobe(erq_index* h, const float* q_rot, size_t nq, size_t k,
                            const uint32_t* nprobe_per_q, uint32_t* results, float* dists,
                            unsigned flags) {
    return guarded([&]() -> int {
        if (!h || (nq && (!q_rot || !results || !nprobe_per_q))) {
            set_error("erq_search_batch_nprobe: null argument");
            return ERQ_EINVAL;
        }
        if (k == 0 || k > 1024 || (flags & ~(unsigned)ERQ_SEARCH_BATCHED)) {
            set_error("erq_search_batch_nprobe: k must be in [1, 1024], flags ERQ_SEARCH_BATCHED");
            return ERQ_EINVAL;
        }
        size_t tot = 0;
        for (size_t i = 0; i < nq; i++) {
            if (nprobe_per_q[i] == 0 || nprobe_per_q[i] > h->K || nprobe_per_q[i] > 2048) {
                set_error("erq_search_batch_nprobe: every nprobe must be in [1, min(K, 2048)]");
                return ERQ_EINVAL;
            }
            tot += nprobe_per_q[i];
        }
        if (nq == 0) return ERQ_OK;
        if (tot >= (1u << 27)) {
            set_error("erq_search_batch_nprobe: the probes of all queries must stay below 2^27");
            return ERQ_EUNSUPPORTED;
        }
        return host_search(h, q_rot, nq, k, 0, nprobe_per_q, flags, results, dists);
    });
}

int erq_search(erq_index* h, const float* q_rot, size_t k, size_t nprobe, uint32_t* results,
               float* dists) {
    return erq_search_batch(h, q_rot, 1, k, nprobe, results, dists);
}

int erq_rotate(erq_index* h, const float* q, size_t nq, float* q_rot) {
    return guarded([&]() -> int {
        if (!h || (nq && (!q || !q_rot))) {
            set_error("erq_rotate: null argument");
            return ERQ_EINVAL;
        }
        if (nq == 0) return ERQ_OK;
        hipStream_t s = h->stream;
        Call call(h, s);
        // the padded input rows and the output share the host-call query buffer's role: the
        // input goes to the workspace tail, the output to hq
        int rc = host_bufs(h, nq, 1);
        if (rc) return rc;
        void* wsp;
        const size_t pad_part = align_up(nq * h->D * 4, 256);  // erq_rotate_dev pads here
        rc = workspace(h, pad_part + nq * h->DIM * 4, &wsp);
        if (rc) return rc;
        float* dq = reinterpret_cast<float*>((char*)wsp + pad_part);
        hipError_t e = hipMemcpyAsync(dq, q, nq * h->DIM * 4, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "copy queries");
        rc = erq_rotate_dev(h, dq, nq, h->hq, s);
        if (rc) return rc;
        e = hipMemcpyAsync(q_rot, h->hq, nq * h->D * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "copy rotated queries");
        return ERQ_OK;
    });
}

int erq_timing_begin(erq_index* h, size_t max_records) {
    return guarded([&]() -> int {
        if (!h || max_records == 0 || max_records > (1u << 20)) {
            set_error("erq_timing_begin: null handle or max_records outside [1, 2^20]");
            return ERQ_EINVAL;
        }
        std::lock_guard<std::recursive_mutex> lk(h->mu);
        timing_release(h);
        h->tim_ev.assign(2 * max_records, nullptr);
        for (auto& ev : h->tim_ev) {
            hipError_t e = hipEventCreate(&ev);
            if (e != hipSuccess) {
                timing_release(h);
                return hip_fail(e, "erq_timing_begin");
            }
        }
        h->tim_stage.reserve(max_records);
        h->tim_on = true;
        return ERQ_OK;
    });
}

int erq_timing_end(erq_index* h, double* ms_per_stage, uint64_t* count_per_stage) {
    return guarded([&]() -> int {
        if (!h || !ms_per_stage) {
            set_error("erq_timing_end: null argument");
            return ERQ_EINVAL;
        }
        std::lock_guard<std::recursive_mutex> lk(h->mu);
        for (int i = 0; i < ERQ_NUM_STAGES; i++) {
            ms_per_stage[i] = 0.0;
            if (count_per_stage) count_per_stage[i] = 0;
        }
        int rc = ERQ_OK;
        for (size_t r = 0; r < h->tim_stage.size() && !rc; r++) {
            float ms = 0.f;
            hipError_t e = hipEventSynchronize(h->tim_ev[2 * r + 1]);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, h->tim_ev[2 * r], h->tim_ev[2 * r + 1]);
            if (e != hipSuccess) {
                rc = hip_fail(e, "erq_timing_end");
                break;
            }
            const int st = h->tim_stage[r];
            ms_per_stage[st] += ms;
            if (count_per_stage) count_per_stage[st]++;
        }
        timing_release(h);
        return rc;
    });
}

// Times the result-invariant launch shapes on the caller's queries (HIP events on the
// caller's stream) and keeps the fastest: waves (queries) per workgroup of the per-query
// scan; workgroups per CU of the persistent batched kernel, tiles per work item and the
// preparation's workgroups per CU.  Holds the handle for the whole sweep, so no other call
// sees an intermediate setting.
int erq_autotune(erq_index* h, const float* q_rot_dev, size_t nq, size_t k, size_t nprobe,
                 float* best_ms_out, void* stream) {
    return guarded([&]() -> int {
        if (!h || !q_rot_dev || nq == 0) {
            set_error("erq_autotune: null argument or empty batch");
            return ERQ_EINVAL;
        }
        hipStream_t s = pick(h, stream);
        Call call(h, s);
        uint32_t* ids = nullptr;
        float* d = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipMalloc(&ids, nq * k * 4);
        if (e == hipSuccess) e = hipMalloc(&d, nq * k * 4);
        if (e == hipSuccess) e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        int rc = e == hipSuccess ? ERQ_OK : hip_fail(e, "erq_autotune setup");
        auto time_it = [&](bool batched, float* ms) -> int {
            int r = batched
                        ? erq_search_batched_dev(h, q_rot_dev, nq, k, nprobe, ids, d, nullptr, s)
                        : erq_search_dev(h, q_rot_dev, nq, k, nprobe, ids, d, nullptr, s);
            if (r) return r;
            hipError_t x = hipEventRecord(e0, s);
            if (x == hipSuccess) {
                r = batched
                        ? erq_search_batched_dev(h, q_rot_dev, nq, k, nprobe, ids, d, nullptr, s)
                        : erq_search_dev(h, q_rot_dev, nq, k, nprobe, ids, d, nullptr, s);
                if (r) return r;
                x = hipEventRecord(e1, s);
            }
            if (x == hipSuccess) x = hipEventSynchronize(e1);
            if (x == hipSuccess) x = hipEventElapsedTime(ms, e0, e1);
            return x == hipSuccess ? ERQ_OK : hip_fail(x, "erq_autotune timing");
        };
        float best_plain = 0, best_batch = 0;
        int bw = h->plain_waves;
        for (int wv = 1; wv <= 4 && !rc; wv *= 2) {
            h->plain_waves = wv;
            float ms;
            rc = time_it(false, &ms);
            if (!rc && (best_plain == 0 || ms < best_plain)) {
                best_plain = ms;
                bw = wv;
            }
        }
        h->plain_waves = bw;
        // the batched search: scan workgroups per CU, then the work items' tiles, then the
        // preparation's workgroups per CU, each kept at its fastest before the next is swept
        auto sweep = [&](int* knob, const int* vals, int nv) {
            int best = *knob;
            float best_ms = 0;
            for (int i = 0; i < nv && !rc && h->factors == 1; i++) {
                *knob = vals[i];
                float ms;
                rc = time_it(true, &ms);
                if (!rc && (best_ms == 0 || ms < best_ms)) {
                    best_ms = ms;
                    best = vals[i];
                }
            }
            *knob = best;
            if (best_ms > 0) best_batch = best_ms;
        };
        static const int kWg[] = {1, 2, 3, 4}, kTiles[] = {16, 8}, kPrep[] = {6, 4, 2};
        sweep(&h->batch_wg_per_cu, kWg, 4);
        sweep(&h->batch_item_tiles, kTiles, 2);
        sweep(&h->prep_wg_per_cu, kPrep, 3);
        if (best_ms_out) {
            best_ms_out[0] = best_plain;
            best_ms_out[1] = best_batch;
        }
        if (ids) (void)hipFree(ids);
        if (d) (void)hipFree(d);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        return rc;
    });
}

}  // extern "C"
