// Compiles the reference's test_search call pattern (src/test_search.cpp:83-134) against the
// drop-in shim, loads an index file given on the command line and runs one search per query.
// Usage: test_search_dropin <index> <queries.f32 (nq x DIM)> <nq> <k> <nprobe> <out.u32>
//                           [<timing.json> [<rounds>]]
// With a timing file: the queries are searched `rounds` times (default 3), each call timed
// on its own as test_search.cpp times IVF::search, and the file receives the rotation time,
// each round's summed search time and QPS = nq / (summed search time + rotation time), and
// the per-call median and mean of the last round (microseconds).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ivf_gpu.hpp"

static double micros_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0)
        .count();
}

int main(int argc, char** argv) {
    if (argc < 7 || argc > 9) {
        std::fprintf(stderr, "usage: %s index queries nq k nprobe out [timing.json [rounds]]\n",
                     argv[0]);
        return 2;
    }
    erq_host::IVF ivf;
    try {
        ivf.load(argv[1]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 3;
    }
    size_t nq = std::strtoul(argv[3], nullptr, 10), k = std::strtoul(argv[4], nullptr, 10);
    size_t nprobe = std::strtoul(argv[5], nullptr, 10);
    const char* timing = argc > 7 ? argv[7] : nullptr;
    const size_t rounds = argc > 8 ? std::strtoul(argv[8], nullptr, 10) : (timing ? 3 : 1);
    size_t D = ivf.padded_dim();
    size_t DIM = erq_dim(ivf.handle());
    std::vector<float> q(nq * DIM), padded(nq * D, 0.f), rotated(nq * D);
    FILE* f = std::fopen(argv[2], "rb");
    if (!f || std::fread(q.data(), 4, q.size(), f) != q.size()) return 4;
    std::fclose(f);
    for (size_t i = 0; i < nq; i++)
        for (size_t j = 0; j < DIM; j++) padded[i * D + j] = q[i * DIM + j];
    auto t0 = std::chrono::steady_clock::now();
    ivf.rotator().rotate(padded.data(), nq, rotated.data());
    const double rotate_us = micros_since(t0);
    std::vector<erq_host::PID> results(nq * k);
    std::vector<double> per(nq), total(rounds, 0.0);
    for (size_t r = 0; r < rounds; r++)
        for (size_t i = 0; i < nq; i++) {
            t0 = std::chrono::steady_clock::now();
            ivf.search(&rotated[i * D], nullptr, k, nprobe, &results[i * k]);
            per[i] = micros_since(t0);
            total[r] += per[i];
        }
    FILE* o = std::fopen(argv[6], "wb");
    if (!o) return 5;
    std::fwrite(results.data(), 4, results.size(), o);
    std::fclose(o);
    if (timing) {
        std::vector<double> s = per;
        std::sort(s.begin(), s.end());
        double mean = 0.0;
        for (double v : per) mean += v;
        mean /= (double)std::max<size_t>(nq, 1);
        FILE* t = std::fopen(timing, "w");
        if (!t) return 6;
        std::fprintf(t, "{\"nq\": %zu, \"nprobe\": %zu, \"k\": %zu, \"rotate_us\": %.1f, \"rounds\": [",
                     nq, nprobe, k, rotate_us);
        for (size_t r = 0; r < rounds; r++)
            std::fprintf(t, "%s{\"search_us\": %.1f, \"qps\": %.1f}", r ? ", " : "", total[r],
                         (double)nq / ((total[r] + rotate_us) / 1e6));
        std::fprintf(t, "], \"median_us\": %.1f, \"mean_us\": %.1f}\n", nq ? s[nq / 2] : 0.0,
                     mean);
        std::fclose(t);
    }
    return 0;
}
