// Compiles the reference's test_search call pattern (src/test_search.cpp:83-111) against the
// drop-in shim, loads an index file given on the command line and runs one search per query.
// Usage: test_search_dropin <index> <queries.f32 (nq x DIM)> <nq> <k> <nprobe> <out.u32>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ivf_gpu.hpp"

int main(int argc, char** argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s index queries nq k nprobe out\n", argv[0]);
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
    size_t D = ivf.padded_dim();
    size_t DIM = erq_dim(ivf.handle());
    std::vector<float> q(nq * DIM), padded(nq * D, 0.f), rotated(nq * D);
    FILE* f = std::fopen(argv[2], "rb");
    if (!f || std::fread(q.data(), 4, q.size(), f) != q.size()) return 4;
    std::fclose(f);
    for (size_t i = 0; i < nq; i++)
        for (size_t j = 0; j < DIM; j++) padded[i * D + j] = q[i * DIM + j];
    ivf.rotator().rotate(padded.data(), nq, rotated.data());
    std::vector<erq_host::PID> results(nq * k);
    for (size_t i = 0; i < nq; i++)
        ivf.search(&rotated[i * D], nullptr, k, nprobe, &results[i * k]);
    FILE* o = std::fopen(argv[6], "wb");
    std::fwrite(results.data(), 4, results.size(), o);
    std::fclose(o);
    return 0;
}
