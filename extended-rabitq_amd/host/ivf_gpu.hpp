// ivf_gpu.hpp -- header-only C++ drop-in for the reference's IVF search interface
// (inc/index/IVF.hpp:26-91) on top of liberq.so's C-ABI (include/erq.h).
//
// Same method names, argument meaning and call pattern as the reference, so that
// src/test_search.cpp-style code compiles against it unchanged apart from the include:
//     IVF ivf; ivf.load(path);
//     ivf.rotator().rotate(padded_query, rotated_query);   // Rotator::rotate
//     ivf.search(&rotated_query(i, 0), data, TOPK, nprobe, results);
// The reference reports failures with assert/abort (IO.hpp:28-31, IVF.hpp:257); this
// shim throws std::runtime_error carrying erq_last_error().
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "erq.h"

namespace erq_host {

using PID = uint32_t;  // defines.hpp:15

inline void check(int rc, const char* what) {
    if (rc != ERQ_OK) throw std::runtime_error(std::string(what) + ": " + erq_last_error());
}

class IVF;

// Rotator::rotate (Rotator.hpp:60-63) for row-major [nq x D] zero-padded queries.
class Rotator {
   public:
    explicit Rotator(IVF* owner) : ivf_(owner) {}
    // A: nq x D padded queries (row-major), RAND_A: nq x D output.
    inline void rotate(const float* A, size_t nq, float* RAND_A) const;
    inline size_t size() const;

   private:
    IVF* ivf_;
};

class IVF {
   public:
    IVF() : rot_(this) {}
    ~IVF() { erq_index_free(h_); }
    IVF(const IVF&) = delete;
    IVF& operator=(const IVF&) = delete;

    // IVF::load (IVF.hpp:254)
    void load(const char* filename) {
        erq_index_free(h_);
        h_ = nullptr;
        check(erq_index_load(filename, &h_), "IVF::load");
    }

    // IVF::search (IVF.hpp:84): query is the rotated, padded D-vector; data is unused
    // (as in the reference); results receives k ids in ascending estimated distance.
    void search(const float* query, const float* /*data*/, size_t k, size_t nprobe,
                PID* results) const {
        check(erq_search(h_, query, k, nprobe, results, nullptr), "IVF::search");
    }

    // Batched form of the same call (one launch for nq queries).
    void search_batch(const float* queries, size_t nq, size_t k, size_t nprobe, PID* results,
                      float* dists = nullptr) const {
        check(erq_search_batch(h_, queries, nq, k, nprobe, results, dists), "IVF::search_batch");
    }

    size_t padded_dim() { return erq_padded_dim(h_); }  // IVF.hpp:86
    Rotator& rotator() { return rot_; }                 // IVF.hpp:88
    size_t k() const { return erq_num_clusters(h_); }   // IVF.hpp:90

    erq_index* handle() const { return h_; }

   private:
    erq_index* h_ = nullptr;
    Rotator rot_;
};

inline size_t Rotator::size() const { return erq_padded_dim(ivf_->handle()); }

inline void Rotator::rotate(const float* A, size_t nq, float* RAND_A) const {
    // erq_rotate takes the un-padded DIM columns; a padded input is compacted first.
    const size_t D = erq_padded_dim(ivf_->handle());
    const size_t DIM = erq_dim(ivf_->handle());
    if (D == DIM) {
        check(erq_rotate(ivf_->handle(), A, nq, RAND_A), "Rotator::rotate");
        return;
    }
    std::vector<float> tmp(nq * DIM);
    for (size_t i = 0; i < nq; i++)
        for (size_t j = 0; j < DIM; j++) tmp[i * DIM + j] = A[i * D + j];
    check(erq_rotate(ivf_->handle(), tmp.data(), nq, RAND_A), "Rotator::rotate");
}

}  // namespace erq_host
