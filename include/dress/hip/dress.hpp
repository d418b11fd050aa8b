// dress/hip/dress.hpp -- the HIP selector of libdress++.
//
// Include this instead of dress/cuda/dress.hpp (libdress++/include/dress/
// cuda/dress.hpp:24-59) to run fit / deltaFit / nablaFit on an MI355X:
//
//   #include "dress/hip/dress.hpp"
//   dress::hip::DRESS g(N, U, V);            // the CPU class's constructors
//   auto [iters, delta] = g.fit(100, 1e-6);  // on the GPU
//
// Everything else (construction, get(), accessors, loadFitResult) is the
// CPU class's.  Result types are the CPU class's, so code written against
// dress::DRESS switches by include.  Unlike the CUDA selector, a failure of
// the GPU path throws std::runtime_error with dress_hip_last_error().
//
// Needs libdress++'s dress/dress.hpp on the include path and links against
// libdress_hip.so.
#ifndef DRESS_HIP_HPP
#define DRESS_HIP_HPP

#include "dress/dress.hpp"
#include "dress/hip/dress_hip.h"

#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace dress {
namespace hip {

namespace detail {

inline void check(const char *what)
{
    const char *e = ::dress_hip_last_error();
    if (e && *e) throw std::runtime_error(std::string(what) + ": " + e);
}

// (value, count) pairs and row-major multisets out of the malloc'd C results
inline void take(dress_hist_pair_t *h, int pairs, double *ms, int64_t rows, int E,
                 std::vector<std::pair<double, int64_t>> &hist, std::vector<double> &multisets)
{
    if (h) {
        hist.reserve(pairs > 0 ? (size_t)pairs : 0);
        for (int i = 0; i < pairs; i++) hist.emplace_back(h[i].value, h[i].count);
        std::free(h);
    }
    if (ms) {
        multisets.assign(ms, ms + (size_t)rows * (size_t)E);
        std::free(ms);
    }
}

}  // namespace detail

class DRESS : public dress::DRESS {
public:
    using dress::DRESS::DRESS;

    // dress_fit on the GPU (dress_fit_hip; FP64 unless the precision was
    // switched with dress_hip_set_precision).
    FitResult fit(int maxIterations, double epsilon)
    {
        if (!raw()) throw std::logic_error("DRESS: accessing a moved-from or null graph");
        int iters = 0;
        double d = 0.0;
        ::dress_fit_hip(raw(), maxIterations, epsilon, &iters, &d);
        detail::check("dress_fit_hip");
        return {iters, d};
    }

    // Delta^k on the GPU (dress_delta_fit_hip_strided, every subgraph).
    DeltaFitResult deltaFit(int k, int maxIterations, double epsilon, int nSamples = 0, unsigned int seed = 0,
                            bool keepMultisets = false, bool computeHistogram = true)
    {
        if (!raw()) throw std::logic_error("DRESS: accessing a moved-from or null graph");
        int pairs = 0;
        double *ms = nullptr;
        int64_t rows = 0;
        dress_hist_pair_t *h = ::dress_delta_fit_hip_strided(raw(), k, maxIterations, epsilon, nSamples, seed,
                                                             computeHistogram ? &pairs : nullptr,
                                                             keepMultisets ? 1 : 0, keepMultisets ? &ms : nullptr,
                                                             &rows, 0, 1);
        DeltaFitResult r;
        detail::take(h, pairs, ms, rows, raw()->E, r.histogram, r.multisets);
        detail::check("dress_delta_fit_hip_strided");
        r.num_subgraphs = rows;
        return r;
    }

    // Nabla^k on the GPU (dress_nabla_fit_hip).
    NablaFitResult nablaFit(int k, int maxIterations, double epsilon, int nSamples = 0, unsigned int seed = 0,
                            bool keepMultisets = false, bool computeHistogram = true)
    {
        if (!raw()) throw std::logic_error("DRESS: accessing a moved-from or null graph");
        int pairs = 0;
        double *ms = nullptr;
        int64_t rows = 0;
        dress_hist_pair_t *h = ::dress_nabla_fit_hip(raw(), k, maxIterations, epsilon, nSamples, seed,
                                                     computeHistogram ? &pairs : nullptr, keepMultisets ? 1 : 0,
                                                     keepMultisets ? &ms : nullptr, &rows);
        NablaFitResult r;
        detail::take(h, pairs, ms, rows, raw()->E, r.histogram, r.multisets);
        detail::check("dress_nabla_fit_hip");
        r.num_tuples = rows;
        return r;
    }
};

// one-shot forms, as in the CPU and CUDA headers
inline DRESS::FitResult fit(int N, const std::vector<int> &U, const std::vector<int> &V, int maxIterations = 100,
                            double epsilon = 1e-6)
{
    DRESS d(N, U, V);
    return d.fit(maxIterations, epsilon);
}

inline DRESS::DeltaFitResult delta_fit(int N, const std::vector<int> &U, const std::vector<int> &V, int k = 0,
                                       int maxIterations = 100, double epsilon = 1e-6, int nSamples = 0,
                                       unsigned int seed = 0, bool keepMultisets = false,
                                       bool computeHistogram = true)
{
    DRESS d(N, U, V);
    return d.deltaFit(k, maxIterations, epsilon, nSamples, seed, keepMultisets, computeHistogram);
}

inline DRESS::NablaFitResult nabla_fit(int N, const std::vector<int> &U, const std::vector<int> &V, int k = 0,
                                       int maxIterations = 100, double epsilon = 1e-6, int nSamples = 0,
                                       unsigned int seed = 0, bool keepMultisets = false,
                                       bool computeHistogram = true)
{
    DRESS d(N, U, V);
    return d.nablaFit(k, maxIterations, epsilon, nSamples, seed, keepMultisets, computeHistogram);
}

}  // namespace hip
}  // namespace dress

#endif  // DRESS_HIP_HPP
