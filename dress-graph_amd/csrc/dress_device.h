// dress_device.h -- device helpers shared by the fit kernels: compensated
// sums, lane-group merges and non-temporal stream loads.
//
// A sum is kept as (s, c): TwoSum adds each term exactly into s with the
// rounding error collected in c, and partial sums of different lanes are
// merged pairwise the same way.  s + c is then the correctly rounded sum
// of the terms whatever order they were added in (libdress gets the same
// value from a sort and a Kahan-Babuska-Neumaier sum, dress.c:51-66).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace dh {
namespace {

__device__ __forceinline__ void two_sum_add(double &s, double &c, double x)
{
    double t = s + x;
    double z = t - s;
    double e = (s - (t - z)) + (x - z);
    s = t;
    c += e;
}

__device__ __forceinline__ void two_sum_merge(double &s, double &c, double s2, double c2)
{
    double t = s + s2;
    double z = t - s;
    double e = (s - (t - z)) + (s2 - z);
    s = t;
    c = (c + c2) + e;
}

// merge over the G lanes of an aligned group (G a power of two, <= 64)
template <int G>
__device__ __forceinline__ void group_reduce(double &s, double &c)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        double s2 = __shfl_xor(s, o, G);
        double c2 = __shfl_xor(c, o, G);
        two_sum_merge(s, c, s2, c2);
    }
}

// merge over the lanes of a wave whose index differs in the bits of
// [LO, 64): lanes l and l ^ o for o = LO, 2 LO, ..., 32 (the lanes that
// share l % LO)
template <int LO>
__device__ __forceinline__ void strided_reduce(double &s, double &c)
{
#pragma unroll
    for (int o = LO; o < 64; o <<= 1) {
        double s2 = __shfl_xor(s, o);
        double c2 = __shfl_xor(c, o);
        two_sum_merge(s, c, s2, c2);
    }
}

template <int LO>
__device__ __forceinline__ double strided_max(double m)
{
#pragma unroll
    for (int o = LO; o < 64; o <<= 1) m = fmax(m, __shfl_xor(m, o));
    return m;
}

template <typename T>
__device__ __forceinline__ T ntl(const T *p) { return __builtin_nontemporal_load(p); }

typedef int v2i __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int2 ntl2(const int2 *p)
{
    v2i v = __builtin_nontemporal_load(reinterpret_cast<const v2i *>(p));
    int2 r;
    r.x = v.x;
    r.y = v.y;
    return r;
}

}  // namespace
}  // namespace dh
