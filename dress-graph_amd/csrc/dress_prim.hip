// dress_prim.hip -- rocPRIM device primitives used by the one-time graph build
// and the histogram assembly (radix sort, exclusive scan).  Kept in a
// translation unit of its own so the template-heavy headers compile once.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>

#include "dress_hip_internal.h"

namespace dh {

bool prim_sort_pairs_u64_i32(const uint64_t *kin, uint64_t *kout, const int *vin, int *vout,
                             size_t n, int end_bit, hipStream_t s)
{
    if (n == 0) return true;
    size_t tmp = 0;
    DH_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, kin, kout, vin, vout, n, 0, end_bit, s));
    void *buf = nullptr;
    if (!dmalloc(&buf, tmp + 16, "radix sort scratch")) return false;
    hipError_t e = rocprim::radix_sort_pairs(buf, tmp, kin, kout, vin, vout, n, 0, end_bit, s);
    dfree(buf);
    DH_CHECK(e);
    return true;
}

bool prim_sort_pairs_i32_u64(const int *kin, int *kout, const uint64_t *vin, uint64_t *vout, size_t n,
                             int end_bit, hipStream_t s)
{
    if (n == 0) return true;
    size_t tmp = 0;
    DH_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, kin, kout, vin, vout, n, 0, end_bit, s));
    void *buf = nullptr;
    if (!dmalloc(&buf, tmp + 16, "radix sort scratch")) return false;
    hipError_t e = rocprim::radix_sort_pairs(buf, tmp, kin, kout, vin, vout, n, 0, end_bit, s);
    dfree(buf);
    DH_CHECK(e);
    return true;
}

bool prim_sort_pairs_u64_u64(const uint64_t *kin, uint64_t *kout, const uint64_t *vin, uint64_t *vout,
                             size_t n, hipStream_t s)
{
    if (n == 0) return true;
    size_t tmp = 0;
    DH_CHECK(rocprim::radix_sort_pairs(nullptr, tmp, kin, kout, vin, vout, n, 0, 64, s));
    void *buf = nullptr;
    if (!dmalloc(&buf, tmp + 16, "radix sort scratch")) return false;
    hipError_t e = rocprim::radix_sort_pairs(buf, tmp, kin, kout, vin, vout, n, 0, 64, s);
    dfree(buf);
    DH_CHECK(e);
    return true;
}

bool prim_sort_u64(const uint64_t *kin, uint64_t *kout, size_t n, hipStream_t s, int end_bit)
{
    if (n == 0) return true;
    size_t tmp = 0;
    DH_CHECK(rocprim::radix_sort_keys(nullptr, tmp, kin, kout, n, 0, end_bit, s));
    void *buf = nullptr;
    if (!dmalloc(&buf, tmp + 16, "radix sort scratch")) return false;
    hipError_t e = rocprim::radix_sort_keys(buf, tmp, kin, kout, n, 0, end_bit, s);
    dfree(buf);
    DH_CHECK(e);
    return true;
}

bool prim_run_length_encode_u64(const uint64_t *in, size_t n, uint64_t *unique, int64_t *counts, int64_t *runs,
                                hipStream_t s, void **keep)
{
    if (keep) *keep = nullptr;
    if (n == 0) return true;
    if (n > 0xffffffffull) {
        set_error("run-length encoding of more than 2^32 keys");
        return false;
    }
    size_t tmp = 0;
    DH_CHECK(rocprim::run_length_encode(nullptr, tmp, in, (unsigned)n, unique, counts, runs, s));
    void *buf = nullptr;
    if (!dmalloc(&buf, tmp + 16, "run-length scratch")) return false;
    hipError_t e = rocprim::run_length_encode(buf, tmp, in, (unsigned)n, unique, counts, runs, s);
    if (keep) *keep = buf;
    else dfree(buf);
    DH_CHECK(e);
    return true;
}

bool prim_exclusive_scan_i64(const int64_t *in, int64_t *out, size_t n, hipStream_t s)
{
    if (n == 0) return true;
    size_t tmp = 0;
    DH_CHECK(rocprim::exclusive_scan(nullptr, tmp, in, out, (int64_t)0, n,
                                     rocprim::plus<int64_t>(), s));
    void *buf = nullptr;
    if (!dmalloc(&buf, tmp + 16, "scan scratch")) return false;
    hipError_t e = rocprim::exclusive_scan(buf, tmp, in, out, (int64_t)0, n,
                                           rocprim::plus<int64_t>(), s);
    dfree(buf);
    DH_CHECK(e);
    return true;
}

bool prim_exclusive_scan_i32_to_i64(const int *in, int64_t *out, size_t n, hipStream_t s)
{
    if (n == 0) return true;
    size_t tmp = 0;
    DH_CHECK(rocprim::exclusive_scan(nullptr, tmp, in, out, (int64_t)0, n,
                                     rocprim::plus<int64_t>(), s));
    void *buf = nullptr;
    if (!dmalloc(&buf, tmp + 16, "scan scratch")) return false;
    hipError_t e = rocprim::exclusive_scan(buf, tmp, in, out, (int64_t)0, n,
                                           rocprim::plus<int64_t>(), s);
    dfree(buf);
    DH_CHECK(e);
    return true;
}

}  // namespace dh
