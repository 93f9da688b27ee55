// sg_sort.hip -- device sort of (63-bit Morton key, particle index) pairs
// with rocPRIM: radix sort, and merge sort for work that may be captured in a
// graph.  Kept in its own translation unit: rocPRIM's templates dominate the
// build time.
#include "sg_internal.h"

#include <rocprim/device/device_merge_sort.hpp>
#include <rocprim/device/device_radix_sort.hpp>

namespace sg {

// Up to SMALL_SORT_N pairs the merge sort below is used: its block sort and
// few merge passes take less time there than the radix sort's eight 8-bit
// passes over the 63-bit keys.  Both are stable, so the order is the same.
constexpr int SMALL_SORT_N = 65536;

hipError_t sort_pairs_temp_bytes(int n, size_t *bytes)
{
    size_t radix = 0, merge = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, radix, (const uint64_t *)nullptr,
                                             (uint64_t *)nullptr, (const int *)nullptr,
                                             (int *)nullptr, (unsigned int)n, 0u, 63u,
                                             (hipStream_t)0);
    if (e == hipSuccess) e = merge_sort_pairs_temp_bytes(n < SMALL_SORT_N ? n : SMALL_SORT_N, &merge);
    *bytes = radix > merge ? radix : merge;
    return e;
}

hipError_t sort_pairs(void *tmp, size_t tmp_bytes, const uint64_t *kin, uint64_t *kout,
                      const int *vin, int *vout, int n, hipStream_t s)
{
    if (n <= SMALL_SORT_N) return merge_sort_pairs(tmp, tmp_bytes, kin, kout, vin, vout, n, s);
    size_t bytes = tmp_bytes;
    return rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, (unsigned int)n, 0u, 63u,
                                     s);
}

// The same stable order from rocPRIM's merge sort, which issues only
// kernels (the radix sort clears its counters with hipMemsetAsync, which
// a caller's stream capture would record as memset nodes).  Keys are below
// 2^63, so comparing all 64 bits orders them as the radix sort's 63 do.
hipError_t merge_sort_pairs_temp_bytes(int n, size_t *bytes)
{
    *bytes = 0;
    return rocprim::merge_sort(nullptr, *bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                               (const int *)nullptr, (int *)nullptr, (size_t)n,
                               rocprim::less<uint64_t>(), (hipStream_t)0);
}

hipError_t merge_sort_pairs(void *tmp, size_t tmp_bytes, const uint64_t *kin, uint64_t *kout,
                            const int *vin, int *vout, int n, hipStream_t s)
{
    size_t bytes = tmp_bytes;
    return rocprim::merge_sort(tmp, bytes, kin, kout, vin, vout, (size_t)n,
                               rocprim::less<uint64_t>(), s);
}

}  // namespace sg
