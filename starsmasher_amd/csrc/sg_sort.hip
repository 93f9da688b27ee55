// sg_sort.hip -- device radix sort of (63-bit Morton key, particle index)
// pairs with rocPRIM.  Kept in its own translation unit: rocPRIM's templates
// dominate the build time.
#include "sg_internal.h"

#include <rocprim/device/device_radix_sort.hpp>

namespace sg {

hipError_t sort_pairs_temp_bytes(int n, size_t *bytes)
{
    *bytes = 0;
    return rocprim::radix_sort_pairs(nullptr, *bytes, (const uint64_t *)nullptr,
                                     (uint64_t *)nullptr, (const int *)nullptr, (int *)nullptr,
                                     (unsigned int)n, 0u, 63u, (hipStream_t)0);
}

hipError_t sort_pairs(void *tmp, size_t tmp_bytes, const uint64_t *kin, uint64_t *kout,
                      const int *vin, int *vout, int n, hipStream_t s)
{
    size_t bytes = tmp_bytes;
    return rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, (unsigned int)n, 0u, 63u,
                                     s);
}

}  // namespace sg
