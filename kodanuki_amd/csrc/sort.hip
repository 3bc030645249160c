// Key-value radix sort of u32 pairs on the device (rocPRIM's device radix
// sort, called directly): the ray queues ordered by direction octant and
// origin cell before a march launch, so that a wave's lanes walk nearby
// cells the same way.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include "ctx.h"

namespace kdn {

bool sortPairsU32(uint32_t* keys, uint32_t* keysAlt, uint32_t* vals, uint32_t* valsAlt, uint32_t n, int endBit, DevBuf& tmp,
                  hipStream_t s, std::string& err, uint32_t** sorted) {
    rocprim::double_buffer<uint32_t> k(keys, keysAlt), v(vals, valsAlt);
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, k, v, n, 0u, (unsigned int)endBit, s) != hipSuccess) {
        err = "radix sort: temporary size query failed";
        return false;
    }
    if (!ensure(tmp, bytes, err, false)) return false;
    if (rocprim::radix_sort_pairs(tmp.ptr, bytes, k, v, n, 0u, (unsigned int)endBit, s) != hipSuccess) {
        err = "radix sort failed";
        return false;
    }
    *sorted = v.current();
    return true;
}

}  // namespace kdn
