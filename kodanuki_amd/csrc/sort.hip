// Key-value radix sort of u32 pairs on the device (rocPRIM's device radix
// sort, called directly): the ray queues ordered by direction octant and
// origin cell before a march launch, so that a wave's lanes walk nearby
// cells the same way. And rocPRIM's exclusive scan of u64 counts: the dense
// numbering of a count pass's paths.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

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

bool scanU64(const uint64_t* in, uint64_t* out, size_t n, DevBuf& tmp, hipStream_t s, std::string& err) {
    size_t bytes = 0;
    if (rocprim::exclusive_scan(nullptr, bytes, in, out, (uint64_t)0, n, rocprim::plus<uint64_t>(), s) != hipSuccess) {
        err = "scan: temporary size query failed";
        return false;
    }
    if (!ensure(tmp, bytes, err, false)) return false;
    if (rocprim::exclusive_scan(tmp.ptr, bytes, in, out, (uint64_t)0, n, rocprim::plus<uint64_t>(), s) != hipSuccess) {
        err = "scan failed";
        return false;
    }
    return true;
}

}  // namespace kdn
