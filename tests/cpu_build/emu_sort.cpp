// The key-value sort on the CPU for the bounds-checking build (the device
// sort is rocPRIM's, which the one-thread-at-a-time launch cannot run): a
// stable sort by the key's low endBit bits, as a radix sort orders them.
#include <algorithm>
#include <numeric>
#include <vector>

#include "ctx.h"

namespace kdn {

bool sortPairsU32(uint32_t* keys, uint32_t* keysAlt, uint32_t* vals, uint32_t* valsAlt, uint32_t n, int endBit, DevBuf&, hipStream_t,
                  std::string&, uint32_t** sorted) {
    const uint32_t mask = endBit >= 32 ? 0xFFFFFFFFu : ((1u << endBit) - 1);
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return (keys[a] & mask) < (keys[b] & mask); });
    for (uint32_t i = 0; i < n; ++i) {
        keysAlt[i] = keys[order[i]];
        valsAlt[i] = vals[order[i]];
    }
    *sorted = valsAlt;
    return true;
}

bool scanU64(const uint64_t* in, uint64_t* out, size_t n, DevBuf&, hipStream_t, std::string&) {
    uint64_t acc = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint64_t v = in[i];
        out[i] = acc;
        acc += v;
    }
    return true;
}

}  // namespace kdn
