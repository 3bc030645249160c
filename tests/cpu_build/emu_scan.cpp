// The exclusive prefix sum on the CPU for the bounds-checking build (the
// device scan exchanges values across a wave, which a one-thread-at-a-time
// launch cannot run).
#include "ctx.h"

namespace kdn {

bool scanU32(const uint32_t* in, uint32_t* out, uint32_t n, uint64_t* total, DevBuf&, hipStream_t, std::string&) {
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t v = in[i];
        out[i] = (uint32_t)acc;
        acc += v;
    }
    if (total) *total = acc;
    return true;
}

}  // namespace kdn
