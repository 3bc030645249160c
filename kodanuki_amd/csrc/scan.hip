// Exclusive prefix sums of u32 counts: the compaction behind every level
// of the octree and every wavefront queue. Three passes: each block of
// 1024 values scans in LDS (a wave-level scan by ballot-free shuffles, then
// the wave totals), the block totals are scanned the same way
// recursively, and the block offsets are added back.
#include <hip/hip_runtime.h>

#include "ctx.h"

namespace kdn {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;  // per thread
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint32_t wave_inclusive(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Scan one tile per block; write the block's total to sums.
__global__ void __launch_bounds__(kScanBlock) k_scan_tile(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* sums) {
    __shared__ uint32_t waves[kScanBlock / 64];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        local += v[k];
    }
    const uint32_t incl = wave_inclusive(local);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) waves[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += waves[w];
    uint32_t run = before + incl - local;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kScanBlock - 1) {
        uint32_t total = 0;
        for (int w = 0; w < kScanBlock / 64; ++w) total += waves[w];
        sums[blockIdx.x] = total;
    }
}

__global__ void __launch_bounds__(kScanBlock) k_scan_add(uint32_t* out, uint32_t n, const uint32_t* offsets) {
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const uint32_t add = offsets[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] += add;
    }
}

static bool scanRec(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* work, uint64_t* total, hipStream_t s, std::string& err) {
    const uint32_t blocks = (n + kScanTile - 1) / kScanTile;
    uint32_t* sums = work;
    uint32_t* sumsOut = work + blocks;
    hipLaunchKernelGGL(k_scan_tile, dim3(blocks), dim3(kScanBlock), 0, s, in, out, n, sums);
    if (blocks == 1) {
        uint32_t t = 0;
        if (hipMemcpyAsync(&t, sums, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
            err = "scan: copy of the total failed";
            return false;
        }
        *total = t;
        return true;
    }
    if (!scanRec(sums, sumsOut, blocks, work + 2 * blocks, total, s, err)) return false;
    hipLaunchKernelGGL(k_scan_add, dim3(blocks), dim3(kScanBlock), 0, s, out, n, sumsOut);
    return true;
}

bool scanU32(const uint32_t* in, uint32_t* out, uint32_t n, uint64_t* total, DevBuf& tmp, hipStream_t s, std::string& err) {
    if (n == 0) {
        *total = 0;
        return true;
    }
    // work space: 2 words per block at each level of the recursion
    size_t words = 0;
    for (uint32_t m = n; m > 1;) {
        m = (m + kScanTile - 1) / kScanTile;
        words += 2 * m + 2;
    }
    words += 4;
    if (!ensure(tmp, words * 4, err, false)) return false;
    return scanRec(in, out, n, (uint32_t*)tmp.ptr, total, s, err);
}

}  // namespace kdn
