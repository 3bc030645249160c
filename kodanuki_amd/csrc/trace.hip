// The extend step: rays in, hits out, one lane per ray.
#include <hip/hip_runtime.h>

#include "ctx.h"
#include "trace_dev.h"

namespace kdn {

__global__ void __launch_bounds__(kTraceBlock, kTraceWaves) k_extend(TreeView tv, const kdn_ray* rays, uint32_t n, kdn_hit* hits) {
    __shared__ float slots[kLaneWords * kTraceBlock];
    __shared__ uint32_t paths[kPathLevels * kTraceBlock];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const LaneRegs lr{SlotRegs{slots + threadIdx.x, kTraceBlock}, PathRegs{paths + threadIdx.x, kTraceBlock}};
    kdn_hit hit;
    trace_ray(tv, rays[i], lr, hit);
    hits[i] = hit;
}

TreeView treeView(const Ctx& c) {
    TreeView tv;
    tv.nodes = (const NodeDev*)c.tree.nodes.ptr;
    tv.entries = (const EntryDev*)c.tree.entries.ptr;
    tv.pool = (const kdn_step*)c.tree.pool.ptr;
    tv.spool = (const SlotStep*)c.tree.spool.ptr;
    tv.mpool = (const SlotStep*)c.tree.mpool.ptr;
    tv.leafMaterial = c.scene.leaf_material;
    for (int k = 0; k < 3; ++k) {
        tv.lower[k] = c.tree.lower[k];
        tv.size[k] = c.tree.size[k];
    }
    tv.depth = c.tree.depth;
    tv.work = nullptr;
    return tv;
}

bool extendRays(Ctx& c, const kdn_ray* rays, uint32_t n, kdn_hit* hits, hipStream_t stream) {
    if (!c.tree.built) {
        c.error = "extend before the octree is built";
        return false;
    }
    if (n == 0) return true;
    hipLaunchKernelGGL(k_extend, dim3((n + kTraceBlock - 1) / kTraceBlock), dim3(kTraceBlock), 0, stream, treeView(c), rays, n, hits);
    KDN_HIP(hipGetLastError());
    return true;
}

}  // namespace kdn
