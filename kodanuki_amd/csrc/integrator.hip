// The wavefront integrator: one pixel pass as queues of paths.
//
//   generate → [extend → shade → shadow] × surfaces → fold
//
// generate starts one path per (pixel, sample) of a chunk; extend traces
// the queued rays through the octree; shade meets lamps and sky, lights
// the vertex toward one lamp (a shadow ray with its MIS-weighted
// contribution) and draws the next direction, appending survivors to the
// next queue; shadow traces the shadow rays and adds the unblocked
// contributions; fold turns each path's four lanes into linear RGB and
// adds the chunk's samples of a pixel, in sample order, to FP64
// accumulators (so the sums do not depend on queue order).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <vector>

#include "ctx.h"
#include "integrator_kernels.h"

namespace kdn {

TreeView treeView(const Ctx& c);

__global__ void k_generate(PassParams P, uint32_t slots) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < slots) generate_one(P, i);
}

__global__ void __launch_bounds__(kTraceBlock, kTraceWaves) k_trace_queue(TreeView tv, const uint32_t* queue, const uint32_t* count,
                                                             const kdn_ray* rays, kdn_hit* hits) {
    __shared__ float slots[kSlots * kTraceBlock];
    __shared__ uint32_t paths[kPathLevels * kTraceBlock];
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= *count) return;
    Dl buf[kMaxRegs];
    const LaneRegs lr{buf, SlotRegs{slots + threadIdx.x, kTraceBlock}, PathRegs{paths + threadIdx.x, kTraceBlock}};
    trace_item(tv, queue[q], rays, hits, lr, nullptr);
}

// The same march counting the FP32 operations of every evaluation (the
// instrumented pass the roofline's algorithmic work comes from).
__global__ void __launch_bounds__(kTraceBlock, kTraceWaves) k_trace_queue_count(TreeView tv, const uint32_t* queue, const uint32_t* count,
                                                                   const kdn_ray* rays, kdn_hit* hits, double* work) {
    __shared__ float slots[kSlots * kTraceBlock];
    __shared__ uint32_t paths[kPathLevels * kTraceBlock];
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= *count) return;
    Dl buf[kMaxRegs];
    const LaneRegs lr{buf, SlotRegs{slots + threadIdx.x, kTraceBlock}, PathRegs{paths + threadIdx.x, kTraceBlock}};
    trace_item(tv, queue[q], rays, hits, lr, work);
}

__global__ void __launch_bounds__(256) k_shade(PassParams P, uint32_t inQueue, uint32_t outQueue) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < P.s.counters[inQueue]) shade_one(P, q, inQueue, outQueue);
}

__global__ void __launch_bounds__(kTraceBlock, kTraceWaves) k_shadow(TreeView tv, PathState s) {
    __shared__ float slots[kSlots * kTraceBlock];
    __shared__ uint32_t paths[kPathLevels * kTraceBlock];
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= s.counters[2]) return;
    Dl buf[kMaxRegs];
    const LaneRegs lr{buf, SlotRegs{slots + threadIdx.x, kTraceBlock}, PathRegs{paths + threadIdx.x, kTraceBlock}};
    shadow_one(tv, s, q, lr);
}

// The paths still in the input queue, each to its end in one lane (the
// wavefront's last bounces carry few paths, and each launch of them costs
// as long as its slowest ray).
__global__ void __launch_bounds__(kTraceBlock, kTraceWaves) k_finish(TreeView tv, TreeView tvShadow, PassParams P, uint32_t inQueue,
                                                                    uint32_t* tally, double* work) {
    __shared__ float slots[kSlots * kTraceBlock];
    __shared__ uint32_t paths[kPathLevels * kTraceBlock];
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P.s.counters[inQueue]) return;
    Dl buf[kMaxRegs];
    const LaneRegs lr{buf, SlotRegs{slots + threadIdx.x, kTraceBlock}, PathRegs{paths + threadIdx.x, kTraceBlock}};
    finish_one(tv, tvShadow, P, q, inQueue, lr, tally, work);
}

__global__ void k_fold(PassParams P, double* accum, double* accumSq, uint32_t* accumN) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    fold_one(P, t, accum, accumSq, accumN);
}

static uint64_t mixed_bits_host(uint64_t v) {
    v += 0x9E3779B97F4A7C15ULL;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
    return v ^ (v >> 31);
}

void freeIntegrator(Ctx& c) {
    if (c.paths) {
        (void)hipFree(c.paths->rays);
        delete c.paths;
        c.paths = nullptr;
    }
    if (c.colour) (void)hipFree(c.colour);
    c.colour = nullptr;
}

static bool allocPaths(Ctx& c, size_t n) {
    if (c.paths && c.paths->capacity >= n) return true;
    if (c.paths) {
        (void)hipFree(c.paths->rays);
        delete c.paths;
        c.paths = nullptr;
    }
    // one allocation, 256-byte aligned pieces
    const size_t pieces[] = {n * sizeof(kdn_ray), n * sizeof(kdn_hit), n * sizeof(kdn_ray), n * 16, n * 16, n * 16, n * 16, n * 16,
                             n * 4, n * 4, n * 4, n * 4, n * 8, n * 4, n * 4, n * 4, 64};
    size_t bytes = 0;
    for (size_t b : pieces) bytes += (b + 255) & ~size_t(255);
    char* base = nullptr;
    if (hipMalloc(&base, bytes) != hipSuccess) {
        c.error = "hipMalloc of the path state failed (" + std::to_string(bytes) + " bytes)";
        return false;
    }
    PathState* s = new PathState();
    size_t at = 0;
    int k = 0;
    auto take = [&]() {
        char* p = base + at;
        at += (pieces[k++] + 255) & ~size_t(255);
        return p;
    };
    s->rays = (kdn_ray*)take();
    s->hits = (kdn_hit*)take();
    s->shadow = (kdn_ray*)take();
    s->srad = (float*)take();
    s->thr = (float*)take();
    s->rad = (float*)take();
    s->nm = (float*)take();
    s->dens = (float*)take();
    s->pdf = (float*)take();
    s->meta = (uint32_t*)take();
    s->pixel = (uint32_t*)take();
    s->sample = (uint32_t*)take();
    s->stream = (uint64_t*)take();
    for (int q = 0; q < 3; ++q) s->queue[q] = (uint32_t*)take();
    s->counters = (uint32_t*)take();
    s->capacity = n;
    c.paths = s;
    return true;
}

bool renderPass(Ctx& c, const kdn_pass_desc& d, kdn_pass_stats* stats) {
    const auto start = std::chrono::steady_clock::now();
    if (!c.integrator_ready) {
        c.error = "kdn_render_pass: scene tables are not on the device";
        return false;
    }
    if (!c.colour) {
        ColourTables t;
        buildColourTables(t);
        KDN_HIP(hipMalloc(&c.colour, sizeof(ColourTables)));
        KDN_HIP(hipMemcpy(c.colour, &t, sizeof(ColourTables), hipMemcpyHostToDevice));
    }
    hipStream_t st = c.stream;
    // accumulators
    const size_t px = (size_t)d.width * d.height;
    if (!c.accum || c.width != d.width || c.height != d.height) {
        if (c.accum) hipFree(c.accum);
        if (c.accum_sq) hipFree(c.accum_sq);
        if (c.accum_n) hipFree(c.accum_n);
        KDN_HIP(hipMalloc(&c.accum, px * 3 * sizeof(double)));
        KDN_HIP(hipMalloc(&c.accum_sq, px * 3 * sizeof(double)));
        KDN_HIP(hipMalloc(&c.accum_n, px * sizeof(uint32_t)));
        c.width = d.width;
        c.height = d.height;
        KDN_HIP(hipMemsetAsync(c.accum, 0, px * 3 * sizeof(double), st));
        KDN_HIP(hipMemsetAsync(c.accum_sq, 0, px * 3 * sizeof(double), st));
        KDN_HIP(hipMemsetAsync(c.accum_n, 0, px * sizeof(uint32_t), st));
    } else if (!(d.flags & 1)) {
        KDN_HIP(hipMemsetAsync(c.accum, 0, px * 3 * sizeof(double), st));
        KDN_HIP(hipMemsetAsync(c.accum_sq, 0, px * 3 * sizeof(double), st));
        KDN_HIP(hipMemsetAsync(c.accum_n, 0, px * sizeof(uint32_t), st));
    }
    const uint32_t tilesX = (d.width + 7) / 8, tilesY = (d.height + 7) / 8;
    const uint32_t plane = tilesX * tilesY * 64;
    const size_t maxPaths = c.max_paths ? c.max_paths : (size_t)1 << 23;
    const uint32_t chunk = (uint32_t)std::max<size_t>(1, std::min<size_t>(d.spp, maxPaths / plane));
    if (!allocPaths(c, (size_t)plane * chunk)) return false;
    PathState* s = c.paths;

    PassParams P;
    P.width = d.width;
    P.height = d.height;
    P.pixels = (uint32_t)px;
    P.seed_bits = 0;
    P.surfaces = d.surfaces;
    P.min_depth = d.min_depth;
    // the lens (core/camera.cpp:9-25)
    const kdn_camera& cam = c.scene.camera;
    double fwd[3], rt[3], upv[3];
    double len = 0;
    for (int k = 0; k < 3; ++k) {
        fwd[k] = (double)cam.target[k] - cam.eye[k];
        len += fwd[k] * fwd[k];
    }
    len = std::sqrt(len);
    for (int k = 0; k < 3; ++k) fwd[k] /= len;
    rt[0] = fwd[1] * cam.up[2] - fwd[2] * cam.up[1];
    rt[1] = fwd[2] * cam.up[0] - fwd[0] * cam.up[2];
    rt[2] = fwd[0] * cam.up[1] - fwd[1] * cam.up[0];
    len = std::sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
    for (int k = 0; k < 3; ++k) rt[k] /= len;
    upv[0] = rt[1] * fwd[2] - rt[2] * fwd[1];
    upv[1] = rt[2] * fwd[0] - rt[0] * fwd[2];
    upv[2] = rt[0] * fwd[1] - rt[1] * fwd[0];
    for (int k = 0; k < 3; ++k) {
        P.eye[k] = cam.eye[k];
        P.forward[k] = (float)fwd[k];
        P.right[k] = (float)rt[k];
        P.up[k] = (float)upv[k];
        P.sky_zenith[k] = c.scene.render.sky_zenith[k];
        P.sky_horizon[k] = c.scene.render.sky_horizon[k];
        P.sky_ground[k] = c.scene.render.sky_ground[k];
    }
    P.reach = (float)std::tan((double)cam.fov / 2);
    P.aspect = (float)((double)d.width / d.height);
    P.n_lamps = c.n_lamps;
    P.n_materials = c.n_materials_dev;
    P.fallback_material = c.n_materials_dev - 1;
    P.lamps = (const LampDev*)c.lamps_dev;
    P.materials = (const MaterialDev*)c.materials_dev;
    P.leaf_material = c.scene.leaf_material;
    P.colour = (const ColourTables*)c.colour;
    P.s = *s;
    P.seed_bits = mixed_bits_host(d.seed);
    const TreeView tv = treeView(c);

    const bool timing = (d.flags & 2) != 0, counting = (d.flags & 4) != 0;
    // kernel timing: an event pair per launch, read after each bounce's sync
    hipEvent_t ev[2];
    if (timing) {
        KDN_HIP(hipEventCreate(&ev[0]));
        KDN_HIP(hipEventCreate(&ev[1]));
    }
    double traceMs = 0, shadeMs = 0, shadowMs = 0, otherMs = 0;
    auto timed = [&](double* acc, auto&& launch) -> bool {
        if (timing) KDN_HIP(hipEventRecord(ev[0], st));
        launch();
        if (timing) {
            KDN_HIP(hipEventRecord(ev[1], st));
            KDN_HIP(hipEventSynchronize(ev[1]));
            float ms = 0;
            KDN_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            *acc += ms;
        }
        return true;
    };
    const bool finishing = (d.flags & 8) == 0;
    double* work = nullptr;
    if (counting) {
        if (!ensure(c.scratch[7], 64, c.error, false)) return false;
        work = (double*)c.scratch[7].ptr;
        KDN_HIP(hipMemsetAsync(work, 0, 24, st));
    }
    TreeView tvShadow = tv;
    tvShadow.work = counting ? work + 1 : nullptr;

    uint64_t segments = 0, shadows = 0;
    uint32_t bounces = 0, launches = 0;
    double finishMs = 0;
    uint32_t shadowLaunches = 0, finishLaunches = 0;
    for (uint32_t first = 0; first < d.spp; first += chunk) {
        const uint32_t n = std::min(chunk, d.spp - first);
        P.first_sample = d.first_sample + first;
        P.chunk = n;
        const uint32_t slots = plane * n;
        KDN_HIP(hipMemsetAsync(s->counters, 0, 64, st));
        timed(&otherMs, [&] { hipLaunchKernelGGL(k_generate, dim3((slots + 255) / 256), dim3(256), 0, st, P, slots); });
        uint32_t in = 0, out = 1;
        for (uint32_t bounce = 0; bounce < d.surfaces; ++bounce) {
            uint32_t count = 0;
            KDN_HIP(hipMemcpyAsync(&count, &s->counters[in], 4, hipMemcpyDeviceToHost, st));
            KDN_HIP(hipStreamSynchronize(st));
            if (count == 0) break;
            if (finishing && (uint64_t)count * kFinishShare <= slots) {
                // the remaining paths to their ends in one launch
                KDN_HIP(hipMemsetAsync(&s->counters[4], 0, 8, st));
                timed(&finishMs, [&] {
                    hipLaunchKernelGGL(k_finish, dim3((count + kTraceBlock - 1) / kTraceBlock), dim3(kTraceBlock), 0, st, tv, tvShadow,
                                       P, in, &s->counters[4], counting ? work + 2 : nullptr);
                });
                uint32_t tally[2] = {0, 0};
                KDN_HIP(hipMemcpyAsync(tally, &s->counters[4], 8, hipMemcpyDeviceToHost, st));
                KDN_HIP(hipStreamSynchronize(st));
                segments += tally[0];
                shadows += tally[1];
                ++bounces;
                ++finishLaunches;
                break;
            }
            segments += count;
            ++bounces;
            ++launches;
            const uint32_t blocks = (count + 255) / 256;
            timed(&traceMs, [&] {
                if (counting) {
                    hipLaunchKernelGGL(k_trace_queue_count, dim3(blocks), dim3(kTraceBlock), 0, st, tv, s->queue[in], &s->counters[in], s->rays,
                                       s->hits, work);
                } else {
                    hipLaunchKernelGGL(k_trace_queue, dim3(blocks), dim3(kTraceBlock), 0, st, tv, s->queue[in], &s->counters[in], s->rays, s->hits);
                }
            });
            KDN_HIP(hipMemsetAsync(&s->counters[out], 0, 4, st));
            KDN_HIP(hipMemsetAsync(&s->counters[2], 0, 4, st));
            timed(&shadeMs, [&] { hipLaunchKernelGGL(k_shade, dim3(blocks), dim3(256), 0, st, P, in, out); });
            uint32_t sc = 0;
            KDN_HIP(hipMemcpyAsync(&sc, &s->counters[2], 4, hipMemcpyDeviceToHost, st));
            KDN_HIP(hipStreamSynchronize(st));
            shadows += sc;
            shadowLaunches += sc ? 1 : 0;
            if (sc) timed(&shadowMs, [&] { hipLaunchKernelGGL(k_shadow, dim3((sc + kTraceBlock - 1) / kTraceBlock), dim3(kTraceBlock), 0, st, tvShadow, *s); });
            std::swap(in, out);
        }
        timed(&otherMs, [&] {
            hipLaunchKernelGGL(k_fold, dim3((plane + 255) / 256), dim3(256), 0, st, P, c.accum, c.accum_sq, c.accum_n);
        });
    }
    KDN_HIP(hipStreamSynchronize(st));
    KDN_HIP(hipGetLastError());
    double counted[3] = {0, 0, 0};
    if (counting) KDN_HIP(hipMemcpy(counted, work, 24, hipMemcpyDeviceToHost));
    if (timing) {
        KDN_HIP(hipEventDestroy(ev[0]));
        KDN_HIP(hipEventDestroy(ev[1]));
    }
    if (stats) {
        *stats = kdn_pass_stats{};
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
        stats->segments = segments;
        stats->shadow_rays = shadows;
        stats->bounces = bounces;
        stats->trace_launches = launches;
        stats->trace_ms = traceMs;
        stats->shade_ms = shadeMs;
        stats->shadow_ms = shadowMs;
        stats->other_ms = otherMs;
        stats->trace_flops = counted[0];
        stats->shadow_flops = counted[1];
        stats->finish_ms = finishMs;
        stats->finish_flops = counted[2];
        stats->shadow_launches = shadowLaunches;
        stats->finish_launches = finishLaunches;
    }
    return true;
}

}  // namespace kdn
