// The wavefront integrator: one pixel pass as queues of paths.
//
//   generate → [extend → shade → shadow] × surfaces → fold
//
// A round runs the paths of one or more slices (samples of one job each:
// a job is a render of one scene at one size) through the same queues.
// generate starts one path per (pixel, sample) of each slice; extend traces
// the queued rays through the octree; shade meets lamps and sky, lights
// the vertex toward one lamp (a shadow ray with its MIS-weighted
// contribution) and draws the next direction, appending survivors to the
// next queue; shadow traces the shadow rays and adds the unblocked
// contributions; fold turns each path's four lanes into linear RGB and
// adds a slice's samples of a pixel, in sample order, to its job's FP64
// accumulators (so the sums do not depend on queue order or on which
// other jobs share the round). Once few paths are left, one launch takes
// each to its end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "ctx.h"
#include "guide_net.h"
#include "integrator_kernels.h"

namespace kdn {

TreeView treeView(const Ctx& c);

__global__ void k_generate(PassParams P) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.paths) generate_one(P, i);
}

#define KDN_LANE_REGS                                              \
    __shared__ float slots[kLaneWords * kTraceBlock];              \
    __shared__ uint32_t paths[kPathLevels * kTraceBlock];          \
    const LaneRegs lr{SlotRegs{slots + threadIdx.x, kTraceBlock}, PathRegs{paths + threadIdx.x, kTraceBlock}}

// (One, here and below: the round has one job, read from the kernel's
// arguments, the same for every lane.)
template <bool One>
__global__ void __launch_bounds__(kTraceBlock, kTraceWaves) k_trace_queue(PassParams P, uint32_t inQueue) {
    KDN_LANE_REGS;
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P.s.counters[inQueue]) return;
    const uint32_t i = P.s.queue[inQueue][q];
    trace_item(job_in<One>(P, i).tv, i, P.s.rays, P.s.hits, lr, nullptr);
}

// The same march counting the FP32 operations of every evaluation (the
// instrumented pass the roofline's algorithmic work comes from).
__global__ void __launch_bounds__(kTraceBlock, kTraceWaves) k_trace_queue_count(PassParams P, uint32_t inQueue, double* work) {
    KDN_LANE_REGS;
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P.s.counters[inQueue]) return;
    trace_path(P, P.s.queue[inQueue][q], lr, work);
}

template <bool One>
__global__ void __launch_bounds__(256) k_shade(PassParams P, uint32_t inQueue, uint32_t outQueue) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < P.s.counters[inQueue]) shade_one<One>(P, q, inQueue, outQueue);
}

template <bool One>
__global__ void __launch_bounds__(kTraceBlock, kTraceWaves) k_shadow(PassParams P, double* work) {
    KDN_LANE_REGS;
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P.s.counters[2]) return;
    shadow_path<One>(P, P.s.queue[2][q], lr, work);
}

// The paths still in the input queue to their ends, in as many lanes as
// the device holds at once (the wavefront's last bounces carry few paths,
// and each launch of them costs as long as its slowest ray): each lane
// takes the queue's entries in turn (tally[2] counts them out).
// (One: the round has one job, read from the kernel's arguments, and no
// operations are counted. Plain: no guide of either kind and no training
// records in the round.)
template <bool One, bool Plain>
__global__ void __launch_bounds__(kTraceBlock, kTraceWaves) k_finish(PassParams P, uint32_t inQueue, uint32_t* tally, double* work) {
    KDN_LANE_REGS;
    finish_lane<One, Plain>(P, inQueue, P.s.counters[inQueue], lr, tally, work);
}

// the sort keys of a queue's rays (the paths' next rays, or their shadow rays)
__global__ void k_sort_keys(PassParams P, const uint32_t* queue, uint32_t count, uint32_t shadow) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const uint32_t i = queue[q];
    P.s.keys[q] = ray_key(job_of(P, i).tv, shadow ? P.s.shadow[i] : P.s.rays[i]);
}

// the job of each queued path: the key that groups a round's queue by job
// (a stable sort on it keeps each job's paths in their sorted order)
__global__ void k_job_keys(PassParams P, const uint32_t* queue, uint32_t count) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    P.s.keys[q] = P.slices[P.s.slice[queue[q]]].job;
}

// where each job's stretch of a queue grouped by job begins (starts[job]
// untouched for a job with no path in it)
__global__ void k_job_starts(PassParams P, const uint32_t* queue, uint32_t count, uint32_t* starts) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const uint32_t job = P.slices[P.s.slice[queue[q]]].job;
    if (q == 0 || P.slices[P.s.slice[queue[q - 1]]].job != job) starts[job] = q;
}

// estimates[pixel] folded with the pixel's spp colours, in sample order
__global__ void k_fold_samples(const double* colours, uint32_t pixels, uint32_t spp, uint32_t batch, kdn_estimate* estimates) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    kdn_estimate e = estimates[p];
    for (uint32_t j = 0; j < spp; ++j) fold_estimate(e, colours + ((size_t)p * spp + j) * 3, batch);
    estimates[p] = e;
}

__global__ void k_sequence_numbers(uint64_t stream, const uint32_t* indices, uint32_t count, uint32_t coordinate, bool sobol,
                                   double* out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) out[k] = sequence_number(sobol, stream, indices[k], coordinate);
}

bool sequenceNumbers(Ctx& c, uint64_t stream, const uint32_t* indices, uint32_t count, uint32_t coordinate, bool sobol, double* out) {
    const size_t ib = (size_t)count * sizeof(uint32_t), ob = (size_t)count * sizeof(double);
    if (!ensure(c.scratch[6], ib, c.error, false) || !ensure(c.scratch[7], ob, c.error, false)) return false;
    KDN_HIP(hipMemcpy(c.scratch[6].ptr, indices, ib, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_sequence_numbers, dim3((count + 255) / 256), dim3(256), 0, c.stream, stream,
                       (const uint32_t*)c.scratch[6].ptr, count, coordinate, sobol, (double*)c.scratch[7].ptr);
    KDN_HIP(hipGetLastError());
    KDN_HIP(hipStreamSynchronize(c.stream));
    KDN_HIP(hipMemcpy(out, c.scratch[7].ptr, ob, hipMemcpyDeviceToHost));
    return true;
}

bool foldSamples(Ctx& c, const double* colours, uint32_t pixels, uint32_t spp, uint32_t batch, kdn_estimate* estimates) {
    const size_t cb = (size_t)pixels * spp * 3 * sizeof(double), eb = (size_t)pixels * sizeof(kdn_estimate);
    if (!ensure(c.scratch[6], cb + 256, c.error, false) || !ensure(c.scratch[7], eb + 256, c.error, false)) return false;
    KDN_HIP(hipMemcpy(c.scratch[6].ptr, colours, cb, hipMemcpyHostToDevice));
    KDN_HIP(hipMemcpy(c.scratch[7].ptr, estimates, eb, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_fold_samples, dim3((pixels + 255) / 256), dim3(256), 0, c.stream, (const double*)c.scratch[6].ptr, pixels, spp,
                       batch, (kdn_estimate*)c.scratch[7].ptr);
    KDN_HIP(hipGetLastError());
    KDN_HIP(hipStreamSynchronize(c.stream));
    KDN_HIP(hipMemcpy(estimates, c.scratch[7].ptr, eb, hipMemcpyDeviceToHost));
    return true;
}

__global__ void k_fold(PassParams P, uint32_t slice) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    fold_one(P, slice, t);
}

#undef KDN_LANE_REGS

void freeIntegrator(Ctx& c) {
    if (c.estimates) (void)hipFree(c.estimates);
    c.estimates = nullptr;
    if (c.paths) {
        (void)hipFree(c.paths->slice);
        delete c.paths;
        c.paths = nullptr;
        c.paths_bytes = 0;
    }
    if (c.colour) (void)hipFree(c.colour);
    c.colour = nullptr;
    if (c.round_dev) (void)hipFree(c.round_dev);
    c.round_dev = nullptr;
    c.round_bytes = 0;
}

static bool allocPaths(Ctx& c, size_t n) {
    if (c.paths && c.paths->capacity >= n) return true;
    if (c.paths) {
        (void)hipFree(c.paths->slice);
        delete c.paths;
        c.paths = nullptr;
        c.paths_bytes = 0;
    }
    // one allocation, 256-byte aligned pieces
    const size_t pieces[] = {n * 4, n * sizeof(kdn_ray), n * sizeof(kdn_hit), n * sizeof(kdn_ray), n * 16, n * 16, n * 16, n * 16,
                             n * 16, n * 4, n * 4, n * 4, n * 4, n * 8, n * 4, n * 4, n * 4, 64, n * 4, n * 4, n * 4};
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
    s->slice = (uint32_t*)take();
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
    s->keys = (uint32_t*)take();
    s->keys_alt = (uint32_t*)take();
    s->vals_alt = (uint32_t*)take();
    s->capacity = n;
    c.paths = s;
    c.paths_bytes = bytes;
    return true;
}

bool reservePaths(Ctx& c, size_t n) { return allocPaths(c, n); }

bool ensureColour(Ctx& c) {
    if (c.colour) return true;
    ColourTables t;
    buildColourTables(t);
    KDN_HIP(hipMalloc(&c.colour, sizeof(ColourTables)));
    KDN_HIP(hipMemcpy(c.colour, &t, sizeof(ColourTables), hipMemcpyHostToDevice));
    return true;
}

static uint64_t mixedBitsHost(uint64_t v) {
    v += 0x9E3779B97F4A7C15ULL;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
    return v ^ (v >> 31);
}

uint32_t planeOf(int width, int height) { return (uint32_t)(((width + 7) / 8) * ((height + 7) / 8) * 64); }

JobDev jobOf(const Ctx& scene, int width, int height, uint32_t seed, uint32_t surfaces, uint32_t minDepth) {
    JobDev J{};
    J.tv = treeView(scene);
    J.width = width;
    J.height = height;
    J.plane = planeOf(width, height);
    J.seed_bits = mixedBitsHost(seed);
    J.surfaces = surfaces;
    J.min_depth = minDepth;
    // the lens (core/camera.cpp:9-25)
    const kdn_camera& cam = scene.scene.camera;
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
        J.eye[k] = cam.eye[k];
        J.forward[k] = (float)fwd[k];
        J.right[k] = (float)rt[k];
        J.up[k] = (float)upv[k];
        J.sky_zenith[k] = scene.scene.render.sky_zenith[k];
        J.sky_horizon[k] = scene.scene.render.sky_horizon[k];
        J.sky_ground[k] = scene.scene.render.sky_ground[k];
    }
    J.reach = (float)std::tan((double)cam.fov / 2);
    J.aspect = (float)((double)width / height);
    J.n_lamps = scene.n_lamps;
    J.n_materials = scene.n_materials_dev;
    J.fallback_material = scene.n_materials_dev - 1;
    J.lamps = (const LampDev*)scene.lamps_dev;
    J.materials = (const MaterialDev*)scene.materials_dev;
    J.leaf_material = scene.scene.leaf_material;
    return J;
}

// trace-kernel blocks the device runs at once: kTraceWaves waves on each
// of a compute unit's four SIMDs
static uint32_t residentTraceBlocks(const Ctx& c) {
    int units = 0;
    if (hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess || units <= 0) units = 256;
    return (uint32_t)units * 4 * kTraceWaves * 64 / kTraceBlock;
}

bool runRound(Ctx& c, const std::vector<JobDev>& jobs, const std::vector<SliceDev>& slices, uint32_t flags, kdn_pass_stats* stats,
              const std::vector<void*>* guides) {
    if (!ensureColour(c)) return false;
    hipStream_t st = c.stream;
    const SliceDev& lastSlice = slices.back();
    const uint32_t total = lastSlice.first_path + slice_paths(lastSlice, jobs[lastSlice.job].plane);
    if (!allocPaths(c, total)) return false;
    // the round's job and slice tables in HBM
    const size_t jobBytes = jobs.size() * sizeof(JobDev), sliceBytes = slices.size() * sizeof(SliceDev);
    const size_t roundBytes = ((jobBytes + 255) & ~size_t(255)) + sliceBytes;
    if (c.round_bytes < roundBytes) {
        if (c.round_dev) KDN_HIP(hipFree(c.round_dev));
        c.round_dev = nullptr;
        KDN_HIP(hipMalloc(&c.round_dev, roundBytes));
        c.round_bytes = roundBytes;
    }
    char* table = (char*)c.round_dev;
    PassParams P{};
    P.jobs = (const JobDev*)table;
    P.slices = (const SliceDev*)(table + ((jobBytes + 255) & ~size_t(255)));
    P.n_slices = (uint32_t)slices.size();
    P.paths = total;
    P.colour = (const ColourTables*)c.colour;
    P.s = *c.paths;
    PathState* s = c.paths;

    const bool timing = (flags & 2) != 0, counting = (flags & 4) != 0, finishing = (flags & 8) == 0, guiding = (flags & 16) != 0;
    // the learned guide: the context's own (a render pass, flags bit 4) or
    // each job's own (serving), with the jobs' fields that place its lobes
    // and training records
    std::vector<JobDev> J(jobs);
    bool ownGuides = false;
    if (guides) {
        if (guides->size() != jobs.size()) {
            c.error = "a round's guides are one for each job";
            return false;
        }
        for (void* g : *guides) ownGuides = ownGuides || g != nullptr;
    }
    if (guiding && ownGuides) {
        c.error = "a round's learned guides are the context's or the jobs' own, not both";
        return false;
    }
    for (size_t j = 0; j < jobs.size(); ++j) {
        if (jobs[j].hg && (guiding || (ownGuides && (*guides)[j]))) {
            c.error = "a round guides its paths by the learned guide or by the plain guide, not both";
            return false;
        }
    }
    if (guiding) {
        if (!guideReady(c) && !guideReset(c, 1, c.tree.lower, c.tree.size)) return false;
        if (!guideBeginRound(c, P, total)) return false;
        for (JobDev& j : J) {
            j.learned = P.lobes != nullptr ? 1u : 0u;
            j.train_stride = P.train_stride;
            j.path_first = 0;
            j.rec_first = 0;
        }
    }
    if (ownGuides && !guidesBeginRound(c, P, J, slices, *guides)) return false;
    const bool learnedRound = guiding || ownGuides;
    KDN_HIP(hipMemcpyAsync(table, J.data(), jobBytes, hipMemcpyHostToDevice, st));
    KDN_HIP(hipMemcpyAsync(table + ((jobBytes + 255) & ~size_t(255)), slices.data(), sliceBytes, hipMemcpyHostToDevice, st));
    P.job0 = J[0];
    // kernel timing: an event pair per launch, read after each bounce's sync
    // (destroyed on every way out of the round)
    struct EventPair {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~EventPair() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } events;
    hipEvent_t* ev = events.e;
    if (timing) {
        KDN_HIP(hipEventCreate(&ev[0]));
        KDN_HIP(hipEventCreate(&ev[1]));
    }
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
    double* work = nullptr;
    if (counting) {
        if (!ensure(c.scratch[7], 64, c.error, false)) return false;
        work = (double*)c.scratch[7].ptr;
        KDN_HIP(hipMemsetAsync(work, 0, 32, st));
    }

    // a round with no guide of either kind and no training records runs the
    // finishing launch with the guided sampling compiled out
    bool plain = !learnedRound;
    for (const JobDev& j : jobs) plain = plain && j.hg == nullptr && j.hg_deep == nullptr;
    // a queue's rays ordered by direction octant and origin cell, so that a
    // wave's lanes walk nearby cells alike (each path's result does not
    // depend on where it stands in a queue)
    auto sortQueue = [&](uint32_t qi, uint32_t n, uint32_t shadow) -> bool {
        if (n < kSortMin) return true;
        bool ok = true;
        timed(&stats->other_ms, [&] {
            uint32_t* queue = s->queue[qi];
            hipLaunchKernelGGL(k_sort_keys, dim3((n + 255) / 256), dim3(256), 0, st, P, (const uint32_t*)queue, n, shadow);
            uint32_t* sorted = nullptr;
            ok = sortPairsU32(s->keys, s->keys_alt, queue, s->vals_alt, n, 30, c.sort_tmp, st, c.error, &sorted);
            if (ok && sorted != queue) ok = hipMemcpyAsync(queue, sorted, (size_t)n * 4, hipMemcpyDeviceToDevice, st) == hipSuccess;
        });
        if (!ok && c.error.empty()) c.error = "sorting a ray queue failed";
        return ok;
    };
    // a queue grouped by job (a stable sort: each job's paths keep their
    // order): each job's first entry and count
    auto groupByJob = [&](uint32_t qi, uint32_t count, std::vector<uint32_t>& first, std::vector<uint32_t>& cnt) -> bool {
        const uint32_t nJobs = (uint32_t)jobs.size();
        first.assign(nJobs, 0);
        cnt.assign(nJobs, 0);
        int bits = 1;
        while ((1u << bits) < nJobs) ++bits;
        uint32_t* queue = s->queue[qi];
        hipLaunchKernelGGL(k_job_keys, dim3((count + 255) / 256), dim3(256), 0, st, P, (const uint32_t*)queue, count);
        uint32_t* sorted = nullptr;
        if (!sortPairsU32(s->keys, s->keys_alt, queue, s->vals_alt, count, bits, c.sort_tmp, st, c.error, &sorted)) return false;
        if (sorted != queue) KDN_HIP(hipMemcpyAsync(queue, sorted, (size_t)count * 4, hipMemcpyDeviceToDevice, st));
        if (!ensure(c.job_counters, (size_t)nJobs * 68 + 64, c.error, false)) return false;
        uint32_t* starts = (uint32_t*)c.job_counters.ptr + (size_t)nJobs * 16;
        KDN_HIP(hipMemsetAsync(starts, 0xFF, (size_t)nJobs * 4, st));
        hipLaunchKernelGGL(k_job_starts, dim3((count + 255) / 256), dim3(256), 0, st, P, (const uint32_t*)queue, count, starts);
        KDN_HIP(hipGetLastError());
        std::vector<uint32_t> at(nJobs);
        KDN_HIP(hipMemcpyAsync(at.data(), starts, (size_t)nJobs * 4, hipMemcpyDeviceToHost, st));
        KDN_HIP(hipStreamSynchronize(st));
        uint32_t end = count;
        for (uint32_t j = nJobs; j-- > 0;) {
            if (at[j] == 0xFFFFFFFFu) continue;
            if (at[j] > end) {
                c.error = "a queue grouped by job is out of order";
                return false;
            }
            first[j] = at[j];
            cnt[j] = end - at[j];
            end = at[j];
        }
        return true;
    };
    // A round of several jobs finishes each job's paths in the launch
    // specialised for one job (its fields from the kernel's arguments, the
    // same for every lane), the jobs' launches side by side on the job
    // streams: the queue, sorted, is grouped by job with a stable sort (so
    // each job's paths keep their order), each job gets its stretch of it
    // and counters of its own. At the camera rays every path of a plain
    // slice is queued (spp x width x height of them), so the stretches are
    // known here; deeper (a learned round finishes after its guided depths)
    // they are read from the grouped queue.
    bool everySlicePlain = true;
    for (const SliceDev& sl : slices) everySlicePlain = everySlicePlain && sl.offsets == nullptr;
    auto finishJobsApart = [&](uint32_t qi, uint32_t count, uint32_t bounce) -> bool {
        const uint32_t nJobs = (uint32_t)jobs.size();
        std::vector<uint64_t> live(nJobs, 0);
        if (bounce > 0) {
            std::vector<uint32_t> first, cnt;
            bool grouped = true;
            timed(&stats->other_ms, [&] { grouped = groupByJob(qi, count, first, cnt); });
            if (!grouped) return false;
            for (uint32_t j = 0; j < nJobs; ++j) live[j] = cnt[j];
        } else {
            for (const SliceDev& sl : slices) live[sl.job] += (uint64_t)sl.spp * jobs[sl.job].width * jobs[sl.job].height;
        }
        uint64_t sum = 0;
        for (uint64_t n : live) sum += n;
        if (sum != count) {
            c.error = "a round's queue holds " + std::to_string(count) + " paths, its jobs' stretches " + std::to_string(sum);
            return false;
        }
        int bits = 1;
        while ((1u << bits) < nJobs) ++bits;
        bool ok = true;
        // (deeper, groupByJob has grouped the queue already)
        if (bounce == 0) timed(&stats->other_ms, [&] {
            uint32_t* queue = s->queue[qi];
            hipLaunchKernelGGL(k_job_keys, dim3((count + 255) / 256), dim3(256), 0, st, P, (const uint32_t*)queue, count);
            uint32_t* sorted = nullptr;
            ok = sortPairsU32(s->keys, s->keys_alt, queue, s->vals_alt, count, bits, c.sort_tmp, st, c.error, &sorted);
            if (ok && sorted != queue) ok = hipMemcpyAsync(queue, sorted, (size_t)count * 4, hipMemcpyDeviceToDevice, st) == hipSuccess;
        });
        if (!ok) return false;
        if (!ensure(c.job_counters, (size_t)nJobs * 64, c.error, false)) return false;
        std::vector<uint32_t> counters((size_t)nJobs * 16, 0);
        for (uint32_t j = 0; j < nJobs; ++j) counters[(size_t)j * 16 + qi] = (uint32_t)live[j];
        uint32_t* jc = (uint32_t*)c.job_counters.ptr;
        KDN_HIP(hipMemcpyAsync(jc, counters.data(), counters.size() * 4, hipMemcpyHostToDevice, st));
        for (int k = 0; k < Ctx::kJobStreams; ++k) {
            if (!c.job_streams[k]) KDN_HIP(hipStreamCreate(&c.job_streams[k]));
        }
        for (hipEvent_t& e : c.job_events) {
            if (!e) KDN_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        KDN_HIP(hipEventRecord(c.job_events[Ctx::kJobStreams], st));
        for (int k = 0; k < Ctx::kJobStreams; ++k) KDN_HIP(hipStreamWaitEvent(c.job_streams[k], c.job_events[Ctx::kJobStreams], 0));
        const uint32_t resident = residentTraceBlocks(c);
        bool launched = true;
        timed(&stats->finish_ms, [&] {
            uint64_t first = 0;
            for (uint32_t j = 0; j < nJobs; ++j) {
                PassParams Pj = P;
                Pj.s.queue[qi] = s->queue[qi] + first;
                Pj.s.counters = jc + (size_t)j * 16;
                Pj.job0 = J[j];
                first += live[j];
                if (live[j] == 0) continue;
                hipStream_t js = c.job_streams[j % Ctx::kJobStreams];
                const dim3 grid(std::min<uint32_t>((uint32_t)((live[j] + kTraceBlock - 1) / kTraceBlock), resident)), block(kTraceBlock);
                if (plain) {
                    hipLaunchKernelGGL((k_finish<true, true>), grid, block, 0, js, Pj, qi, Pj.s.counters + 4, nullptr);
                } else {
                    hipLaunchKernelGGL((k_finish<true, false>), grid, block, 0, js, Pj, qi, Pj.s.counters + 4, nullptr);
                }
                launched = launched && hipGetLastError() == hipSuccess;
            }
            for (int k = 0; k < Ctx::kJobStreams; ++k) {
                (void)hipEventRecord(c.job_events[k], c.job_streams[k]);
                (void)hipStreamWaitEvent(st, c.job_events[k], 0);
            }
        });
        if (!launched) {
            c.error = "a job's finishing launch failed";
            return false;
        }
        std::vector<uint32_t> tallies(counters.size(), 0);
        KDN_HIP(hipMemcpyAsync(tallies.data(), jc, tallies.size() * 4, hipMemcpyDeviceToHost, st));
        KDN_HIP(hipStreamSynchronize(st));
        for (uint32_t j = 0; j < nJobs; ++j) {
            stats->segments += tallies[(size_t)j * 16 + 4];
            stats->shadow_rays += tallies[(size_t)j * 16 + 5];
        }
        ++stats->bounces;
        stats->finish_launches += nJobs;
        return true;
    };
    KDN_HIP(hipMemsetAsync(s->counters, 0, 64, st));
    timed(&stats->other_ms, [&] { hipLaunchKernelGGL(k_generate, dim3((total + 255) / 256), dim3(256), 0, st, P); });
    uint32_t in = 0, out = 1;
    uint32_t maxSurfaces = 0;
    for (const JobDev& J : jobs) maxSurfaces = std::max(maxSurfaces, J.surfaces);
    for (uint32_t bounce = 0; bounce < maxSurfaces; ++bounce) {
        uint32_t count = 0;
        KDN_HIP(hipMemcpyAsync(&count, &s->counters[in], 4, hipMemcpyDeviceToHost, st));
        KDN_HIP(hipStreamSynchronize(st));
        if (count == 0) break;
        // (a guided round finishes paths only past the guided depths: the
        // finishing launch shades without the network's lobes)
        if (finishing && (!learnedRound || bounce >= (uint32_t)kGuideDepths) && (uint64_t)count * kFinishShare <= total) {
            // the remaining paths to their ends in one launch
            // (camera rays too: grouped by direction octant, the finishing
            // lanes ran 1.05x faster at 1080p than in the 8x8 tiles' order)
            if (!sortQueue(in, count, 0)) return false;
            if (jobs.size() > 1 && !counting && everySlicePlain) {
                if (!finishJobsApart(in, count, bounce)) return false;
                break;
            }
            KDN_HIP(hipMemsetAsync(&s->counters[4], 0, 12, st));
            const uint32_t finishBlocks = std::min<uint32_t>((count + kTraceBlock - 1) / kTraceBlock, residentTraceBlocks(c));
            timed(&stats->finish_ms, [&] {
                const dim3 grid(finishBlocks), block(kTraceBlock);
                if (jobs.size() == 1 && !counting) {
                    if (plain) {
                        hipLaunchKernelGGL((k_finish<true, true>), grid, block, 0, st, P, in, &s->counters[4], nullptr);
                    } else {
                        hipLaunchKernelGGL((k_finish<true, false>), grid, block, 0, st, P, in, &s->counters[4], nullptr);
                    }
                } else if (plain) {
                    hipLaunchKernelGGL((k_finish<false, true>), grid, block, 0, st, P, in, &s->counters[4], counting ? work + 2 : nullptr);
                } else {
                    hipLaunchKernelGGL((k_finish<false, false>), grid, block, 0, st, P, in, &s->counters[4], counting ? work + 2 : nullptr);
                }
            });
            uint32_t tally[2] = {0, 0};
            KDN_HIP(hipMemcpyAsync(tally, &s->counters[4], 8, hipMemcpyDeviceToHost, st));
            KDN_HIP(hipStreamSynchronize(st));
            stats->segments += tally[0];
            stats->shadow_rays += tally[1];
            ++stats->bounces;
            ++stats->finish_launches;
            break;
        }
        stats->segments += count;
        ++stats->bounces;
        ++stats->trace_launches;
        if (bounce > 0 && !sortQueue(in, count, 0)) return false;  // camera rays come in 8x8 tiles already
        const uint32_t blocks = (count + 255) / 256, traceBlocks = (count + kTraceBlock - 1) / kTraceBlock;
        timed(&stats->trace_ms, [&] {
            if (counting) {
                hipLaunchKernelGGL(k_trace_queue_count, dim3(traceBlocks), dim3(kTraceBlock), 0, st, P, in, work);
            } else {
                if (jobs.size() == 1) {
                    hipLaunchKernelGGL(k_trace_queue<true>, dim3(traceBlocks), dim3(kTraceBlock), 0, st, P, in);
                } else {
                    hipLaunchKernelGGL(k_trace_queue<false>, dim3(traceBlocks), dim3(kTraceBlock), 0, st, P, in);
                }
            }
        });
        if (learnedRound && bounce < (uint32_t)kGuideDepths) {
            bool ok = true;
            double ms = 0;
            if (guiding) {
                timed(&ms, [&] { ok = guideInferQueue(c, P, in, count); });
            } else if (P.lobes) {
                // each learned job's stretch of the queue by its own guide
                std::vector<uint32_t> first(jobs.size(), 0), cnt(jobs.size(), 0);
                if (jobs.size() == 1) {
                    cnt[0] = count;
                } else {
                    double sortMs = 0;
                    timed(&sortMs, [&] { ok = groupByJob(in, count, first, cnt); });
                    stats->other_ms += sortMs;
                    if (!ok) return false;
                }
                timed(&ms, [&] {
                    for (size_t j = 0; j < jobs.size() && ok; ++j) {
                        if ((*guides)[j] && J[j].learned) ok = guidesInferQueue(c, P, in, first[j], cnt[j], (*guides)[j]);
                    }
                });
            }
            if (!ok) return false;
            stats->guide_ms += ms;
            stats->guide_infer_ms += ms;
            if (P.lobes) stats->guide_queries += count;
        }
        KDN_HIP(hipMemsetAsync(&s->counters[out], 0, 4, st));
        KDN_HIP(hipMemsetAsync(&s->counters[2], 0, 4, st));
        timed(&stats->shade_ms, [&] {
            if (jobs.size() == 1) {
                hipLaunchKernelGGL(k_shade<true>, dim3(blocks), dim3(256), 0, st, P, in, out);
            } else {
                hipLaunchKernelGGL(k_shade<false>, dim3(blocks), dim3(256), 0, st, P, in, out);
            }
        });
        uint32_t sc = 0;
        KDN_HIP(hipMemcpyAsync(&sc, &s->counters[2], 4, hipMemcpyDeviceToHost, st));
        KDN_HIP(hipStreamSynchronize(st));
        stats->shadow_rays += sc;
        if (sc) {
            if (!sortQueue(2, sc, 1)) return false;
            ++stats->shadow_launches;
            timed(&stats->shadow_ms, [&] {
                const dim3 grid((sc + kTraceBlock - 1) / kTraceBlock);
                if (jobs.size() == 1) {
                    hipLaunchKernelGGL(k_shadow<true>, grid, dim3(kTraceBlock), 0, st, P, counting ? work + 1 : nullptr);
                } else {
                    hipLaunchKernelGGL(k_shadow<false>, grid, dim3(kTraceBlock), 0, st, P, counting ? work + 1 : nullptr);
                }
            });
        }
        std::swap(in, out);
    }
    // each slice's samples into its job's accumulators, slices in order
    for (uint32_t k = 0; k < (uint32_t)slices.size(); ++k) {
        const uint32_t plane = jobs[slices[k].job].plane;
        timed(&stats->other_ms, [&] { hipLaunchKernelGGL(k_fold, dim3((plane + 255) / 256), dim3(256), 0, st, P, k); });
    }
    if (learnedRound) {
        bool ok = true;
        timed(&stats->guide_ms, [&] { ok = guiding ? guideEndRound(c, P, total, stats) : guidesEndRound(c, P, J, slices, *guides, stats); });
        if (!ok) return false;
    }
    KDN_HIP(hipStreamSynchronize(st));
    KDN_HIP(hipGetLastError());
    if (counting) {
        // trace, shadow, and the finishing launch's path marches and shadow
        // marches (it counts them at work + 2 and work + 3)
        double counted[4] = {0, 0, 0, 0};
        KDN_HIP(hipMemcpy(counted, work, 32, hipMemcpyDeviceToHost));
        stats->trace_flops += counted[0];
        stats->shadow_flops += counted[1];
        stats->finish_flops += counted[2] + counted[3];
    }
    return true;
}

bool renderPass(Ctx& c, const kdn_pass_desc& d, kdn_pass_stats* stats) {
    const auto start = std::chrono::steady_clock::now();
    if (!c.integrator_ready) {
        c.error = "kdn_render_pass: scene tables are not on the device";
        return false;
    }
    hipStream_t st = c.stream;
    // accumulators
    const size_t px = (size_t)d.width * d.height;
    const bool fresh = !c.accum || c.width != d.width || c.height != d.height;
    if (fresh) {
        if (c.accum) hipFree(c.accum);
        if (c.accum_sq) hipFree(c.accum_sq);
        if (c.accum_n) hipFree(c.accum_n);
        if (c.estimates) hipFree(c.estimates);
        c.estimates = nullptr;
        KDN_HIP(hipMalloc(&c.accum, px * 3 * sizeof(double)));
        KDN_HIP(hipMalloc(&c.accum_sq, px * 3 * sizeof(double)));
        KDN_HIP(hipMalloc(&c.accum_n, px * sizeof(uint32_t)));
        c.width = d.width;
        c.height = d.height;
    }
    // the estimates are kept from the first pass that asks for them, zeroed
    // with the accumulators
    const bool newEstimates = (d.flags & 64) && !c.estimates;
    if (newEstimates) KDN_HIP(hipMalloc(&c.estimates, px * sizeof(kdn_estimate)));
    if (fresh || !(d.flags & 1)) {
        KDN_HIP(hipMemsetAsync(c.accum, 0, px * 3 * sizeof(double), st));
        KDN_HIP(hipMemsetAsync(c.accum_sq, 0, px * 3 * sizeof(double), st));
        KDN_HIP(hipMemsetAsync(c.accum_n, 0, px * sizeof(uint32_t), st));
        if (c.estimates) KDN_HIP(hipMemsetAsync(c.estimates, 0, px * sizeof(kdn_estimate), st));
    } else if (newEstimates) {
        c.error = "kdn_render_pass: a pass that keeps the accumulators (bit 0) cannot start the estimates (bit 6): "
                  "the passes before it kept none";
        return false;
    }
    // the count pass: every pixel its allotment less its trial
    const bool counted = (d.flags & 256) != 0;
    if (counted) {
        if (c.stop_width != d.width || c.stop_height != d.height || fresh) {
            c.error = "kdn_render_pass: no allotment for this image (kdn_stopping_allot after its trial pass)";
            return false;
        }
        if ((d.flags & 64) && (d.flags & 1) && !stoppingSumStart(c)) return false;
    }
    std::vector<JobDev> jobs{jobOf(c, d.width, d.height, d.seed, d.surfaces, d.min_depth)};
    jobs[0].sobol = (d.flags & 128) ? 1u : 0u;
    if (d.flags & 512) {
        // the frame record: each pixel's first surface of its last sample
        if (counted || d.spp == 0) {
            c.error = "kdn_render_pass: the frame record (bit 9) needs a plain pass of at least one sample";
            return false;
        }
        if (!c.firsts || c.firsts_width != d.width || c.firsts_height != d.height) {
            if (c.firsts) KDN_HIP(hipFree(c.firsts));
            c.firsts = nullptr;
            KDN_HIP(hipMalloc(&c.firsts, px * sizeof(kdn_surface)));
            c.firsts_width = d.width;
            c.firsts_height = d.height;
        }
        jobs[0].firsts = c.firsts;
        jobs[0].record_sample = d.first_sample + d.spp - 1;
    }
    if (d.flags & 32) {
        // the plain guide, of the image the guides were made for
        if ((d.flags & 16) || c.hguide_width != d.width || c.hguide_height != d.height) {
            c.error = "kdn_render_pass: no plain guide for this image (kdn_highlight_guides first; not with the learned guide)";
            return false;
        }
        jobs[0].hg = (const float*)c.hguides.ptr;
        jobs[0].hg_deep = c.hguide_deep ? (const HgDeepHeader*)((const char*)c.hguides.ptr + c.hguide_deep) : nullptr;
    }
    const uint32_t plane = jobs[0].plane;
    const size_t maxPaths = d.round_paths ? std::max<size_t>(d.round_paths, 4096) : c.max_paths ? c.max_paths : (size_t)1 << 23;
    const uint32_t chunk = (uint32_t)std::max<size_t>(1, std::min<size_t>(d.spp, maxPaths / plane));
    kdn_pass_stats local{};
    if (counted) {
        // rounds of at most maxPaths of the dense paths, pixel by pixel
        for (uint64_t base = 0; base < c.stop_total; base += maxPaths) {
            SliceDev sl{0, d.first_sample, 0, 0, c.accum, c.accum_sq, c.accum_n, (d.flags & 64) ? c.estimates : nullptr, d.batch};
            sl.offsets = (const uint64_t*)c.stop_offsets.ptr;
            sl.base = base;
            sl.paths = (uint32_t)std::min<uint64_t>(maxPaths, c.stop_total - base);
            if (!runRound(c, jobs, std::vector<SliceDev>{sl}, d.flags, &local)) return false;
        }
    }
    for (uint32_t first = 0; !counted && first < d.spp; first += chunk) {
        const uint32_t n = std::min(chunk, d.spp - first);
        const std::vector<SliceDev> slices{SliceDev{0, d.first_sample + first, n, 0, c.accum, c.accum_sq, c.accum_n,
                                                    (d.flags & 64) ? c.estimates : nullptr, d.batch}};
        if (!runRound(c, jobs, slices, d.flags, &local)) return false;
    }
    local.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
    if (stats) *stats = local;
    return true;
}

}  // namespace kdn
