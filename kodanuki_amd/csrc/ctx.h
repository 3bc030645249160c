// The device context behind kdn_ctx: every buffer the hot path keeps
// resident in HBM, and the launchers the C ABI calls.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "kodanuki.h"
#include "layout.h"

namespace kdn {

// A device buffer that grows by doubling and keeps its contents.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct SceneDev {
    kdn_step* steps = nullptr;       // the scene tape
    kdn_object* objects = nullptr;
    uint32_t n_steps = 0, n_objects = 0;
    kdn_material* materials = nullptr;
    uint32_t n_materials = 0;
    kdn_light* lights = nullptr;
    uint32_t n_lights = 0;
    uint32_t* leaf_material = nullptr;  // per leaf
    uint32_t n_leaves = 0;
    kdn_camera camera{};
    kdn_render_settings render{};
};

struct OctreeDev {
    bool built = false;
    float lower[3] = {0, 0, 0};
    float size[3] = {0, 0, 0};
    int depth = 0;
    std::vector<uint32_t> level_first;   // first node of each level in nodes
    std::vector<uint32_t> level_count;
    DevBuf nodes;        // NodeDev, levels one after another
    DevBuf entries;      // EntryDev
    DevBuf pool;         // kdn_step programs with a header step each
    DevBuf spool;        // the same programs on slot registers (the landing's)
    DevBuf mpool;        // the same with their literal steps folded away (the march's)
    uint64_t n_nodes = 0, n_entries = 0, n_pool = 0;
    uint32_t* level_first_dev = nullptr;  // device copy of level_first
};

struct PathState;

struct Ctx {
    int device = 0;
    std::string error;
    hipStream_t stream = nullptr;
    // the streams a round of several jobs runs its jobs' finishing launches
    // on side by side (default priority, made on first use), and their
    // events; the job's own counters (count and tally) per job
    static constexpr int kJobStreams = 4;
    hipStream_t job_streams[kJobStreams] = {};
    hipEvent_t job_events[kJobStreams + 1] = {};
    DevBuf job_counters;
    SceneDev scene;
    OctreeDev tree;
    // integrator state
    int width = 0, height = 0;
    double* accum = nullptr;        // rgb sums, FP64, per pixel
    double* accum_sq = nullptr;     // luminance squares, FP64, per pixel
    uint32_t* accum_n = nullptr;    // samples per pixel
    kdn_estimate* estimates = nullptr;  // per pixel, the reference's Estimate (passes with flags bit 6)
    PathState* paths = nullptr;     // host record of the path SoA in HBM
    size_t paths_bytes = 0;         // its one allocation
    size_t max_paths = 0;           // paths in flight per chunk (0: 2^23)
    DevBuf scratch[8];              // reusable scratch for launches
    void* colour = nullptr;         // ColourTables
    void* materials_dev = nullptr;  // MaterialDev per material, the fallback last
    uint32_t n_materials_dev = 0;
    void* lamps_dev = nullptr;      // LampDev per quad lamp
    uint32_t n_lamps = 0;
    bool integrator_ready = false;
    void* round_dev = nullptr;      // the running round's job and slice tables
    size_t round_bytes = 0;
    bool scene_has_glass = false;   // a leaf is painted with a glass
    void* guide = nullptr;          // the learned guide (GuideState), made on first use
    void* guide_round = nullptr;    // lobes and records of rounds whose jobs own their guides (serving)
    void* leaf_steps_dev = nullptr;  // every leaf's own program (LeafStepDev), the survey's chains walk them
    void* leaf_progs_dev = nullptr;  // LeafProgDev per leaf
    uint32_t n_leaf_progs = 0;
    DevBuf sort_tmp;                // the ray queues' sort's temporary storage
    DevBuf hguides;                 // the plain guide: a packed guide per pixel
    int hguide_width = 0, hguide_height = 0;  // the image they were made for (0: none)
    size_t hguide_deep = 0;         // byte offset of the deep guides' filing in hguides (0: none)
    float* denoise_static = nullptr;  // this device's static denoiser working memory (denoise.hip)
    // adaptive stopping (stopping.hip): each pixel's judged trial, its
    // allotment, and the count pass's per-slot counts and dense offsets
    DevBuf stop_wanted, stop_spreads, stop_allotted, stop_counts, stop_offsets, stop_tmp;
    uint64_t stop_total = 0;        // the count pass's paths
    int stop_width = 0, stop_height = 0;  // the image allotted (0: none)
    int32_t stop_trial = 0;
    // the frame record's first surfaces (passes with flags bit 9) and the
    // temporal member's resident history (temporal.hip)
    kdn_surface* firsts = nullptr;
    int firsts_width = 0, firsts_height = 0;
    DevBuf tm_history[2], tm_gbuffer;
    int tm_current = 0;
    int tm_width = 0, tm_height = 0;  // the history's image (0: none)
    kdn_lens tm_lens{};               // the lens of the frame before
};

// Glass scattering is held off the device: renders with glass faulted the
// device and no cause is shown yet. kdn_render_pass refuses such scenes
// unless this is set (the bounds-checking CPU build of the kernels sets
// it; the device build does not).
#ifndef KDN_GLASS_ON_DEVICE
#define KDN_GLASS_ON_DEVICE 0
#endif

bool ensure(DevBuf& b, size_t bytes, std::string& err, bool keep);
void releaseAll(Ctx& c);

// octree.hip
bool buildOctree(Ctx& c, const float lower[3], const float size[3], int depth, kdn_octree_stats* stats);

// trace.hip
bool extendRays(Ctx& c, const kdn_ray* rays, uint32_t n, kdn_hit* hits, hipStream_t stream);

// integrator.hip
bool renderPass(Ctx& c, const kdn_pass_desc& d, kdn_pass_stats* stats);
bool foldSamples(Ctx& c, const double* colours, uint32_t pixels, uint32_t spp, uint32_t batch, kdn_estimate* estimates);
// an eye path's numbers (numberAt) of `count` sample indices at one coordinate
bool sequenceNumbers(Ctx& c, uint64_t stream, const uint32_t* indices, uint32_t count, uint32_t coordinate, bool sobol, double* out);
void freeIntegrator(Ctx& c);

// guide.hip: the learned guide's parameters over a region, from a seed
bool guideReset(Ctx& c, uint32_t seed, const float lower[3], const float size[3]);
bool guideReady(const Ctx& c);
void freeGuide(Ctx& c);
// a served job's own learned guide: made over its scene's region from a
// seed (null on failure, c.error set), its device bytes, its Adam steps
void* guideOpen(Ctx& c, uint32_t seed, const float lower[3], const float size[3]);
void guideClose(void* guide);
uint64_t guideBytes();
uint32_t guideSteps(const void* guide);
// whether the next round trains the guide (a served job's last slice does not)
void guideTrainNext(void* guide, bool on);
// the device bytes of a context's rounds' lobes and records (serving)
uint64_t guideRoundBytes(const Ctx& c);

// stopping.hip: effortFor, confidenceTrial, and the device's judge and
// allotment of every pixel's count from the trial estimates
void stoppingEffort(double tolerance, int32_t* samples, int32_t* batch);
int32_t stoppingTrial(const kdn_stopping& s);
bool stoppingCheck(const kdn_stopping& s, std::string& err);
bool stoppingAllot(Ctx& c, const kdn_stopping& s, int32_t trial, uint64_t* total);
bool stoppingSumStart(Ctx& c);
bool readStopping(Ctx& c, float* wanted, float* spreads, int32_t* allotted);

// temporal.hip: the temporal member's frame and step, and a camera's lens
kdn_lens lensOfCamera(const kdn_camera& cam, int width, int height);
bool temporalFrame(Ctx& c, const kdn_temporal_desc& d, float* shown, float* denoised, kdn_temporal_stats* stats);
bool temporalHistory(Ctx& c, kdn_denoise_pixel* pixels);
bool temporalStep(Ctx& c, int width, int height, const kdn_lens& lens, const kdn_lens& before, const kdn_temporal_desc& d,
                  const double* parts, const kdn_surface* record, const kdn_surface* gbuffer, const kdn_denoise_pixel* beforePixels,
                  int beforeWidth, int beforeHeight, kdn_denoise_pixel* pixels, float* shown);

// denoise.hip: look up this device's static denoiser working memory; the
// a-trous passes over a history on the device
bool denoiseStaticScratch(Ctx& c);
bool denoiseDevice(Ctx& c, const kdn_denoise_frame& f, const kdn_denoise_pixel* pixels, float* rgb, hipStream_t st);

// hguide.hip: the plain guide of every pixel of a width x height image
bool hguideBuild(Ctx& c, int width, int height, kdn_hguide_stats* stats, DevBuf& out, DevBuf& work, size_t* deepOffset);

// sort.hip: n (key, value) pairs sorted by the key's low endBit bits;
// *sorted is vals or valsAlt, whichever holds the sorted values
bool sortPairsU32(uint32_t* keys, uint32_t* keysAlt, uint32_t* vals, uint32_t* valsAlt, uint32_t n, int endBit, DevBuf& tmp,
                  hipStream_t s, std::string& err, uint32_t** sorted);

// sort.hip: exclusive prefix sum of n u64 values into out (rocPRIM's)
bool scanU64(const uint64_t* in, uint64_t* out, size_t n, DevBuf& tmp, hipStream_t s, std::string& err);

// scan.hip: exclusive prefix sum of n u32 values into out, returns the total
bool scanU32(const uint32_t* in, uint32_t* out, uint32_t n, uint64_t* total, DevBuf& tmp, hipStream_t s, std::string& err);

}  // namespace kdn

// The context behind the C ABI's opaque handle.
struct kdn_ctx {
    kdn::Ctx c;
};

namespace kdn {

// integrator.hip: a round of slices (the render pass's and the serving
// loop's unit of work)
struct JobDev;
struct SliceDev;
// guides: a learned guide of each job's own (null entries: none; a job
// with one has one plain slice in the round)
bool runRound(Ctx& c, const std::vector<JobDev>& jobs, const std::vector<SliceDev>& slices, uint32_t flags, kdn_pass_stats* stats,
              const std::vector<void*>* guides = nullptr);
uint32_t planeOf(int width, int height);
// the path state for n paths in flight, allocated now (it grows on demand)
bool reservePaths(Ctx& c, size_t n);
// the device memory a context holds (scene, octree, integrator state)
uint64_t residentBytes(const Ctx& c);
// a number that tells loaded scenes apart for the life of the process
// (a freed scene's address may come back)
uint64_t sceneSerial(const kdn_scene* scene);
// free the octree build's scratch buffers
void releaseScratch(Ctx& c);

#define KDN_HIP(call)                                                                 \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            c.error = std::string(#call) + ": " + hipGetErrorString(e_);              \
            return false;                                                             \
        }                                                                             \
    } while (0)

}  // namespace kdn
