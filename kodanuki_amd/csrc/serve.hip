// The serving loop: render jobs of mixed scenes, sizes and sample counts,
// admitted as they arrive and run together.
//
// A job is a render of one scene at one size: its tape and octree stay
// resident (one per scene and depth cap, shared by its jobs and kept while
// any job uses it), its FP64 accumulators live until the job is released.
// Each step is one round of the integrator: the admitted jobs' next slices
// (a fixed number of samples of every pixel, chosen from the job alone)
// fill the round's path budget, their paths share the wavefront queues,
// and each slice folds into its own job's accumulators in sample order.
// A job that arrives while a round runs joins the next round, so the work
// of jobs overlaps on the device; because a job's slices and their folds
// do not depend on the other jobs, its result is the same alone and in a
// mix, to the bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "ctx.h"
#include "integrator_kernels.h"

namespace kdn {

JobDev jobOf(const Ctx& scene, int width, int height, uint32_t seed, uint32_t surfaces, uint32_t minDepth);

namespace {

// Paths a slice of one job takes at most (slices of one sample may take
// more, for images beyond 2^21 pixels).
constexpr uint64_t kSlicePaths = 1ull << 21;

struct Resident {
    kdn_ctx* ctx = nullptr;  // the scene's own context: tape, octree, tables
    int users = 0;
    uint64_t bytes = 0;
    uint64_t lastRound = 0;
};

struct Job {
    uint64_t id = 0;
    kdn_job_desc desc{};
    std::pair<uint64_t, int> key;
    uint32_t state = KDN_JOB_WAITING;
    uint32_t sliceSpp = 1;
    uint32_t next = 0;     // first sample of its next slice
    uint32_t folded = 0;   // samples folded
    double submitted = 0, admitted = 0, finished = 0;
    std::string error;     // why it failed
    double* accum = nullptr;
    double* accumSq = nullptr;
    uint32_t* accumN = nullptr;
    DevBuf hg;             // the plain guide's pixel guides and deep filing (flags bit 0)
    size_t hgDeep = 0;     // the filing's byte offset in hg (0: none)
    void* guide = nullptr; // the job's own learned guide (flags bit 1)
    uint64_t bytes = 0;
};

}  // namespace
}  // namespace kdn

struct kdn_server {
    kdn_ctx* ctx = nullptr;  // the rounds' own context: path state, round tables, colour tables
    uint64_t budget = 0;
    uint32_t maxPaths = 0;
    std::chrono::steady_clock::time_point opened;
    uint64_t nextId = 1;
    uint64_t rounds = 0;
    std::map<uint64_t, kdn::Job> jobs;
    std::deque<uint64_t> waiting;     // in arrival order
    std::vector<uint64_t> running;    // in admission order
    std::map<std::pair<uint64_t, int>, kdn::Resident> scenes;
    std::map<std::pair<uint64_t, int>, uint64_t> sceneBytes;  // sizes of scenes built before
    std::string error;

    double now() const { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - opened).count(); }
};

namespace {

thread_local std::string g_serveError;

int failed(kdn_server* s, const std::string& message, int code) {
    if (s) s->error = message;
    g_serveError = message;
    return code;
}

uint64_t residentTotal(const kdn_server& s) {
    uint64_t bytes = 0;
    for (const auto& [key, r] : s.scenes) bytes += r.bytes;
    for (const auto& [id, j] : s.jobs) bytes += j.bytes;
    return bytes + kdn::residentBytes(s.ctx->c);
}

// Drop resident scenes no job uses, least recently used first, until
// `need` more bytes fit in the budget.
void evict(kdn_server& s, uint64_t need) {
    while (residentTotal(s) + need > s.budget) {
        auto victim = s.scenes.end();
        for (auto it = s.scenes.begin(); it != s.scenes.end(); ++it) {
            if (it->second.users == 0 && (victim == s.scenes.end() || it->second.lastRound < victim->second.lastRound)) victim = it;
        }
        if (victim == s.scenes.end()) return;
        kdn_close(victim->second.ctx);
        s.scenes.erase(victim);
    }
}

// The scene of a job resident on the device: upload and build once per
// (scene, depth cap).
int makeResident(kdn_server& s, kdn::Job& j) {
    auto it = s.scenes.find(j.key);
    if (it != s.scenes.end()) return KDN_OK;
    kdn::Resident r;
    int rc = kdn_open(s.ctx->c.device, &r.ctx);
    if (rc) return failed(&s, std::string("serving: ") + kdn_global_error(), rc);
    auto fail = [&](int code) {
        char message[512];
        kdn_last_error(r.ctx, message, sizeof message);
        kdn_close(r.ctx);
        return failed(&s, std::string("serving: ") + message, code);
    };
    if ((rc = kdn_upload_scene(r.ctx, j.desc.scene))) return fail(rc);
    if (r.ctx->c.scene_has_glass && !KDN_GLASS_ON_DEVICE) {
        kdn_close(r.ctx);
        return failed(&s, "serving: glass scattering is held off the device (renders with glass faulted it; the cause is not found)",
                      KDN_E_STATE);
    }
    float lower[3], size[3];
    if ((rc = kdn_scene_region(j.desc.scene, j.desc.depth, lower, size))) {
        kdn_close(r.ctx);
        return failed(&s, std::string("serving: ") + kdn_global_error(), rc);
    }
    if ((rc = kdn_build_octree(r.ctx, lower, size, j.desc.depth, nullptr))) return fail(rc);
    kdn::releaseScratch(r.ctx->c);
    r.bytes = kdn::residentBytes(r.ctx->c);
    s.scenes[j.key] = r;
    s.sceneBytes[j.key] = r.bytes;
    return KDN_OK;
}

void releaseJob(kdn_server& s, kdn::Job& j);

// A job that does not fit: it waits while other jobs run (KDN_E_MEMORY
// with no message), and fails when it would not fit even alone.
int tooBig(kdn_server& s, uint64_t need) {
    if (!s.running.empty()) return KDN_E_MEMORY;
    return failed(&s, "serving: the job does not fit the memory budget even alone (" + std::to_string(residentTotal(s) + need) +
                          " bytes needed, " + std::to_string(s.budget) + " budgeted)",
                  KDN_E_MEMORY);
}

// Admit a waiting job: its scene resident, its accumulators allocated and
// cleared. A scene's size is known once it has been built: a first build
// that does not fit beside the running jobs is dropped and its size kept,
// so that the job then waits without building again (the build itself
// may pass the budget while it runs).
int admit(kdn_server& s, kdn::Job& j) {
    const uint64_t px = (uint64_t)j.desc.width * j.desc.height;
    const bool guided = (j.desc.flags & 1) != 0;
    const uint32_t plane = kdn::planeOf(j.desc.width, j.desc.height);
    // samples a slice, fixed by the job and the server alone, never by the
    // other jobs
    uint64_t slicePaths = std::min<uint64_t>(kdn::kSlicePaths, s.maxPaths);
    // a learned job's slices are up to half a round where a round holds four
    // plain slices or more: it takes one slice a round, and fewer, larger
    // slices (each trained on before the next) take it through fewer rounds
    if ((j.desc.flags & 2) != 0 && s.maxPaths >= 4 * kdn::kSlicePaths)
        slicePaths = std::min<uint64_t>(2 * kdn::kSlicePaths, s.maxPaths / 2);
    const uint32_t sliceSpp = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(j.desc.spp, slicePaths / plane));
    // a learned job gets its guide when it has more than one slice (the
    // guide of a job of one slice would never steer a sample)
    const bool learned = (j.desc.flags & 2) != 0 && sliceSpp < j.desc.spp;
    // (a guided job's pixel guides, and its survey points' filing: about
    // 1 KB a point, one point to 64 pixels, and the 32^3 cell table; a
    // learned job's network, its Adam moments and gradients)
    const uint64_t bytes = px * (6 * sizeof(double) + sizeof(uint32_t) + (guided ? kdn::kHgFloats * sizeof(float) + 16 : 0)) +
                           (guided ? (1u << 18) : 0) + (learned ? kdn::guideBytes() : 0);
    bool built = false;
    if (!s.scenes.count(j.key)) {
        auto known = s.sceneBytes.find(j.key);
        if (known != s.sceneBytes.end()) {
            evict(s, known->second + bytes);
            if (residentTotal(s) + known->second + bytes > s.budget) return tooBig(s, known->second + bytes);
        }
        const int rc = makeResident(s, j);
        if (rc) return rc;
        built = true;
    }
    kdn::Resident& res = s.scenes[j.key];
    res.lastRound = s.rounds + 1;  // newest: evicted last
    res.users += 1;                // held while the job is admitted
    evict(s, bytes);
    if (residentTotal(s) + bytes > s.budget) {
        res.users -= 1;
        const int rc = tooBig(s, bytes);
        if (built && res.users == 0) {
            kdn_close(res.ctx);
            s.scenes.erase(j.key);
        }
        return rc;
    }
    kdn::Ctx& c = s.ctx->c;
    hipSetDevice(c.device);
    j.bytes = bytes;
    if (hipMalloc(&j.accum, px * 3 * sizeof(double)) != hipSuccess || hipMalloc(&j.accumSq, px * 3 * sizeof(double)) != hipSuccess ||
        hipMalloc(&j.accumN, px * sizeof(uint32_t)) != hipSuccess) {
        releaseJob(s, j);
        res.users -= 1;
        return failed(&s, "serving: accumulator allocation failed", KDN_E_MEMORY);
    }
    if (hipMemsetAsync(j.accum, 0, px * 3 * sizeof(double), c.stream) != hipSuccess ||
        hipMemsetAsync(j.accumSq, 0, px * 3 * sizeof(double), c.stream) != hipSuccess ||
        hipMemsetAsync(j.accumN, 0, px * sizeof(uint32_t), c.stream) != hipSuccess) {
        releaseJob(s, j);
        res.users -= 1;
        return failed(&s, "serving: accumulator clear failed", KDN_E_DEVICE);
    }
    if (guided) {
        // the survey of the job's image on its scene's context, in the
        // server's own scratch (kept for the next survey, and counted)
        kdn::Ctx& sc = res.ctx->c;
        const bool ok = kdn::hguideBuild(sc, j.desc.width, j.desc.height, nullptr, j.hg, c.scratch[5], &j.hgDeep);
        hipSetDevice(c.device);
        if (!ok) {
            const std::string message = "serving: the plain guide's survey failed: " + sc.error;
            releaseJob(s, j);
            res.users -= 1;
            return failed(&s, message, KDN_E_DEVICE);
        }
    }
    j.sliceSpp = sliceSpp;
    if (learned) {
        // the job's own guide over its scene's region, from the job's seed
        const kdn::Ctx& sc = res.ctx->c;
        j.guide = kdn::guideOpen(c, j.desc.seed, sc.tree.lower, sc.tree.size);
        if (!j.guide) {
            const std::string message = "serving: the learned guide could not be made: " + c.error;
            releaseJob(s, j);
            res.users -= 1;
            return failed(&s, message, KDN_E_MEMORY);
        }
    }
    j.state = KDN_JOB_RUNNING;
    j.admitted = s.now();
    return KDN_OK;
}

void releaseJob(kdn_server& s, kdn::Job& j) {
    if (j.accum) hipFree(j.accum);
    if (j.accumSq) hipFree(j.accumSq);
    if (j.accumN) hipFree(j.accumN);
    j.accum = j.accumSq = nullptr;
    j.accumN = nullptr;
    if (j.hg.ptr) hipFree(j.hg.ptr);
    j.hg = kdn::DevBuf{};
    j.hgDeep = 0;
    if (j.guide) kdn::guideClose(j.guide);
    j.guide = nullptr;
    j.bytes = 0;
}

}  // namespace

extern "C" {

const char* kdn_serve_error(const kdn_server* s) { return s ? s->error.c_str() : g_serveError.c_str(); }

const char* kdn_serve_job_error(const kdn_server* s, uint64_t id) {
    if (!s) return "";
    auto it = s->jobs.find(id);
    return it == s->jobs.end() ? "" : it->second.error.c_str();
}

int kdn_serve_open(kdn_ctx* ctx, uint64_t budget_bytes, uint32_t max_paths, kdn_server** out) {
    if (!ctx || !out) return failed(nullptr, "kdn_serve_open: null argument", KDN_E_ARGUMENT);
    if (max_paths < (1u << 16)) return failed(nullptr, "kdn_serve_open: a round needs room for 65536 paths at least", KDN_E_ARGUMENT);
    kdn_ctx* round = nullptr;
    int rc = kdn_open(ctx->c.device, &round);
    if (rc) return failed(nullptr, std::string("kdn_serve_open: ") + kdn_global_error(), rc);
    if (!kdn::reservePaths(round->c, max_paths)) {
        const std::string message = "kdn_serve_open: " + round->c.error;
        kdn_close(round);
        return failed(nullptr, message, KDN_E_MEMORY);
    }
    if (kdn::residentBytes(round->c) > budget_bytes) {
        const std::string message = "kdn_serve_open: the path state of a round (" + std::to_string(kdn::residentBytes(round->c)) +
                                    " bytes) does not fit the budget";
        kdn_close(round);
        return failed(nullptr, message, KDN_E_MEMORY);
    }
    kdn_server* s = new kdn_server();
    s->ctx = round;
    s->budget = budget_bytes;
    s->maxPaths = max_paths;
    s->opened = std::chrono::steady_clock::now();
    *out = s;
    return KDN_OK;
}

int kdn_serve_submit(kdn_server* s, const kdn_job_desc* job, uint64_t* id) {
    if (!s || !job || !job->scene || !id) return failed(s, "kdn_serve_submit: null argument", KDN_E_ARGUMENT);
    if (job->width <= 0 || job->height <= 0 || job->width > 16384 || job->height > 16384 || job->spp == 0 || job->surfaces == 0 ||
        job->depth < 1 || job->depth > kdn::kMaxDepth || (job->flags & ~3u) != 0) {
        return failed(s, "kdn_serve_submit: image size out of 1..16384, no samples, no surfaces, depth cap out of 1..11 or "
                         "unknown flags", KDN_E_ARGUMENT);
    }
    if ((job->flags & 3u) == 3u) {
        return failed(s, "kdn_serve_submit: a job is guided by the plain guide or by the learned guide, not both", KDN_E_ARGUMENT);
    }
    if ((uint64_t)kdn::planeOf(job->width, job->height) > s->maxPaths) {
        return failed(s, "kdn_serve_submit: one sample of the image needs more paths than a round holds", KDN_E_ARGUMENT);
    }
    kdn::Job j;
    j.id = s->nextId++;
    j.desc = *job;
    j.key = {kdn::sceneSerial(job->scene), job->depth};
    j.submitted = s->now();
    s->jobs[j.id] = j;
    s->waiting.push_back(j.id);
    *id = j.id;
    return KDN_OK;
}

int kdn_serve_step(kdn_server* s, kdn_serve_stats* stats) {
    if (!s) return failed(nullptr, "kdn_serve_step: null server", KDN_E_ARGUMENT);
    const auto start = std::chrono::steady_clock::now();
    kdn::Ctx& c = s->ctx->c;
    hipSetDevice(c.device);
    // admit waiting jobs in arrival order while memory allows
    while (!s->waiting.empty()) {
        kdn::Job& j = s->jobs[s->waiting.front()];
        const int rc = admit(*s, j);
        if (rc == KDN_E_MEMORY && !s->running.empty()) break;  // waits for running jobs to free memory
        if (rc) {
            // the job cannot run: it fails, the queue goes on
            j.state = KDN_JOB_FAILED;
            j.finished = s->now();
            j.error = s->error;
            s->waiting.pop_front();
            continue;
        }
        s->running.push_back(j.id);
        s->waiting.pop_front();
    }
    kdn_serve_stats local{};
    if (!s->running.empty()) {
        // the round: a slice of every running job in turn while paths last
        std::vector<kdn::JobDev> jobs;
        std::vector<kdn::SliceDev> slices;
        std::vector<uint64_t> owners;   // job id of each JobDev
        std::vector<void*> guides;      // each JobDev's own learned guide (null: none)
        std::vector<uint32_t> nextOf(s->running.size());
        for (size_t r = 0; r < s->running.size(); ++r) nextOf[r] = s->jobs[s->running[r]].next;
        std::vector<int> jobIndex(s->running.size(), -1);
        uint64_t paths = 0;
        // a slice of each job in turn, from a job that rotates with the
        // rounds, until the round is full or every job has its samples
        const size_t n = s->running.size(), first = s->rounds % n;
        for (bool added = true; added;) {
            added = false;
            for (size_t turn = 0; turn < n; ++turn) {
                const size_t r = (first + turn) % n;
                kdn::Job& j = s->jobs[s->running[r]];
                if (nextOf[r] >= j.desc.spp) continue;
                // a learned job's slices one a round: each slice is steered
                // by the guide as trained on the slices before it, alone or
                // in a mix
                if (j.guide && jobIndex[r] >= 0) continue;
                const uint32_t plane = kdn::planeOf(j.desc.width, j.desc.height);
                const uint32_t spp = std::min(j.sliceSpp, j.desc.spp - nextOf[r]);
                if (paths + (uint64_t)spp * plane > s->maxPaths) continue;
                if (jobIndex[r] < 0) {
                    const kdn::Resident& res = s->scenes[j.key];
                    jobIndex[r] = (int)jobs.size();
                    jobs.push_back(kdn::jobOf(res.ctx->c, j.desc.width, j.desc.height, j.desc.seed, j.desc.surfaces, j.desc.min_depth));
                    jobs.back().hg = (const float*)j.hg.ptr;
                    jobs.back().hg_deep = j.hgDeep ? (const kdn::HgDeepHeader*)((const char*)j.hg.ptr + j.hgDeep) : nullptr;
                    owners.push_back(j.id);
                    guides.push_back(j.guide);
                }
                slices.push_back(kdn::SliceDev{(uint32_t)jobIndex[r], nextOf[r], spp, (uint32_t)paths, j.accum, j.accumSq, j.accumN, nullptr, 0});
                paths += (uint64_t)spp * plane;
                nextOf[r] += spp;
                // a learned job's last slice: its guide is not trained after it
                if (j.guide) kdn::guideTrainNext(j.guide, nextOf[r] < j.desc.spp);
                added = true;
            }
        }
        if (!slices.empty()) {
            if (!kdn::runRound(c, jobs, slices, 0, &local.pass, &guides)) return failed(s, "serving: " + c.error, KDN_E_DEVICE);
            ++s->rounds;
            local.jobs_in_round = (uint32_t)jobs.size();
            local.slices_in_round = (uint32_t)slices.size();
            local.paths = paths;
            const double t = s->now();
            for (size_t r = 0; r < s->running.size(); ++r) {
                kdn::Job& j = s->jobs[s->running[r]];
                j.folded += nextOf[r] - j.next;
                j.next = nextOf[r];
                s->scenes[j.key].lastRound = s->rounds;
                if (j.folded >= j.desc.spp) {
                    j.state = KDN_JOB_DONE;
                    j.finished = t;
                }
            }
            std::vector<uint64_t> still;
            for (uint64_t id : s->running) {
                kdn::Job& j = s->jobs[id];
                if (j.state == KDN_JOB_DONE) {
                    s->scenes[j.key].users -= 1;
                } else {
                    still.push_back(id);
                }
            }
            s->running.swap(still);
        }
    }
    local.rounds = (uint32_t)s->rounds;
    local.waiting = (uint32_t)s->waiting.size();
    local.running = (uint32_t)s->running.size();
    local.resident_bytes = residentTotal(*s);
    local.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
    if (stats) *stats = local;
    return KDN_OK;
}

int kdn_serve_status(kdn_server* s, uint64_t id, kdn_job_status* status) {
    if (!s || !status) return failed(s, "kdn_serve_status: null argument", KDN_E_ARGUMENT);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return failed(s, "kdn_serve_status: no such job", KDN_E_ARGUMENT);
    const kdn::Job& j = it->second;
    status->state = j.state;
    status->samples_done = j.folded;
    status->submitted_ms = j.submitted;
    status->admitted_ms = j.admitted;
    status->finished_ms = j.finished;
    return KDN_OK;
}

int kdn_serve_result(kdn_server* s, uint64_t id, double* rgb_sum, double* sq_sum, uint32_t* count) {
    if (!s) return failed(nullptr, "kdn_serve_result: null server", KDN_E_ARGUMENT);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return failed(s, "kdn_serve_result: no such job", KDN_E_ARGUMENT);
    const kdn::Job& j = it->second;
    if (!j.accum) return failed(s, "kdn_serve_result: the job has not been admitted or was released", KDN_E_STATE);
    kdn::Ctx& c = s->ctx->c;
    hipSetDevice(c.device);
    const size_t px = (size_t)j.desc.width * j.desc.height;
    hipStreamSynchronize(c.stream);
    if ((rgb_sum && hipMemcpy(rgb_sum, j.accum, px * 3 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
        (sq_sum && hipMemcpy(sq_sum, j.accumSq, px * 3 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
        (count && hipMemcpy(count, j.accumN, px * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)) {
        return failed(s, "kdn_serve_result: copy failed", KDN_E_DEVICE);
    }
    return KDN_OK;
}

int kdn_serve_release(kdn_server* s, uint64_t id) {
    if (!s) return failed(nullptr, "kdn_serve_release: null server", KDN_E_ARGUMENT);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return failed(s, "kdn_serve_release: no such job", KDN_E_ARGUMENT);
    if (it->second.state == KDN_JOB_RUNNING) return failed(s, "kdn_serve_release: the job is still running", KDN_E_STATE);
    hipSetDevice(s->ctx->c.device);
    hipStreamSynchronize(s->ctx->c.stream);
    if (it->second.state == KDN_JOB_WAITING) {
        s->waiting.erase(std::remove(s->waiting.begin(), s->waiting.end(), id), s->waiting.end());
    }
    releaseJob(*s, it->second);
    s->jobs.erase(it);
    return KDN_OK;
}

void kdn_serve_close(kdn_server* s) {
    if (!s) return;
    hipSetDevice(s->ctx->c.device);
    hipStreamSynchronize(s->ctx->c.stream);
    for (auto& [id, j] : s->jobs) releaseJob(*s, j);
    for (auto& [key, r] : s->scenes) kdn_close(r.ctx);
    kdn_close(s->ctx);
    delete s;
}

}  // extern "C"
