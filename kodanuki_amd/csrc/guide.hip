// The learned guide on the device: its parameters (a hash grid of 16
// levels × 2^19 entries × 2 features, four 64x64 layers), their Adam
// state, the per-path lobes of a round, training records and samples; the
// round's part (inference before shading, collection and training after
// the folds) and the C ABI of the guide.
//
// Training: each round keeps records of the guided vertices of one path in
// a stride (at most 2^20 paths a round by default; point, outgoing direction, normal, hero wavelength, the
// direction taken and its density, the throughput after the vertex, the
// radiance the path held when its next vertex began). After the round a
// record's sample weighs the radiance that came back along the direction
// (final radiance less the snapshot, over the throughput, averaged over
// the live lanes) by its density: the Monte Carlo estimate of the
// gradient of the KL divergence from the incident radiance to the
// mixture (Neural Importance Sampling, Mueller et al. 2019). Adam steps
// (16 to 64 a round) on batches of up to 65536 samples, weights normalised
// by their mean, a loss
// scale on the backward pass (halved when a step meets a non-finite
// gradient, which skips the step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"
#include "guide_net.h"

namespace kdn {

constexpr uint32_t kGuideBatch = 1u << 16;
constexpr uint32_t kGuideStepsPerRound = 64;  // Adam steps of one round, at most
constexpr uint32_t kGuideMinSamples = 64;
constexpr float kGuideLearningRate = 1e-2f;
constexpr float kGuideBeta1 = 0.9f, kGuideBeta2 = 0.99f, kGuideEps = 1e-8f;

constexpr kdn_guide_config defaultConfig() {
    return kdn_guide_config{kGuideLearningRate, kGuideShare, 2.0f, 16, kGuideStepsPerRound, kGuideMaxRes, kGuideTrainPaths, 0, 1, 0};
}

struct GuideState {
    GuideView view{};
    uint32_t seed = 0;
    float lower[3] = {0, 0, 0}, size[3] = {1, 1, 1};
    float* grid = nullptr;  // params, then Adam m, v, then gradients
    uint64_t* grid_fixed = nullptr;  // the grid gradients as training sums them (fixed point, kept cleared)
    float* mlp = nullptr;   // the same for the layers
    uint16_t* mlp_bf = nullptr;
    float* scalars = nullptr;  // [0] weight scale, [1] loss, [2..] per-step skip flags (uint)
    DevBuf partial, lobes, recs, samples, flags, offsets, scan_tmp;
    uint32_t steps = 0;
    bool train_next = true;  // the next round keeps training records for it (serving: not after a job's last slice)
    float loss_scale = 128.0f;
    kdn_guide_config config = defaultConfig();
    double last_loss = 0;
    uint32_t last_samples = 0, last_steps = 0;
    float* gridM() const { return grid + kGuideGridParams; }
    float* gridV() const { return grid + 2 * kGuideGridParams; }
    float* gridG() const { return grid + 3 * kGuideGridParams; }
    float* mlpM() const { return mlp + kGuideMlpParams; }
    float* mlpV() const { return mlp + 2 * kGuideMlpParams; }
    float* mlpG() const { return mlp + 3 * kGuideMlpParams; }
};

// a round's lobes and records when its jobs' learned guides are their own
// (serving)
struct GuideRound {
    DevBuf lobes, recs, samples, flags, offsets, scan_tmp;
};

constexpr int kScalarFlags = 2;
constexpr int kScalarParts = kScalarFlags + 2 * kGuideStepsPerRound + 2;  // 256 doubles of weight sums from here
constexpr int kScalarWords = kScalarParts + 2 * 256;

namespace {

__device__ float init_unit(uint32_t seed, uint64_t index) {
    return (float)(mixed_bits((uint64_t)seed * 0x100000001B3ull ^ mixed_bits(index)) >> 40) * 0x1p-24f;
}

__global__ void k_guide_init(float* grid, float* mlp, uint16_t* bf, uint32_t seed) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kGuideGridParams) grid[i] = (init_unit(seed, i) * 2 - 1) * 1e-4f;
    if (i < (uint64_t)kGuideMlpParams) {
        const int layer = (int)(i >> 12);
        const float bound = 0.30618621784789724f * (layer == kGuideLayers - 1 ? 0.1f : 1.0f);  // sqrt(6 / 64)
        const float w = (init_unit(seed ^ 0x9E3779B9u, i) * 2 - 1) * bound;
        mlp[i] = w;
        bf[i] = to_bf16(w);
    }
}

// the sum of each weight's gradient over the training blocks' partial
// sums, in block order; a non-finite sum raises the step's flag
__global__ void k_guide_reduce(const float* partial, uint32_t tiles, float* grad, uint32_t* flag) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (uint32_t)kGuideMlpParams) return;
    float sum = 0;
    for (uint32_t t = 0; t < tiles; ++t) sum += partial[(size_t)t * kGuideMlpParams + e];
    grad[e] = sum;
    if (!(__builtin_fabsf(sum) <= 3.0e38f)) atomicOr(flag, 1u);
}

__global__ void k_guide_check(const float* g, uint64_t n, uint32_t* flag) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (!(__builtin_fabsf(g[i]) <= 3.0e38f)) {
            atomicOr(flag, 1u);
            return;
        }
    }
}

struct AdamStep {
    float lr, inv_scale, c1, c2;  // c1 = 1 - beta1^t, c2 = 1 - beta2^t
};

// the fixed-point grid gradients as floats, the sums cleared for the next
// step
__global__ void k_guide_unfix(uint64_t* fixed, float* grad, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        grad[i] = float_of_fixed(fixed[i]);
        fixed[i] = 0;
    }
}

// Adam on n parameters; gradients are cleared; with the step's flag
// raised nothing moves. bf, when set, receives the BF16 copy.
__global__ void k_guide_adam(float* p, float* m, float* v, float* g, uint64_t n, AdamStep a, const uint32_t* flag, uint16_t* bf) {
    const bool skip = *flag != 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const float grad = g[i] * a.inv_scale;
        g[i] = 0;
        if (skip) continue;
        const float mi = kGuideBeta1 * m[i] + (1 - kGuideBeta1) * grad;
        const float vi = kGuideBeta2 * v[i] + (1 - kGuideBeta2) * grad * grad;
        m[i] = mi;
        v[i] = vi;
        const float step = a.lr * (mi / a.c1) / (sqrtf(vi / a.c2) + kGuideEps);
        const float w = p[i] - step;
        p[i] = w;
        if (bf) bf[i] = to_bf16(w);
    }
}

// the samples of a round's records: weight = radiance back along the
// direction (final less snapshot, over the throughput, mean of the live
// lanes) over its density; valid[r] = 1 for a sample that carries some
__global__ void k_guide_collect(const float* recs, uint32_t nRec, uint32_t first, uint32_t stride, const float* rad, size_t N,
                                uint32_t product, float* samples, uint32_t* valid) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nRec) return;
    const float* rec = recs + (size_t)r * kRecFloats;
    uint32_t ok = 0;
    if (rec[kRecState] == 3.0f) {
        const size_t i = first + (size_t)(r / kGuideDepths) * stride;
        float sum = 0;
        int live = 0;
        // the incident radiance over the density, or (product) the
        // integrand f cos L over the density: radiance back over the
        // throughput before the vertex
        for (int l = 0; l < kLanes; ++l) {
            const float t = rec[(product ? kRecBefore : kRecThr) + l];
            if (!(t > 0) || !(rec[kRecThr + l] > 0)) continue;
            const float back = rad[l * N + i] - rec[kRecSnap + l];
            sum += (back > 0 ? back : 0.0f) / t;
            ++live;
        }
        const float pdf = rec[kRecPdf];
        const float w = live > 0 && pdf > 0 ? (product ? sum / live : sum / live / pdf) : 0.0f;
        if (w > 0 && w <= 1e30f) {
            float* s = samples + (size_t)r * kSamFloats;
            for (int k = 0; k < 13; ++k) s[k] = rec[k];
            s[kSamWeight] = w;
            s[kSamBsdf] = rec[kRecBsdf];
            s[15] = 0;
            ok = 1;
        }
    }
    valid[r] = ok;
}

__global__ void k_guide_compact(const float* in, const uint32_t* valid, const uint32_t* offsets, uint32_t n, float* out) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n || !valid[r]) return;
    const float* a = in + (size_t)r * kSamFloats;
    float* b = out + (size_t)offsets[r] * kSamFloats;
    for (int k = 0; k < kSamFloats; ++k) b[k] = a[k];
}

// the sum of the samples' weights in 256 parts (part t: samples t, t +
// 256, ...), then scale = loss scale / (mean weight × batch) from the
// parts in order: the same sum on every run
__global__ void k_guide_weight_parts(const float* samples, uint32_t n, double* parts) {
    const uint32_t t = threadIdx.x;
    double sum = 0;
    for (uint32_t i = t; i < n; i += 256) sum += samples[(size_t)i * kSamFloats + kSamWeight];
    parts[t] = sum;
}

__global__ void k_guide_scale(const double* parts, uint32_t n, uint32_t batch, float lossScale, float* scale) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double sum = 0;
    for (int t = 0; t < 256; ++t) sum += parts[t];
    const double mean = n ? sum / n : 0.0;
    *scale = mean > 0 ? (float)(lossScale / (mean * batch)) : 0.0f;
}

__global__ void k_guide_set_scale(float* scale, float value) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *scale = value;
}

GuideState* stateOf(Ctx& c) { return (GuideState*)c.guide; }

void viewOf(GuideState& G, uint32_t maxRes) {
    G.view.grid = G.grid;
    G.view.mlp = G.mlp_bf;
    const double growth = std::exp((std::log((double)maxRes) - std::log((double)kGuideMinRes)) / (kGuideLevels - 1));
    G.view.dense = 0;
    for (int l = 0; l < kGuideLevels; ++l) {
        const uint32_t res = (uint32_t)std::floor(kGuideMinRes * std::pow(growth, l) + 0.5);
        G.view.res[l] = res;
        const uint64_t side = res + 1;
        if (side * side * side <= kGuideTable) G.view.dense |= 1u << l;
    }
    for (int k = 0; k < 3; ++k) {
        G.view.lower[k] = G.lower[k];
        G.view.inv_size[k] = 1.0f / G.size[k];
    }
}

void deleteGuide(GuideState* G) {
    if (!G) return;
    (void)hipFree(G->grid);
    (void)hipFree(G->grid_fixed);
    (void)hipFree(G->mlp);
    (void)hipFree(G->mlp_bf);
    (void)hipFree(G->scalars);
    for (DevBuf* b : {&G->partial, &G->lobes, &G->recs, &G->samples, &G->flags, &G->offsets, &G->scan_tmp}) {
        if (b->ptr) (void)hipFree(b->ptr);
    }
    delete G;
}

GuideState* newGuide(std::string& error) {
    GuideState* G = new GuideState();
    if (hipMalloc(&G->grid, 4 * kGuideGridParams * sizeof(float)) != hipSuccess ||
        hipMalloc(&G->grid_fixed, kGuideGridParams * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc(&G->mlp, 4 * (size_t)kGuideMlpParams * sizeof(float)) != hipSuccess ||
        hipMalloc(&G->mlp_bf, (size_t)kGuideMlpParams * sizeof(uint16_t)) != hipSuccess ||
        hipMalloc(&G->scalars, kScalarWords * sizeof(float)) != hipSuccess) {
        error = "hipMalloc of the learned guide failed";
        deleteGuide(G);
        return nullptr;
    }
    return G;
}

bool allocGuide(Ctx& c) {
    if (c.guide) return true;
    c.guide = newGuide(c.error);
    return c.guide != nullptr;
}

// fresh parameters from a seed over a region, Adam's moments cleared
bool resetGuide(GuideState& G, hipStream_t st, uint32_t seed, const float lower[3], const float size[3], std::string& error) {
    G.seed = seed;
    for (int k = 0; k < 3; ++k) {
        G.lower[k] = lower[k];
        G.size[k] = size[k];
    }
    viewOf(G, G.config.max_resolution);
    G.view.share_cap = G.config.share;
    G.view.learn_select = G.config.learn_select;
    if (hipMemsetAsync(G.gridM(), 0, 3 * kGuideGridParams * sizeof(float), st) != hipSuccess ||
        hipMemsetAsync(G.grid_fixed, 0, kGuideGridParams * sizeof(uint64_t), st) != hipSuccess ||
        hipMemsetAsync(G.mlpM(), 0, 3 * (size_t)kGuideMlpParams * sizeof(float), st) != hipSuccess) {
        error = "clearing the learned guide failed";
        return false;
    }
    hipLaunchKernelGGL(k_guide_init, dim3((uint32_t)((kGuideGridParams + 255) / 256)), dim3(256), 0, st, G.grid, G.mlp, G.mlp_bf, seed);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
        error = "the learned guide's initialisation failed";
        return false;
    }
    G.steps = 0;
    G.config = defaultConfig();
    G.loss_scale = 128.0f;
    G.last_loss = 0;
    return true;
}


// Adam steps over the samples in HBM: batches of `batch` (modulo n) with
// weights normalised by their mean; the last step's loss in G.last_loss.
bool trainSteps(Ctx& c, GuideState& G, const float* samples, uint32_t n, uint32_t start, uint32_t steps, uint32_t batch,
                std::vector<double>* losses) {
    hipStream_t st = c.stream;
    batch = std::min(batch, n);
    steps = std::min(steps, kGuideStepsPerRound);
    if (steps == 0 || batch == 0) return true;
    const uint32_t tiles = guideTrainPartials(batch);
    if (!ensure(G.partial, (size_t)tiles * kGuideMlpParams * sizeof(float), c.error, false)) return false;
    KDN_HIP(hipMemsetAsync(G.scalars, 0, kScalarWords * sizeof(float), st));
    double* parts = (double*)(G.scalars + kScalarParts);
    hipLaunchKernelGGL(k_guide_weight_parts, dim3(1), dim3(256), 0, st, samples, n, parts);
    hipLaunchKernelGGL(k_guide_scale, dim3(1), dim3(1), 0, st, (const double*)parts, n, batch, G.loss_scale, G.scalars);
    uint32_t* flags = (uint32_t*)(G.scalars + kScalarFlags);
    float* lossOf = G.scalars + kScalarFlags + kGuideStepsPerRound;  // per-step losses
    for (uint32_t k = 0; k < steps; ++k) {
        const uint32_t first = (uint32_t)(((uint64_t)(start + k) * batch) % n);
        if (!launchGuideTrain(G.view, samples, n, first, batch, G.scalars, (float*)G.partial.ptr, G.grid_fixed, lossOf + k, flags + k,
                              st)) {
            c.error = "the guide's training launch failed";
            return false;
        }
        hipLaunchKernelGGL(k_guide_unfix, dim3(4096), dim3(256), 0, st, G.grid_fixed, G.gridG(), kGuideGridParams);
        hipLaunchKernelGGL(k_guide_reduce, dim3(kGuideMlpParams / 256), dim3(256), 0, st, (const float*)G.partial.ptr, tiles, G.mlpG(),
                           flags + k);
        hipLaunchKernelGGL(k_guide_check, dim3(2048), dim3(256), 0, st, (const float*)G.gridG(), kGuideGridParams, flags + k);
        const uint32_t t = G.steps + k + 1;
        const AdamStep a{G.config.learning_rate, 1.0f / G.loss_scale, (float)(1 - std::pow((double)kGuideBeta1, t)),
                         (float)(1 - std::pow((double)kGuideBeta2, t))};
        hipLaunchKernelGGL(k_guide_adam, dim3(kGuideMlpParams / 256), dim3(256), 0, st, G.mlp, G.mlpM(), G.mlpV(), G.mlpG(),
                           (uint64_t)kGuideMlpParams, a, (const uint32_t*)(flags + k), G.mlp_bf);
        hipLaunchKernelGGL(k_guide_adam, dim3(4096), dim3(256), 0, st, G.grid, G.gridM(), G.gridV(), G.gridG(), kGuideGridParams, a,
                           (const uint32_t*)(flags + k), (uint16_t*)nullptr);
    }
    KDN_HIP(hipGetLastError());
    float host[kScalarParts];
    KDN_HIP(hipMemcpyAsync(host, G.scalars, sizeof host, hipMemcpyDeviceToHost, st));
    KDN_HIP(hipStreamSynchronize(st));
    uint32_t skipped = 0;
    for (uint32_t k = 0; k < steps; ++k) {
        uint32_t f;
        std::memcpy(&f, &host[kScalarFlags + k], 4);
        if (f) {
            ++skipped;
        } else {
            ++G.steps;
        }
        // the loss of the step, without the loss scale
        const double loss = host[kScalarFlags + kGuideStepsPerRound + k] / G.loss_scale;
        if (losses) losses->push_back(f ? NAN : loss);
        G.last_loss = loss;
    }
    if (skipped) G.loss_scale = std::max(1.0f, G.loss_scale / 2);
    G.last_steps = steps - skipped;
    return true;
}

}  // namespace

bool guideReset(Ctx& c, uint32_t seed, const float lower[3], const float size[3]) {
    if (!allocGuide(c)) return false;
    return resetGuide(*stateOf(c), c.stream, seed, lower, size, c.error);
}

bool guideReady(const Ctx& c) { return c.guide != nullptr; }

void freeGuide(Ctx& c) {
    deleteGuide((GuideState*)c.guide);
    c.guide = nullptr;
    GuideRound* R = (GuideRound*)c.guide_round;
    if (R) {
        for (DevBuf* b : {&R->lobes, &R->recs, &R->samples, &R->flags, &R->offsets, &R->scan_tmp}) {
            if (b->ptr) (void)hipFree(b->ptr);
        }
        delete R;
        c.guide_round = nullptr;
    }
}

// Before a guided round: lobes for every path, records cleared.
bool guideBeginRound(Ctx& c, PassParams& P, uint32_t paths) {
    GuideState& G = *stateOf(c);
    const size_t N = c.paths->capacity;
    if (!ensure(G.lobes, N * kGuideOutputs * sizeof(float), c.error, false)) return false;
    uint32_t stride = 1;
    while ((uint64_t)stride * G.config.train_paths < paths) stride *= 2;
    const size_t nRec = (size_t)((paths + stride - 1) / stride) * kGuideDepths;
    if (!ensure(G.recs, nRec * kRecFloats * sizeof(float), c.error, false)) return false;
    KDN_HIP(hipMemsetAsync(G.recs.ptr, 0, nRec * kRecFloats * sizeof(float), c.stream));
    P.lobes = G.steps > 0 ? (const float*)G.lobes.ptr : nullptr;
    P.recs = (float*)G.recs.ptr;
    P.train_stride = stride;
    return true;
}

// The guide's lobes for the vertices of a queue about to be shaded.
bool guideInferQueue(Ctx& c, const PassParams& P, uint32_t queue, uint32_t count) {
    if (!P.lobes) return true;
    GuideState& G = *stateOf(c);
    if (!launchGuideInferPaths(G.view, P, queue, count, (float*)G.lobes.ptr, c.stream)) {
        c.error = "the guide's inference launch failed";
        return false;
    }
    return true;
}

// After a guided round: its records as samples, then Adam steps on them.
// The samples of `nRec` records (those of the paths first, first +
// stride, ...) packed in the buffers given, then Adam steps on them as the
// guide's config asks; the counts added to stats.
bool trainOnRecords(Ctx& c, GuideState& G, const float* recs, uint32_t nRec, uint32_t first, uint32_t stride, const float* rad,
                    size_t N, DevBuf& flagsBuf, DevBuf& offsetsBuf, DevBuf& samplesBuf, DevBuf& scanTmp, kdn_pass_stats* stats);

bool guideEndRound(Ctx& c, const PassParams& P, uint32_t paths, kdn_pass_stats* stats) {
    GuideState& G = *stateOf(c);
    const uint32_t stride = P.train_stride;
    const uint32_t nRec = ((paths + stride - 1) / stride) * kGuideDepths;
    return trainOnRecords(c, G, (const float*)G.recs.ptr, nRec, 0, stride, (const float*)P.s.rad, P.s.capacity, G.flags, G.offsets,
                          G.samples, G.scan_tmp, stats);
}

bool trainOnRecords(Ctx& c, GuideState& G, const float* recs, uint32_t nRec, uint32_t first, uint32_t stride, const float* rad,
                    size_t N, DevBuf& flagsBuf, DevBuf& offsetsBuf, DevBuf& samplesBuf, DevBuf& scanTmp, kdn_pass_stats* stats) {
    hipStream_t st = c.stream;
    if (!ensure(flagsBuf, (size_t)nRec * 4 + 64, c.error, false)) return false;
    if (!ensure(offsetsBuf, (size_t)nRec * 4 + 64, c.error, false)) return false;
    if (!ensure(samplesBuf, (size_t)nRec * 2 * kSamFloats * sizeof(float), c.error, false)) return false;
    float* raw = (float*)samplesBuf.ptr;
    float* packed = raw + (size_t)nRec * kSamFloats;
    hipLaunchKernelGGL(k_guide_collect, dim3((nRec + 255) / 256), dim3(256), 0, st, recs, nRec, first, stride, rad, N, G.config.product,
                       raw, (uint32_t*)flagsBuf.ptr);
    uint64_t total = 0;
    if (!scanU32((const uint32_t*)flagsBuf.ptr, (uint32_t*)offsetsBuf.ptr, nRec, &total, scanTmp, st, c.error)) return false;
    const uint32_t n = (uint32_t)total;
    hipLaunchKernelGGL(k_guide_compact, dim3((nRec + 255) / 256), dim3(256), 0, st, (const float*)raw, (const uint32_t*)flagsBuf.ptr,
                       (const uint32_t*)offsetsBuf.ptr, nRec, packed);
    KDN_HIP(hipGetLastError());
    G.last_samples = n;
    G.last_steps = 0;
    if (n >= kGuideMinSamples) {
        // config.epochs passes over the samples, in min_steps to max_steps steps
        const uint32_t batch = std::min(kGuideBatch, n);
        const uint32_t want = (uint32_t)std::ceil(G.config.epochs * n / (double)batch);
        const uint32_t steps = std::max(G.config.min_steps, std::min(G.config.max_steps, want));
        if (!trainSteps(c, G, packed, n, 0, steps, batch, nullptr)) return false;
    }
    if (stats) {
        stats->guide_samples += n;
        stats->guide_steps += G.last_steps;
        stats->guide_loss = G.last_loss;
    }
    return true;
}

// ---- the learned guides of served jobs --------------------------------------
// Each served job with the learned guide owns a GuideState over its scene's
// region, trained after every round on its own paths' records only; a
// round's lobes and records sit in the round context's GuideRound.

void* guideOpen(Ctx& c, uint32_t seed, const float lower[3], const float size[3]) {
    GuideState* G = newGuide(c.error);
    if (!G) return nullptr;
    if (!resetGuide(*G, c.stream, seed, lower, size, c.error)) {
        deleteGuide(G);
        return nullptr;
    }
    return G;
}

void guideClose(void* guide) { deleteGuide((GuideState*)guide); }

uint64_t guideBytes() {
    return 4 * kGuideGridParams * sizeof(float) + kGuideGridParams * sizeof(uint64_t) + 4 * (uint64_t)kGuideMlpParams * sizeof(float) +
           (uint64_t)kGuideMlpParams * sizeof(uint16_t) + kScalarWords * sizeof(float) +
           (uint64_t)guideTrainPartials(kGuideBatch) * kGuideMlpParams * sizeof(float);
}

uint32_t guideSteps(const void* guide) { return ((const GuideState*)guide)->steps; }

void guideTrainNext(void* guide, bool on) { ((GuideState*)guide)->train_next = on; }

uint64_t guideRoundBytes(const Ctx& c) {
    const GuideRound* R = (const GuideRound*)c.guide_round;
    if (!R) return 0;
    return R->lobes.bytes + R->recs.bytes + R->samples.bytes + R->flags.bytes + R->offsets.bytes + R->scan_tmp.bytes;
}

bool guidesBeginRound(Ctx& c, PassParams& P, std::vector<JobDev>& jobs, const std::vector<SliceDev>& slices,
                      const std::vector<void*>& guides) {
    if (!c.guide_round) c.guide_round = new GuideRound();
    GuideRound& R = *(GuideRound*)c.guide_round;
    const size_t N = c.paths->capacity;
    uint64_t blocks = 0;
    bool anyLearned = false;
    for (uint32_t j = 0; j < (uint32_t)jobs.size(); ++j) {
        if (!guides[j]) continue;
        const GuideState& G = *(const GuideState*)guides[j];
        const SliceDev* one = nullptr;
        for (const SliceDev& sl : slices) {
            if (sl.job != j) continue;
            if (one || sl.offsets) {
                c.error = "a job with the learned guide has one plain slice a round";
                return false;
            }
            one = &sl;
        }
        if (!one) continue;
        const uint32_t n = slice_paths(*one, jobs[j].plane);
        uint32_t stride = 1;
        while ((uint64_t)stride * G.config.train_paths < n) stride *= 2;
        jobs[j].learned = G.steps > 0 ? 1u : 0u;
        anyLearned = anyLearned || G.steps > 0;
        if (!G.train_next) continue;  // no records: nothing trains on this round
        jobs[j].train_stride = stride;
        jobs[j].path_first = one->first_path;
        jobs[j].rec_first = (uint32_t)blocks;
        blocks += (n + stride - 1) / stride;
    }
    if (anyLearned && !ensure(R.lobes, N * kGuideOutputs * sizeof(float), c.error, false)) return false;
    const size_t recBytes = (size_t)std::max<uint64_t>(blocks, 1) * kGuideDepths * kRecFloats * sizeof(float);
    if (!ensure(R.recs, recBytes, c.error, false)) return false;
    KDN_HIP(hipMemsetAsync(R.recs.ptr, 0, recBytes, c.stream));
    P.lobes = anyLearned ? (const float*)R.lobes.ptr : nullptr;
    P.recs = (float*)R.recs.ptr;
    P.train_stride = 0;
    return true;
}

bool guidesInferQueue(Ctx& c, const PassParams& P, uint32_t queue, uint32_t first, uint32_t count, void* guide) {
    if (!P.lobes || count == 0) return true;
    GuideState& G = *(GuideState*)guide;
    if (G.steps == 0) return true;  // its job samples by the BSDF this round
    PassParams Q = P;
    Q.s.queue[queue] = P.s.queue[queue] + first;
    if (!launchGuideInferPaths(G.view, Q, queue, count, (float*)((GuideRound*)c.guide_round)->lobes.ptr, c.stream)) {
        c.error = "the guide's inference launch failed";
        return false;
    }
    return true;
}

bool guidesEndRound(Ctx& c, const PassParams& P, const std::vector<JobDev>& jobs, const std::vector<SliceDev>& slices,
                    const std::vector<void*>& guides, kdn_pass_stats* stats) {
    GuideRound& R = *(GuideRound*)c.guide_round;
    for (uint32_t j = 0; j < (uint32_t)jobs.size(); ++j) {
        if (!guides[j] || jobs[j].train_stride == 0) continue;
        uint32_t n = 0;
        for (const SliceDev& sl : slices)
            if (sl.job == j) n = slice_paths(sl, jobs[j].plane);
        if (n == 0) continue;
        const uint32_t stride = jobs[j].train_stride;
        const uint32_t nRec = ((n + stride - 1) / stride) * kGuideDepths;
        const float* recs = (const float*)R.recs.ptr + (size_t)jobs[j].rec_first * kGuideDepths * kRecFloats;
        if (!trainOnRecords(c, *(GuideState*)guides[j], recs, nRec, jobs[j].path_first, stride, (const float*)P.s.rad, P.s.capacity,
                            R.flags, R.offsets, R.samples, R.scan_tmp, stats)) {
            return false;
        }
    }
    return true;
}

}  // namespace kdn

using namespace kdn;

namespace {

int guideFail(kdn_ctx* ctx, const std::string& m, int code) {
    ctx->c.error = m;
    return code;
}

}  // namespace

extern "C" {

int kdn_guide_reset(kdn_ctx* ctx, uint32_t seed, const float* lower3, const float* size3) {
    if (!ctx) return KDN_E_ARGUMENT;
    Ctx& c = ctx->c;
    hipSetDevice(c.device);
    float lower[3], size[3];
    if (lower3 && size3) {
        for (int k = 0; k < 3; ++k) {
            lower[k] = lower3[k];
            size[k] = size3[k];
            if (!(size[k] > 0)) return guideFail(ctx, "kdn_guide_reset: the region's sizes must be positive", KDN_E_ARGUMENT);
        }
    } else if (c.tree.built) {
        for (int k = 0; k < 3; ++k) {
            lower[k] = c.tree.lower[k];
            size[k] = c.tree.size[k];
        }
    } else {
        return guideFail(ctx, "kdn_guide_reset: no region given and no octree built", KDN_E_STATE);
    }
    return guideReset(c, seed, lower, size) ? KDN_OK : KDN_E_DEVICE;
}

int kdn_guide_describe(kdn_ctx* ctx, kdn_guide_info* info) {
    if (!ctx || !info) return KDN_E_ARGUMENT;
    GuideState* G = stateOf(ctx->c);
    if (!G) return guideFail(ctx, "kdn_guide_describe: the guide is not initialised (kdn_guide_reset)", KDN_E_STATE);
    std::memset(info, 0, sizeof *info);
    info->levels = kGuideLevels;
    info->features = kGuideFeatures;
    info->table_size = kGuideTable;
    info->width = kGuideWidth;
    info->layers = kGuideLayers;
    info->lobes = kGuideLobes;
    info->depths = kGuideDepths;
    info->train_paths = G->config.train_paths;
    for (int l = 0; l < kGuideLevels; ++l) info->resolution[l] = G->view.res[l];
    info->dense_levels = G->view.dense;
    info->steps = G->steps;
    info->loss_scale = G->loss_scale;
    info->share = G->config.share;
    info->last_loss = G->last_loss;
    info->last_samples = G->last_samples;
    info->grid_params = kGuideGridParams;
    info->mlp_params = kGuideMlpParams;
    for (int k = 0; k < 3; ++k) {
        info->lower[k] = G->lower[k];
        info->size[k] = G->size[k];
    }
    return KDN_OK;
}

int kdn_guide_configure(kdn_ctx* ctx, const kdn_guide_config* config) {
    if (!ctx || !config) return KDN_E_ARGUMENT;
    GuideState* G = stateOf(ctx->c);
    if (!G) return guideFail(ctx, "kdn_guide_configure: the guide is not initialised (kdn_guide_reset)", KDN_E_STATE);
    const kdn_guide_config& k = *config;
    if (!(k.learning_rate > 0 && k.learning_rate < 1) || !(k.share >= 0 && k.share <= 1) || !(k.epochs > 0 && k.epochs <= 64) ||
        k.min_steps > k.max_steps || k.max_steps > kGuideStepsPerRound || k.max_resolution < kGuideMinRes ||
        k.max_resolution > 65536 || k.train_paths < 1024 || k.train_paths > (1u << 22) || k.learn_select > 1 || k.product > 1) {
        return guideFail(ctx, "kdn_guide_configure: learning rate in (0, 1), share in [0, 1], epochs in (0, 64], "
                              "min_steps <= max_steps <= 64, max_resolution in [16, 65536], train_paths in [2^10, 2^22]",
                         KDN_E_ARGUMENT);
    }
    G->config = k;
    viewOf(*G, k.max_resolution);
    G->view.share_cap = k.share;
    G->view.learn_select = k.learn_select;
    return KDN_OK;
}

int kdn_guide_params(kdn_ctx* ctx, float* grid, float* mlp) {
    if (!ctx) return KDN_E_ARGUMENT;
    Ctx& c = ctx->c;
    GuideState* G = stateOf(c);
    if (!G) return guideFail(ctx, "kdn_guide_params: the guide is not initialised", KDN_E_STATE);
    hipSetDevice(c.device);
    if (hipStreamSynchronize(c.stream) != hipSuccess ||
        (grid && hipMemcpy(grid, G->grid, kGuideGridParams * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) ||
        (mlp && hipMemcpy(mlp, G->mlp, kGuideMlpParams * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)) {
        return guideFail(ctx, "kdn_guide_params: copy failed", KDN_E_DEVICE);
    }
    return KDN_OK;
}

int kdn_guide_set_params(kdn_ctx* ctx, const float* grid, const float* mlp) {
    if (!ctx) return KDN_E_ARGUMENT;
    Ctx& c = ctx->c;
    GuideState* G = stateOf(c);
    if (!G) return guideFail(ctx, "kdn_guide_set_params: the guide is not initialised", KDN_E_STATE);
    hipSetDevice(c.device);
    if (grid && hipMemcpy(G->grid, grid, kGuideGridParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        return guideFail(ctx, "kdn_guide_set_params: copy failed", KDN_E_DEVICE);
    }
    if (mlp) {
        std::vector<uint16_t> bf(kGuideMlpParams);
        for (int i = 0; i < kGuideMlpParams; ++i) {
            uint32_t u;
            std::memcpy(&u, &mlp[i], 4);
            if ((u & 0x7FFFFFFFu) > 0x7F800000u) {
                bf[i] = (uint16_t)((u >> 16) | 0x40u);
            } else {
                u += 0x7FFFu + ((u >> 16) & 1u);
                bf[i] = (uint16_t)(u >> 16);
            }
        }
        if (hipMemcpy(G->mlp, mlp, kGuideMlpParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(G->mlp_bf, bf.data(), kGuideMlpParams * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess) {
            return guideFail(ctx, "kdn_guide_set_params: copy failed", KDN_E_DEVICE);
        }
    }
    return KDN_OK;
}

int kdn_guide_infer(kdn_ctx* ctx, const kdn_guide_query* queries, uint32_t n, kdn_guide_lobe* lobes, float* select) {
    if (!ctx || (n && (!queries || !lobes))) return KDN_E_ARGUMENT;
    Ctx& c = ctx->c;
    GuideState* G = stateOf(c);
    if (!G) return guideFail(ctx, "kdn_guide_infer: the guide is not initialised", KDN_E_STATE);
    if (n > (1u << 24)) return guideFail(ctx, "kdn_guide_infer: at most 2^24 queries a call", KDN_E_ARGUMENT);
    if (n == 0) return KDN_OK;
    hipSetDevice(c.device);
    const size_t lobeBytes = (size_t)n * kGuideLobes * sizeof(kdn_guide_lobe);
    if (!ensure(c.scratch[6], (size_t)n * (sizeof(kdn_guide_query) + sizeof(float)) + lobeBytes, c.error, false)) return KDN_E_MEMORY;
    char* base = (char*)c.scratch[6].ptr;
    float* q = (float*)base;
    float* out = (float*)(base + (size_t)n * sizeof(kdn_guide_query));
    float* sel = (float*)((char*)out + lobeBytes);
    if (hipMemcpyAsync(q, queries, (size_t)n * sizeof(kdn_guide_query), hipMemcpyHostToDevice, c.stream) != hipSuccess ||
        !launchGuideInferQueries(G->view, q, n, out, sel, c.stream) ||
        hipMemcpyAsync(lobes, out, lobeBytes, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        (select && hipMemcpyAsync(select, sel, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c.stream) != hipSuccess) ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
        return guideFail(ctx, "kdn_guide_infer: device call failed", KDN_E_DEVICE);
    }
    return KDN_OK;
}

int kdn_guide_gradients(kdn_ctx* ctx, const kdn_guide_sample* samples, uint32_t n, float* grid_grad, float* mlp_grad, double* loss) {
    if (!ctx || !samples || n == 0) return KDN_E_ARGUMENT;
    Ctx& c = ctx->c;
    GuideState* G = stateOf(c);
    if (!G) return guideFail(ctx, "kdn_guide_gradients: the guide is not initialised", KDN_E_STATE);
    if (n > (1u << 20)) return guideFail(ctx, "kdn_guide_gradients: at most 2^20 samples a call", KDN_E_ARGUMENT);
    hipSetDevice(c.device);
    hipStream_t st = c.stream;
    const uint32_t tiles = guideTrainPartials(n);
    if (!ensure(c.scratch[6], (size_t)n * sizeof(kdn_guide_sample), c.error, false) ||
        !ensure(G->partial, (size_t)tiles * kGuideMlpParams * sizeof(float), c.error, false)) {
        return KDN_E_MEMORY;
    }
    float* s = (float*)c.scratch[6].ptr;
    float host[kScalarParts];
    if (hipMemcpyAsync(s, samples, (size_t)n * sizeof(kdn_guide_sample), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemsetAsync(G->scalars, 0, kScalarWords * sizeof(float), st) != hipSuccess) {
        return guideFail(ctx, "kdn_guide_gradients: device call failed", KDN_E_DEVICE);
    }
    hipLaunchKernelGGL(k_guide_set_scale, dim3(1), dim3(64), 0, st, G->scalars, 1.0f);
    if (!launchGuideTrain(G->view, s, n, 0, n, G->scalars, (float*)G->partial.ptr, G->grid_fixed, G->scalars + 1,
                          (uint32_t*)(G->scalars + kScalarFlags), st)) {
        return guideFail(ctx, "kdn_guide_gradients: launch failed", KDN_E_DEVICE);
    }
    hipLaunchKernelGGL(k_guide_unfix, dim3(4096), dim3(256), 0, st, G->grid_fixed, G->gridG(), kGuideGridParams);
    hipLaunchKernelGGL(k_guide_reduce, dim3(kGuideMlpParams / 256), dim3(256), 0, st, (const float*)G->partial.ptr, tiles, G->mlpG(),
                       (uint32_t*)(G->scalars + kScalarFlags));
    if (hipGetLastError() != hipSuccess ||
        (mlp_grad && hipMemcpyAsync(mlp_grad, G->mlpG(), kGuideMlpParams * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess) ||
        (grid_grad && hipMemcpyAsync(grid_grad, G->gridG(), kGuideGridParams * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipMemcpyAsync(host, G->scalars, sizeof host, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemsetAsync(G->gridG(), 0, kGuideGridParams * sizeof(float), st) != hipSuccess ||
        hipMemsetAsync(G->mlpG(), 0, kGuideMlpParams * sizeof(float), st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
        return guideFail(ctx, "kdn_guide_gradients: device call failed", KDN_E_DEVICE);
    }
    if (loss) *loss = host[1];
    return KDN_OK;
}

int kdn_guide_train(kdn_ctx* ctx, const kdn_guide_sample* samples, uint32_t n, uint32_t steps, uint32_t batch, double* losses) {
    if (!ctx || !samples || n == 0 || steps == 0 || batch == 0) return KDN_E_ARGUMENT;
    Ctx& c = ctx->c;
    GuideState* G = stateOf(c);
    if (!G) return guideFail(ctx, "kdn_guide_train: the guide is not initialised", KDN_E_STATE);
    if (n > (1u << 22) || batch > (1u << 20)) {
        return guideFail(ctx, "kdn_guide_train: at most 2^22 samples and batches of 2^20", KDN_E_ARGUMENT);
    }
    hipSetDevice(c.device);
    if (!ensure(c.scratch[6], (size_t)n * sizeof(kdn_guide_sample), c.error, false)) return KDN_E_MEMORY;
    float* s = (float*)c.scratch[6].ptr;
    if (hipMemcpyAsync(s, samples, (size_t)n * sizeof(kdn_guide_sample), hipMemcpyHostToDevice, c.stream) != hipSuccess) {
        return guideFail(ctx, "kdn_guide_train: copy failed", KDN_E_DEVICE);
    }
    for (uint32_t done = 0; done < steps;) {
        const uint32_t now = std::min(steps - done, kGuideStepsPerRound);
        std::vector<double> l;
        if (!trainSteps(c, *G, s, n, done, now, batch, &l)) return KDN_E_DEVICE;
        for (uint32_t k = 0; k < now && losses; ++k) losses[done + k] = l[k];
        done += now;
    }
    return KDN_OK;
}

}  // extern "C"
