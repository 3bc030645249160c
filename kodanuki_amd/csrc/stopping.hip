// Adaptive stopping on the device: each pixel's trial judged against the
// tolerance and every pixel's count allotted from its neighbours' trials,
// as the reference's confidence-interval member decides them
// (plugins/stopping/confidence_interval/confidence_interval.cpp), then the
// counts scanned into the dense numbering the count pass draws its paths in.
//
// The judge is one lane per pixel: three channels of a geometric bisection
// for the widest interval whose graded width stays within the tolerance
// (each step two FP64 powers), so a 1080p frame is a few microseconds of
// FP64 work; the allotment is a 3x3 stencil over the judged trials. Both
// read and write a few bytes per pixel and are bound by HBM.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ctx.h"

namespace kdn {

namespace {

constexpr double kConfidence = 1.96;     // confidence_interval.cpp:15
constexpr double kSmallest = 1e-30;
constexpr double kLargest = 1e30;
constexpr int kRefinements = 16;
constexpr double kSharedCost = 1.5;      // confidence_interval.h:33
constexpr int kTrialBatches = 2;         // confidence_interval.h:18
constexpr int kMostSplit = 8;            // core/stopping.h:12
constexpr double kFilmicCeiling = 2.51 / 2.43;  // shared/image.cpp:13-15
constexpr double kFilmicSaturated = 16.0;
constexpr double kByteLevels = 255.0;

struct StopArgs {
    int32_t samples, batch;
    double tolerance;
    float exposure;
    int32_t curve;
    int32_t trial;        // samples each trial drew
    int32_t trial_split;  // samples that shared a first vertex in the trials
    int32_t split;        // the stopping's own split (0: the member's)
};

struct SpreadPair {
    float shared, own;
};

// displayed and stepsBetween (shared/image.cpp:25-50)
__device__ double displayed(double value, const StopArgs& a) {
    if (!(value > 0)) return 0;
    const double exposed = value * (double)a.exposure;
    double mapped = exposed;
    if (a.curve == KDN_CURVE_FILMIC)
        mapped = exposed >= kFilmicSaturated ? kFilmicCeiling
                                             : exposed * (2.51 * exposed + 0.03) / (exposed * (2.43 * exposed + 0.59) + 0.14);
    const double bounded = !(mapped > 0) ? 0.0 : (mapped < 1 ? mapped : 1.0);
    return pow(bounded, 1 / 2.2);
}
__device__ double steps_between(double low, double high, const StopArgs& a) {
    return fabs(displayed(high, a) - displayed(low, a)) * kByteLevels;
}

// The widest half width around a mean whose graded span stays within the
// bar: doubling or halving from the mean, then 16 geometric bisections
// (roomFits .. roomAround, confidence_interval.cpp:31-76).
__device__ bool room_fits(double mean, double bar, const StopArgs& a, double width) {
    return steps_between(mean - width, mean + width, a) <= bar;
}
__device__ double bisected_room(double mean, double bar, const StopArgs& a, double low, double high) {
    for (int step = 0; step < kRefinements; ++step) {
        const double middle = sqrt(low * high);
        const bool fits = room_fits(mean, bar, a, middle);
        low = fits ? middle : low;
        high = fits ? high : middle;
    }
    return low;
}
__device__ double room_around(double mean, double bar, const StopArgs& a) {
    const double start = fmax(fabs(mean), kSmallest);
    double low = start, high = start;
    if (room_fits(mean, bar, a, start)) {
        while (room_fits(mean, bar, a, high)) {
            if (high > kLargest) return INFINITY;
            low = high;
            high = high * 2;
        }
    } else {
        while (!room_fits(mean, bar, a, low)) {
            if (low < kSmallest) return 0;
            high = low;
            low = low * 0.5;
        }
    }
    return bisected_room(mean, bar, a, low, high);
}

// batchesNeeded with the prepared light's spread `unsure`
// (confidence_interval.cpp:85-111)
__device__ double batches_needed(const StopArgs& a, double mean, double spread, double unsure) {
    if (!(spread > 0)) return 0;
    const double bar = 2 * a.tolerance;
    const double whole = room_around(mean, bar, a);
    const double theirs = kConfidence * sqrt(unsure);
    double own = whole;
    if (theirs > 0) {
        const bool alone = steps_between(mean - theirs, mean + theirs, a) > bar;
        own = alone ? whole : sqrt(fmax(whole * whole - theirs * theirs, 0.0));
    }
    if (!(own > 0)) return INFINITY;
    return spread * (kConfidence / own) * (kConfidence / own);
}

// the spread of batch means (batchSpread / meansSpread)
__device__ double means_spread(double means, double squares, int batches) {
    const double count = (double)batches;
    const double mean = means / count;
    const double second = squares / count;
    return fmax(second - mean * mean, 0.0) * (count / (count - 1));
}

__host__ __device__ int split_ceiling(int batch) {
    int split = 1;
    while (split < kMostSplit && batch > 0 && batch % (split * 2) == 0) split = split * 2;
    return split;
}

// confidenceWanted and splitSpread of one pixel's trial (no prepared light)
__global__ void k_judge(const kdn_estimate* estimates, uint32_t pixels, StopArgs a, float* wanted, SpreadPair* spreads) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    const kdn_estimate e = estimates[p];
    double batches = 0;
    for (int c = 0; c < 3; ++c) {
        const double mean = e.total[c] * (1.0 / e.taken);
        const double spread = means_spread(e.means[c], e.squares[c], e.batches);
        const double needed = batches_needed(a, mean, spread, 0.0);
        batches = batches < needed ? needed : batches;
    }
    wanted[p] = (float)(batches * a.batch);
    SpreadPair s{0, 0};
    if (e.batches >= 2) {
        const double own = (e.means[0] + e.means[1] + e.means[2]) / 3 - e.shared_means;
        const double groups = (double)a.batch / (double)a.trial_split;
        s.shared = (float)(means_spread(e.shared_means, e.shared_squares, e.batches) * groups);
        s.own = (float)(means_spread(own, e.own_squares, e.batches) * (double)a.batch);
    }
    spreads[p] = s;
}

// confidenceSplit's power of two nearest the ideal split (splitNear)
__device__ int split_near(double ideal, int ceiling) {
    int split = 1;
    while (split < ceiling && split * sqrt(2.0) <= ideal) split = split * 2;
    return split;
}

// confidenceAllotted and splitFor from the eight neighbours' trials; the
// pixel's count pass draws the allotment less its trial, in whole groups,
// written at its slot of the 8x8-tiled plane the passes number paths by.
__global__ void k_allot(const float* wanted, const SpreadPair* spreads, int width, int height, StopArgs a, int32_t* allotted,
                        uint64_t* counts) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)width * (uint32_t)height) return;
    const int x = (int)(p % (uint32_t)width), y = (int)(p / (uint32_t)width);
    double sum = 0, shared = 0, own = 0;
    int count = 0;
    for (int row = y - 1 > 0 ? y - 1 : 0; row <= (y + 1 < height - 1 ? y + 1 : height - 1); ++row) {
        for (int column = x - 1 > 0 ? x - 1 : 0; column <= (x + 1 < width - 1 ? x + 1 : width - 1); ++column) {
            if (row == y && column == x) continue;
            const size_t n = (size_t)row * (size_t)width + (size_t)column;
            sum += wanted[n];
            shared += spreads[n].shared;
            own += spreads[n].own;
            count += 1;
        }
    }
    const int denominator = count > 1 ? count : 1;
    const SpreadPair around{(float)(shared / denominator), (float)(own / denominator)};
    // confidenceSplit
    const int ceiling = split_ceiling(a.batch);
    const int chosen = !(around.shared > 0) ? ceiling
                                            : split_near(sqrt((double)around.own * kSharedCost / (double)around.shared), ceiling);
    int samples = a.samples;
    if (count > 0) {
        // splitNoise
        const double tried = (double)around.shared * a.trial_split + (double)around.own;
        const double noise = !(tried > 0) ? 1.0 : ((double)around.shared * chosen + (double)around.own) / tried;
        const double w = sum / count * noise;
        const double batches = ceil(sqrt(w * a.samples) / a.batch);
        if (batches * a.batch < a.samples) {
            const int whole = (int)batches * a.batch;
            samples = whole > a.trial ? whole : a.trial;
        }
    }
    allotted[p] = samples;
    const int split = a.split > 0 ? a.split : chosen;  // splitFor
    const uint32_t tilesX = ((uint32_t)width + 7) / 8;
    const uint32_t slot = (((uint32_t)y / 8) * tilesX + (uint32_t)x / 8) * 64 + ((uint32_t)y % 8) * 8 + (uint32_t)x % 8;
    counts[slot] = (uint64_t)((samples - a.trial) / split * split);
}

// sumStart (core/frame.cpp:18-23): the count pass's estimates start from
// the trial's total and count alone
__global__ void k_sum_start(kdn_estimate* estimates, uint32_t pixels) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    kdn_estimate e = estimates[p];
    kdn_estimate s{};
    for (int c = 0; c < 3; ++c) s.total[c] = e.total[c];
    s.taken = e.taken;
    estimates[p] = s;
}

StopArgs argsOf(const kdn_stopping& s, int32_t trial) {
    StopArgs a{};
    a.samples = s.samples;
    a.batch = s.batch;
    a.tolerance = s.tolerance;
    a.exposure = s.exposure;
    a.curve = s.curve;
    a.trial = trial;
    a.trial_split = s.split;
    a.split = s.split;
    return a;
}

}  // namespace

// effortFor (core/stopping.cpp:18-34)
void stoppingEffort(double tolerance, int32_t* samples, int32_t* batch) {
    const double loosest = 62.5, tightest = 0.5, leastCeiling = 8, mostCeiling = 4096, deepestStride = 8, adaptiveRoom = 4;
    const double mostEffort = std::pow(2.0, 0.66);
    double effort = 1.0;
    if (tolerance > 0) {
        effort = std::log(loosest / tolerance) / std::log(loosest / tightest);
        effort = effort < 0 ? 0 : (effort > mostEffort ? mostEffort : effort);
    }
    const int stride = (int)std::lround(1 + (deepestStride - 1) * effort);
    const int b = (stride * stride + kMostSplit - 1) / kMostSplit * kMostSplit;
    const double room = tolerance > 0 ? adaptiveRoom : 1;
    const int ceiling = (int)std::lround(room * leastCeiling * std::pow(mostCeiling / leastCeiling, effort));
    *samples = b > ceiling ? b : ceiling;
    *batch = b;
}

// confidenceTrial (confidence_interval.cpp:26-29)
int32_t stoppingTrial(const kdn_stopping& s) {
    const int32_t samples = kTrialBatches * s.batch;
    return s.tolerance > 0 && samples < s.samples ? samples : 0;
}

bool stoppingCheck(const kdn_stopping& s, std::string& err) {
    if (s.samples <= 0 || s.batch <= 0) {
        err = "stopping: samples and batch must be positive";
        return false;
    }
    if (s.split != 1) {
        err = "stopping: the split must be 1 (every sample draws its own first vertex here)";
        return false;
    }
    if (s.curve != KDN_CURVE_FILMIC && s.curve != KDN_CURVE_CLAMPED) {
        err = "stopping: a curve is filmic or clamped";
        return false;
    }
    return true;
}

bool stoppingAllot(Ctx& c, const kdn_stopping& s, int32_t trial, uint64_t* total) {
    if (!stoppingCheck(s, c.error)) return false;
    if (!c.estimates || c.width <= 0 || c.height <= 0) {
        c.error = "kdn_stopping_allot: no trial estimates (a pass with flags bit 6 first)";
        return false;
    }
    if (trial <= 0 || trial % s.batch != 0 || trial >= s.samples) {
        c.error = "kdn_stopping_allot: the trial must be whole batches below the ceiling (kdn_stopping_trial)";
        return false;
    }
    hipStream_t st = c.stream;
    const uint32_t w = (uint32_t)c.width, h = (uint32_t)c.height, px = w * h;
    const uint32_t plane = (uint32_t)(((w + 7) / 8) * ((h + 7) / 8) * 64);
    if (!ensure(c.stop_wanted, px * sizeof(float), c.error, false)) return false;
    if (!ensure(c.stop_spreads, px * sizeof(SpreadPair), c.error, false)) return false;
    if (!ensure(c.stop_allotted, px * sizeof(int32_t), c.error, false)) return false;
    if (!ensure(c.stop_counts, (plane + 1) * sizeof(uint64_t), c.error, false)) return false;
    if (!ensure(c.stop_offsets, (plane + 1) * sizeof(uint64_t), c.error, false)) return false;
    const StopArgs a = argsOf(s, trial);
    hipLaunchKernelGGL(k_judge, dim3((px + 255) / 256), dim3(256), 0, st, (const kdn_estimate*)c.estimates, px, a,
                       (float*)c.stop_wanted.ptr, (SpreadPair*)c.stop_spreads.ptr);
    KDN_HIP(hipGetLastError());
    KDN_HIP(hipMemsetAsync(c.stop_counts.ptr, 0, (plane + 1) * sizeof(uint64_t), st));
    hipLaunchKernelGGL(k_allot, dim3((px + 255) / 256), dim3(256), 0, st, (const float*)c.stop_wanted.ptr,
                       (const SpreadPair*)c.stop_spreads.ptr, (int)w, (int)h, a, (int32_t*)c.stop_allotted.ptr,
                       (uint64_t*)c.stop_counts.ptr);
    KDN_HIP(hipGetLastError());
    // offsets[p] = the dense paths before slot p; offsets[plane] = them all
    uint64_t* counts = (uint64_t*)c.stop_counts.ptr;
    uint64_t* offsets = (uint64_t*)c.stop_offsets.ptr;
    if (!scanU64(counts, offsets, (size_t)plane + 1, c.stop_tmp, st, c.error)) return false;
    uint64_t all = 0;
    KDN_HIP(hipMemcpyAsync(&all, offsets + plane, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    KDN_HIP(hipStreamSynchronize(st));
    c.stop_total = all;
    c.stop_width = c.width;
    c.stop_height = c.height;
    c.stop_trial = trial;
    if (total) *total = all;
    return true;
}

bool stoppingSumStart(Ctx& c) {
    if (!c.estimates) return true;
    const uint32_t px = (uint32_t)c.width * (uint32_t)c.height;
    hipLaunchKernelGGL(k_sum_start, dim3((px + 255) / 256), dim3(256), 0, c.stream, c.estimates, px);
    KDN_HIP(hipGetLastError());
    return true;
}

bool readStopping(Ctx& c, float* wanted, float* spreads, int32_t* allotted) {
    if (c.stop_width <= 0 || c.stop_width != c.width || c.stop_height != c.height) {
        c.error = "kdn_read_stopping: no allotment for this image (kdn_stopping_allot first)";
        return false;
    }
    const size_t px = (size_t)c.width * (size_t)c.height;
    if (wanted) KDN_HIP(hipMemcpyAsync(wanted, c.stop_wanted.ptr, px * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    if (spreads) KDN_HIP(hipMemcpyAsync(spreads, c.stop_spreads.ptr, px * 2 * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    if (allotted) KDN_HIP(hipMemcpyAsync(allotted, c.stop_allotted.ptr, px * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    KDN_HIP(hipStreamSynchronize(c.stream));
    return true;
}

}  // namespace kdn
