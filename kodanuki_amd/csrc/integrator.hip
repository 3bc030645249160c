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
#include "integrator_dev.h"

namespace kdn {

TreeView treeView(const Ctx& c);

struct PathState {
    kdn_ray* rays;
    kdn_hit* hits;
    kdn_ray* shadow;
    float* srad;       // 4 lanes, lane-major
    float* thr;        // 4 lanes
    float* rad;        // 4 lanes
    float* nm;         // 4 lanes
    float* dens;       // 4 lanes
    float* pdf;
    uint32_t* meta;    // depth | flags << 8
    uint32_t* pixel;
    uint32_t* sample;
    uint64_t* stream;
    uint32_t* queue[3];
    uint32_t* counters;
    size_t capacity;
};

enum : uint32_t { FlagExact = 1u << 8, FlagDispersed = 1u << 9 };

struct PassParams {
    int width, height;
    uint32_t pixels;
    uint32_t first_sample, chunk;
    uint64_t seed_bits;
    uint32_t surfaces, min_depth;
    float eye[3], forward[3], right[3], up[3];
    float reach, aspect;
    float sky_zenith[3], sky_horizon[3], sky_ground[3];
    uint32_t n_lamps, n_materials, fallback_material;
    const LampDev* lamps;
    const MaterialDev* materials;
    const uint32_t* leaf_material;
    const ColourTables* colour;
    PathState s;
};

KDN_D float lane(const float* a, uint32_t l, size_t i, size_t n) { return a[l * n + i]; }

// Pixel of path slot i: the chunk's sample planes one after another, each
// plane in 8x8 tiles so that a wavefront's camera rays are neighbours.
KDN_D void pixel_of(const PassParams& P, uint32_t p, uint32_t* x, uint32_t* y) {
    const uint32_t tilesX = (P.width + 7) / 8;
    const uint32_t tile = p / 64, in = p % 64;
    const uint32_t tx = tile % tilesX, ty = tile / tilesX;
    *x = tx * 8 + (in % 8);
    *y = ty * 8 + (in / 8);
}

__global__ void k_generate(PassParams P, uint32_t slots) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    PathState& s = P.s;
    const size_t N = s.capacity;
    const uint32_t tilesX = (P.width + 7) / 8, tilesY = (P.height + 7) / 8;
    const uint32_t plane = tilesX * tilesY * 64;
    const uint32_t sIdx = i / plane, p = i % plane;
    uint32_t x, y;
    pixel_of(P, p, &x, &y);
    const bool live = x < (uint32_t)P.width && y < (uint32_t)P.height;
    s.pixel[i] = live ? y * P.width + x : 0xFFFFFFFFu;
    if (!live) {
        for (int l = 0; l < kLanes; ++l) s.rad[l * N + i] = 0;
        return;
    }
    const uint32_t sample = P.first_sample + sIdx;
    // pixelNumbers and pixelWavelengths (trace.cpp:65-84)
    const uint64_t place = ((uint64_t)x << 32) | (uint64_t)y;
    const uint64_t stream = P.seed_bits ^ mixed_bits(place);
    const double start = hashed(stream, ~0ull, 0);
    const Lanes w = sampled_wavelengths(wavelength_share(start, sample), P.colour);
    // filmDirection (core/camera.cpp:32-38) at the jittered film point
    const double fx = x + hashed(stream, sample, PixelX), fy = y + hashed(stream, sample, PixelY);
    const float u = (float)(2 * fx / P.width - 1), v = (float)(1 - 2 * fy / P.height);
    V3 d = v3(P.forward[0], P.forward[1], P.forward[2]) + v3(P.right[0], P.right[1], P.right[2]) * (u * P.reach * P.aspect) +
           v3(P.up[0], P.up[1], P.up[2]) * (v * P.reach);
    d = normalized(d);
    kdn_ray r;
    r.origin[0] = P.eye[0];
    r.origin[1] = P.eye[1];
    r.origin[2] = P.eye[2];
    r.direction[0] = d.x;
    r.direction[1] = d.y;
    r.direction[2] = d.z;
    r.far_ = kInf;
    r.inside = 0;
    r.within = KDN_NO_MATERIAL;
    s.rays[i] = r;
    for (int l = 0; l < kLanes; ++l) {
        s.thr[l * N + i] = 1.0f;
        s.rad[l * N + i] = 0.0f;
        s.nm[l * N + i] = w.nm[l];
        s.dens[l * N + i] = w.density[l];
    }
    s.pdf[i] = 0;
    s.meta[i] = FlagExact;
    s.sample[i] = sample;
    s.stream[i] = stream;
    const uint32_t at = atomicAdd(&s.counters[0], 1u);
    s.queue[0][at] = i;
}

__global__ void __launch_bounds__(256) k_trace_queue(TreeView tv, const uint32_t* queue, const uint32_t* count, const kdn_ray* rays,
                                                     kdn_hit* hits) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= *count) return;
    const uint32_t i = queue[q];
    Dl buf[kMaxRegs];
    kdn_hit h;
    trace_ray(tv, rays[i], buf, h);
    hits[i] = h;
}

// The same march, adding the FP32 operations of every evaluation to *work
// (the instrumented pass the roofline's algorithmic work comes from).
__global__ void __launch_bounds__(256) k_trace_queue_count(TreeView tv, const uint32_t* queue, const uint32_t* count,
                                                           const kdn_ray* rays, kdn_hit* hits, double* work) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= *count) return;
    const uint32_t i = queue[q];
    Dl buf[kMaxRegs];
    kdn_hit h;
    double w = 0;
    TreeView counted = tv;
    counted.work = &w;
    trace_ray(counted, rays[i], buf, h);
    hits[i] = h;
    atomicAdd(work, w);
}

KDN_D float lift_at(V3 x) {
    return 8e-5f + 8e-6f * fmax_(fmax_(__builtin_fabsf(x.x), __builtin_fabsf(x.y)), __builtin_fabsf(x.z));
}

__global__ void __launch_bounds__(256) k_shade(PassParams P, uint32_t inQueue, uint32_t outQueue) {
    PathState& s = P.s;
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= s.counters[inQueue]) return;
    const size_t N = s.capacity;
    const uint32_t i = s.queue[inQueue][q];
    const kdn_hit hit = s.hits[i];
    const kdn_ray ray = s.rays[i];
    const uint64_t stream = s.stream[i];
    const uint32_t sample = s.sample[i];
    uint32_t meta = s.meta[i];
    const int depth = (int)(meta & 0xFF);
    Lanes w;
    float thr[kLanes], rad[kLanes];
    for (int l = 0; l < kLanes; ++l) {
        w.nm[l] = s.nm[l * N + i];
        w.density[l] = s.dens[l * N + i];
        thr[l] = s.thr[l * N + i];
        rad[l] = s.rad[l * N + i];
    }
    const V3 o = v3(ray.origin[0], ray.origin[1], ray.origin[2]);
    const V3 d = v3(ray.direction[0], ray.direction[1], ray.direction[2]);
    const bool inside = ray.inside != 0;
    const float tHit = hit.kind == KDN_HIT_MISSED ? kInf : hit.distance;
    const bool exactBefore = (meta & FlagExact) != 0;
    const float pdfBefore = s.pdf[i];

    // lamps met along the ray (metAlong, trace.cpp:352-374), MIS against aiming
    if (!inside) {
        for (uint32_t j = 0; j < P.n_lamps; ++j) {
            const LampDev& L = P.lamps[j];
            LampHit lh;
            if (!meet_quad(L, o, d, tHit, &lh)) continue;
            float Le[kLanes];
            spectrum_of(L.radiance, w, Le);
            const float pdfL = L.pick * lh.t * lh.t / (lh.cosine * L.area);
            const float wt = exactBefore ? 1.0f : power_weight(pdfBefore, pdfL);
            for (int l = 0; l < kLanes; ++l) rad[l] += thr[l] * Le[l] * wt;
        }
    }
    bool alive = false;
    if (hit.kind == KDN_HIT_MISSED) {
        if (!inside) {
            float sky[kLanes];
            sky_radiance(P.sky_zenith, P.sky_horizon, P.sky_ground, d, w, sky);
            for (int l = 0; l < kLanes; ++l) rad[l] += thr[l] * sky[l];
        }
    } else if (hit.kind == KDN_HIT_SURFACE) {
        const uint32_t mi = hit.leaf != KDN_NO_LEAF ? P.leaf_material[hit.leaf] : P.fallback_material;
        const MaterialDev& m = P.materials[mi < P.n_materials ? mi : P.fallback_material];
        const V3 x = v3(hit.point[0], hit.point[1], hit.point[2]);
        const V3 out = -d;
        V3 n = v3(hit.normal[0], hit.normal[1], hit.normal[2]);
        if (dot(n, out) < 0) n = -n;
        if (m.glow[0] > 0 || m.glow[1] > 0 || m.glow[2] > 0) {
            float g[kLanes];
            spectrum_of(m.glow, w, g);
            for (int l = 0; l < kLanes; ++l) rad[l] += thr[l] * g[l];
        }
        const float lift = lift_at(x);
        const bool exact = m.kind == KDN_GLASS && m.exact;
        // aim at one lamp (aimedFrom, trace.cpp:896-947), MIS against meeting
        if (!exact && P.n_lamps > 0) {
            const float uPick = number(stream, sample, depth, Pick);
            uint32_t j = 0;
            while (j + 1 < P.n_lamps && uPick >= P.lamps[j].pick_cdf) ++j;
            const LampDev& L = P.lamps[j];
            const float u = number(stream, sample, depth, LightU), v = number(stream, sample, depth, LightV);
            const V3 y = v3(L.at[0], L.at[1], L.at[2]) + v3(L.along[0], L.along[1], L.along[2]) * u + v3(L.across[0], L.across[1], L.across[2]) * v;
            const V3 toward = y - x;
            const float dist = length(toward);
            const V3 in = toward * (1.0f / dist);
            const float cosL = -dot(v3(L.facing[0], L.facing[1], L.facing[2]), in);
            const float cosS = dot(n, in);
            if (cosL > 0 && cosS > 0 && dist > 2 * lift) {
                float f[kLanes], Le[kLanes];
                rough_value(m, w, n, out, in, f);
                spectrum_of(L.radiance, w, Le);
                const float pdfL = L.pick * dist * dist / (cosL * L.area);
                const float wt = power_weight(pdfL, rough_density(m, n, out, in));
                float c[kLanes];
                float any = 0;
                for (int l = 0; l < kLanes; ++l) {
                    c[l] = thr[l] * f[l] * cosS * Le[l] / pdfL * wt;
                    any += c[l];
                }
                if (any > 0) {
                    kdn_ray sr;
                    const V3 so = x + n * lift;
                    sr.origin[0] = so.x;
                    sr.origin[1] = so.y;
                    sr.origin[2] = so.z;
                    sr.direction[0] = in.x;
                    sr.direction[1] = in.y;
                    sr.direction[2] = in.z;
                    sr.far_ = dist - lift;
                    sr.inside = 0;
                    sr.within = KDN_NO_MATERIAL;
                    s.shadow[i] = sr;
                    for (int l = 0; l < kLanes; ++l) s.srad[l * N + i] = c[l];
                    const uint32_t at = atomicAdd(&s.counters[2], 1u);
                    s.queue[2][at] = i;
                }
            }
        }
        // scatter (scatteredFrom, trace.cpp:1020-1074)
        if (depth + 1 < (int)P.surfaces) {
            const float uLobe = number(stream, sample, depth, Lobe), uSide = number(stream, sample, depth, Side);
            const float du = number(stream, sample, depth, DirectionU), dv = number(stream, sample, depth, DirectionV);
            const Scatter sc = m.kind == KDN_GLASS ? glass_scatter(m, w, n, out, inside, uSide) : rough_scatter(m, w, n, out, uLobe, du, dv);
            if (sc.found) {
                const float side = dot(sc.in, n);
                const bool below = sc.through ? side >= 0 : side <= 0;
                if (!below) {
                    float carried[kLanes];
                    for (int l = 0; l < kLanes; ++l) carried[l] = thr[l] * sc.weight[l];
                    if (sc.collapse && !(meta & FlagDispersed)) {
                        carried[0] *= kLanes;
                        carried[1] = carried[2] = carried[3] = 0;
                        meta |= FlagDispersed;
                    }
                    bool survives = true;
                    if (depth >= (int)P.min_depth) {
                        const float survival = fmin_(0.95f, fmax_(fmax_(carried[0], carried[1]), fmax_(carried[2], carried[3])));
                        if (!(number(stream, sample, depth, Survival) < survival)) {
                            survives = false;
                        } else {
                            for (int l = 0; l < kLanes; ++l) carried[l] /= survival;
                        }
                    }
                    if (survives) {
                        const bool nowInside = sc.through ? !inside : inside;
                        const V3 no = x + n * (sc.through ? -lift : lift);
                        kdn_ray nr;
                        nr.origin[0] = no.x;
                        nr.origin[1] = no.y;
                        nr.origin[2] = no.z;
                        nr.direction[0] = sc.in.x;
                        nr.direction[1] = sc.in.y;
                        nr.direction[2] = sc.in.z;
                        nr.far_ = kInf;
                        nr.inside = nowInside ? 1u : 0u;
                        nr.within = nowInside ? mi : KDN_NO_MATERIAL;
                        s.rays[i] = nr;
                        for (int l = 0; l < kLanes; ++l) thr[l] = carried[l];
                        s.pdf[i] = sc.exact ? 0.0f : sc.density;
                        meta = (meta & ~(0xFFu | FlagExact)) | (uint32_t)(depth + 1) | (sc.exact ? FlagExact : 0u);
                        alive = true;
                    }
                }
            }
        }
    }
    for (int l = 0; l < kLanes; ++l) {
        s.rad[l * N + i] = rad[l];
        s.thr[l * N + i] = thr[l];
    }
    s.meta[i] = meta;
    if (alive) {
        const uint32_t at = atomicAdd(&s.counters[outQueue], 1u);
        s.queue[outQueue][at] = i;
    }
}

__global__ void __launch_bounds__(256) k_shadow(TreeView tv, PathState s) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= s.counters[2]) return;
    const uint32_t i = s.queue[2][q];
    const kdn_ray r = s.shadow[i];
    Dl buf[kMaxRegs];
    kdn_hit h;
    TreeView t = tv;
    double w = 0;
    if (tv.work) t.work = &w;
    trace_ray(t, r, buf, h);
    if (tv.work) atomicAdd(tv.work, w);
    if (h.kind == KDN_HIT_MISSED || !(h.distance < r.far_)) {
        const size_t N = s.capacity;
        for (int l = 0; l < kLanes; ++l) s.rad[l * N + i] += s.srad[l * N + i];
    }
}

// Fold the chunk's samples of each pixel into the FP64 accumulators, in
// sample order.
__global__ void k_fold(PassParams P, double* accum, double* accumSq, uint32_t* accumN) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t tilesX = (P.width + 7) / 8, tilesY = (P.height + 7) / 8;
    const uint32_t plane = tilesX * tilesY * 64;
    if (t >= plane) return;
    PathState& s = P.s;
    const size_t N = s.capacity;
    const uint32_t pixel = s.pixel[t];
    if (pixel == 0xFFFFFFFFu) return;
    double sum[3] = {0, 0, 0}, sq[3] = {0, 0, 0};
    for (uint32_t k = 0; k < P.chunk; ++k) {
        const size_t i = (size_t)k * plane + t;
        Lanes w;
        float rad[kLanes];
        for (int l = 0; l < kLanes; ++l) {
            w.nm[l] = s.nm[l * N + i];
            w.density[l] = s.dens[l * N + i];
            rad[l] = s.rad[l * N + i];
        }
        double rgb[3];
        colour_of(w, rad, P.colour, rgb);
        for (int c = 0; c < 3; ++c) {
            sum[c] += rgb[c];
            sq[c] += rgb[c] * rgb[c];
        }
    }
    for (int c = 0; c < 3; ++c) {
        accum[3 * (size_t)pixel + c] += sum[c];
        accumSq[3 * (size_t)pixel + c] += sq[c];
    }
    accumN[pixel] += P.chunk;
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
    double* work = nullptr;
    if (counting) {
        if (!ensure(c.scratch[7], 64, c.error, false)) return false;
        work = (double*)c.scratch[7].ptr;
        KDN_HIP(hipMemsetAsync(work, 0, 16, st));
    }
    TreeView tvShadow = tv;
    tvShadow.work = counting ? work + 1 : nullptr;

    uint64_t segments = 0, shadows = 0;
    uint32_t bounces = 0, launches = 0;
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
            segments += count;
            ++bounces;
            ++launches;
            const uint32_t blocks = (count + 255) / 256;
            timed(&traceMs, [&] {
                if (counting) {
                    hipLaunchKernelGGL(k_trace_queue_count, dim3(blocks), dim3(256), 0, st, tv, s->queue[in], &s->counters[in], s->rays,
                                       s->hits, work);
                } else {
                    hipLaunchKernelGGL(k_trace_queue, dim3(blocks), dim3(256), 0, st, tv, s->queue[in], &s->counters[in], s->rays, s->hits);
                }
            });
            KDN_HIP(hipMemsetAsync(&s->counters[out], 0, 4, st));
            KDN_HIP(hipMemsetAsync(&s->counters[2], 0, 4, st));
            timed(&shadeMs, [&] { hipLaunchKernelGGL(k_shade, dim3(blocks), dim3(256), 0, st, P, in, out); });
            uint32_t sc = 0;
            KDN_HIP(hipMemcpyAsync(&sc, &s->counters[2], 4, hipMemcpyDeviceToHost, st));
            KDN_HIP(hipStreamSynchronize(st));
            shadows += sc;
            if (sc) timed(&shadowMs, [&] { hipLaunchKernelGGL(k_shadow, dim3((sc + 255) / 256), dim3(256), 0, st, tvShadow, *s); });
            std::swap(in, out);
        }
        timed(&otherMs, [&] {
            hipLaunchKernelGGL(k_fold, dim3((plane + 255) / 256), dim3(256), 0, st, P, c.accum, c.accum_sq, c.accum_n);
        });
    }
    KDN_HIP(hipStreamSynchronize(st));
    KDN_HIP(hipGetLastError());
    double counted[2] = {0, 0};
    if (counting) KDN_HIP(hipMemcpy(counted, work, 16, hipMemcpyDeviceToHost));
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
    }
    return true;
}

}  // namespace kdn
