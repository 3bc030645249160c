// The integrator's per-item kernel bodies: one path slot, queue entry or
// pixel each. The HIP kernels call them one lane per item; a CPU build of
// this header (bounds-checked) runs the same code item by item.
#pragma once

#include "guide.h"
#include "hguide.h"
#include "integrator_dev.h"

namespace kdn {

struct PathState {
    uint32_t* slice;   // the round's slice each path belongs to
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
    uint32_t* keys;      // a queue's sort keys, and the sort's other buffers
    uint32_t* keys_alt;
    uint32_t* vals_alt;
    size_t capacity;
};

enum : uint32_t { FlagExact = 1u << 8, FlagDispersed = 1u << 9 };

// The scene and camera of one job (a render of one scene at one size), as
// the kernels of its paths read them.
struct JobDev {
    TreeView tv;
    int width, height;
    uint32_t plane;        // paths of one sample of every pixel (8x8 tiles)
    uint64_t seed_bits;
    uint32_t surfaces, min_depth;
    uint32_t sobol;        // the first 28 coordinates off the Sobol sequence (0: all hashed)
    // each pixel's first surface of sample record_sample (null: none kept)
    kdn_surface* firsts;
    uint32_t record_sample;
    float eye[3], forward[3], right[3], up[3];
    float reach, aspect;
    float sky_zenith[3], sky_horizon[3], sky_ground[3];
    uint32_t n_lamps, n_materials, fallback_material;
    const LampDev* lamps;
    const MaterialDev* materials;
    const uint32_t* leaf_material;
    // the plain guide: every pixel's packed guide of Gaussian lumps
    // (kHgFloats each, pixel-major; null: none), steering the first vertex,
    // and the survey's highlights filed by cell, steering deeper vertices
    const float* hg;
    const HgDeepHeader* hg_deep;
    // the learned guide: its lobes steer this job's guided vertices
    // (learned 0: BSDF sampling), and every train_stride-th of the job's
    // paths from path_first on keeps training records, the path
    // path_first + k * train_stride in record block rec_first + k
    // (train_stride 0: no records)
    uint32_t learned, train_stride, path_first, rec_first;
};

// A slice of a round: samples [first_sample, first_sample + spp) of every
// pixel of one job, on the paths [first_path, first_path + spp * plane),
// folded into that job's accumulators.
struct SliceDev {
    uint32_t job;
    uint32_t first_sample, spp;
    uint32_t first_path;
    double* accum;
    double* accum_sq;
    uint32_t* accum_n;
    kdn_estimate* estimate;  // the job's per-pixel estimates (null: none kept)
    uint32_t batch;          // samples a batch of an estimate holds
    // A counted slice (offsets set): pixel slot p draws offsets[p + 1] -
    // offsets[p] samples from first_sample on, its own count, and the
    // pass's paths are those samples pixel by pixel, densely numbered; this
    // slice holds the dense paths [base, base + paths) of them (spp unused).
    const uint64_t* offsets;
    uint64_t base;
    uint32_t paths;
};

// The paths of a slice.
__host__ __device__ inline uint32_t slice_paths(const SliceDev& sl, uint32_t plane) { return sl.offsets ? sl.paths : sl.spp * plane; }

// A round: the slices of one or more jobs whose paths share the queues.
struct PassParams {
    const JobDev* jobs;
    const SliceDev* slices;
    uint32_t n_slices;
    uint32_t paths;        // paths of the round (the last slice's end)
    const ColourTables* colour;
    PathState s;
    // the learned guide: lobes and selection probability of the guided
    // vertices (lane-major, kGuideOutputs per path, read for the paths of
    // jobs with learned set; null: BSDF sampling only) and training records
    // (blocks of kGuideDepths, placed by each job's train_stride,
    // path_first and rec_first; null: none kept); train_stride: the one
    // job's stride of a render pass's guided round
    const float* lobes;
    float* recs;
    uint32_t train_stride;
    // the round's one job, where it has one (the kernels specialised for
    // that case read it from their arguments, the same for every lane)
    JobDev job0;
};

KDN_D const JobDev& job_of(const PassParams& P, uint32_t i) { return P.jobs[P.slices[P.s.slice[i]].job]; }

// Path i's job: with One, the round's one job (P.job0).
template <bool One>
KDN_D const JobDev& job_in(const PassParams& P, uint32_t i) {
    if constexpr (One) {
        return P.job0;
    } else {
        return job_of(P, i);
    }
}

KDN_D float lane(const float* a, uint32_t l, size_t i, size_t n) { return a[l * n + i]; }

// Pixel of plane slot p: 8x8 tiles in rows, so that a wavefront's camera
// rays are neighbours.
KDN_D void pixel_of(const JobDev& J, uint32_t p, uint32_t* x, uint32_t* y) {
    const uint32_t tilesX = (J.width + 7) / 8;
    const uint32_t tile = p / 64, in = p % 64;
    const uint32_t tx = tile % tilesX, ty = tile / tilesX;
    *x = tx * 8 + (in % 8);
    *y = ty * 8 + (in / 8);
}

// generate: the path of slot i of the round (its slice's sample planes one
// after another)
KDN_D void generate_one(PassParams& P, uint32_t i) {
    PathState& s = P.s;
    const size_t N = s.capacity;
    uint32_t k = 0;
    while (k + 1 < P.n_slices && P.slices[k + 1].first_path <= i) ++k;
    const SliceDev& sl = P.slices[k];
    const JobDev& J = P.jobs[sl.job];
    const uint32_t local = i - sl.first_path;
    uint32_t sIdx, p;
    if (sl.offsets) {
        // the slot whose samples hold dense path g: the last p with
        // offsets[p] <= g (offsets[0] = 0 <= g < offsets[plane])
        const uint64_t g = sl.base + local;
        uint32_t lo = 0, hi = J.plane;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sl.offsets[mid] <= g) lo = mid;
            else hi = mid;
        }
        p = lo;
        sIdx = (uint32_t)(g - sl.offsets[lo]);
    } else {
        sIdx = local / J.plane;
        p = local % J.plane;
    }
    s.slice[i] = k;
    uint32_t x, y;
    pixel_of(J, p, &x, &y);
    const bool live = x < (uint32_t)J.width && y < (uint32_t)J.height;
    s.pixel[i] = live ? y * J.width + x : 0xFFFFFFFFu;
    if (!live) {
        for (int l = 0; l < kLanes; ++l) s.rad[l * N + i] = 0;
        return;
    }
    const uint32_t sample = sl.first_sample + sIdx;
    // pixelNumbers and pixelWavelengths (trace.cpp:65-84)
    const uint64_t place = ((uint64_t)x << 32) | (uint64_t)y;
    const uint64_t stream = J.seed_bits ^ mixed_bits(place);
    const double start = hashed(stream, ~0ull, 0);
    const Lanes w = sampled_wavelengths(wavelength_share(start, sample), P.colour);
    // filmDirection (core/camera.cpp:32-38) at the jittered film point
    const double fx = x + sequence_number(J.sobol, stream, sample, PixelX), fy = y + sequence_number(J.sobol, stream, sample, PixelY);
    const float u = (float)(2 * fx / J.width - 1), v = (float)(1 - 2 * fy / J.height);
    V3 d = v3(J.forward[0], J.forward[1], J.forward[2]) + v3(J.right[0], J.right[1], J.right[2]) * (u * J.reach * J.aspect) +
           v3(J.up[0], J.up[1], J.up[2]) * (v * J.reach);
    d = normalized(d);
    kdn_ray r;
    r.origin[0] = J.eye[0];
    r.origin[1] = J.eye[1];
    r.origin[2] = J.eye[2];
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

// extend: trace path i's ray; with work set, count the march's operations
KDN_D void trace_item(const TreeView& tv, uint32_t i, const kdn_ray* rays, kdn_hit* hits, const LaneRegs& lr, double* work);

// extend path i of a round through its job's octree
KDN_D void trace_path(const PassParams& P, uint32_t i, const LaneRegs& lr, double* work) {
    trace_item(job_of(P, i).tv, i, P.s.rays, P.s.hits, lr, work);
}

KDN_D void trace_item(const TreeView& tv, uint32_t i, const kdn_ray* rays, kdn_hit* hits, const LaneRegs& lr, double* work) {
    kdn_hit h;
    if (work) {
        double w = 0;
        TreeView counted = tv;
        counted.work = &w;
        trace_ray(counted, rays[i], lr, h);
        atomicAdd(work, w);
    } else {
        trace_ray(tv, rays[i], lr, h);
    }
    hits[i] = h;
}

// the material a hit stands on
KDN_D const MaterialDev& material_at(const JobDev& J, const kdn_hit& hit, uint32_t* index) {
    const uint32_t mi = hit.leaf != KDN_NO_LEAF ? J.leaf_material[hit.leaf] : J.fallback_material;
    *index = mi;
    return J.materials[mi < J.n_materials ? mi : J.fallback_material];
}

// A vertex the guide may sample at: a surface hit of a rough material at
// a depth below kGuideDepths. Gives the guide's query: point, outgoing
// direction, normal facing it, hero wavelength.
KDN_D bool guide_vertex_at(const PassParams& P, uint32_t i, float x[3], float out[3], float normal[3], float* nm) {
    const PathState& s = P.s;
    if ((s.meta[i] & 0xFF) >= (uint32_t)kGuideDepths) return false;
    const kdn_hit hit = s.hits[i];
    if (hit.kind != KDN_HIT_SURFACE) return false;
    uint32_t mi;
    if (material_at(job_of(P, i), hit, &mi).kind == KDN_GLASS) return false;
    const kdn_ray& ray = s.rays[i];
    V3 n = v3(hit.normal[0], hit.normal[1], hit.normal[2]);
    const V3 o = v3(-ray.direction[0], -ray.direction[1], -ray.direction[2]);
    if (dot(n, o) < 0) n = -n;
    for (int k = 0; k < 3; ++k) x[k] = hit.point[k];
    out[0] = o.x, out[1] = o.y, out[2] = o.z;
    normal[0] = n.x, normal[1] = n.y, normal[2] = n.z;
    *nm = s.nm[i];
    return true;
}

KDN_D bool guided_vertex(const PassParams& P, uint32_t i, float x[3], float out[3], float normal[3], float* nm) {
    return P.lobes != nullptr && job_of(P, i).learned != 0 && guide_vertex_at(P, i, x, out, normal, nm);
}

KDN_D void store_lobes(float* lobes, size_t N, uint32_t i, const GuideLobes& L) {
    for (int k = 0; k < kGuideLobes; ++k) {
        lobes[(5 * k) * N + i] = L.axis[k][0];
        lobes[(5 * k + 1) * N + i] = L.axis[k][1];
        lobes[(5 * k + 2) * N + i] = L.axis[k][2];
        lobes[(5 * k + 3) * N + i] = L.sharp[k];
        lobes[(5 * k + 4) * N + i] = L.share[k];
    }
    lobes[(size_t)kGuideSelect * N + i] = L.select;
}

KDN_D void load_lobes(const float* lobes, size_t N, uint32_t i, GuideLobes& L) {
    for (int k = 0; k < kGuideLobes; ++k) {
        L.axis[k][0] = lobes[(5 * k) * N + i];
        L.axis[k][1] = lobes[(5 * k + 1) * N + i];
        L.axis[k][2] = lobes[(5 * k + 2) * N + i];
        L.sharp[k] = lobes[(5 * k + 3) * N + i];
        L.share[k] = lobes[(5 * k + 4) * N + i];
    }
    L.select = lobes[(size_t)kGuideSelect * N + i];
}

// numbers of the guided sampling, apart from the reference's dimensions
KDN_D float guide_number(uint64_t stream, uint32_t sample, int event, int draw) {
    return (float)hashed(stream, sample, (1ull << 40) + (uint64_t)event * 4 + (uint64_t)draw);
}

// One-sample MIS of the guide (share) and the BSDF (1 - share): the
// direction from one of them, its density the mixture of both.
KDN_D Scatter guided_scatter(const MaterialDev& m, const Lanes& w, V3 n, V3 out, const GuideLobes& L, float share, float uSelect,
                             float uGuideLobe, float uLobe, float u, float v) {
    Scatter s;
    s.found = false;
    s.exact = false;
    s.through = false;
    s.collapse = false;
    if (dot(n, out) <= 0) return s;
    V3 in;
    if (uSelect < share) {
        in = mixture_sample(L, uGuideLobe, u, v);
    } else {
        const Scatter b = rough_scatter(m, w, n, out, uLobe, u, v);
        if (!b.found) return s;
        in = b.in;
    }
    const float arriving = dot(n, in);
    if (arriving <= 0) return s;
    const float density = share * mixture_pdf(L, in) + (1 - share) * rough_density(m, n, out, in);
    if (!(density > 0)) return s;
    float f[kLanes];
    rough_value(m, w, n, out, in, f);
    s.found = true;
    s.in = in;
    s.density = density;
    for (int l = 0; l < kLanes; ++l) s.weight[l] = f[l] * (arriving / density);
    return s;
}

// The plain guide's mixture at the first vertex (mixtureDirection,
// mixture.cpp:57-116): with chance share a direction from the lumps, else
// the material's; either way its density is the mixture of both.
KDN_D Scatter highlight_scatter(const MaterialDev& m, const Lanes& w, V3 n, V3 out, const float* g, V3 place, float share, float uPart,
                                float uPick, float uLobe, float u, float v) {
    Scatter s;
    s.found = false;
    s.exact = false;
    s.through = false;
    s.collapse = false;
    if (dot(n, out) <= 0) return s;
    V3 in;
    if (uPart < share) {
        in = hg_draw(g, place, uPick, u, v);
        if (!__builtin_isfinite(in.x + in.y + in.z)) return s;
    } else {
        const Scatter b = rough_scatter(m, w, n, out, uLobe, u, v);
        if (!b.found) return s;
        in = b.in;
    }
    const float arriving = dot(n, in);
    if (arriving <= 0) return s;
    const float density = share * hg_density(g, place, n, out, in) + (1 - share) * rough_density(m, n, out, in);
    if (!(density > 0)) return s;
    float f[kLanes];
    rough_value(m, w, n, out, in, f);
    s.found = true;
    s.in = in;
    s.density = density;
    for (int l = 0; l < kLanes; ++l) s.weight[l] = f[l] * (arriving / density);
    return s;
}

// The same mixture at a deeper vertex, with its deep guide's lumps.
KDN_D Scatter deep_scatter(const MaterialDev& m, const Lanes& w, V3 n, V3 out, const HgDeepView& dv, V3 x, uint32_t mi,
                           const LampDev* lamps, const MaterialDev* materials, float uPart, float uPick, float uLobe, float u, float v) {
    Scatter s;
    s.found = false;
    s.exact = false;
    s.through = false;
    s.collapse = false;
    if (dot(n, out) <= 0) return s;
    V3 in;
    if (uPart < dv.share) {
        in = hg_deep_draw(dv, x, n, out, mi, m, lamps, materials, uPick, u, v);
        if (!__builtin_isfinite(in.x + in.y + in.z) || dot(in, in) == 0) return s;
    } else {
        const Scatter b = rough_scatter(m, w, n, out, uLobe, u, v);
        if (!b.found) return s;
        in = b.in;
    }
    const float arriving = dot(n, in);
    if (arriving <= 0) return s;
    const float density = dv.share * hg_deep_density(dv, x, n, out, mi, m, lamps, materials, in) + (1 - dv.share) * rough_density(m, n, out, in);
    if (!(density > 0)) return s;
    float f[kLanes];
    rough_value(m, w, n, out, in, f);
    s.found = true;
    s.in = in;
    s.density = density;
    for (int l = 0; l < kLanes; ++l) s.weight[l] = f[l] * (arriving / density);
    return s;
}

// A shadow test of path i toward a lamp, carrying c, when c brings light.
KDN_D bool offer_shadow(PathState& s, size_t N, uint32_t i, V3 from, V3 in, float far, const float c[kLanes]) {
    if (!(c[0] + c[1] + c[2] + c[3] > 0)) return false;
    kdn_ray sr;
    sr.origin[0] = from.x;
    sr.origin[1] = from.y;
    sr.origin[2] = from.z;
    sr.direction[0] = in.x;
    sr.direction[1] = in.y;
    sr.direction[2] = in.z;
    sr.far_ = far;
    sr.inside = 0;
    sr.within = KDN_NO_MATERIAL;
    s.shadow[i] = sr;
    for (int l = 0; l < kLanes; ++l) s.srad[l * N + i] = c[l];
    return true;
}

KDN_D float lift_at(V3 x) {
    return 8e-5f + 8e-6f * fmax_(fmax_(__builtin_fabsf(x.x), __builtin_fabsf(x.y)), __builtin_fabsf(x.z));
}

// The sort key of a ray: its direction's octant over the Morton code of its
// origin's cell in a 512^3 grid over the region (9 bits an axis), so that
// sorted rays leave nearby cells the same way.
KDN_D uint32_t spread_bits(uint32_t v) {
    v &= 0x1FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

KDN_D uint32_t ray_key(const TreeView& tv, const kdn_ray& r) {
    uint32_t cell[3];
    for (int k = 0; k < 3; ++k) {
        const float f = (r.origin[k] - tv.lower[k]) / tv.size[k] * 512.0f;
        cell[k] = f > 0 ? (f < 511.0f ? (uint32_t)f : 511u) : 0u;
    }
    const uint32_t octant = (r.direction[0] < 0 ? 1u : 0u) | (r.direction[1] < 0 ? 2u : 0u) | (r.direction[2] < 0 ? 4u : 0u);
    return (octant << 27) | spread_bits(cell[0]) | (spread_bits(cell[1]) << 1) | (spread_bits(cell[2]) << 2);
}

// shade path i at its hit: emission met, a shadow test toward one lamp
// (its ray in s.shadow[i] and what it carries in s.srad; *shadow set) and
// the scattered ray that continues the path (in s.rays[i]; returns true).
// (Plain: the round has no guide of either kind and keeps no training
// records, so the guided sampling is compiled out: the kernel keeps fewer
// values live and spills fewer registers.)
// (Lobes false: the caller shades no vertex the learned guide steers, as
// the finishing launch, which runs paths past the guided depths only; the
// lobes' sampling is compiled out.)
template <bool One = false, bool Plain = false, bool Lobes = true>
KDN_D bool shade_path(PassParams& P, uint32_t i, bool* shadow) {
    PathState& s = P.s;
    const size_t N = s.capacity;
    const JobDev& J = job_in<One>(P, i);
    *shadow = false;
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
    const bool training = !Plain && P.recs != nullptr && J.train_stride != 0 && i >= J.path_first &&
                          (i - J.path_first) % J.train_stride == 0;
    const size_t recBlock = training ? (size_t)J.rec_first + (i - J.path_first) / J.train_stride : 0;
    // the frame record's first surface of the pixel's last sample
    if (depth == 0 && J.firsts && sample == J.record_sample && s.pixel[i] != 0xFFFFFFFFu)
        J.firsts[s.pixel[i]] = surface_met(hit, ray, J.leaf_material, J.materials, J.n_materials);

    // lamps met along the ray (metAlong, trace.cpp:352-374), MIS against
    // aiming: quads and spheres before the hit, suns where the ray escapes
    if (!inside) {
        for (uint32_t j = 0; j < J.n_lamps; ++j) {
            const LampDev& L = J.lamps[j];
            float pdfL;
            if (L.kind == KDN_LIGHT_QUAD) {
                LampHit lh;
                if (!meet_quad(L, o, d, tHit, &lh)) continue;
                pdfL = L.pick * lh.t * lh.t / (lh.cosine * L.area);
            } else if (L.kind == KDN_LIGHT_SPHERE) {
                float t;
                if (!meet_sphere(L, o, d, tHit, &t)) continue;
                const float away = length(v3(L.at[0], L.at[1], L.at[2]) - o);
                pdfL = away > L.radius ? L.pick * cone_density(cone_cosine(L.radius, away)) : 0.0f;
            } else if (L.kind == KDN_LIGHT_SUN) {
                if (hit.kind != KDN_HIT_MISSED || dot(d, v3(L.direction[0], L.direction[1], L.direction[2])) < L.cosine) continue;
                pdfL = L.pick * cone_density(L.cosine);
            } else {
                continue;  // points and directional lights are met by aiming only
            }
            float Le[kLanes];
            spectrum_of(L.radiance, w, Le);
            const float wt = exactBefore ? 1.0f : power_weight(pdfBefore, pdfL);
            for (int l = 0; l < kLanes; ++l) rad[l] += thr[l] * Le[l] * wt;
        }
    }
    // training records: the radiance the path holds as this vertex begins
    // (lamps met on the way counted in: lamps are aimed at from every
    // guided vertex, so the guide learns the light that aiming misses)
    // closes the record of the vertex before
    if (training && depth >= 1 && depth - 1 < kGuideDepths) {
        float* rec = P.recs + (recBlock * kGuideDepths + depth - 1) * kRecFloats;
        if (rec[kRecState] == 1.0f) {
            for (int l = 0; l < kLanes; ++l) rec[kRecSnap + l] = rad[l];
            rec[kRecState] = 3.0f;
        }
    }
    bool alive = false;
    if (hit.kind == KDN_HIT_MISSED) {
        if (!inside) {
            float sky[kLanes];
            sky_radiance(J.sky_zenith, J.sky_horizon, J.sky_ground, d, w, sky);
            for (int l = 0; l < kLanes; ++l) rad[l] += thr[l] * sky[l];
        }
    } else if (hit.kind == KDN_HIT_SURFACE) {
        uint32_t mi;
        const MaterialDev& m = material_at(J, hit, &mi);
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
        // the guide's lobes at this vertex (the same test as its inference)
        const bool guidable = depth < kGuideDepths && m.kind != KDN_GLASS;
        const bool guided = !Plain && Lobes && P.lobes != nullptr && J.learned != 0 && guidable;
        GuideLobes lobes;
        if (guided) load_lobes(P.lobes, N, i, lobes);
        const float share = guided ? lobes.select : 0.0f;
        // the pixel's plain guide steers the first vertex on its material
        // (steersPlace, guide.cpp:162-164)
        const float* hg = nullptr;
        if (!Plain && J.hg != nullptr && depth == 0 && !inside && s.pixel[i] != 0xFFFFFFFFu) {
            const float* g = J.hg + (size_t)s.pixel[i] * kHgFloats;
            if (g[0] > 0 && (uint32_t)g[1] == mi && m.kind != KDN_GLASS) hg = g;
        }
        const float hgShare = hg ? hg[2] : 0.0f;
        // deeper vertices of a broad surface: the survey's highlights filed
        // in their cell (the deep guide)
        HgDeepView deepView;
        const bool deep = !Plain && !hg && !guided && J.hg_deep != nullptr && depth >= 1 && !inside && m.kind != KDN_GLASS &&
                          m.width > kHgShinyWidth &&
                          hg_deep_view(J.hg_deep, x, n, out, mi, m, J.lamps, J.n_lamps, J.materials, &deepView);
        // aim at one lamp (aimedFrom, trace.cpp:896-947; aimLamp,
        // light.cpp:359-370), MIS against meeting
        if (!exact && J.n_lamps > 0) {
            const float uPick = number(J.sobol, stream, sample, depth, Pick);
            uint32_t j = 0;
            while (j + 1 < J.n_lamps && uPick >= J.lamps[j].pick_cdf) ++j;
            const LampDev& L = J.lamps[j];
            const float u = number(J.sobol, stream, sample, depth, LightU), v = number(J.sobol, stream, sample, depth, LightV);
            const V3 at = v3(L.at[0], L.at[1], L.at[2]);
            if (L.kind == KDN_LIGHT_QUAD) {
                const V3 y = at + v3(L.along[0], L.along[1], L.along[2]) * u + v3(L.across[0], L.across[1], L.across[2]) * v;
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
                    const float pdfS = guided ? share * mixture_pdf(lobes, in) + (1 - share) * rough_density(m, n, out, in)
                                       : hg   ? hgShare * hg_density(hg, x, n, out, in) + (1 - hgShare) * rough_density(m, n, out, in)
                                       : deep ? deepView.share * hg_deep_density(deepView, x, n, out, mi, m, J.lamps, J.materials, in) +
                                                    (1 - deepView.share) * rough_density(m, n, out, in)
                                              : rough_density(m, n, out, in);
                    const float wt = power_weight(pdfL, pdfS);
                    float c[kLanes];
                    for (int l = 0; l < kLanes; ++l) c[l] = thr[l] * f[l] * cosS * Le[l] / pdfL * wt;
                    *shadow = offer_shadow(s, N, i, x + n * lift, in, dist - lift, c);
                }
            } else {
                // the direction, the distance to the lamp along it, what
                // arrives over the chance of the direction (radiance over
                // density, or the singular lights' irradiance), and the
                // density (0: singular)
                V3 in = v3(0, 0, 1);
                float dist = 0, density = 0, arriving[3] = {0, 0, 0};
                bool aimed = false;
                if (L.kind == KDN_LIGHT_POINT) {
                    const V3 toward = at - x;
                    dist = length(toward);
                    if (dist > 1e-9f) {
                        in = toward * (1.0f / dist);
                        for (int k = 0; k < 3; ++k) arriving[k] = L.power[k] / (dist * dist);
                        aimed = true;
                    }
                } else if (L.kind == KDN_LIGHT_DIRECTIONAL) {
                    in = v3(L.direction[0], L.direction[1], L.direction[2]);
                    dist = kInf;
                    for (int k = 0; k < 3; ++k) arriving[k] = L.power[k];
                    aimed = true;
                } else if (L.kind == KDN_LIGHT_SPHERE) {
                    const V3 toward = at - x;
                    const float away = length(toward);
                    if (away > 1e-9f && away > L.radius) {
                        const V3 centre = toward * (1.0f / away);
                        const float cosine = cone_cosine(L.radius, away);
                        in = uniform_cone(centre, cosine, u, v);
                        const float height = dot(in, centre);
                        const float ring = __builtin_sqrtf(fmax_(1 - height * height, 0.0f));
                        dist = fmax_(away * height - __builtin_sqrtf(fmax_(L.radius * L.radius - away * away * ring * ring, 0.0f)), 0.0f);
                        density = cone_density(cosine);
                        for (int k = 0; k < 3; ++k) arriving[k] = L.power[k] * (2 * (1 - cosine) / (L.radius * L.radius));
                        aimed = density > 0;
                    }
                } else if (L.kind == KDN_LIGHT_SUN) {
                    in = uniform_cone(v3(L.direction[0], L.direction[1], L.direction[2]), L.cosine, u, v);
                    dist = kInf;
                    density = cone_density(L.cosine);
                    for (int k = 0; k < 3; ++k) arriving[k] = L.radiance[k] / density;
                    aimed = density > 0;
                }
                const float cosS = dot(n, in);
                if (aimed && cosS > 0 && dist > 2 * lift) {
                    float f[kLanes], Le[kLanes];
                    rough_value(m, w, n, out, in, f);
                    spectrum_of(arriving, w, Le);
                    float wt = 1.0f;  // singular lights: aiming alone finds them
                    if (density > 0) {
                        const float pdfS = guided ? share * mixture_pdf(lobes, in) + (1 - share) * rough_density(m, n, out, in)
                                           : hg   ? hgShare * hg_density(hg, x, n, out, in) + (1 - hgShare) * rough_density(m, n, out, in)
                                           : deep ? deepView.share * hg_deep_density(deepView, x, n, out, mi, m, J.lamps, J.materials, in) +
                                                        (1 - deepView.share) * rough_density(m, n, out, in)
                                                  : rough_density(m, n, out, in);
                        wt = power_weight(L.pick * density, pdfS);
                    }
                    float c[kLanes];
                    for (int l = 0; l < kLanes; ++l) c[l] = thr[l] * f[l] * cosS * Le[l] / L.pick * wt;
                    *shadow = offer_shadow(s, N, i, x + n * lift, in, dist < kInf ? dist - lift : kInf, c);
                }
            }
        }
        // scatter (scatteredFrom, trace.cpp:1020-1074)
        if (depth + 1 < (int)J.surfaces) {
            const float uLobe = number(J.sobol, stream, sample, depth, Lobe), uSide = number(J.sobol, stream, sample, depth, Side);
            const float du = number(J.sobol, stream, sample, depth, DirectionU), dv = number(J.sobol, stream, sample, depth, DirectionV);
            const Scatter sc = m.kind == KDN_GLASS ? glass_scatter(m, w, n, out, inside, uSide)
                               : guided ? guided_scatter(m, w, n, out, lobes, share, guide_number(stream, sample, depth, 0),
                                                         guide_number(stream, sample, depth, 1), uLobe, du, dv)
                               : hg     ? highlight_scatter(m, w, n, out, hg, x, hgShare, guide_number(stream, sample, depth, 0),
                                                            guide_number(stream, sample, depth, 1), uLobe, du, dv)
                               : deep   ? deep_scatter(m, w, n, out, deepView, x, mi, J.lamps, J.materials, guide_number(stream, sample, depth, 0),
                                                       guide_number(stream, sample, depth, 1), uLobe, du, dv)
                                        : rough_scatter(m, w, n, out, uLobe, du, dv);
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
                    if (depth >= (int)J.min_depth) {
                        const float survival = fmin_(0.95f, fmax_(fmax_(carried[0], carried[1]), fmax_(carried[2], carried[3])));
                        if (!(number(J.sobol, stream, sample, depth, Survival) < survival)) {
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
                        s.pdf[i] = sc.exact ? 0.0f : sc.density;
                        meta = (meta & ~(0xFFu | FlagExact)) | (uint32_t)(depth + 1) | (sc.exact ? FlagExact : 0u);
                        alive = true;
                        if (training && guidable) {
                            float* rec = P.recs + (recBlock * kGuideDepths + depth) * kRecFloats;
                            const float q[14] = {x.x, x.y, x.z, out.x, out.y, out.z, n.x, n.y, n.z, w.nm[0], sc.in.x, sc.in.y, sc.in.z, sc.density};
                            for (int k = 0; k < 14; ++k) rec[k] = q[k];
                            for (int l = 0; l < kLanes; ++l) {
                                rec[kRecThr + l] = carried[l];
                                rec[kRecBefore + l] = thr[l];
                            }
                            rec[kRecBsdf] = rough_density(m, n, out, sc.in);
                            rec[kRecState] = 1.0f;
                        }
                        for (int l = 0; l < kLanes; ++l) thr[l] = carried[l];
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
    return alive;
}

// shade: queue entry q of the input queue; shadow tests and live paths
// are appended to their queues
template <bool One = false>
KDN_D void shade_one(PassParams& P, uint32_t q, uint32_t inQueue, uint32_t outQueue) {
    PathState& s = P.s;
    const uint32_t i = s.queue[inQueue][q];
    bool shadow;
    const bool alive = shade_path<One>(P, i, &shadow);
    if (shadow) {
        const uint32_t at = atomicAdd(&s.counters[2], 1u);
        s.queue[2][at] = i;
    }
    if (alive) {
        const uint32_t at = atomicAdd(&s.counters[outQueue], 1u);
        s.queue[outQueue][at] = i;
    }
}

// the shadow test of path i: what it carries reaches the path when nothing
// lies on its ray (with work set, the march's operations are counted)
template <bool One = false>
KDN_D void shadow_path(const PassParams& P, uint32_t i, const LaneRegs& lr, double* work) {
    const PathState& s = P.s;
    const kdn_ray r = s.shadow[i];
    kdn_hit h;
    TreeView t = job_in<One>(P, i).tv;
    double w = 0;
    t.work = work ? &w : nullptr;
    occlusion_ray(t, r, lr, h);
    if (work) atomicAdd(work, w);
    if (h.kind == KDN_HIT_MISSED || !(h.distance < r.far_)) {
        const size_t N = s.capacity;
        for (int l = 0; l < kLanes; ++l) s.rad[l * N + i] += s.srad[l * N + i];
    }
}

// shadow: entry q of the shadow queue
KDN_D void shadow_one(const PassParams& P, uint32_t q, const LaneRegs& lr, double* work) {
    shadow_path(P, P.s.queue[2][q], lr, work);
}

// finish: the lane takes entries of the input queue in turn (tally[2]
// counts them out; count of them in all) and marches one ray of its path
// an iteration: the path's next segment (then shading, which stops every
// path at its last surface) or, by turns, the shadow test shading asked
// for; it takes the next entry when its path stops. A wave's lanes thus
// run one march between them an iteration, whichever rays they hold, and
// stay busy until the queue drains rather than wait for the wave's longest
// path. The same per-path arithmetic, in the same order, as the
// wavefront's launches, so the estimates do not change; counts segments
// and shadow tests in tally[0] and tally[1].
// (With One: the round has one job and counts no operations.)
template <bool One, bool Plain>
KDN_D void finish_lane(PassParams& P, uint32_t inQueue, uint32_t count, const LaneRegs& lr, uint32_t* tally, double* work) {
    PathState& s = P.s;
    const size_t N = s.capacity;
    uint32_t segments = 0, shadows = 0;
    uint32_t q = atomicAdd(&tally[2], 1u);
    bool shadowTurn = false, aliveAfter = false;
    while (q < count) {
        const uint32_t i = s.queue[inQueue][q];
        const kdn_ray r = shadowTurn ? s.shadow[i] : s.rays[i];
        kdn_hit h;
        int iterations = 0;
        if constexpr (One) {
            march_any(P.job0.tv, r, lr, h, iterations, shadowTurn);
        } else {
            TreeView t = job_of(P, i).tv;
            double w = 0;
            t.work = work ? &w : nullptr;
            march_any(t, r, lr, h, iterations, shadowTurn);
            if (work) atomicAdd(shadowTurn ? work + 1 : work, w);
        }
        KDN_TRACE_ITERATIONS(iterations);
        bool next = false;
        if (shadowTurn) {
            if (h.kind == KDN_HIT_MISSED || !(h.distance < r.far_)) {
                for (int l = 0; l < kLanes; ++l) s.rad[l * N + i] += s.srad[l * N + i];
            }
            shadowTurn = false;
            next = !aliveAfter;
        } else {
            s.hits[i] = h;
            ++segments;
            bool shadow;
            const bool alive = shade_path<One, Plain, false>(P, i, &shadow);
            if (shadow) {
                ++shadows;
                shadowTurn = true;
                aliveAfter = alive;
            } else {
                next = !alive;
            }
        }
        if (next) q = atomicAdd(&tally[2], 1u);
    }
    atomicAdd(&tally[0], segments);
    atomicAdd(&tally[1], shadows);
}

// A sample drawn alone folded into a pixel's estimate (foldedSample and
// closedBatch, core/stopping.cpp:36-67): its colour added to the total and
// the batch, its shared part (its whole colour's channel mean) to the
// batch's; a whole batch closes into the sums of means and squares.
KDN_D void fold_estimate(kdn_estimate& e, const double c[3], uint32_t batch) {
    for (int k = 0; k < 3; ++k) {
        e.total[k] = e.total[k] + c[k];
        e.batch[k] = e.batch[k] + c[k];
    }
    e.taken += 1;
    e.shared = e.shared + (c[0] + c[1] + c[2]) / 3;
    if (batch > 0 && (uint32_t)e.taken % batch == 0) {
        const double scale = 1.0 / (double)batch;
        double mean[3];
        for (int k = 0; k < 3; ++k) mean[k] = e.batch[k] * scale;
        const double shared = e.shared / (double)batch;
        const double own = (mean[0] + mean[1] + mean[2]) / 3 - shared;
        for (int k = 0; k < 3; ++k) {
            e.means[k] = e.means[k] + mean[k];
            e.squares[k] = e.squares[k] + mean[k] * mean[k];
            e.batch[k] = 0;
        }
        e.batches += 1;
        e.shared_means = e.shared_means + shared;
        e.shared_squares = e.shared_squares + shared * shared;
        e.own_squares = e.own_squares + own * own;
        e.shared = 0;
    }
}

// fold: pixel slot t of slice k's first sample plane; its samples folded
// into the job's FP64 accumulators in sample order (and into its estimate
// where the slice keeps them)
KDN_D void fold_one(const PassParams& P, uint32_t k, uint32_t t) {
    const SliceDev& sl = P.slices[k];
    const JobDev& J = P.jobs[sl.job];
    if (t >= J.plane) return;
    const PathState& s = P.s;
    const size_t N = s.capacity;
    // the slot's paths in this slice: n of them from first, stride apart
    size_t first = (size_t)sl.first_path + t, stride = J.plane;
    uint32_t n = sl.spp;
    if (sl.offsets) {
        const uint64_t end = sl.base + sl.paths;
        const uint64_t lo = sl.offsets[t] > sl.base ? sl.offsets[t] : sl.base;
        const uint64_t hi = sl.offsets[t + 1] < end ? sl.offsets[t + 1] : end;
        if (lo >= hi) return;
        first = (size_t)sl.first_path + (size_t)(lo - sl.base);
        stride = 1;
        n = (uint32_t)(hi - lo);
    }
    const uint32_t pixel = s.pixel[first];
    if (pixel == 0xFFFFFFFFu) return;
    double sum[3] = {0, 0, 0}, sq[3] = {0, 0, 0};
    kdn_estimate est;
    if (sl.estimate) est = sl.estimate[pixel];
    for (uint32_t j = 0; j < n; ++j) {
        const size_t i = first + (size_t)j * stride;
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
        if (sl.estimate) fold_estimate(est, rgb, sl.batch);
    }
    if (sl.estimate) sl.estimate[pixel] = est;
    for (int c = 0; c < 3; ++c) {
        sl.accum[3 * (size_t)pixel + c] += sum[c];
        sl.accum_sq[3 * (size_t)pixel + c] += sq[c];
    }
    sl.accum_n[pixel] += n;
}

}  // namespace kdn
