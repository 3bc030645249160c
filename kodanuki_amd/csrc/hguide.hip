// The plain (highlight) guide on the device: the survey of highlights (a
// shiny surface mirroring a lamp toward a surface the camera sees) at
// every eighth pixel each way, and every pixel's guide of up to 12
// Gaussian lumps toward them with the share of draws it takes.
//
// Follows the reference's structure (plugins/sampler/highlight_guide/
// survey.h:27-70, survey.cpp): survey points every eighth pixel from the
// fourth cast their middle ray and draw 128 directions through the
// material (counter-based numbers keyed apart from any sample); the draws
// that meet a sharp surface (metal of facet width at most 0.125) are the
// point's seeds, at most three to a leaf, 0.05 radians apart, sixteen in
// all (surveySeeds). Each seed and lamp walks the reference's chain (a
// damped Newton walk over the leaf's own program in FP64, chain_dev.h) to
// the point whose normal halves the directions to the stand and the lamp;
// the most important settled one (glare times the spread's area) is kept,
// and a ray checks that the surface is there and mirrors the lamp within
// its outline widened by the facets' lobe. A pixel's guide takes the highlights of its
// four nearest survey points on its own material, each walked again from
// the pixel's own stand (one mirror, one step) and one lump per highlight,
// weighs each by nearness, the lamp's glare, the lump's size and the
// material's density toward it (weighedLump, survey.cpp:427-430), keeps
// the 12 heaviest and
// drops those under a hundredth of the heaviest, and shares its draws by
// the light the lumps bring against the direct light, within [0.02, 0.5]
// (guideShare, survey.cpp:818-828). The guide's lumps are the reference's
// and so is their use in shading (hguide.h), and so are the seeds and the
// chains; the pixel guides' walk (one mirror in one step from the pixel's
// stand) and the lumps' round spread (the flat glossy mirror's, (2 width +
// the lamp's angular radius) d / (D + d) radians across, d from the point
// to the lamp, D from the viewer) are this build's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "chain_dev.h"
#include "ctx.h"
#include "hguide.h"
#include "integrator_kernels.h"

namespace kdn {
namespace {

constexpr int kHgBlock = 256;
constexpr int kSurveyDraws = 128;
constexpr int kSpots = kHgSpots;
constexpr int kSpotFloats = kHgSpotFloats;
constexpr float kShinyWidth = kHgShinyWidth;
constexpr float kSameSpot = kHgSameSpot;

struct HgLens {
    double corner[3], across[3], down[3];
    float eye[3];
};

struct HgScene {
    const uint32_t* leafMaterial;
    const MaterialDev* materials;
    uint32_t nMaterials;
    const LampDev* lamps;
    uint32_t nLamps;
    const LeafStepDev* leafSteps;  // every leaf's own program (the chains walk them)
    const LeafProgDev* leafProgs;
    uint32_t nLeaves;
};

KDN_D V3 lens_dir(const HgLens& L, double x, double y) {
    double h[3];
    for (int k = 0; k < 3; ++k) h[k] = L.corner[k] + L.across[k] * x + L.down[k] * y;
    const double len = __builtin_sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    return v3((float)(h[0] / len), (float)(h[1] / len), (float)(h[2] / len));
}

KDN_D const MaterialDev& mat_of(const HgScene& S, uint32_t leaf, uint32_t* mi) {
    *mi = leaf != KDN_NO_LEAF ? S.leafMaterial[leaf] : S.nMaterials - 1;
    if (*mi >= S.nMaterials) *mi = S.nMaterials - 1;
    return S.materials[*mi];
}

// A stand: where a ray from the eye through (x, y) first meets a rough
// surface; record: valid, point 3, normal 3 (facing the eye), out 3,
// material.
constexpr int kStandFloats = 11;

__global__ void k_hg_rays(HgLens L, int width, int height, int stride, int offset, int cols, int rows, kdn_ray* rays) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= (uint32_t)(cols * rows)) return;
    const int x = offset + stride * (int)(at % cols), y = offset + stride * (int)(at / cols);
    const V3 d = lens_dir(L, x + 0.5, y + 0.5);
    kdn_ray r;
    for (int k = 0; k < 3; ++k) r.origin[k] = L.eye[k];
    r.direction[0] = d.x;
    r.direction[1] = d.y;
    r.direction[2] = d.z;
    r.far_ = kInf;
    r.inside = 0;
    r.within = KDN_NO_MATERIAL;
    rays[at] = r;
}

__global__ void k_hg_stands(HgScene S, const kdn_ray* rays, const kdn_hit* hits, uint32_t n, float* stands) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= n) return;
    float* s = stands + (size_t)at * kStandFloats;
    s[0] = 0;
    const kdn_hit h = hits[at];
    if (h.kind != KDN_HIT_SURFACE) return;
    uint32_t mi;
    const MaterialDev& m = mat_of(S, h.leaf, &mi);
    if (m.kind == KDN_GLASS) return;
    const V3 d = v3(rays[at].direction[0], rays[at].direction[1], rays[at].direction[2]);
    V3 nv = v3(h.normal[0], h.normal[1], h.normal[2]);
    if (dot(nv, d) > 0) nv = -nv;
    s[0] = 1;
    s[1] = h.point[0];
    s[2] = h.point[1];
    s[3] = h.point[2];
    s[4] = nv.x;
    s[5] = nv.y;
    s[6] = nv.z;
    s[7] = -d.x;
    s[8] = -d.y;
    s[9] = -d.z;
    s[10] = (float)mi;
}

KDN_D float survey_number(uint32_t point, int draw, int role) {
    return (float)hashed(mixed_bits(0x5EEDull ^ ((uint64_t)point << 20)), (uint64_t)draw, (uint64_t)role);
}

// the survey's draws through each stand's material
__global__ void k_hg_draws(HgScene S, const float* stands, uint32_t nPoints, kdn_ray* rays) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= nPoints * kSurveyDraws) return;
    const uint32_t point = at / kSurveyDraws;
    const int draw = (int)(at % kSurveyDraws);
    const float* s = stands + (size_t)point * kStandFloats;
    kdn_ray r;
    r.far_ = -1;  // no draw: a ray that meets nothing (an empty span)
    r.inside = 0;
    r.within = KDN_NO_MATERIAL;
    for (int k = 0; k < 3; ++k) {
        r.origin[k] = 0;
        r.direction[k] = k == 2 ? 1.0f : 0.0f;
    }
    if (s[0] == 1) {
        const MaterialDev& m = S.materials[(uint32_t)s[10]];
        const V3 x = v3(s[1], s[2], s[3]), n = v3(s[4], s[5], s[6]), out = v3(s[7], s[8], s[9]);
        const Scatter sc = rough_scatter(m, hg_grey_lanes(), n, out, survey_number(point, draw, 0), survey_number(point, draw, 1),
                                         survey_number(point, draw, 2));
        if (sc.found) {
            const V3 o = x + n * lift_at(x);
            r.origin[0] = o.x;
            r.origin[1] = o.y;
            r.origin[2] = o.z;
            r.direction[0] = sc.in.x;
            r.direction[1] = sc.in.y;
            r.direction[2] = sc.in.z;
            r.far_ = kInf;
        }
    }
    rays[at] = r;
}

// Each survey point's seeds among its draws, in draw order (surveySeeds,
// survey.cpp:487-529): a draw that meets a sharp metal is a seed unless
// three seeds of its leaf are kept already or one of them lies within
// 0.05 radians of its direction; sixteen at most. The seeds' draws are
// appended to seeds (*seedCount of them, in no fixed order: each walk
// writes its own draw's candidate and aim alone). Every draw's candidate and
// aim start empty.
constexpr int kMostSeeds = 16;
__global__ void k_hg_seeds(HgScene S, const kdn_ray* rays, const kdn_hit* hits, uint32_t nPoints, uint32_t* seeds,
                           uint32_t* seedCount, kdn_ray* aims, float* cand) {
    const uint32_t point = blockIdx.x * blockDim.x + threadIdx.x;
    if (point >= nPoints) return;
    constexpr int kSeedsPerLeaf = 3;
    const float close = 0.99875026f;  // cos 0.05
    uint32_t leafOf[kMostSeeds];
    V3 dirOf[kMostSeeds];
    int kept = 0;
    kdn_ray none;
    none.far_ = -1;  // no aim: an empty span
    none.inside = 0;
    none.within = KDN_NO_MATERIAL;
    for (int k = 0; k < 3; ++k) {
        none.origin[k] = 0;
        none.direction[k] = k == 2 ? 1.0f : 0.0f;
    }
    for (int draw = 0; draw < kSurveyDraws; ++draw) {
        const size_t at = (size_t)point * kSurveyDraws + draw;
        cand[at * kSpotFloats] = 0;
        aims[at] = none;
        const kdn_ray r = rays[at];
        if (!(r.far_ > 0) || kept == kMostSeeds) continue;
        const kdn_hit h = hits[at];
        if (h.kind != KDN_HIT_SURFACE) continue;
        const V3 d = v3(r.direction[0], r.direction[1], r.direction[2]);
        int same = 0;
        bool near = false;
        for (int k = 0; k < kept; ++k) {
            if (leafOf[k] != h.leaf) continue;
            ++same;
            near = near || dot(dirOf[k], d) > close;
        }
        uint32_t mi;
        const MaterialDev& m = mat_of(S, h.leaf, &mi);
        if (same < kSeedsPerLeaf && !near && m.kind == KDN_METAL && m.width <= kShinyWidth) {
            leafOf[kept] = h.leaf;
            dirOf[kept] = d;
            seeds[atomicAdd(seedCount, 1u)] = (uint32_t)at;
            ++kept;
        }
    }
}

// A seed starts a highlight: for each lamp, the chain from the stand (its
// surface point lifted along its normal, as standAt, survey.cpp:138-156)
// by the metal's leaf to the lamp is walked from the seed's hit
// (walkedChain, chain_dev.h; walkedFrom, survey.cpp:312-326); of the
// settled ones with the lamp in front of the point, the most important
// (the lamp's glare times the square root of the spread's determinant,
// spotImportance, survey.cpp:582-585) is kept, and a ray from the stand
// toward its point checks that the surface is there. One lane a seed (a
// launch sized for the most seeds; the lanes past *seedCount end at once).
__global__ void k_hg_walks(HgScene S, const float* stands, const kdn_hit* hits, const uint32_t* seeds, const uint32_t* seedCount,
                           kdn_ray* aims, float* cand) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= *seedCount) return;
    const uint32_t at = seeds[slot];
    float* c = cand + (size_t)at * kSpotFloats;
    const kdn_hit h = hits[at];
    uint32_t mi;
    const MaterialDev& m = mat_of(S, h.leaf, &mi);
    const float* s = stands + (size_t)(at / kSurveyDraws) * kStandFloats;
    const V3 stand = v3(s[1], s[2], s[3]), sn = v3(s[4], s[5], s[6]);
    const V3 x = v3(h.point[0], h.point[1], h.point[2]);
    const V3 from = stand + sn * lift_at(stand);
    D3 side0, side1;
    basis_of(d3(sn.x, sn.y, sn.z), &side0, &side1);
    LeafRegs regs;
    float best = 0;
    for (uint32_t j = 0; j < S.nLamps && h.leaf < S.nLeaves; ++j) {
        const LampDev& L = S.lamps[j];
        if (!hg_mirrored(L)) continue;
        const ChainOut walked = walked_chain(S.leafSteps, S.leafProgs[h.leaf], d3(from.x, from.y, from.z), side0, side1, m.width,
                                             chain_lamp(L), d3(x.x, x.y, x.z), regs);
        if (!walked.found) continue;
        const V3 point = v3((float)walked.point.x, (float)walked.point.y, (float)walked.point.z);
        const V3 toward = hg_lamp_centre(L) - point;
        const float dl = length(toward), D = length(point - stand);
        if (!(dl > 1e-4f && D > 1e-4f)) continue;
        if (!hg_lamp_faces(L, toward * (1.0f / dl))) continue;
        const double det = walked.spread[0] * walked.spread[2] - walked.spread[1] * walked.spread[1];
        if (!(det > 0)) continue;
        const float importance = hg_lamp_glare(L) * (float)__builtin_sqrt(det);
        if (!(importance > best)) continue;
        best = importance;
        const V3 d = point - from;
        const float reach = length(d);
        kdn_ray a;
        a.inside = 0;
        a.within = KDN_NO_MATERIAL;
        a.origin[0] = from.x;
        a.origin[1] = from.y;
        a.origin[2] = from.z;
        a.direction[0] = d.x / reach;
        a.direction[1] = d.y / reach;
        a.direction[2] = d.z / reach;
        a.far_ = reach * 1.05f + 0.01f;
        aims[at] = a;
        c[0] = 1;
        c[1] = (float)j;
        c[5] = (float)__builtin_sqrt(det);  // the spread's area as a round lump's sigma^2
        c[6] = m.width;
        c[8] = s[10];
        c[9] = importance;
        c[10] = __builtin_bit_cast(float, h.leaf);
        c[11] = (float)mi;
    }
}

// A start whose aim meets the same sharp metal with the lamp within its
// mirror direction's reach (the lamp's outline widened by the lobe): a
// candidate highlight there, of its chain's spread and importance.
__global__ void k_hg_candidates(HgScene S, const kdn_ray* aims, const kdn_hit* hits, uint32_t n, float* cand) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= n) return;
    float* c = cand + (size_t)at * kSpotFloats;
    if (c[0] == 0) return;
    c[0] = 0;
    const kdn_hit h = hits[at];
    if (h.kind != KDN_HIT_SURFACE) return;
    uint32_t mi;
    mat_of(S, h.leaf, &mi);
    if (mi != (uint32_t)c[11]) return;
    const MaterialDev& m = S.materials[mi];
    const LampDev& L = S.lamps[(uint32_t)c[1]];
    const V3 o = v3(aims[at].origin[0], aims[at].origin[1], aims[at].origin[2]);
    const V3 d = v3(aims[at].direction[0], aims[at].direction[1], aims[at].direction[2]);
    const V3 x = v3(h.point[0], h.point[1], h.point[2]);
    V3 nv = v3(h.normal[0], h.normal[1], h.normal[2]);
    if (dot(nv, d) > 0) nv = -nv;
    const V3 mirror = mirrored(-d, nv);
    const V3 toward = hg_lamp_centre(L) - x;
    const float dl = length(toward), D = length(x - o);
    if (!(dl > 1e-4f)) return;
    const V3 dir = toward * (1.0f / dl);
    if (!hg_lamp_faces(L, dir)) return;
    const float rho = hg_lamp_reach(L, dl);
    const float reach = rho + 2 * m.width;
    if (!(dot(mirror, dir) > f_cos(fmin_(reach, 1.5f)))) return;
    c[0] = 1;
    c[2] = x.x;
    c[3] = x.y;
    c[4] = x.z;
    c[7] = D;
    c[10] = __builtin_bit_cast(float, h.leaf);
    c[12] = nv.x;
    c[13] = nv.y;
    c[14] = nv.z;
}

// each survey point's highlights: its candidates less the ones of a
// leaf and lamp already kept within 0.05, the most important kSpots
__global__ void k_hg_spots(const float* cand, uint32_t nPoints, float* spots) {
    const uint32_t point = blockIdx.x * blockDim.x + threadIdx.x;
    if (point >= nPoints) return;
    float* out = spots + (size_t)point * (1 + kSpots * kSpotFloats);
    float* rec = out + 1;
    int kept = 0;
    for (int draw = 0; draw < kSurveyDraws; ++draw) {
        const float* c = cand + ((size_t)point * kSurveyDraws + draw) * kSpotFloats;
        if (c[0] == 0) continue;
        bool known = false;
        for (int k = 0; k < kept && !known; ++k) {
            const float* r = rec + k * kSpotFloats;
            const float dx = r[2] - c[2], dy = r[3] - c[3], dz = r[4] - c[4];
            known = __builtin_bit_cast(uint32_t, r[10]) == __builtin_bit_cast(uint32_t, c[10]) && r[1] == c[1] && dx * dx + dy * dy + dz * dz < kSameSpot * kSameSpot;
        }
        if (known) continue;
        int place = kept < kSpots ? kept : kSpots - 1;
        if (kept == kSpots && !(c[9] > rec[place * kSpotFloats + 9])) continue;
        while (place > 0 && rec[(place - 1) * kSpotFloats + 9] < c[9]) {
            for (int f = 0; f < kSpotFloats; ++f) rec[place * kSpotFloats + f] = rec[(place - 1) * kSpotFloats + f];
            --place;
        }
        for (int f = 0; f < kSpotFloats; ++f) rec[place * kSpotFloats + f] = c[f];
        if (kept < kSpots) ++kept;
    }
    out[0] = (float)kept;
}

// every pixel's guide (pixelGuideAt, survey.cpp:916-966, as above)
__global__ void k_hg_pixels(HgScene S, const float* stands, const float* spots, int width, int height, int cols, int rows,
                            float* guides) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= (uint32_t)(width * height)) return;
    float* g = guides + (size_t)at * kHgFloats;
    g[0] = 0;
    g[1] = 0;
    g[2] = 0;
    g[3] = 0;
    const float* s = stands + (size_t)at * kStandFloats;
    if (s[0] != 1 || cols == 0 || rows == 0) return;
    const uint32_t mi = (uint32_t)s[10];
    const MaterialDev& m = S.materials[mi];
    const V3 x = v3(s[1], s[2], s[3]), n = v3(s[4], s[5], s[6]), out = v3(s[7], s[8], s[9]);
    const int px = (int)(at % width), py = (int)(at / width);
    // the four nearest survey points and their nearness
    const float fx = fmin_(fmax_((px + 0.5f - 4.5f) / 8, 0.0f), (float)(cols - 1));
    const float fy = fmin_(fmax_((py + 0.5f - 4.5f) / 8, 0.0f), (float)(rows - 1));
    const int i0 = (int)fx, j0 = (int)fy;
    const int i1 = i0 + 1 < cols ? i0 + 1 : i0, j1 = j0 + 1 < rows ? j0 + 1 : j0;
    const float ax = fx - i0, ay = fy - j0;
    const int pi[4] = {i0, i1, i0, i1}, pj[4] = {j0, j0, j1, j1};
    const float blend[4] = {(1 - ax) * (1 - ay), ax * (1 - ay), (1 - ax) * ay, ax * ay};
    float heavy[kHgLumps];
    float lump[kHgLumps][kHgLumpFloats];
    float radiance[kHgLumps];
    uint32_t lampOf[kHgLumps];
    int count = 0;
    for (int q = 0; q < 4; ++q) {
        if (!(blend[q] > 0)) continue;
        if (q == 1 && i1 == i0) continue;
        if (q == 2 && j1 == j0) continue;
        if (q == 3 && (i1 == i0 || j1 == j0)) continue;
        const float* spot = spots + (size_t)(pj[q] * cols + pi[q]) * (1 + kSpots * kSpotFloats);
        const int kept = (int)spot[0];
        for (int k = 0; k < kept; ++k) {
            const float* h = spot + 1 + k * kSpotFloats;
            if ((uint32_t)h[8] != mi) continue;
            HgSeen seen;
            if (!hg_seen(h, x, n, out, m, S.lamps, S.materials, &seen)) continue;
            const uint32_t lamp = seen.lamp;
            const V3 point = seen.point, axis = seen.axis;
            const float f0 = seen.sigma;
            const float w = seen.weight * blend[q];
            if (!(w > 0)) continue;
            // the same highlight seen through another survey point adds
            // its weight to the lump already kept
            int same = -1;
            for (int k = 0; k < count && same < 0; ++k) {
                const V3 d = v3(lump[k][6], lump[k][7], lump[k][8]) - point;
                if (lampOf[k] == lamp && dot(d, d) < kSameSpot * kSameSpot) same = k;
            }
            if (same >= 0) {
                heavy[same] += w;
                lump[same][14] += w;
                while (same > 0 && heavy[same - 1] < heavy[same]) {
                    const float hw = heavy[same];
                    heavy[same] = heavy[same - 1];
                    heavy[same - 1] = hw;
                    const float r = radiance[same];
                    radiance[same] = radiance[same - 1];
                    radiance[same - 1] = r;
                    const uint32_t lo = lampOf[same];
                    lampOf[same] = lampOf[same - 1];
                    lampOf[same - 1] = lo;
                    for (int f = 0; f < kHgLumpFloats; ++f) {
                        const float v = lump[same][f];
                        lump[same][f] = lump[same - 1][f];
                        lump[same - 1][f] = v;
                    }
                    --same;
                }
                continue;
            }
            // keep the kHgLumps heaviest, heaviest first
            int place = count < kHgLumps ? count : kHgLumps - 1;
            if (count == kHgLumps && !(w > heavy[place])) continue;
            while (place > 0 && heavy[place - 1] < w) {
                heavy[place] = heavy[place - 1];
                radiance[place] = radiance[place - 1];
                lampOf[place] = lampOf[place - 1];
                for (int f = 0; f < kHgLumpFloats; ++f) lump[place][f] = lump[place - 1][f];
                --place;
            }
            const Basis B = basis_around(axis);
            const float rec[kHgLumpFloats] = {axis.x, axis.y, axis.z, B.t.x, B.t.y, B.t.z, point.x, point.y, point.z,
                                              0, 0, f0, 0, f0, w};
            for (int f = 0; f < kHgLumpFloats; ++f) lump[place][f] = rec[f];
            heavy[place] = w;
            lampOf[place] = lamp;
            radiance[place] = seen.light;
            if (count < kHgLumps) ++count;
        }
    }
    // drop the lumps under a hundredth of the heaviest
    int live = 0;
    while (live < count && heavy[live] >= 0.01f * heavy[0]) ++live;
    if (live == 0) return;
    // the share: the light the lumps bring against the direct light
    double spotsLight = 0;
    for (int k = 0; k < live; ++k) {
        const V3 axis = v3(lump[k][0], lump[k][1], lump[k][2]);
        spotsLight += (double)hg_responded(m, n, out, axis) * 2 * kPiD * lump[k][11] * lump[k][13] * radiance[k];
    }
    const float share = hg_share(spotsLight, m, x, n, out, S.lamps, S.nLamps);
    g[0] = (float)live;
    g[1] = (float)mi;
    g[2] = share;
    for (int k = 0; k < live; ++k)
        for (int f = 0; f < kHgLumpFloats; ++f) g[kHgHead + k * kHgLumpFloats + f] = lump[k][f];
}

}  // namespace

// The survey points that found highlights filed by cell (HgDeepHeader and
// what it points at), built on the host from the survey's stands and
// highlights: at most kHgDeepFiled points a cell, those with the most
// highlights first.
struct DeepFiling {
    std::vector<uint32_t> table, list;
    std::vector<float> place;
};

static DeepFiling fileDeep(const Ctx& c, const std::vector<float>& stands, const std::vector<float>& spots, uint32_t nPoints) {
    DeepFiling f;
    const uint32_t G = kHgDeepGrid;
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> cells((size_t)G * G * G);  // (highlights, point)
    f.place.resize((size_t)nPoints * 3);
    for (uint32_t p = 0; p < nPoints; ++p) {
        const float* s = &stands[(size_t)p * kStandFloats];
        for (int k = 0; k < 3; ++k) f.place[3 * p + k] = s[1 + k];
        const uint32_t kept = (uint32_t)spots[(size_t)p * kHgPointFloats];
        if (s[0] != 1 || kept == 0) continue;
        uint32_t at[3];
        bool inside = true;
        for (int k = 0; k < 3; ++k) {
            const float q = (s[1 + k] - c.tree.lower[k]) * (G / c.tree.size[k]);
            inside = inside && q >= 0 && q < (float)G;
            at[k] = inside ? std::min<uint32_t>((uint32_t)q, G - 1) : 0;
        }
        if (inside) cells[((size_t)at[2] * G + at[1]) * G + at[0]].push_back({kept, p});
    }
    f.table.assign((size_t)G * G * G * 2, 0);
    for (size_t cell = 0; cell < cells.size(); ++cell) {
        auto& v = cells[cell];
        std::stable_sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        const uint32_t n = (uint32_t)std::min<size_t>(v.size(), kHgDeepFiled);
        f.table[2 * cell] = (uint32_t)f.list.size();
        f.table[2 * cell + 1] = n;
        for (uint32_t i = 0; i < n; ++i) f.list.push_back(v[i].second);
    }
    return f;
}

// Survey and pixel guides for an image of width x height of the context's
// scene and camera, into out (kHgFloats a pixel, then the deep guides'
// filing at byte *deepOffset, 0 for none), with work as the survey's
// scratch.
bool hguideBuild(Ctx& c, int width, int height, kdn_hguide_stats* stats, DevBuf& out, DevBuf& work, size_t* deepOffset) {
    *deepOffset = 0;
    hipStream_t st = c.stream;
    const kdn_camera& cam = c.scene.camera;
    double fwd[3], rt[3], upv[3], len = 0;
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
    HgLens L;
    const double reach = std::tan((double)cam.fov / 2), sideways = reach * ((double)width / height);
    for (int k = 0; k < 3; ++k) {
        L.corner[k] = fwd[k] - rt[k] * sideways + upv[k] * reach;
        L.across[k] = rt[k] * (2 * sideways / width);
        L.down[k] = upv[k] * (-2 * reach / height);
        L.eye[k] = cam.eye[k];
    }
    HgScene S{c.scene.leaf_material, (const MaterialDev*)c.materials_dev, c.n_materials_dev, (const LampDev*)c.lamps_dev, c.n_lamps,
              (const LeafStepDev*)c.leaf_steps_dev, (const LeafProgDev*)c.leaf_progs_dev, c.n_leaf_progs};
    const int cols = width > 4 ? (width - 4 + 7) / 8 : 0, rows = height > 4 ? (height - 4 + 7) / 8 : 0;
    const uint32_t nPoints = (uint32_t)(cols * rows), nDraws = nPoints * kSurveyDraws, nPixels = (uint32_t)(width * height);
    // buffers: rays and hits (the largest of draws and pixels), stands, candidates, spots, guides
    const size_t nRays = std::max<size_t>(nDraws, nPixels);
    const size_t bytes = nRays * (sizeof(kdn_ray) + sizeof(kdn_hit)) + (size_t)nDraws * sizeof(kdn_ray) + (size_t)std::max<size_t>(nPoints, nPixels) * kStandFloats * 4 +
                         (size_t)nDraws * kSpotFloats * 4 + (size_t)nPoints * (1 + kSpots * kSpotFloats) * 4 +
                         (size_t)nPoints * kMostSeeds * 4 + 10 * 256;
    if (!ensure(work, bytes, c.error, false)) return false;
    char* at = (char*)work.ptr;
    auto take = [&](size_t b) {
        char* p = at;
        at += (b + 255) & ~size_t(255);
        return p;
    };
    kdn_ray* rays = (kdn_ray*)take(nRays * sizeof(kdn_ray));
    kdn_hit* hits = (kdn_hit*)take(nRays * sizeof(kdn_hit));
    kdn_ray* aims = (kdn_ray*)take((size_t)nDraws * sizeof(kdn_ray));
    float* stands = (float*)take((size_t)std::max<size_t>(nPoints, nPixels) * kStandFloats * 4);
    float* cand = (float*)take((size_t)nDraws * kSpotFloats * 4);
    float* spots = (float*)take((size_t)nPoints * (1 + kSpots * kSpotFloats) * 4);
    uint32_t* seeds = (uint32_t*)take((size_t)nPoints * kMostSeeds * 4);
    uint32_t* seedCount = (uint32_t*)take(4);
    const bool surveyed = nPoints > 0 && c.n_lamps > 0;
    const size_t pixelBytes = ((size_t)nPixels * kHgFloats * 4 + 255) & ~size_t(255);
    const size_t G3 = (size_t)kHgDeepGrid * kHgDeepGrid * kHgDeepGrid;
    // the filing's words after its header: table, list, places, highlights
    constexpr size_t headWords = 16;
    static_assert(sizeof(HgDeepHeader) <= headWords * 4, "the filing's header fits its 64 bytes");
    const size_t tableWords = G3 * 2, listWords = nPoints, placeWords = (size_t)nPoints * 3, spotWords = (size_t)nPoints * kHgPointFloats;
    const size_t deepBytes = surveyed ? (headWords + tableWords + listWords + placeWords + spotWords) * 4 : 0;
    if (!ensure(out, pixelBytes + deepBytes, c.error, false)) return false;
    auto grid = [](uint32_t n) { return dim3((n + kHgBlock - 1) / kHgBlock); };
    hipEvent_t ev[2];
    KDN_HIP(hipEventCreate(&ev[0]));
    KDN_HIP(hipEventCreate(&ev[1]));
    KDN_HIP(hipEventRecord(ev[0], st));
    uint32_t found = 0;
    if (nPoints > 0 && c.n_lamps > 0) {
        hipLaunchKernelGGL(k_hg_rays, grid(nPoints), dim3(kHgBlock), 0, st, L, width, height, 8, 4, cols, rows, rays);
        if (!extendRays(c, rays, nPoints, hits, st)) return false;
        hipLaunchKernelGGL(k_hg_stands, grid(nPoints), dim3(kHgBlock), 0, st, S, (const kdn_ray*)rays, (const kdn_hit*)hits, nPoints, stands);
        hipLaunchKernelGGL(k_hg_draws, grid(nDraws), dim3(kHgBlock), 0, st, S, (const float*)stands, nPoints, rays);
        if (!extendRays(c, rays, nDraws, hits, st)) return false;
        KDN_HIP(hipMemsetAsync(seedCount, 0, 4, st));
        hipLaunchKernelGGL(k_hg_seeds, grid(nPoints), dim3(kHgBlock), 0, st, S, (const kdn_ray*)rays, (const kdn_hit*)hits, nPoints,
                           seeds, seedCount, aims, cand);
        hipLaunchKernelGGL(k_hg_walks, dim3((nPoints * kMostSeeds + 63) / 64), dim3(64), 0, st, S, (const float*)stands,
                           (const kdn_hit*)hits, (const uint32_t*)seeds, (const uint32_t*)seedCount, aims, cand);
        if (!extendRays(c, aims, nDraws, hits, st)) return false;
        hipLaunchKernelGGL(k_hg_candidates, grid(nDraws), dim3(kHgBlock), 0, st, S, (const kdn_ray*)aims, (const kdn_hit*)hits,
                           nDraws, cand);
        hipLaunchKernelGGL(k_hg_spots, grid(nPoints), dim3(kHgBlock), 0, st, (const float*)cand, nPoints, spots);
        // the deep guides' filing, from the survey's stands (before the
        // pixels' stands take their place) and highlights
        std::vector<float> hostStands((size_t)nPoints * kStandFloats), hostSpots((size_t)nPoints * kHgPointFloats);
        KDN_HIP(hipMemcpyAsync(hostStands.data(), stands, hostStands.size() * 4, hipMemcpyDeviceToHost, st));
        KDN_HIP(hipMemcpyAsync(hostSpots.data(), spots, hostSpots.size() * 4, hipMemcpyDeviceToHost, st));
        KDN_HIP(hipStreamSynchronize(st));
        const DeepFiling filing = fileDeep(c, hostStands, hostSpots, nPoints);
        HgDeepHeader head{};
        head.grid = kHgDeepGrid;
        for (int k = 0; k < 3; ++k) {
            head.lower[k] = c.tree.lower[k];
            head.inv[k] = kHgDeepGrid / c.tree.size[k];
        }
        head.points = nPoints;
        head.table = (uint32_t)headWords;
        head.list = (uint32_t)(headWords + tableWords);
        head.place = (uint32_t)(headWords + tableWords + listWords);
        head.spots = (uint32_t)(headWords + tableWords + listWords + placeWords);
        uint32_t* words = (uint32_t*)((char*)out.ptr + pixelBytes);
        KDN_HIP(hipMemcpy(words, &head, sizeof head, hipMemcpyHostToDevice));
        KDN_HIP(hipMemcpy(words + head.table, filing.table.data(), filing.table.size() * 4, hipMemcpyHostToDevice));
        if (!filing.list.empty()) KDN_HIP(hipMemcpy(words + head.list, filing.list.data(), filing.list.size() * 4, hipMemcpyHostToDevice));
        KDN_HIP(hipMemcpy(words + head.place, filing.place.data(), filing.place.size() * 4, hipMemcpyHostToDevice));
        KDN_HIP(hipMemcpyAsync(words + head.spots, spots, spotWords * 4, hipMemcpyDeviceToDevice, st));
        *deepOffset = pixelBytes;
    }
    // the pixels' stands, then their guides
    hipLaunchKernelGGL(k_hg_rays, grid(nPixels), dim3(kHgBlock), 0, st, L, width, height, 1, 0, width, height, rays);
    if (!extendRays(c, rays, nPixels, hits, st)) return false;
    hipLaunchKernelGGL(k_hg_stands, grid(nPixels), dim3(kHgBlock), 0, st, S, (const kdn_ray*)rays, (const kdn_hit*)hits, nPixels, stands);
    hipLaunchKernelGGL(k_hg_pixels, grid(nPixels), dim3(kHgBlock), 0, st, S, (const float*)stands, (const float*)spots, width, height,
                       (nPoints > 0 && c.n_lamps > 0) ? cols : 0, rows, (float*)out.ptr);
    KDN_HIP(hipEventRecord(ev[1], st));
    KDN_HIP(hipEventSynchronize(ev[1]));
    KDN_HIP(hipGetLastError());
    float ms = 0;
    KDN_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    KDN_HIP(hipEventDestroy(ev[0]));
    KDN_HIP(hipEventDestroy(ev[1]));
    if (stats) {
        // guided pixels and highlights, counted on the host
        std::vector<float> heads((size_t)nPixels * kHgFloats);
        KDN_HIP(hipMemcpy(heads.data(), out.ptr, heads.size() * 4, hipMemcpyDeviceToHost));
        uint32_t guided = 0;
        double lumps = 0, share = 0;
        for (uint32_t p = 0; p < nPixels; ++p) {
            if (heads[(size_t)p * kHgFloats] > 0) {
                ++guided;
                lumps += heads[(size_t)p * kHgFloats];
                share += heads[(size_t)p * kHgFloats + 2];
            }
        }
        std::vector<float> sp((size_t)nPoints * (1 + kSpots * kSpotFloats));
        if (surveyed) {
            KDN_HIP(hipMemcpy(sp.data(), spots, sp.size() * 4, hipMemcpyDeviceToHost));
            for (uint32_t q = 0; q < nPoints; ++q) found += (uint32_t)sp[(size_t)q * (1 + kSpots * kSpotFloats)];
        }
        stats->ms = ms;
        stats->survey_points = nPoints;
        stats->highlights = found;
        stats->guided_pixels = guided;
        stats->mean_lumps = guided ? lumps / guided : 0;
        stats->mean_share = guided ? share / guided : 0;
    }
    return true;
}

// chains walked for queries (kdn_highlight_chains): one lane a chain, its
// leaf's hyper-dual registers in the lane's own memory
__global__ void k_hg_chains(const LeafStepDev* steps, const LeafProgDev* progs, uint32_t nLeaves, const LampDev* lamps, uint32_t nLamps,
                            const kdn_chain_query* queries, uint32_t n, kdn_chain* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    LeafRegs regs;
    const kdn_chain_query q = queries[i];
    kdn_chain r;
    r.found = 0;
    r.pad = 0;
    for (int k = 0; k < 3; ++k) {
        r.point[k] = r.axis[k] = r.across[k] = r.up[k] = r.spread[k] = 0;
        r.motion[0][k] = r.motion[1][k] = 0;
    }
    if (q.leaf < nLeaves && q.light < nLamps) {
        const ChainOut o = walked_chain(steps, progs[q.leaf], d3(q.from[0], q.from[1], q.from[2]),
                                        d3(q.sideways[0][0], q.sideways[0][1], q.sideways[0][2]),
                                        d3(q.sideways[1][0], q.sideways[1][1], q.sideways[1][2]), q.width, chain_lamp(lamps[q.light]),
                                        d3(q.seed[0], q.seed[1], q.seed[2]), regs);
        if (o.found) {
            r.found = 1;
            const D3 v[4] = {o.point, o.axis, o.across, o.up};
            for (int k = 0; k < 3; ++k) {
                r.point[k] = compd(v[0], k);
                r.axis[k] = compd(v[1], k);
                r.across[k] = compd(v[2], k);
                r.up[k] = compd(v[3], k);
                r.spread[k] = o.spread[k];
                r.motion[0][k] = compd(o.motion[0], k);
                r.motion[1][k] = compd(o.motion[1], k);
            }
        }
    }
    out[i] = r;
}

}  // namespace kdn

extern "C" int kdn_highlight_chains(kdn_ctx* ctx, const kdn_chain_query* queries, uint32_t n, kdn_chain* chains) {
    if (!ctx || (n && (!queries || !chains))) return KDN_E_ARGUMENT;
    kdn::Ctx& c = ctx->c;
    if (!c.integrator_ready || !c.leaf_progs_dev) {
        c.error = "kdn_highlight_chains: no scene on the device";
        return KDN_E_STATE;
    }
    if (n > (1u << 22)) {
        c.error = "kdn_highlight_chains: at most 2^22 chains a call";
        return KDN_E_ARGUMENT;
    }
    if (n == 0) return KDN_OK;
    hipSetDevice(c.device);
    const size_t qb = (size_t)n * sizeof(kdn_chain_query), cb = (size_t)n * sizeof(kdn_chain);
    if (!kdn::ensure(c.scratch[6], qb + cb, c.error, false)) return KDN_E_MEMORY;
    kdn_chain_query* q = (kdn_chain_query*)c.scratch[6].ptr;
    kdn_chain* out = (kdn_chain*)((char*)c.scratch[6].ptr + qb);
    if (hipMemcpyAsync(q, queries, qb, hipMemcpyHostToDevice, c.stream) != hipSuccess) {
        c.error = "kdn_highlight_chains: copy failed";
        return KDN_E_DEVICE;
    }
    hipLaunchKernelGGL(kdn::k_hg_chains, dim3((n + 63) / 64), dim3(64), 0, c.stream, (const kdn::LeafStepDev*)c.leaf_steps_dev,
                       (const kdn::LeafProgDev*)c.leaf_progs_dev, c.n_leaf_progs, (const kdn::LampDev*)c.lamps_dev, c.n_lamps, q, n, out);
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(chains, out, cb, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
        c.error = "kdn_highlight_chains: device call failed";
        return KDN_E_DEVICE;
    }
    return KDN_OK;
}

extern "C" int kdn_highlight_guides(kdn_ctx* ctx, int32_t width, int32_t height, kdn_hguide_stats* stats) {
    if (!ctx) return KDN_E_ARGUMENT;
    kdn::Ctx& c = ctx->c;
    if (width <= 0 || height <= 0 || width > 16384 || height > 16384) {
        c.error = "kdn_highlight_guides: image size out of 1..16384";
        return KDN_E_ARGUMENT;
    }
    if (!c.tree.built || !c.integrator_ready) {
        c.error = "kdn_highlight_guides: no scene and octree on the device";
        return KDN_E_STATE;
    }
    if (c.scene_has_glass && !KDN_GLASS_ON_DEVICE) {
        c.error = "kdn_highlight_guides: glass is held off the device";
        return KDN_E_STATE;
    }
    hipSetDevice(c.device);
    c.hguide_width = c.hguide_height = 0;
    if (!kdn::hguideBuild(c, width, height, stats, c.hguides, c.scratch[5], &c.hguide_deep)) return KDN_E_DEVICE;
    c.hguide_width = width;
    c.hguide_height = height;
    return KDN_OK;
}

extern "C" int kdn_highlight_guide_at(kdn_ctx* ctx, int32_t x, int32_t y, float* packed) {
    if (!ctx || !packed) return KDN_E_ARGUMENT;
    kdn::Ctx& c = ctx->c;
    if (!c.hguides.ptr || x < 0 || y < 0 || x >= c.hguide_width || y >= c.hguide_height) {
        c.error = "kdn_highlight_guide_at: no guide built there";
        return KDN_E_STATE;
    }
    hipSetDevice(c.device);
    const size_t at = ((size_t)y * c.hguide_width + x) * kdn::kHgFloats;
    if (hipMemcpy(packed, (const float*)c.hguides.ptr + at, kdn::kHgFloats * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        c.error = "kdn_highlight_guide_at: copy failed";
        return KDN_E_DEVICE;
    }
    return KDN_OK;
}

namespace kdn {
namespace {

// the mixture's lump density and draws of one packed guide at one vertex
__global__ void k_hg_density(const float* g, V3 place, V3 n, V3 out, const float* dirs, uint32_t count, float* density) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= count) return;
    density[at] = hg_density(g, place, n, out, v3(dirs[3 * at], dirs[3 * at + 1], dirs[3 * at + 2]));
}

__global__ void k_hg_draw(const float* g, V3 place, const float* numbers, uint32_t count, float* drawn) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= count) return;
    const V3 d = hg_draw(g, place, numbers[3 * at], numbers[3 * at + 1], numbers[3 * at + 2]);
    drawn[3 * at] = d.x;
    drawn[3 * at + 1] = d.y;
    drawn[3 * at + 2] = d.z;
}

bool checkedGuide(Ctx& c, const float* packed, const char* who) {
    const float count = packed[0];
    if (!(count >= 1 && count <= kHgLumps && count == (float)(int)count)) {
        c.error = std::string(who) + ": a guide holds 1 to 12 lumps";
        return false;
    }
    double total = 0;
    for (int k = 0; k < (int)count; ++k) {
        const float* p = packed + kHgHead + k * kHgLumpFloats;
        if (!(p[14] >= 0)) {
            c.error = std::string(who) + ": a lump's weight is negative";
            return false;
        }
        total += p[14];
    }
    if (!(total > 0)) {
        c.error = std::string(who) + ": the lumps weigh nothing";
        return false;
    }
    return true;
}

// host arrays through the device: the guide, n * 3 inputs, n or n * 3 outputs
template <typename Launch>
bool hgEvaluate(Ctx& c, const float* packed, const float* in, uint32_t n, float* out, uint32_t outFloats, Launch&& launch) {
    hipStream_t st = c.stream;
    const size_t gBytes = (size_t)kHgFloats * 4, inBytes = (size_t)n * 3 * 4, outBytes = (size_t)n * outFloats * 4;
    const size_t bytes = ((gBytes + 255) & ~size_t(255)) + ((inBytes + 255) & ~size_t(255)) + outBytes;
    if (!ensure(c.scratch[6], bytes, c.error, false)) return false;
    char* base = (char*)c.scratch[6].ptr;
    float* g = (float*)base;
    float* i = (float*)(base + ((gBytes + 255) & ~size_t(255)));
    float* o = (float*)((char*)i + ((inBytes + 255) & ~size_t(255)));
    KDN_HIP(hipMemcpyAsync(g, packed, gBytes, hipMemcpyHostToDevice, st));
    KDN_HIP(hipMemcpyAsync(i, in, inBytes, hipMemcpyHostToDevice, st));
    launch((const float*)g, (const float*)i, o, dim3((n + kHgBlock - 1) / kHgBlock), st);
    KDN_HIP(hipGetLastError());
    KDN_HIP(hipMemcpyAsync(out, o, outBytes, hipMemcpyDeviceToHost, st));
    KDN_HIP(hipStreamSynchronize(st));
    return true;
}

}  // namespace
}  // namespace kdn

extern "C" int kdn_highlight_density(kdn_ctx* ctx, const float* packed, const float place[3], const float normal[3], const float out[3],
                                     const float* dirs, uint32_t count, float* density) {
    if (!ctx || !packed || !place || !normal || !out || (count && (!dirs || !density))) return KDN_E_ARGUMENT;
    kdn::Ctx& c = ctx->c;
    if (!kdn::checkedGuide(c, packed, "kdn_highlight_density")) return KDN_E_ARGUMENT;
    if (count == 0) return KDN_OK;
    hipSetDevice(c.device);
    const kdn::V3 x{place[0], place[1], place[2]}, n{normal[0], normal[1], normal[2]}, o{out[0], out[1], out[2]};
    const bool ok = kdn::hgEvaluate(c, packed, dirs, count, density, 1, [&](const float* g, const float* i, float* r, dim3 grid, hipStream_t st) {
        hipLaunchKernelGGL(kdn::k_hg_density, grid, dim3(kdn::kHgBlock), 0, st, g, x, n, o, i, count, r);
    });
    return ok ? KDN_OK : KDN_E_DEVICE;
}

extern "C" int kdn_highlight_draw(kdn_ctx* ctx, const float* packed, const float place[3], const float* numbers, uint32_t count,
                                  float* drawn) {
    if (!ctx || !packed || !place || (count && (!numbers || !drawn))) return KDN_E_ARGUMENT;
    kdn::Ctx& c = ctx->c;
    if (!kdn::checkedGuide(c, packed, "kdn_highlight_draw")) return KDN_E_ARGUMENT;
    if (count == 0) return KDN_OK;
    hipSetDevice(c.device);
    const kdn::V3 x{place[0], place[1], place[2]};
    const bool ok = kdn::hgEvaluate(c, packed, numbers, count, drawn, 3, [&](const float* g, const float* i, float* r, dim3 grid, hipStream_t st) {
        hipLaunchKernelGGL(kdn::k_hg_draw, grid, dim3(kdn::kHgBlock), 0, st, g, x, i, count, r);
    });
    return ok ? KDN_OK : KDN_E_DEVICE;
}
