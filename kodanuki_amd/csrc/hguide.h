// The plain (highlight) guide's per-lane arithmetic: Gaussian lumps over
// the plane one unit along an axis, their density and draws, and the
// one-sample mixture of the lumps (share s) with the material (1 - s).
//
// Restates the reference's highlight guide (plugins/sampler/
// highlight_guide): the lump (guide.h:62-99), unpackedLump and
// unpackGuide's weights over their total (guide.cpp:52-92), shiftedCentre
// (guide.cpp:131-139), lumpAt / lumpDensity / pickedLump / drawnFromLump
// (guide.cpp:210-252, cutoff 60, kept mass 1 - e^-30), mixtureDirection /
// mixtureDensity (mixture.cpp:57-124). A lump is unpacked where it is
// used, from the pixel's packed guide in HBM, rather than held unpacked
// per path: twelve unpacked lumps would not fit a lane's registers.
#pragma once

#include "integrator_dev.h"

namespace kdn {

constexpr int kHgLumps = 12;               // the most lumps a guide holds (guide.h:17)
constexpr int kHgLumpFloats = 15;          // axis 3, across 3, point 3, mean 2, factor 3, weight
constexpr int kHgHead = 4;                 // count, material, share, unused
constexpr int kHgFloats = kHgHead + kHgLumps * kHgLumpFloats;  // a pixel's packed guide
constexpr float kHgCutoff = 60;
constexpr double kHgKeptMass = 0.99999999999990641;  // 1 - e^(-30)

struct HgLump {
    V3 axis, across, up;
    float centre[2];     // the mean as seen from the vertex (shiftedCentre)
    float precision[3];
    float factor[3];
    float height;
    float weight;        // over the guide's total
};

KDN_D V3 hg_normalized(V3 a) {
    const double l = __builtin_sqrt((double)a.x * a.x + (double)a.y * a.y + (double)a.z * a.z);
    return v3((float)(a.x / l), (float)(a.y / l), (float)(a.z / l));
}

KDN_D double hg_total(const float* g) {
    double total = 0;
    for (int k = 0; k < (int)g[0]; ++k) total += g[kHgHead + k * kHgLumpFloats + 14];
    return total;
}

// lump k of packed guide g as seen from place
KDN_D HgLump hg_lump(const float* g, int k, double total, V3 place) {
    const float* p = g + kHgHead + k * kHgLumpFloats;
    HgLump L;
    L.axis = hg_normalized(v3(p[0], p[1], p[2]));
    const V3 raw = v3(p[3], p[4], p[5]);
    L.across = hg_normalized(raw - L.axis * dot(raw, L.axis));
    L.up = cross(L.axis, L.across);
    for (int j = 0; j < 3; ++j) L.factor[j] = p[11 + j];
    const double f0 = L.factor[0], f1 = L.factor[1], f2 = L.factor[2];
    const double c00 = f0 * f0, c01 = f0 * f1, c11 = f1 * f1 + f2 * f2;
    const double root = f0 * f2, det = root * root;
    L.precision[0] = (float)(c11 / det);
    L.precision[1] = (float)(-c01 / det);
    L.precision[2] = (float)(c00 / det);
    L.height = (float)(1 / (2 * kPiD * root));
    L.weight = (float)(p[14] / total);
    const V3 to = hg_normalized(v3(p[6], p[7], p[8]) - place);
    const float along = dot(to, L.axis);
    if (!(along > 1e-3f)) {
        L.centre[0] = p[9];
        L.centre[1] = p[10];
    } else {
        L.centre[0] = p[9] + dot(to, L.across) / along;
        L.centre[1] = p[10] + dot(to, L.up) / along;
    }
    return L;
}

KDN_D float hg_lump_at(const HgLump& L, V3 d) {
    const float along = dot(d, L.axis);
    if (along <= 1e-9f) return 0.0f;
    const float u = dot(d, L.across) / along - L.centre[0];
    const float v = dot(d, L.up) / along - L.centre[1];
    const float spread = L.precision[0] * u * u + 2 * L.precision[1] * u * v + L.precision[2] * v * v;
    if (spread > kHgCutoff) return 0.0f;
    return (float)(L.height / kHgKeptMass) * expf(-0.5f * spread) / (along * along * along);
}

// lumpDensity: zero unless both directions leave the surface's side
KDN_D float hg_density(const float* g, V3 place, V3 n, V3 out, V3 in) {
    if (!(dot(n, in) > 0 && dot(n, out) > 0)) return 0.0f;
    const double total = hg_total(g);
    float sum = 0;
    for (int k = 0; k < (int)g[0]; ++k) {
        const HgLump L = hg_lump(g, k, total, place);
        sum += L.weight * hg_lump_at(L, in);
    }
    return sum;
}

// pickedLump and drawnFromLump
KDN_D V3 hg_draw(const float* g, V3 place, float pick, float u, float v) {
    const double total = hg_total(g);
    const int count = (int)g[0];
    int index = 0;
    float below = (float)(g[kHgHead + 14] / total);
    while (index + 1 < count && pick >= below) {
        ++index;
        below += (float)(g[kHgHead + index * kHgLumpFloats + 14] / total);
    }
    const HgLump L = hg_lump(g, index, total, place);
    const float keep = 1 - u * (float)kHgKeptMass;
    const float radius = __builtin_sqrtf(-2 * logf(keep > 1e-30f ? keep : 1e-30f));
    const float angle = 2 * kPi * v;
    const float a = radius * f_cos(angle), b = radius * f_sin(angle);
    const float uu = L.centre[0] + L.factor[0] * a;
    const float vv = L.centre[1] + L.factor[1] * a + L.factor[2] * b;
    return hg_normalized(L.axis + L.across * uu + L.up * vv);
}

// ---- highlights as the survey keeps them, seen from a vertex -------------

constexpr int kHgSpots = 16;           // highlights a survey point keeps
// a highlight's record: kept, lamp, point (3), spread, facet width, distance
// from its stand, the stand's material, importance, leaf bits, the sharp
// surface's material, its normal there (3, facing the stand)
constexpr int kHgSpotFloats = 15;
constexpr int kHgPointFloats = 1 + kHgSpots * kHgSpotFloats;  // a survey point's highlights
constexpr float kHgShinyWidth = 0.125f;  // the widest facets a highlight is seen in
constexpr float kHgSameSpot = 0.05f;
constexpr float kHgMoved = 0.1f;         // the farthest a highlight walked again may move
constexpr double kHgEfficiency = 0.3;    // guideEfficiency (survey.cpp:818-828)
constexpr float kHgLeastShare = 0.02f, kHgMostShare = 0.5f;

// lampGlare (guide.cpp:269-281): the power's luminance over the area a
// kind spreads it on (a quad's, pi r^2 for a sphere, pi times the half
// angle squared for a sun), times 1e4 for a point or directional light
KDN_D float hg_lamp_glare(const LampDev& L) {
    const float bright = 0.2126f * L.power[0] + 0.7152f * L.power[1] + 0.0722f * L.power[2];
    switch (L.kind) {
        case KDN_LIGHT_SPHERE: return bright / (kPi * fmax_(L.radius * L.radius, 1e-12f));
        case KDN_LIGHT_QUAD: return bright / fmax_(L.area, 1e-12f);
        case KDN_LIGHT_SUN: {
            const float half = acosf(fmin_(fmax_(L.cosine, -1.0f), 1.0f));
            return bright / (kPi * fmax_(half * half, 1e-12f));
        }
        default: return bright * 1e4f;
    }
}

// The lamps a highlight mirrors: quads (from their centre, on their lit
// side), spheres and points (from anywhere); a far light has no point a
// mirror could image.
KDN_D bool hg_mirrored(const LampDev& L) {
    return L.kind == KDN_LIGHT_QUAD || L.kind == KDN_LIGHT_SPHERE || L.kind == KDN_LIGHT_POINT;
}

KDN_D V3 hg_lamp_centre(const LampDev& L) {
    if (L.kind != KDN_LIGHT_QUAD) return v3(L.at[0], L.at[1], L.at[2]);
    return v3(L.at[0] + 0.5f * (L.along[0] + L.across[0]), L.at[1] + 0.5f * (L.along[1] + L.across[1]),
              L.at[2] + 0.5f * (L.along[2] + L.across[2]));
}

// whether the lamp sends light along dir (unit, from a point toward it)
KDN_D bool hg_lamp_faces(const LampDev& L, V3 dir) {
    return L.kind != KDN_LIGHT_QUAD || -dot(v3(L.facing[0], L.facing[1], L.facing[2]), dir) > 0;
}

// the lamp's angular radius seen from distance dl
KDN_D float hg_lamp_reach(const LampDev& L, float dl) {
    return L.kind == KDN_LIGHT_QUAD ? 0.5f * __builtin_sqrtf(L.area) / dl : L.kind == KDN_LIGHT_SPHERE ? L.radius / dl : 0.0f;
}

KDN_D Lanes hg_grey_lanes() {
    Lanes w;
    for (int l = 0; l < kLanes; ++l) {
        w.nm[l] = 550;
        w.density[l] = 1;
    }
    return w;
}

// the material's grey response toward in, times the cosine
KDN_D float hg_responded(const MaterialDev& m, V3 n, V3 out, V3 in) {
    const float facing = dot(n, in);
    if (facing <= 0) return 0.0f;
    float f[kLanes];
    rough_value(m, hg_grey_lanes(), n, out, in, f);
    return 0.25f * (f[0] + f[1] + f[2] + f[3]) * facing;
}

// A highlight seen from a stand x (normal n, outgoing out, material m):
// walked again from x (the point of the highlight's tangent plane that
// mirrors the lamp's centre toward x, kept when it moves less than
// kHgMoved), the lump's axis toward it, its spread (the spread a flat
// glossy mirror of the highlight's facet width gives, seen from x), its
// weight (the lamp's glare times the lump's size times the material's
// density toward it: weighedLump, survey.cpp:427-430) and the light it
// brings (glare times the sharp surface's albedo).
struct HgSeen {
    V3 axis, point;
    float sigma, weight, light;
    uint32_t lamp;
};

KDN_D bool hg_seen(const float* h, V3 x, V3 n, V3 out, const MaterialDev& m, const LampDev* lamps, const MaterialDev* materials,
                   HgSeen* seen) {
    const uint32_t lamp = (uint32_t)h[1];
    const LampDev& L = lamps[lamp];
    const V3 centre = hg_lamp_centre(L);
    V3 point = v3(h[2], h[3], h[4]);
    {
        const V3 mirrorN = v3(h[12], h[13], h[14]);
        const float above = dot(centre - point, mirrorN);
        const V3 image = centre - mirrorN * (2 * above);
        const float t = dot(point - x, mirrorN) / dot(image - x, mirrorN);
        if (above > 0 && t > 0 && t < 1) {
            const V3 moved = x + (image - x) * t;
            if (length(moved - point) < kHgMoved) point = moved;
        }
    }
    const V3 towards = point - x;
    const float dist = length(towards), dl = length(centre - point);
    if (!(dist > 1e-4f && dl > 1e-4f)) return false;
    const V3 axis = towards * (1.0f / dist);
    const float rho = hg_lamp_reach(L, dl);
    const float sigma = fmax_((2 * h[6] + rho) * dl / (dist + dl), 1e-3f);
    if (!(sigma < 1)) return false;
    const float weight = hg_lamp_glare(L) * sigma * sigma * rough_density(m, n, out, axis);
    if (!(weight > 0)) return false;
    const MaterialDev& sm = materials[(uint32_t)h[11]];
    seen->axis = axis;
    seen->point = point;
    seen->sigma = sigma;
    seen->weight = weight;
    seen->light = hg_lamp_glare(L) * (0.2126f * sm.albedo[0] + 0.7152f * sm.albedo[1] + 0.0722f * sm.albedo[2]);
    seen->lamp = lamp;
    return true;
}

// The share of draws a guide takes: the light its lumps bring (spotsLight)
// against the direct light of the lamps, the former over sqrt(0.3), within
// [0.02, 0.5] (guideShare, survey.cpp:818-828).
KDN_D float hg_share(double spotsLight, const MaterialDev& m, V3 x, V3 n, V3 out, const LampDev* lamps, uint32_t nLamps) {
    double rest = 0;
    for (uint32_t j = 0; j < nLamps; ++j) {
        const LampDev& L = lamps[j];
        const float power = 0.2126f * L.power[0] + 0.7152f * L.power[1] + 0.0722f * L.power[2];
        if (L.kind == KDN_LIGHT_QUAD) {
            const V3 toward = hg_lamp_centre(L) - x;
            const float dl = length(toward);
            const V3 dir = toward * (1.0f / dl);
            if (!(-dot(v3(L.facing[0], L.facing[1], L.facing[2]), dir) > 0)) continue;
            const float size = fmax_(0.5f * __builtin_sqrtf(L.area) / dl, 1e-3f);
            rest += (double)hg_responded(m, n, out, dir) * kPiD * size * size * hg_lamp_glare(L);
        } else if (L.kind == KDN_LIGHT_POINT || L.kind == KDN_LIGHT_SPHERE) {
            // a point's or a sphere's irradiance: power / distance^2
            const V3 toward = v3(L.at[0], L.at[1], L.at[2]) - x;
            const float dl2 = fmax_(dot(toward, toward), 1e-12f);
            rest += (double)hg_responded(m, n, out, toward * (1.0f / __builtin_sqrtf(dl2))) * power / dl2;
        } else {
            // a directional light's or a sun's irradiance: its power
            rest += (double)hg_responded(m, n, out, v3(L.direction[0], L.direction[1], L.direction[2])) * power;
        }
    }
    const double spotted = spotsLight / __builtin_sqrt(kHgEfficiency);
    const double all = spotted + rest;
    double share = all > 1e-30 ? spotted / all : 0.0;
    share = share < kHgLeastShare ? kHgLeastShare : (share > kHgMostShare ? kHgMostShare : share);
    return (float)share;
}

// A round lump of spread sigma about an axis, centred (a lump made at the
// vertex that reads it: its point lies on its axis)
KDN_D HgLump hg_round_lump(V3 axis, float sigma, float weight) {
    HgLump L;
    L.axis = axis;
    L.across = basis_around(axis).t;
    L.up = cross(axis, L.across);
    L.centre[0] = L.centre[1] = 0;
    const double s2 = (double)sigma * sigma;
    L.precision[0] = L.precision[2] = (float)(1 / s2);
    L.precision[1] = 0;
    L.factor[0] = L.factor[2] = sigma;
    L.factor[1] = 0;
    L.height = (float)(1 / (2 * kPiD * s2));
    L.weight = weight;
    return L;
}

// ---- deep guides: the survey's highlights filed by cell ------------------
// The survey points that found highlights are filed by the cell of a
// kHgDeepGrid^3 grid over the scene's region they stand in, at most
// kHgDeepFiled a cell (those with the most highlights). A deeper vertex of
// a broad surface (facets wider than kHgShinyWidth) takes the highlights
// of the two filed points of its cell nearest to it, on its own material,
// each walked again from it as above, as a mixture of round lumps weighed
// as above: the reference's deep guides (guide.cpp:337-450, steerDeep,
// mixture.cpp:133) restated by this build's survey.

constexpr int kHgDeepGrid = 32;
constexpr int kHgDeepFiled = 8;
constexpr int kHgDeepNearest = 2;

struct HgDeepHeader {
    uint32_t grid;
    float lower[3];
    float inv[3];        // cells per unit along each axis
    uint32_t points;     // survey points
    uint32_t table;      // offsets in 4-byte words from the header
    uint32_t list;
    uint32_t place;      // survey points' stands (3 floats each)
    uint32_t spots;      // survey points' highlights (kHgPointFloats each)
};

struct HgDeepView {
    const float* spots;
    uint32_t point[kHgDeepNearest];
    int count;
    float total;         // the lumps' weights
    float share;
};

// the deep guide of a vertex x of material mi (false: none)
KDN_D bool hg_deep_view(const HgDeepHeader* D, V3 x, V3 n, V3 out, uint32_t mi, const MaterialDev& m,
                                          const LampDev* lamps, uint32_t nLamps, const MaterialDev* materials, HgDeepView* v) {
    const uint32_t* words = (const uint32_t*)D;
    const float xs[3] = {x.x, x.y, x.z};
    uint32_t cell[3];
    for (int k = 0; k < 3; ++k) {
        const float f = (xs[k] - D->lower[k]) * D->inv[k];
        if (!(f >= 0 && f < (float)D->grid)) return false;
        cell[k] = (uint32_t)f;
        if (cell[k] >= D->grid) cell[k] = D->grid - 1;
    }
    const uint32_t at = (cell[2] * D->grid + cell[1]) * D->grid + cell[0];
    const uint32_t first = words[D->table + 2 * at], filed = words[D->table + 2 * at + 1];
    if (filed == 0) return false;
    const float* place = (const float*)(words + D->place);
    // the nearest kHgDeepNearest filed points
    float best[kHgDeepNearest];
    v->count = 0;
    for (uint32_t f = 0; f < filed && f < (uint32_t)kHgDeepFiled; ++f) {
        const uint32_t p = words[D->list + first + f];
        const V3 d = v3(place[3 * p], place[3 * p + 1], place[3 * p + 2]) - x;
        const float dd = dot(d, d);
        int k = v->count < kHgDeepNearest ? v->count : kHgDeepNearest - 1;
        if (v->count == kHgDeepNearest && !(dd < best[k])) continue;
        while (k > 0 && best[k - 1] > dd) {
            best[k] = best[k - 1];
            v->point[k] = v->point[k - 1];
            --k;
        }
        best[k] = dd;
        v->point[k] = p;
        if (v->count < kHgDeepNearest) ++v->count;
    }
    v->spots = (const float*)(words + D->spots);
    double total = 0, spotsLight = 0;
    for (int q = 0; q < v->count; ++q) {
        const float* spot = v->spots + (size_t)v->point[q] * kHgPointFloats;
        for (int k = 0; k < (int)spot[0]; ++k) {
            const float* h = spot + 1 + k * kHgSpotFloats;
            HgSeen s;
            if ((uint32_t)h[8] != mi || !hg_seen(h, x, n, out, m, lamps, materials, &s)) continue;
            total += s.weight;
            spotsLight += (double)hg_responded(m, n, out, s.axis) * 2 * kPiD * s.sigma * s.sigma * s.light;
        }
    }
    if (!(total > 0)) return false;
    v->total = (float)total;
    v->share = hg_share(spotsLight, m, x, n, out, lamps, nLamps);
    return true;
}

// the deep guide's density of a direction in (lumpDensity over its lumps)
KDN_D float hg_deep_density(const HgDeepView& v, V3 x, V3 n, V3 out, uint32_t mi, const MaterialDev& m,
                                              const LampDev* lamps, const MaterialDev* materials, V3 in) {
    if (!(dot(n, in) > 0 && dot(n, out) > 0)) return 0.0f;
    float sum = 0;
    for (int q = 0; q < v.count; ++q) {
        const float* spot = v.spots + (size_t)v.point[q] * kHgPointFloats;
        for (int k = 0; k < (int)spot[0]; ++k) {
            const float* h = spot + 1 + k * kHgSpotFloats;
            HgSeen s;
            if ((uint32_t)h[8] != mi || !hg_seen(h, x, n, out, m, lamps, materials, &s)) continue;
            sum += hg_lump_at(hg_round_lump(s.axis, s.sigma, 1.0f), in) * (s.weight / v.total);
        }
    }
    return sum;
}

// a direction drawn from the deep guide's lumps (pickedLump, drawnFromLump)
KDN_D V3 hg_deep_draw(const HgDeepView& v, V3 x, V3 n, V3 out, uint32_t mi, const MaterialDev& m,
                                        const LampDev* lamps, const MaterialDev* materials, float pick, float u, float w) {
    float below = 0;
    HgSeen last;
    bool any = false, picked = false;
    for (int q = 0; q < v.count && !picked; ++q) {
        const float* spot = v.spots + (size_t)v.point[q] * kHgPointFloats;
        for (int k = 0; k < (int)spot[0] && !picked; ++k) {
            const float* h = spot + 1 + k * kHgSpotFloats;
            HgSeen s;
            if ((uint32_t)h[8] != mi || !hg_seen(h, x, n, out, m, lamps, materials, &s)) continue;
            last = s;
            any = true;
            below += s.weight / v.total;
            picked = pick < below;  // the first lump whose running sum passes pick
        }
    }
    if (!any) return v3(0, 0, 0);
    const HgLump L = hg_round_lump(last.axis, last.sigma, 1.0f);
    const float keep = 1 - u * (float)kHgKeptMass;
    const float radius = __builtin_sqrtf(-2 * logf(keep > 1e-30f ? keep : 1e-30f));
    const float angle = 2 * kPi * w;
    const float a = radius * f_cos(angle), b = radius * f_sin(angle);
    return hg_normalized(L.axis + L.across * (L.factor[0] * a) + L.up * (L.factor[1] * a + L.factor[2] * b));
}

}  // namespace kdn
