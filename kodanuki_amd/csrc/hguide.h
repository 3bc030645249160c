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

}  // namespace kdn
