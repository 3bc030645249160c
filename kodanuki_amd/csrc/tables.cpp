// Host-built tables: colour matching, wavelength importance cuts and the
// GGX escape tables. They restate the reference's own constructions so
// that the integrator samples and weights wavelengths and rough metals as
// the reference does.
#include "tables.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

namespace kdn {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kShortest = 380.0, kLongest = 700.0, kStep = 5.0, kEvenShare = 0.3;
constexpr int kSlices = kTableSize - 1;

// Fitted CIE 1931 curves as sums of skewed Gaussian lobes
// (source/math/colour.cpp:105-129).
double skewedLobe(double nm, double peak, double below, double above) {
    const double width = nm < peak ? below : above;
    const double off = (nm - peak) / width;
    return std::exp(-0.5 * off * off);
}
double fittedX(double nm) {
    return 1.056 * skewedLobe(nm, 599.8, 37.9, 31.0) + 0.362 * skewedLobe(nm, 442.0, 16.0, 26.7) - 0.065 * skewedLobe(nm, 501.1, 20.4, 26.2);
}
double fittedY(double nm) { return 0.821 * skewedLobe(nm, 568.8, 46.9, 40.5) + 0.286 * skewedLobe(nm, 530.9, 16.3, 31.1); }
double fittedZ(double nm) { return 1.217 * skewedLobe(nm, 437.0, 11.8, 36.0) + 0.681 * skewedLobe(nm, 459.0, 26.0, 13.8); }

void linearOf(double x, double y, double z, double rgb[3]) {
    rgb[0] = 3.2406 * x - 1.5372 * y - 0.4986 * z;
    rgb[1] = -0.9689 * x + 1.8758 * y + 0.0415 * z;
    rgb[2] = 0.0557 * x - 0.2040 * y + 1.0570 * z;
}

// smithVisibility (microfacet.cpp:22-29) with cosines (leaving, arriving)
double smithVisibility(double leaving, double arriving, double width) {
    const double sq = width * width;
    const double first = arriving * std::sqrt(sq + (1 - sq) * leaving * leaving);
    const double second = leaving * std::sqrt(sq + (1 - sq) * arriving * arriving);
    const double total = first + second;
    return total > 0 ? 0.5 / total : 0;
}

}  // namespace

void buildColourTables(ColourTables& t) {
    double X[kTableSize], Y[kTableSize], Z[kTableSize];
    double sx = 0, sy = 0, sz = 0;
    for (int slot = 0; slot < kTableSize; ++slot) {
        const double nm = kShortest + kStep * slot;
        X[slot] = fittedX(nm);
        Y[slot] = fittedY(nm);
        Z[slot] = fittedZ(nm);
        const bool edge = slot == 0 || slot == kTableSize - 1;
        const double w = (edge ? 0.5 : 1.0) * kStep;
        sx += X[slot] * w;
        sy += Y[slot] * w;
        sz += Z[slot] * w;
    }
    double white[3];
    linearOf(sx, sy, sz, white);
    // slice shares: what the eye sees of each 5 nm slice, mixed with an even
    // share (colour.cpp:163-186)
    double seen[kTableSize];
    for (int slot = 0; slot < kTableSize; ++slot) {
        double rgb[3];
        linearOf(X[slot], Y[slot], Z[slot], rgb);
        seen[slot] = std::abs(rgb[0] / white[0]) + std::abs(rgb[1] / white[1]) + std::abs(rgb[2] / white[2]);
    }
    double total = 0;
    for (int s = 0; s < kSlices; ++s) total += seen[s] + seen[s + 1];
    double shares[kSlices];
    for (int s = 0; s < kSlices; ++s) shares[s] = (1 - kEvenShare) * (seen[s] + seen[s + 1]) / total + kEvenShare / kSlices;
    // the cuts: equal shares of the summed slices (colour.cpp:188-205)
    double before[kSlices + 1] = {0};
    for (int s = 0; s < kSlices; ++s) before[s + 1] = before[s] + shares[s];
    double cuts[kWavelengthCuts + 1];
    int slice = 0;
    for (int index = 0; index <= kWavelengthCuts; ++index) {
        const double u = double(index) / kWavelengthCuts * before[kSlices];
        while (slice < kSlices - 1 && before[slice + 1] <= u) ++slice;
        const double along = std::clamp((u - before[slice]) / shares[slice], 0.0, 1.0);
        cuts[index] = kShortest + kStep * (slice + along);
    }
    cuts[0] = kShortest;
    cuts[kWavelengthCuts] = kLongest;
    for (int i = 0; i < kTableSize; ++i) {
        t.x[i] = (float)X[i];
        t.y[i] = (float)Y[i];
        t.z[i] = (float)Z[i];
    }
    for (int k = 0; k < 3; ++k) t.white[k] = (float)white[k];
    for (int i = 0; i <= kWavelengthCuts; ++i) t.cuts[i] = (float)cuts[i];
}

double facetWidthOf(double roughness) {
    const double c = std::clamp(roughness, 0.0, 1.0);
    return std::max(c * c, 1e-3);
}

// What one GGX bounce returns at 32 view cosines, by a fixed grid of 24
// rings x 12 spokes of facets laid by the distribution (energy.cpp:45-118).
void escapeTable(double width, float table[kEnergyViews], float* average) {
    const int rings = 24, spokes = 12;
    std::vector<std::array<double, 3>> facets;
    for (int ring = 0; ring < rings; ++ring) {
        for (int spoke = 0; spoke < spokes; ++spoke) {
            const double along = (ring + 0.5) / rings;
            const double tilt = std::sqrt((1 - along) / (1 + (width * width - 1) * along));
            const double spread = std::sqrt(std::max(1 - tilt * tilt, 0.0));
            const double angle = 2 * kPi * (spoke + 0.5) / spokes;
            facets.push_back({spread * std::cos(angle), spread * std::sin(angle), tilt});
        }
    }
    double avg = 0;
    for (int view = 0; view < kEnergyViews; ++view) {
        const double leaving = (view + 0.5) / kEnergyViews;
        const double out[3] = {std::sqrt(std::max(1 - leaving * leaving, 0.0)), 0, leaving};
        double total = 0;
        for (const auto& f : facets) {
            const double onto = out[0] * f[0] + out[1] * f[1] + out[2] * f[2];
            if (onto <= 0) continue;
            const double inz = f[2] * (2 * onto) - out[2];
            if (inz <= 0) continue;
            total += 4 * smithVisibility(out[2], inz, width) * inz * onto / f[2];
        }
        const double e = total / (rings * spokes);
        table[view] = (float)e;
        avg += 2 * e * leaving / kEnergyViews;
    }
    *average = (float)avg;
}

std::vector<MaterialDev> materialTable(const kdn_material* mats, size_t count) {
    std::vector<MaterialDev> md;
    for (size_t i = 0; i < count; ++i) {
        const kdn_material& m = mats[i];
        MaterialDev x{};
        x.kind = m.kind;
        for (int k = 0; k < 3; ++k) {
            x.albedo[k] = m.albedo[k];
            x.glow[k] = m.glow[k] * m.strength;
        }
        const double width = facetWidthOf(m.kind == KDN_MATTE ? 1.0 : m.roughness);
        x.width = (float)width;
        if (m.kind == KDN_METAL) {
            escapeTable(width, x.escape, &x.escape_avg);
        } else {
            for (float& e : x.escape) e = 1.0f;
            x.escape_avg = 1.0f;
        }
        x.facet_share = std::min(std::max(x.escape_avg, 0.25f), 1.0f);  // material.cpp:196
        x.exact = m.kind == KDN_GLASS && m.roughness <= 0 ? 1u : 0u;
        x.index = m.refraction;
        x.spread = m.dispersion;
        md.push_back(x);
    }
    return md;
}

std::vector<LampDev> lampTable(const kdn_light* lights, size_t count) {
    std::vector<LampDev> lamps;
    double totalPower = 0;
    for (size_t i = 0; i < count; ++i) {
        const kdn_light& l = lights[i];
        LampDev x{};
        x.kind = l.kind == KDN_LIGHT_MOON ? KDN_LIGHT_SUN : l.kind;  // a moon is aimed at and met as a sun
        const double f[3] = {(double)l.along[1] * l.across[2] - (double)l.along[2] * l.across[1],
                             (double)l.along[2] * l.across[0] - (double)l.along[0] * l.across[2],
                             (double)l.along[0] * l.across[1] - (double)l.along[1] * l.across[0]};
        const double quad = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        const double dlen = std::sqrt((double)l.direction[0] * l.direction[0] + (double)l.direction[1] * l.direction[1] +
                                      (double)l.direction[2] * l.direction[2]);
        const double half = 0.5 * (double)l.angle, r = (double)l.radius;
        // radianceOf and areaOf (light.cpp:32-61); a light that spreads its
        // power on nothing is left out (as startsPaths and the aims do)
        double spread = 1, area = 0;
        switch (x.kind) {
            case KDN_LIGHT_QUAD:
                if (!(quad > 1e-12)) continue;
                spread = quad;
                area = quad;
                break;
            case KDN_LIGHT_SPHERE:
                if (!(r > 0)) continue;
                spread = kPi * r * r;
                area = 4 * kPi * r * r;
                break;
            case KDN_LIGHT_SUN:
                if (!(std::cos(half) < 1) || !(dlen > 0)) continue;
                spread = kPi * std::sin(half) * std::sin(half);
                break;
            case KDN_LIGHT_DIRECTIONAL:
                if (!(dlen > 0)) continue;
                break;
            case KDN_LIGHT_POINT:
                break;
            default:
                continue;
        }
        for (int k = 0; k < 3; ++k) {
            x.at[k] = l.at[k];
            x.along[k] = l.along[k];
            x.across[k] = l.across[k];
            x.facing[k] = quad > 1e-12 ? (float)(f[k] / quad) : 0.0f;
            x.radiance[k] = (float)(l.power[k] / spread);
            x.direction[k] = dlen > 0 ? (float)(l.direction[k] / dlen) : 0.0f;
            x.power[k] = l.power[k];
        }
        x.area = (float)area;
        x.radius = (float)r;
        x.cosine = (float)std::cos(half);
        // emittedOf (light.cpp:45-48): pi for quads, 4 pi otherwise, times the mean power
        const double emitted = std::max((l.power[0] + l.power[1] + l.power[2]) / 3.0, 0.0) * (x.kind == KDN_LIGHT_QUAD ? kPi : 4 * kPi);
        x.pick = (float)emitted;
        totalPower += emitted;
        lamps.push_back(x);
    }
    double cdf = 0;
    for (LampDev& x : lamps) {
        const double p = totalPower > 0 ? x.pick / totalPower : 1.0 / lamps.size();
        x.pick = (float)p;
        cdf += p;
        x.pick_cdf = (float)cdf;
    }
    if (!lamps.empty()) lamps.back().pick_cdf = 1.0f;
    return lamps;
}

}  // namespace kdn
