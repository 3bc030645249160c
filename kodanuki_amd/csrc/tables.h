// Host-built tables the integrator reads from HBM.
#pragma once

#include <cstdint>
#include <vector>

#include "kodanuki.h"

namespace kdn {

constexpr int kTableSize = 65;        // CIE table entries, 380..700 nm by 5 nm
constexpr int kWavelengthCuts = 256;  // importance cuts of the wavelength draw
constexpr int kEnergyViews = 32;      // escape table entries of a GGX finish

// The colour-matching tables and the wavelength importance cuts, as the
// reference builds them (source/math/colour.cpp:134-211).
struct ColourTables {
    float x[kTableSize], y[kTableSize], z[kTableSize];
    float white[3];
    float cuts[kWavelengthCuts + 1];
};

// A material as the integrator reads it: the scene's material with its
// GGX width, facet share and escape table (material.cpp:175-209,
// energy.cpp:100-127).
struct MaterialDev {
    uint32_t kind;
    uint32_t exact;        // smooth glass: a true mirror and passage
    float albedo[3];
    float width;           // GGX alpha
    float facet_share;     // the chance a rough draw follows the facets
    float escape_avg;
    float escape[kEnergyViews];
    float index;           // glass: Cauchy A
    float spread;          // glass: Cauchy B (um^2), 0 for none
    float glow[3];         // glow * strength
};

// A lamp as the integrator reads it (lampOf, light.cpp:77-94): a point,
// a directional light, a sphere, a quad or a sun.
// A leaf program's step (opcodes of kodanuki.h) and a leaf's program in the
// pool of every leaf's own program (the highlight survey's chains walk them).
struct LeafStepDev {
    uint32_t op, a, b, c;
    double value;
};
struct LeafProgDev {
    uint32_t first, count, result, pad;
};

struct LampDev {
    float at[3], along[3], across[3];
    float facing[3];
    float radiance[3];     // radianceOf (light.cpp:32-43): power over the area a kind spreads it on
    float area;
    float pick;            // the chance NEE picks it
    float pick_cdf;        // cumulative pick chance up to and with it
    uint32_t kind;
    float direction[3];    // unit, toward a sun or a directional light
    float radius;          // a sphere's
    float cosine;          // a sun's half angle's cosine
    float power[3];
};

void buildColourTables(ColourTables& t);
// The integrator's material and lamp tables from the scene's.
std::vector<MaterialDev> materialTable(const kdn_material* materials, size_t count);
std::vector<LampDev> lampTable(const kdn_light* lights, size_t count);
void escapeTable(double width, float table[kEnergyViews], float* average);
double facetWidthOf(double roughness);

}  // namespace kdn
