// The plain spectral integrator on the device: counter-based numbers,
// hero wavelengths, camera, materials, lamps and sky.
//
// Restates the reference transport (plugins/transport/paths) for the
// plain configuration: numbers keyed (stream, sample, dimension)
// (trace.cpp:47-74, path.cpp:15-46, random.cpp:19-21); four wavelength
// lanes by the CIE-importance cuts (colour.cpp:236-255); RGB to spectrum
// by smooth-step bases (spectrum.cpp:49-67, colour.cpp:88-103); matte,
// GGX metal with energy compensation and exact dispersive glass
// (material.cpp, microfacet.cpp); quad lamps met along rays and aimed at
// (light.cpp); the gradient sky (light.cpp:582-591); power-heuristic MIS
// between aiming and meeting; roulette from depth 4 with survival
// min(0.95, max lane) (trace.cpp:1051-1055).
#pragma once

#include "tables.h"
#include "trace_dev.h"

namespace kdn {

constexpr float kPi = 3.14159265358979f;
constexpr int kLanes = 4;

// ---- numbers --------------------------------------------------------------
KDN_D uint64_t mixed_bits(uint64_t v) {
    v += 0x9E3779B97F4A7C15ULL;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
    return v ^ (v >> 31);
}
KDN_D double hashed(uint64_t stream, uint64_t index, uint64_t coordinate) {
    const uint64_t bits = mixed_bits(stream + mixed_bits(index + mixed_bits(coordinate)));
    return (double)(bits >> 11) * 0x1p-53;
}
// dimensionOf for the per-event draws LightU..Pick (path.cpp:22-37)
enum Draw : int { PixelX = 0, PixelY = 1, LensU = 2, LensV = 3, LightU = 4, LightV = 5, DirectionU = 6, DirectionV = 7,
                  Lobe = 8, Side = 9, Survival = 10, Pick = 11 };
KDN_D uint64_t dimension(int event, int draw) {
    if (draw < LightU) return (uint64_t)draw;
    return (uint64_t)(4 + event * 8 + draw - LightU);
}

// ---- the Sobol member (plugins/sampler/sobol/sobol.cpp) -------------------------
// The first 28 coordinates of an eye path (the pixel, the lens and three
// events) come off an Owen-scrambled Sobol sequence, the rest stay hashed
// (sampler.cpp:15-30). Each coordinate is its own Sobol dimension with Joe
// and Kuo's direction numbers over 16 bits; the 6 shared coordinates take
// dimensions 0..5 and the later ones take dimensions 0.. again on a second
// index shuffle (sobol.h:13-21). The sample's place is shuffled by a
// Laine-Karras hash of its reversed bits and every value scrambled the same
// way, seeded by the upper half of the stream (sobol.cpp:129-148).
constexpr uint32_t kSobolDimensions = 28;
constexpr uint32_t kSharedCoordinates = 6;

// The unscrambled values by place nibble: entry [d][q][k] XORs the
// bit-reversed direction columns 4q..4q+3 of dimension d that the nibble k
// selects, so a 16-bit place takes four loads from a 7 KB table that stays
// in the vector L1.
struct SobolNibbles {
    uint32_t v[kSobolDimensions][4][16];
};
constexpr uint32_t reversed_bits(uint32_t v) {
    v = (v << 16) | (v >> 16);
    v = ((v & 0x00FF00FFu) << 8) | ((v & 0xFF00FF00u) >> 8);
    v = ((v & 0x0F0F0F0Fu) << 4) | ((v & 0xF0F0F0F0u) >> 4);
    v = ((v & 0x33333333u) << 2) | ((v & 0xCCCCCCCCu) >> 2);
    return ((v & 0x55555555u) << 1) | ((v & 0xAAAAAAAAu) >> 1);
}
constexpr SobolNibbles sobol_nibbles() {
    // degree, primitive polynomial's inner coefficients and the first
    // direction numbers of each dimension (Joe and Kuo, new-joe-kuo-6.21201)
    constexpr uint32_t rules[kSobolDimensions][9] = {
        {0, 0, 1, 1, 1, 1, 1, 1, 1},      {1, 0, 1, 0, 0, 0, 0, 0, 0},      {2, 1, 1, 3, 0, 0, 0, 0, 0},
        {3, 1, 1, 3, 1, 0, 0, 0, 0},      {3, 2, 1, 1, 1, 0, 0, 0, 0},      {4, 1, 1, 1, 3, 3, 0, 0, 0},
        {4, 4, 1, 3, 5, 13, 0, 0, 0},     {5, 2, 1, 1, 5, 5, 17, 0, 0},     {5, 4, 1, 1, 5, 5, 5, 0, 0},
        {5, 7, 1, 1, 7, 11, 19, 0, 0},    {5, 11, 1, 1, 5, 1, 1, 0, 0},     {5, 13, 1, 1, 1, 3, 11, 0, 0},
        {5, 14, 1, 3, 5, 5, 31, 0, 0},    {6, 1, 1, 3, 3, 9, 7, 49, 0},     {6, 13, 1, 1, 1, 15, 21, 21, 0},
        {6, 16, 1, 3, 1, 13, 27, 49, 0},  {6, 19, 1, 1, 1, 15, 7, 5, 0},    {6, 22, 1, 3, 1, 15, 13, 25, 0},
        {6, 25, 1, 1, 5, 5, 19, 61, 0},   {7, 1, 1, 3, 7, 11, 23, 15, 103}, {7, 4, 1, 3, 7, 13, 13, 15, 69},
        {7, 7, 1, 1, 3, 13, 7, 35, 63},   {7, 8, 1, 3, 5, 9, 1, 25, 53},    {7, 14, 1, 3, 1, 13, 9, 35, 107},
        {7, 19, 1, 3, 1, 5, 27, 61, 31},  {7, 21, 1, 1, 5, 11, 19, 41, 61}, {7, 28, 1, 3, 5, 3, 3, 13, 69},
        {7, 31, 1, 1, 7, 13, 1, 19, 1}};
    SobolNibbles t{};
    for (uint32_t d = 0; d < kSobolDimensions; ++d) {
        const uint32_t degree = rules[d][0], poly = rules[d][1];
        uint32_t m[16] = {};
        uint32_t column[16] = {};
        for (uint32_t b = 0; b < 16; ++b) {
            if (degree == 0 || b < degree) {
                m[b] = rules[d][2 + (degree == 0 ? 0 : b)];
            } else {
                uint32_t n = m[b - degree] ^ (m[b - degree] << degree);
                for (uint32_t k = 1; k < degree; ++k)
                    if ((poly >> (degree - 1 - k)) & 1u) n ^= m[b - k] << k;
                m[b] = n;
            }
            column[b] = reversed_bits(m[b] << (31 - b));
        }
        for (uint32_t q = 0; q < 4; ++q)
            for (uint32_t k = 0; k < 16; ++k) {
                uint32_t x = 0;
                for (uint32_t j = 0; j < 4; ++j)
                    if ((k >> j) & 1u) x ^= column[4 * q + j];
                t.v[d][q][k] = x;
            }
    }
    return t;
}
static __constant__ SobolNibbles kSobol = sobol_nibbles();

KDN_D uint32_t stirred(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7FEB352Du;
    v ^= v >> 15;
    v *= 0x846CA68Bu;
    return v ^ (v >> 16);
}
KDN_D uint32_t laine_karras(uint32_t v, uint32_t seed) {
    v += seed;
    v ^= v * 0x6C50B47Cu;
    v ^= v * 0xB82F1E52u;
    v ^= v * 0xC7AFE638u;
    return v ^ (v * 0x8D22F6E6u);
}
KDN_D uint32_t reversed_sixteen(uint32_t v) { return reversed_bits(v & 0xFFFFu) >> 16; }
// sobolNumber for a coordinate below kSobolDimensions, as a double
KDN_D double sobol_number(uint64_t stream, uint32_t index, uint32_t coordinate) {
    const uint32_t scramble = (uint32_t)(stream >> 32);
    const bool shared = coordinate < kSharedCoordinates;
    const uint32_t place =
        reversed_sixteen(laine_karras(reversed_sixteen(index), stirred(scramble + (shared ? 0u : 0x68E31DA4u))));
    const uint32_t seed = stirred(scramble + coordinate * 0xA136AAADu + 0x2545F491u);
    const uint32_t (*t)[16] = kSobol.v[shared ? coordinate : coordinate - kSharedCoordinates];
    const uint32_t value = t[0][place & 15u] ^ t[1][(place >> 4) & 15u] ^ t[2][(place >> 8) & 15u] ^ t[3][(place >> 12) & 15u];
    return (double)reversed_bits(laine_karras(value, seed)) * 0x1p-32;
}
// numberAt (sampler.cpp:15-30): an eye path's coordinate, off the Sobol
// sequence where the job draws with it and the coordinate is one of its
// 28, hashed otherwise
KDN_D double sequence_number(bool sobol, uint64_t stream, uint32_t sample, uint64_t coordinate) {
    return sobol && coordinate < kSobolDimensions ? sobol_number(stream, sample, (uint32_t)coordinate)
                                                 : hashed(stream, sample, coordinate);
}
KDN_D float number(bool sobol, uint64_t stream, uint32_t sample, int event, int draw) {
    return (float)sequence_number(sobol, stream, sample, dimension(event, draw));
}

// ---- what a camera ray met ---------------------------------------------------------
// The Surface a frame records of a camera ray (core/batch.h:244-279):
// distance along the unit ray, the unit normal facing the ray, the leaf and
// the albedo there (glass white; a miss infinitely far, no leaf, zero
// normal and albedo). The G-buffer and the integrator's frame record both
// take it from here.
KDN_D kdn_surface surface_met(const kdn_hit& h, const kdn_ray& r, const uint32_t* leafMaterial, const MaterialDev* materials,
                              uint32_t nMaterials) {
    kdn_surface s;
    if (h.kind != KDN_HIT_SURFACE) {
        s.distance = kInf;
        s.leaf = KDN_NO_LEAF;
        for (int c = 0; c < 3; ++c) s.normal[c] = s.albedo[c] = 0;
        return s;
    }
    s.distance = h.distance;
    s.leaf = h.leaf;
    double nv[3] = {h.normal[0], h.normal[1], h.normal[2]};
    if (nv[0] * r.direction[0] + nv[1] * r.direction[1] + nv[2] * r.direction[2] > 0) {
        for (int c = 0; c < 3; ++c) nv[c] = -nv[c];
    }
    const double nl = __builtin_sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    for (int c = 0; c < 3; ++c) s.normal[c] = nl > 0 ? (float)(nv[c] / nl) : 0.0f;
    const uint32_t mi = h.leaf != KDN_NO_LEAF ? leafMaterial[h.leaf] : nMaterials - 1;
    const MaterialDev& m = materials[mi < nMaterials ? mi : nMaterials - 1];
    for (int c = 0; c < 3; ++c) s.albedo[c] = m.kind == KDN_GLASS ? 1.0f : m.albedo[c];
    return s;
}

// ---- wavelengths (colour.cpp:236-257, spectrum.cpp) ----------------------------
struct Lanes {
    float nm[kLanes];
    float density[kLanes];
};
KDN_D float wavelength_share(double start, uint32_t index) {
    const double turned = (double)index * 0.6180339887498949 + start;
    return (float)(turned - __builtin_floor(turned));
}
KDN_D Lanes sampled_wavelengths(float u, const ColourTables* t) {
    Lanes w;
    for (int lane = 0; lane < kLanes; ++lane) {
        const float raw = u + 0.25f * lane;
        const float turned = raw >= 1 ? raw - 1 : raw;
        const float place = turned * kWavelengthCuts;
        const int index = (int)place < kWavelengthCuts - 1 ? (int)place : kWavelengthCuts - 1;
        const float width = t->cuts[index + 1] - t->cuts[index];
        w.nm[lane] = fmin_(t->cuts[index] + (place - index) * width, 699.99994f);
        w.density[lane] = 1.0f / (kWavelengthCuts * width);
    }
    return w;
}
KDN_D float smooth_rise(float nm, float lower, float upper) {
    float b = (nm - lower) / (upper - lower);
    b = b < 0 ? 0.0f : (b > 1 ? 1.0f : b);
    return b * b * (3 - 2 * b);
}
KDN_D void bases(float nm, float* r, float* g, float* b) {
    *r = smooth_rise(nm, 575.0f, 605.0f);
    *b = 1 - smooth_rise(nm, 445.0f, 530.0f);
    *g = (1 - *b) * (1 - *r);
}
// spectrumOf (colour.cpp:88-103)
KDN_D void spectrum_of(const float rgb[3], const Lanes& w, float out[kLanes]) {
    int least = (rgb[0] <= rgb[1] && rgb[0] <= rgb[2]) ? 0 : ((rgb[1] <= rgb[0] && rgb[1] <= rgb[2]) ? 1 : 2);
    const int first = least == 0 ? 1 : 0, second = least == 2 ? 1 : 2;
    const int most = rgb[first] <= rgb[second] ? second : first;
    const int mid = 3 - least - most;
    const float base = rgb[least], low = rgb[mid] - base, high = rgb[most] - rgb[mid];
    for (int lane = 0; lane < kLanes; ++lane) {
        float bb[3];
        bases(w.nm[lane], &bb[0], &bb[1], &bb[2]);
        out[lane] = base + low * (1 - bb[least]) + high * bb[most];
    }
}
KDN_D float table_at(const float* curve, float nm) {
    const float slot = (nm - 380.0f) / 5.0f;
    int below = (int)slot;
    below = below < 0 ? 0 : (below > kTableSize - 2 ? kTableSize - 2 : below);
    float blend = slot - below;
    blend = blend < 0 ? 0.0f : (blend > 1 ? 1.0f : blend);
    return curve[below] + blend * (curve[below + 1] - curve[below]);
}
// colourOf (colour.cpp:213-234): lanes → XYZ → linear sRGB over white
KDN_D void colour_of(const Lanes& w, const float s[kLanes], const ColourTables* t, double rgb[3]) {
    double x = 0, y = 0, z = 0;
    for (int lane = 0; lane < kLanes; ++lane) {
        if (w.density[lane] <= 0) continue;
        const double weight = (double)s[lane] / w.density[lane];
        x += table_at(t->x, w.nm[lane]) * weight;
        y += table_at(t->y, w.nm[lane]) * weight;
        z += table_at(t->z, w.nm[lane]) * weight;
    }
    x /= kLanes;
    y /= kLanes;
    z /= kLanes;
    rgb[0] = (3.2406 * x - 1.5372 * y - 0.4986 * z) / t->white[0];
    rgb[1] = (-0.9689 * x + 1.8758 * y + 0.0415 * z) / t->white[1];
    rgb[2] = (0.0557 * x - 0.2040 * y + 1.0570 * z) / t->white[2];
}

// ---- vectors -------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
KDN_D V3 v3(float x, float y, float z) { return V3{x, y, z}; }
KDN_D V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
KDN_D V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
KDN_D V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
KDN_D V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
KDN_D float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KDN_D V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
KDN_D float length(V3 a) { return __builtin_sqrtf(dot(a, a)); }
KDN_D V3 normalized(V3 a) { return a * (1.0f / length(a)); }
KDN_D V3 mirrored(V3 out, V3 n) { return n * (2 * dot(out, n)) - out; }

// An orthonormal basis around n (Duff et al. 2017, branchless).
struct Basis {
    V3 t, b, n;
};
KDN_D Basis basis_around(V3 n) {
    const float sign = n.z >= 0 ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    Basis o;
    o.t = v3(1 + sign * n.x * n.x * a, sign * b, -sign * n.x);
    o.b = v3(b, sign + n.y * n.y * a, -n.y);
    o.n = n;
    return o;
}
KDN_D V3 to_world(const Basis& B, V3 v) { return B.t * v.x + B.b * v.y + B.n * v.z; }
KDN_D V3 to_local(const Basis& B, V3 v) { return v3(dot(v, B.t), dot(v, B.b), dot(v, B.n)); }

// Sine and cosine of a float angle through the tape's double polynomials,
// rounded once: the same bits on every IEEE machine, so a CPU build of
// these kernels follows the device's paths exactly.
KDN_D float f_cos(float x) { return (float)d_cos((double)x); }
KDN_D float f_sin(float x) { return (float)d_sin((double)x); }

KDN_D V3 cosine_hemisphere(const Basis& B, float u, float v) {
    const float reach = __builtin_sqrtf(u);
    const float angle = 2 * kPi * v;
    const float height = __builtin_sqrtf(fmax_(1 - u, 0.0f));
    return to_world(B, v3(reach * f_cos(angle), reach * f_sin(angle), height));
}

// ---- microfacets (microfacet.cpp) --------------------------------------------
KDN_D float facet_distribution(float cosine, float width) {
    if (cosine <= 0) return 0.0f;
    const float sq = width * width;
    const float lower = cosine * cosine * (sq - 1) + 1;
    return sq / (kPi * lower * lower);
}
KDN_D float smith_visibility(float leaving, float arriving, float width) {
    const float sq = width * width;
    const float first = arriving * __builtin_sqrtf(sq + (1 - sq) * leaving * leaving);
    const float second = leaving * __builtin_sqrtf(sq + (1 - sq) * arriving * arriving);
    const float total = first + second;
    return total > 0 ? 0.5f / total : 0.0f;
}
KDN_D float smith_masking(float cosine, float width) {
    const float sq = width * width;
    const float under = cosine + __builtin_sqrtf(sq + (1 - sq) * cosine * cosine);
    return under > 0 ? 2 * cosine / under : 0.0f;
}
KDN_D V3 halfway_of(V3 out, V3 in, V3 n) {
    const V3 total = out + in;
    const float size = length(total);
    return size > 1e-12f ? total * (1.0f / size) : n;
}
KDN_D V3 raised_facet(V3 s, float u, float v) {
    const float flat = s.x * s.x + s.y * s.y;
    const V3 first = flat > 1e-12f ? v3(-s.y, s.x, 0) * (1.0f / __builtin_sqrtf(flat)) : v3(1, 0, 0);
    const V3 second = cross(s, first);
    const float reach = __builtin_sqrtf(u);
    const float angle = 2 * kPi * v;
    const float across = reach * f_cos(angle);
    const float lean = 0.5f * (1 + s.z);
    const float along = (1 - lean) * __builtin_sqrtf(fmax_(1 - across * across, 0.0f)) + lean * reach * f_sin(angle);
    const float height = __builtin_sqrtf(fmax_(1 - across * across - along * along, 0.0f));
    return first * across + second * along + s * height;
}
// visibleFacet (microfacet.cpp:54-62)
KDN_D V3 visible_facet(V3 n, V3 out, float width, float u, float v) {
    const Basis B = basis_around(n);
    const V3 local = to_local(B, out);
    const V3 stretched = normalized(v3(width * local.x, width * local.y, local.z));
    const V3 raised = raised_facet(stretched, u, v);
    const V3 un = v3(width * raised.x, width * raised.y, fmax_(raised.z, 1e-12f));
    return to_world(B, normalized(un));
}
KDN_D float dielectric_reflectance(float cosine, float ratio) {
    const float leftover = 1 - ratio * ratio * (1 - cosine * cosine);
    if (leftover <= 0) return 1.0f;
    const float onward = __builtin_sqrtf(leftover);
    const float perp = (ratio * cosine - onward) / (ratio * cosine + onward);
    const float para = (ratio * onward - cosine) / (ratio * onward + cosine);
    return 0.5f * (perp * perp + para * para);
}

// ---- materials ---------------------------------------------------------------
KDN_D float escape_at(const MaterialDev& m, float cosine) {
    const int last = kEnergyViews - 1;
    const float place = (cosine < 0 ? 0.0f : (cosine > 1 ? 1.0f : cosine)) * kEnergyViews - 0.5f;
    int lower = (int)__builtin_floorf(place);
    lower = lower < 0 ? 0 : (lower > last ? last : lower);
    const int upper = lower + 1 < last ? lower + 1 : last;
    float share = place - lower;
    share = share < 0 ? 0.0f : (share > 1 ? 1.0f : share);
    return m.escape[lower] * (1 - share) + m.escape[upper] * share;
}

// f(out, in) per lane for matte and metal; in and out unit, n faces out.
KDN_D void rough_value(const MaterialDev& m, const Lanes& w, V3 n, V3 out, V3 in, float f[kLanes]) {
    const float leaving = dot(n, out), arriving = dot(n, in);
    for (int l = 0; l < kLanes; ++l) f[l] = 0;
    if (leaving <= 0 || arriving <= 0) return;
    if (m.kind == KDN_MATTE || m.kind == KDN_PLASTIC) {
        const float a[3] = {m.albedo[0] / kPi, m.albedo[1] / kPi, m.albedo[2] / kPi};
        spectrum_of(a, w, f);
        return;
    }
    // metal: Schlick facets plus the energy lost to multiple bounces
    // (material.cpp:592-660)
    const V3 h = halfway_of(out, in, n);
    const float onto = fmax_(dot(out, h), 0.0f);
    const float facets = facet_distribution(dot(n, h), m.width) * smith_visibility(leaving, arriving, m.width);
    const float apart = fmin_(fmax_(1 - onto, 0.0f), 1.0f);
    const float apart2 = apart * apart;
    const float rest = apart2 * apart2 * apart;
    float schlick[3], avg[3];
    for (int k = 0; k < 3; ++k) {
        schlick[k] = m.albedo[k] + (1 - m.albedo[k]) * rest;
        avg[k] = m.albedo[k] + (1 - m.albedo[k]) * (1.0f / 21);
    }
    float face[kLanes], flat[kLanes];
    spectrum_of(schlick, w, face);
    spectrum_of(avg, w, flat);
    const float room = 1 - m.escape_avg;
    const float lost = room <= 1e-6f ? 0.0f : (1 - escape_at(m, leaving)) * (1 - escape_at(m, arriving)) / (kPi * room);
    for (int l = 0; l < kLanes; ++l) {
        const float twice = flat[l] * flat[l] * m.escape_avg / (1 - flat[l] * (1 - m.escape_avg));
        f[l] = face[l] * facets + twice * lost;
    }
}

KDN_D float rough_density(const MaterialDev& m, V3 n, V3 out, V3 in) {
    const float leaving = dot(n, out), arriving = dot(n, in);
    if (leaving <= 0 || arriving <= 0) return 0.0f;
    const float cosine = arriving / kPi;
    if (m.kind != KDN_METAL) return cosine;
    const V3 h = halfway_of(out, in, n);
    float facets = 0;
    if (dot(out, h) > 0) facets = facet_distribution(dot(n, h), m.width) * smith_masking(leaving, m.width) / (4 * leaving);
    return m.facet_share * facets + (1 - m.facet_share) * cosine;
}

struct Scatter {
    bool found;
    bool through;    // passes into or out of the solid
    bool exact;      // a delta lobe (no density)
    V3 in;
    float weight[kLanes];
    float density;
    bool collapse;   // the passage bends the lanes apart: keep the hero
};

// Per-lane indices of the two sides (material.cpp:329-339 with the
// dispersion feature, dispersion.cpp:44-73).
KDN_D void sides_of(const MaterialDev& m, const Lanes& w, bool inside, float travelling[kLanes], float beyond[kLanes], bool* parted) {
    *parted = m.spread != 0;
    for (int l = 0; l < kLanes; ++l) {
        const float um = w.nm[l] * 1e-3f;
        const float n = m.spread != 0 ? m.index + m.spread / (um * um) : m.index;
        travelling[l] = inside ? n : 1.0f;
        beyond[l] = inside ? 1.0f : n;
    }
}

// exactGlassScatter (material.cpp:1296-1325)
KDN_D Scatter glass_scatter(const MaterialDev& m, const Lanes& w, V3 n, V3 out, bool inside, float uSide) {
    Scatter s;
    s.found = false;
    s.exact = true;
    s.density = 0;
    s.collapse = false;
    const float leaving = dot(n, out);
    if (leaving <= 0) return s;
    float tr[kLanes], be[kLanes];
    bool parted;
    sides_of(m, w, inside, tr, be, &parted);
    float R[kLanes];
    for (int l = 0; l < kLanes; ++l) R[l] = dielectric_reflectance(leaving, tr[l] / be[l]);
    if (!parted) R[1] = R[2] = R[3] = R[0];
    const float chance = parted ? 0.25f * (R[0] + R[1] + R[2] + R[3]) : R[0];
    const float ratio = tr[0] / be[0];
    const float leftover = 1 - ratio * ratio * (1 - leaving * leaving);
    const bool passed = leftover > 0;
    s.found = true;
    if (uSide < chance || !passed) {
        s.in = mirrored(out, n);
        s.through = false;
        const float c = passed ? chance : 1.0f;
        for (int l = 0; l < kLanes; ++l) s.weight[l] = R[l] / c;
        return s;
    }
    s.in = (n * leaving - out) * ratio - n * __builtin_sqrtf(leftover);
    s.through = true;
    for (int l = 0; l < kLanes; ++l) s.weight[l] = (1 - R[l]) / (1 - chance);
    s.collapse = parted && ratio != 1.0f;
    return s;
}

KDN_D Scatter rough_scatter(const MaterialDev& m, const Lanes& w, V3 n, V3 out, float uLobe, float u, float v) {
    Scatter s;
    s.found = false;
    s.exact = false;
    s.through = false;
    s.collapse = false;
    const float leaving = dot(n, out);
    if (leaving <= 0) return s;
    V3 in;
    if (m.kind != KDN_METAL) {
        in = cosine_hemisphere(basis_around(n), u, v);
    } else if (uLobe < m.facet_share) {
        const V3 facet = visible_facet(n, out, m.width, u, v);
        if (dot(out, facet) <= 0) return s;
        in = mirrored(out, facet);
    } else {
        in = cosine_hemisphere(basis_around(n), u, v);
    }
    const float arriving = dot(n, in);
    if (arriving <= 0) return s;
    const float density = rough_density(m, n, out, in);
    if (!(density > 0)) return s;
    float f[kLanes];
    rough_value(m, w, n, out, in, f);
    s.found = true;
    s.in = in;
    s.density = density;
    for (int l = 0; l < kLanes; ++l) s.weight[l] = f[l] * (arriving / density);
    return s;
}

// ---- lamps and sky -----------------------------------------------------------------
struct LampHit {
    float t, cosine;
};
// Where a ray meets a quad lamp from its lit side (meetQuad, light.cpp:458-480).
KDN_D bool meet_quad(const LampDev& L, V3 o, V3 d, float tMax, LampHit* h) {
    const V3 facing = v3(L.facing[0], L.facing[1], L.facing[2]);
    const float toward = -dot(facing, d);
    if (!(toward > 0)) return false;
    const V3 at = v3(L.at[0], L.at[1], L.at[2]);
    const float t = dot(at - o, facing) / dot(d, facing);
    if (!(t > 0) || !(t < tMax)) return false;
    const V3 p = o + d * t - at;
    const V3 a = v3(L.along[0], L.along[1], L.along[2]), c = v3(L.across[0], L.across[1], L.across[2]);
    // solve p = a s + c r on the plane
    const float aa = dot(a, a), ac = dot(a, c), cc = dot(c, c), pa = dot(p, a), pc = dot(p, c);
    const float det = aa * cc - ac * ac;
    const float sA = (pa * cc - pc * ac) / det, sC = (pc * aa - pa * ac) / det;
    if (sA < 0 || sA > 1 || sC < 0 || sC > 1) return false;
    h->t = t;
    h->cosine = toward;
    return true;
}

// Where a ray meets a sphere lamp before tMax (meetSphere,
// light.cpp:436-452): its nearer crossing, or the farther from inside.
KDN_D bool meet_sphere(const LampDev& L, V3 o, V3 d, float tMax, float* t) {
    const V3 off = o - v3(L.at[0], L.at[1], L.at[2]);
    const float along = dot(off, d);
    const float under = along * along - (dot(off, off) - L.radius * L.radius);
    if (!(under > 0)) return false;
    const float root = __builtin_sqrtf(under);
    const float first = -along - root;
    const float reach = first <= 1e-9f ? -along + root : first;
    if (!(reach > 1e-9f) || !(reach < tMax)) return false;
    *t = reach;
    return true;
}

// the cosine of the cone a sphere of radius r fills seen from distance
// (coneCosine, light.cpp:303-306), and a cone's uniform density
// (coneDensity, :308-311)
KDN_D float cone_cosine(float r, float distance) {
    const float sideways = r / distance;
    return __builtin_sqrtf(fmax_(1 - sideways * sideways, 0.0f));
}
KDN_D float cone_density(float cosine) {
    const float cone = 2 * kPi * (1 - cosine);
    return cone > 1e-12f ? 1 / cone : 0.0f;
}

// uniformCone (math/random.cpp:41-46)
KDN_D V3 uniform_cone(V3 axis, float cosine, float u, float v) {
    const float height = 1 - u * (1 - cosine);
    const float ring = __builtin_sqrtf(fmax_(1 - height * height, 0.0f));
    const float angle = 2 * kPi * v;
    return to_world(basis_around(axis), v3(ring * f_cos(angle), ring * f_sin(angle), height));
}

KDN_D void sky_radiance(const float* zenith, const float* horizon, const float* ground, V3 d, const Lanes& w, float out[kLanes]) {
    const float height = d.z;
    const float* far = height >= 0 ? zenith : ground;
    const float share = __builtin_sqrtf(__builtin_fabsf(height));
    float c[3];
    for (int k = 0; k < 3; ++k) c[k] = horizon[k] * (1 - share) + far[k] * share;
    spectrum_of(c, w, out);
}

KDN_D float power_weight(float a, float b) {
    const float a2 = a * a, b2 = b * b;
    return a2 + b2 > 0 ? a2 / (a2 + b2) : 0.0f;
}

}  // namespace kdn
