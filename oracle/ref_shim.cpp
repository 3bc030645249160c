// A C entry point onto the reference's own functions, for the oracle's
// pins: TEST INFRASTRUCTURE ONLY. It is compiled, together with the few
// reference source files it calls, straight from the reference tree into
// oracle/_ref/libkdn_ref.so (oracle/Makefile, target `ref`); nothing of the
// reference is copied into this repository, and the product never loads it.
//
// Every function below calls the reference function it names and converts
// plain C arguments to the reference's structs and back; none restates an
// algorithm.
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "core/camera.h"
#include "core/shown.h"
#include "core/stopping.h"
#include "math/colour.h"
#include "math/random.h"
#include "math/spectrum.h"
#include "plugins/reconstruction/atrous/atrous.h"
#include "plugins/reconstruction/temporal/temporal.h"
#include "plugins/sampler/highlight_guide/chain.h"
#include "plugins/sampler/sobol/sobol.h"
#include "plugins/transport/paths/material/energy.h"
#include "sdf/program.h"
#include "shared/image.h"

using namespace kodanuki;

namespace {

Grade gradeOf(float exposure, int curve) {
    Grade g;
    g.exposure = exposure;
    g.curve = curve == 0 ? Curve::Filmic : Curve::Clamped;
    return g;
}

math::Key keyOf(uint64_t stream, uint64_t index, uint64_t coordinate) {
    math::Key k;
    k.stream = stream;
    k.index = index;
    k.coordinate = coordinate;
    return k;
}

// An Estimate as 18 doubles: total[3], taken, batch[3], batches, means[3],
// squares[3], shared, sharedMeans, sharedSquares, ownSquares.
void packed(const Estimate& e, double* out) {
    const Colour* c3[] = {&e.total, &e.batch, &e.means, &e.squares};
    int o = 0;
    for (int k = 0; k < 4; ++k) {
        out[o++] = c3[k]->red;
        out[o++] = c3[k]->green;
        out[o++] = c3[k]->blue;
        if (k == 0) out[o++] = e.taken;
        if (k == 1) out[o++] = e.batches;
    }
    out[o++] = e.shared;
    out[o++] = e.sharedMeans;
    out[o++] = e.sharedSquares;
    out[o++] = e.ownSquares;
}

Estimate unpacked(const double* in) {
    Estimate e = emptyEstimate();
    Colour* c3[] = {&e.total, &e.batch, &e.means, &e.squares};
    int o = 0;
    for (int k = 0; k < 4; ++k) {
        c3[k]->red = in[o++];
        c3[k]->green = in[o++];
        c3[k]->blue = in[o++];
        if (k == 0) e.taken = (int)in[o++];
        if (k == 1) e.batches = (int)in[o++];
    }
    e.shared = in[o++];
    e.sharedMeans = in[o++];
    e.sharedSquares = in[o++];
    e.ownSquares = in[o++];
    return e;
}

}  // namespace

// Plain C views of a lens, a first surface and a history pixel.
struct kr_lens {
    double eye[3], forward[3], right[3], up[3];
    double reach, aspect;
    int32_t width, height;
};
struct kr_surface {
    float distance;
    float normal[3];
    uint32_t leaf;
    float albedo[3];
};
struct kr_pixel {
    double colour[3];
    double moment;
    float length;
    float distance;
    double normal[3];
    uint32_t leaf;
    uint32_t pad;
    double albedo[3];
};

namespace {

math::Vector vectorOf(const double v[3]) { return {v[0], v[1], v[2]}; }

Lens lensOf(const kr_lens& l) {
    Lens lens{};
    lens.eye = vectorOf(l.eye);
    lens.forward = vectorOf(l.forward);
    lens.right = vectorOf(l.right);
    lens.up = vectorOf(l.up);
    lens.reach = l.reach;
    lens.aspect = l.aspect;
    lens.aperture = 0;
    lens.focus = 1;
    lens.width = l.width;
    lens.height = l.height;
    return lens;
}

Surface surfaceOf(const kr_surface& s) {
    Surface out{};
    out.distance = s.distance;
    out.normal = {s.normal[0], s.normal[1], s.normal[2]};
    out.leaf = s.leaf;
    out.material = s.leaf == noLeaf ? noMaterial : 0;
    out.albedo = {s.albedo[0], s.albedo[1], s.albedo[2]};
    out.roughness = 1;
    out.onward = std::numeric_limits<float>::infinity();
    return out;
}

TemporalPixel pixelOf(const kr_pixel& p) {
    TemporalPixel t{};
    t.colour = {p.colour[0], p.colour[1], p.colour[2]};
    t.moment = p.moment;
    t.length = p.length;
    t.distance = p.distance;
    t.normal = vectorOf(p.normal);
    t.leaf = p.leaf;
    t.albedo = {p.albedo[0], p.albedo[1], p.albedo[2]};
    return t;
}

kr_pixel plainOf(const TemporalPixel& t) {
    kr_pixel p{};
    p.colour[0] = t.colour.red;
    p.colour[1] = t.colour.green;
    p.colour[2] = t.colour.blue;
    p.moment = t.moment;
    p.length = t.length;
    p.distance = t.distance;
    p.normal[0] = t.normal.x;
    p.normal[1] = t.normal.y;
    p.normal[2] = t.normal.z;
    p.leaf = t.leaf;
    p.albedo[0] = t.albedo.red;
    p.albedo[1] = t.albedo.green;
    p.albedo[2] = t.albedo.blue;
    return p;
}

// A frame of a viewer sequence as the temporal and à-trous members read it,
// at factor 1 (every pixel drawn each frame, through its centre).
struct ViewerFrame {
    Lens lens, before;
    FrameRecord record;
    std::vector<PixelSum> estimates;
    std::vector<Surface> surfaces;
};

ViewerFrame viewerFrame(int width, int height, const kr_lens* lens, const kr_lens* before, const double* parts,
                        const kr_surface* record, const kr_surface* gbuffer) {
    ViewerFrame f;
    f.lens = lensOf(*lens);
    f.before = lensOf(*before);
    const size_t n = (size_t)width * (size_t)height;
    f.record.width = width;
    f.record.height = height;
    f.record.parts.resize(n);
    f.record.surfaces.resize(n);
    f.surfaces.resize(gbuffer ? n : 0);
    for (size_t k = 0; k < n; ++k) {
        f.record.parts[k].total = {parts[3 * k], parts[3 * k + 1], parts[3 * k + 2]};
        f.record.surfaces[k] = surfaceOf(record[k]);
        if (gbuffer) f.surfaces[k] = surfaceOf(gbuffer[k]);
    }
    return f;
}

TemporalState stateOf(int width, int height, int demodulated, const kr_pixel* pixels) {
    TemporalState s = emptyTemporal();
    s.width = width;
    s.height = height;
    s.demodulated = demodulated != 0;
    s.film = stillFilm();
    s.pixels.resize((size_t)width * (size_t)height);
    for (size_t k = 0; k < s.pixels.size(); ++k) s.pixels[k] = pixelOf(pixels[k]);
    return s;
}

}  // namespace

extern "C" {

// temporalFrame (temporal.h:237) of one frame at factor 1: the history
// before (before_width x before_height pixels; 0 x 0 for none), this
// frame's record (each pixel's light and its sample's first surface) and
// G-buffer (null: none, the history then not demodulated) -> the history
// after and the shown image.
void kr_temporal_frame(int width, int height, const kr_lens* lens, const kr_lens* before, int moved, int restarted, int recounted,
                       const double* parts, const kr_surface* record, const kr_surface* gbuffer, int before_width, int before_height,
                       int before_demodulated, const kr_pixel* before_pixels, kr_pixel* out_pixels, float* out_image) {
    const ViewerFrame f = viewerFrame(width, height, lens, before, parts, record, gbuffer);
    const ShownFrame frame{width, height, f.lens, f.before, stillFilm(), moved != 0, restarted != 0, recounted != 0, 0u,
                           f.record, f.estimates, f.surfaces, 1};
    const Temporal t = temporalFrame(stateOf(before_width, before_height, before_demodulated, before_pixels), frame);
    for (size_t k = 0; k < t.state.pixels.size(); ++k) out_pixels[k] = plainOf(t.state.pixels[k]);
    for (size_t k = 0; k < t.image.rgb.size(); ++k) out_image[k] = t.image.rgb[k];
}

// denoisedFrame (atrous.h:462) of a history (the temporal state after a
// frame), with no denoiser history before it -> the shown image.
void kr_denoised_frame(int width, int height, const kr_lens* lens, const double* parts, const kr_surface* record,
                       const kr_surface* gbuffer, int demodulated, const kr_pixel* pixels, float* out_image) {
    const ViewerFrame f = viewerFrame(width, height, lens, lens, parts, record, gbuffer);
    const ShownFrame frame{width, height, f.lens, f.before, stillFilm(), false, false, false, 0u, f.record, f.estimates, f.surfaces, 1};
    const Denoised d = denoisedFrame(emptyDenoise(), stateOf(width, height, demodulated, pixels), frame);
    for (size_t k = 0; k < d.image.rgb.size(); ++k) out_image[k] = d.image.rgb[k];
}


uint64_t kr_mixed_bits(uint64_t v) { return math::mixedBits(v); }

// walkedChain (chain.h:97) over one leaf program (count steps: opcode,
// operands a b c, value; its result register) from a start with its two
// sideways directions, for a lamp of this build's kind numbering (0 point,
// 1 quad, 2 directional, 3 sphere, 4 sun, 5 moon); out: point, axis,
// across, up, spread (3), motion (6). Returns found.
int kr_walked_chain(const uint8_t* ops, const uint32_t* abc, const double* values, uint32_t count, uint32_t result, const double* from,
                    const double* sideways, double width, int kind, const double* at, const double* direction, const double* along,
                    const double* across, double radius, double cosine, const double* seed, double* out) {
    Program leaf;
    leaf.name = "leaf";
    leaf.result = result;
    leaf.arguments = 0;
    for (uint32_t i = 0; i < count; ++i) {
        Step st{};
        st.opcode = Opcode(ops[i]);
        st.a = abc[3 * i];
        st.b = abc[3 * i + 1];
        st.c = abc[3 * i + 2];
        st.value = values[i];
        st.leaf = 0;
        leaf.steps.push_back(st);
    }
    const LightKind kinds[6] = {LightKind::Point, LightKind::Quad, LightKind::Directional, LightKind::Sphere, LightKind::Sun,
                                LightKind::Moon};
    Lamp lamp{};
    lamp.kind = kinds[kind];
    lamp.at = {at[0], at[1], at[2]};
    lamp.direction = {direction[0], direction[1], direction[2]};
    lamp.along = {along[0], along[1], along[2]};
    lamp.across = {across[0], across[1], across[2]};
    lamp.radius = radius;
    lamp.cosine = cosine;
    const BounceChain chain = {
        .from = {from[0], from[1], from[2]},
        .sideways = {math::Vector{sideways[0], sideways[1], sideways[2]}, math::Vector{sideways[3], sideways[4], sideways[5]}},
        .leaf = leaf,
        .width = width,
        .lamp = lamp,
    };
    const SettledChain r = walkedChain(chain, {seed[0], seed[1], seed[2]});
    const math::Vector v[4] = {r.point, r.axis, r.across, r.up};
    for (int k = 0; k < 4; ++k) {
        out[3 * k] = v[k].x;
        out[3 * k + 1] = v[k].y;
        out[3 * k + 2] = v[k].z;
    }
    for (int k = 0; k < 3; ++k) out[12 + k] = r.spread[k];
    for (int m = 0; m < 2; ++m) {
        out[15 + 3 * m] = r.motion[m].x;
        out[16 + 3 * m] = r.motion[m].y;
        out[17 + 3 * m] = r.motion[m].z;
    }
    return r.found ? 1 : 0;
}

double kr_hashed_number(uint64_t stream, uint64_t index, uint64_t coordinate) {
    return math::hashedNumber(keyOf(stream, index, coordinate));
}

double kr_sobol_number(uint64_t stream, uint64_t index, uint64_t coordinate) {
    return sobolNumber(keyOf(stream, index, coordinate));
}

void kr_sobol_pair(uint64_t stream, uint64_t index, uint64_t coordinate, double out[2]) {
    const math::UnitPair p = sobolPair(keyOf(stream, index, coordinate));
    out[0] = p.u;
    out[1] = p.v;
}

int kr_sobol_dimensions(void) { return (int)sobolDimensions; }
int kr_shared_coordinates(void) { return (int)sharedCoordinates; }

void kr_sampled_wavelengths(double u, float nanometres[4], float density[4]) {
    const Wavelengths w = sampledWavelengths(u);
    std::memcpy(nanometres, w.nanometres, sizeof(float) * 4);
    std::memcpy(density, w.density, sizeof(float) * 4);
}

double kr_wavelength_share(double start, uint32_t index) { return wavelengthShare(start, index); }

void kr_colour_of(const double nanometres[4], const double density[4], const double spectrum[4], double rgb[3]) {
    const Wavelengths w = wavelengthsAt(nanometres, density);
    Spectrum4 s;
    for (int l = 0; l < 4; ++l) s.values[l] = spectrum[l];
    const Colour c = colourOf(w, s);
    rgb[0] = c.red;
    rgb[1] = c.green;
    rgb[2] = c.blue;
}

void kr_spectrum_of(const double rgb[3], const double nanometres[4], const double density[4], double out[4]) {
    const Wavelengths w = wavelengthsAt(nanometres, density);
    const Colour c = {rgb[0], rgb[1], rgb[2]};
    const Spectrum4 s = spectrumOf(c, w);
    for (int l = 0; l < 4; ++l) out[l] = s.values[l];
}

void kr_escape_of(double width, double table[32], double* average) {
    const Escape e = escapeOf(width);
    for (int k = 0; k < energyViews && k < 32; ++k) table[k] = e.table[(size_t)k];
    *average = e.average;
}

void kr_effort_for(double tolerance, int* samples, int* batch) {
    const Effort e = effortFor(tolerance);
    *samples = e.samples;
    *batch = e.batch;
}

int kr_split_ceiling(int batch) { return splitCeiling(batch); }

// Fold count samples (colour rgb, shared part rgb: 6 doubles each) into the
// estimate `state` (18 doubles, read and written) through foldedSample.
void kr_fold_samples(const double* samples, int count, int batch, double* state) {
    Estimate e = unpacked(state);
    for (int k = 0; k < count; ++k) {
        Sampled s;
        s.colour = {samples[6 * k], samples[6 * k + 1], samples[6 * k + 2]};
        s.shared = {samples[6 * k + 3], samples[6 * k + 4], samples[6 * k + 5]};
        e = foldedSample(e, s, batch);
    }
    packed(e, state);
}

void kr_split_spread(const double* state, int batch, int split, float out[2]) {
    const SplitSpread s = splitSpread(unpacked(state), batch, split);
    out[0] = s.shared;
    out[1] = s.own;
}

double kr_displayed(double value, float exposure, int curve) { return displayed(value, gradeOf(exposure, curve)); }

double kr_steps_between(double low, double high, float exposure, int curve) {
    return stepsBetween(low, high, gradeOf(exposure, curve));
}

unsigned kr_encoded_byte(double value, float exposure, int curve) { return encodedByte(value, gradeOf(exposure, curve)); }

void kr_write_ppm(int width, int height, const float* rgb, float exposure, int curve, const char* path) {
    Image image = blankImage(width, height);
    image.rgb.assign(rgb, rgb + (size_t)width * (size_t)height * 3);
    writePpm(image, gradeOf(exposure, curve), std::string(path));
}

void kr_write_pfm(int width, int height, const float* rgb, const char* path) {
    Image image = blankImage(width, height);
    image.rgb.assign(rgb, rgb + (size_t)width * (size_t)height * 3);
    writePfm(image, std::string(path));
}

}  // extern "C"
