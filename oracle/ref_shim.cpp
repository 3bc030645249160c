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
#include <string>
#include <vector>

#include "core/stopping.h"
#include "math/colour.h"
#include "math/random.h"
#include "math/spectrum.h"
#include "plugins/sampler/sobol/sobol.h"
#include "plugins/transport/paths/material/energy.h"
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

extern "C" {

uint64_t kr_mixed_bits(uint64_t v) { return math::mixedBits(v); }

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
