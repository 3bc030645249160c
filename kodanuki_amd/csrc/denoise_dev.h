// The à-trous denoiser's per-pixel arithmetic (SVGF-like), shared by the
// HIP kernels and the CPU build of the library.
//
// Restates the reference filter (plugins/reconstruction/atrous/atrous.cpp,
// behaviour documented at atrous.h:430-458): light divided by the albedo
// (temporal.cpp:208-214), a variance per pixel from its history or, for
// a young history, from the 7x7 pixels around it (atrous.cpp:333-380),
// then 4 passes of a 5x5 B3 spline at steps 1, 2, 4, 8 whose taps weigh
// the normals' cosine to the 128th, exp(-plane distance - luminance
// difference over the 3x3-prefiltered deviation) by a quadratic fit of
// 2^f (atrous.cpp:27-35, 178-296), and a fade of the filtered light into
// the held light as the history grows (atrous.cpp:446-470). FP32 where
// the reference computes in float, FP64 where it computes in double, each
// operation in the reference's order (the kernels are built without FP
// contraction), so the filter gives the reference's bits.
#pragma once

#include <cmath>
#include <cstdint>

#include "kdn_math.h"

namespace kdn {

constexpr int kDenoisePasses = 4;
constexpr int kVarianceRadius = 3;
constexpr float kYoungHistory = 4;
constexpr double kSettledLength = 8;
constexpr double kPlaneShare = 0.02;
constexpr float kDeviationSpread = 4;
constexpr float kLuminanceFloor = 1e-4f;
constexpr float kLuminanceShare = 1e-3f;
constexpr float kExponentFloor = -30;
constexpr float kFacingFloor = 0.5f;
constexpr double kAlbedoFloor = 1e-3;

// The guides and the light of every pixel, one float array a field.
struct DenoiseBuffers {
    int width, height;
    float *x, *y, *z, *nx, *ny, *nz, *offset, *inverseReach, *sky;  // guides
    float *red[2], *green[2], *blue[2], *luminance[2], *variance[2];  // light, ping-pong
};

KDN_D float spline_weight(int i) {
    return i == 0 ? 3.0f / 8 : (i == 1 || i == -1) ? 1.0f / 4 : 1.0f / 16;
}
KDN_D float gaussian_weight(int i) { return i == 1 ? 0.5f : 0.25f; }  // i in 0..2

// exp(x) for x <= 0 by 2^f ~ 1 + f (0.66023 + 0.33977 f) on the fraction
KDN_D float quick_exp(float x) {
    const float t = fmax_(x, kExponentFloor) * 1.44269504f;
    const float whole = __builtin_floorf(t);
    const float f = t - whole;
    const float power = 1 + f * (0.66023f + f * 0.33977f);
    // the exponent added to power's bits (two's complement, as the
    // reference's int32 sum)
    const float value = __uint_as_float(__float_as_uint(power) + ((uint32_t)(int32_t)whole << 23));
    return x > kExponentFloor ? value : 0.0f;
}

KDN_D double fmax_d(double a, double b) { return a < b ? b : a; }  // std::max
KDN_D float rgb_luminance(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }
KDN_D double luminance_of(double r, double g, double b) { return 0.2126 * r + 0.7152 * g + 0.0722 * b; }
KDN_D double channel_divisor(double albedo) { return albedo < kAlbedoFloor ? 1 : albedo; }

// The lens's heading to a pixel's centre, FP64.
struct DenoiseLensDev {
    double corner[3], across[3], down[3];
    double cx, cy;
};

// A pixel's guide and light from its temporal pixel (preparedDenoise,
// atrous.cpp:60-160).
KDN_D void prepare_pixel(const DenoiseBuffers& B, const DenoiseLensDev& L, const kdn_denoise_pixel& px, int xpix, int ypix,
                         bool demodulated) {
    const size_t at = (size_t)ypix * B.width + xpix;
    const bool sky = px.leaf == KDN_NO_LEAF || !__builtin_isfinite(px.distance) || !(px.distance > 0);
    if (sky) {
        B.x[at] = B.y[at] = B.z[at] = 0;
        B.nx[at] = B.ny[at] = B.nz[at] = 0;
        B.offset[at] = 0;
        B.inverseReach[at] = 0;
        B.sky[at] = 1;
    } else {
        double h[3];
        for (int k = 0; k < 3; ++k) h[k] = L.corner[k] + L.across[k] * ((double)xpix + L.cx) + L.down[k] * ((double)ypix + L.cy);
        const double len = __builtin_sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
        const double scale = (double)px.distance / len;
        const double p[3] = {h[0] * scale, h[1] * scale, h[2] * scale};
        B.x[at] = (float)p[0];
        B.y[at] = (float)p[1];
        B.z[at] = (float)p[2];
        B.nx[at] = (float)px.normal[0];
        B.ny[at] = (float)px.normal[1];
        B.nz[at] = (float)px.normal[2];
        B.offset[at] = (float)(p[0] * px.normal[0] + p[1] * px.normal[1] + p[2] * px.normal[2]);
        B.inverseReach[at] = (float)(1 / (kPlaneShare * (double)px.distance));
        B.sky[at] = 0;
    }
    double div[3];
    for (int k = 0; k < 3; ++k) div[k] = demodulated ? 1.0 : channel_divisor(px.albedo[k]);
    const float r = (float)(px.colour[0] / div[0]), g = (float)(px.colour[1] / div[1]), b = (float)(px.colour[2] / div[2]);
    float variance = -1;
    if (px.length >= kYoungHistory) {
        const double lum = luminance_of(px.colour[0], px.colour[1], px.colour[2]);
        const double s = luminance_of(div[0], div[1], div[2]);
        const double spread = fmax_d(px.moment - lum * lum, 0.0);
        variance = (float)(spread / ((double)px.length * s * s));
    }
    B.red[0][at] = r;
    B.green[0][at] = g;
    B.blue[0][at] = b;
    B.luminance[0][at] = rgb_luminance(r, g, b);
    B.variance[0][at] = variance;
}

// What the filter reads of a tap: its guides and its light.
struct TapFields {
    float x, y, z, nx, ny, nz, sky, red, green, blue, luminance, variance;
};
KDN_D TapFields tap_of(const DenoiseBuffers& B, int src, size_t q) {
    return TapFields{B.x[q],   B.y[q],      B.z[q],        B.nx[q],        B.ny[q],        B.nz[q],
                     B.sky[q], B.red[src][q], B.green[src][q], B.blue[src][q], B.luminance[src][q], B.variance[src][q]};
}

// The weight of tap q for pixel p (without the luminance term): the
// normals' cosine to the 128th times exp(-plane - apart).
KDN_D float facing_of(float pnx, float pny, float pnz, float psky, const TapFields& q) {
    const float facing = pnx * q.nx + pny * q.ny + pnz * q.nz + psky * q.sky;
    float power = facing > kFacingFloor ? facing : 0.0f;
    for (int n = 0; n < 7; ++n) power *= power;
    return power;
}
KDN_D float facing_weight(const DenoiseBuffers& B, size_t p, size_t q) {
    return facing_of(B.nx[p], B.ny[p], B.nz[p], B.sky[p], tap_of(B, 0, q));
}

KDN_D float plane_of(float pnx, float pny, float pnz, float poffset, const TapFields& q, float inversePlane) {
    const float along = q.x * pnx + q.y * pny + q.z * pnz;
    return __builtin_fabsf(along - poffset) * inversePlane;
}
KDN_D float plane_exponent(const DenoiseBuffers& B, size_t p, size_t q, float inversePlane) {
    return plane_of(B.nx[p], B.ny[p], B.nz[p], B.offset[p], tap_of(B, 0, q), inversePlane);
}

// The variance of a young pixel from the 7x7 pixels around it, weighed by
// the plane and the normal (estimatedChunk, atrous.cpp:333-380).
KDN_D void estimate_pixel(const DenoiseBuffers& B, int src, int xpix, int ypix) {
    const size_t p = (size_t)ypix * B.width + xpix;
    const int dst = src ^ 1;
    B.red[dst][p] = B.red[src][p];
    B.green[dst][p] = B.green[src][p];
    B.blue[dst][p] = B.blue[src][p];
    B.luminance[dst][p] = B.luminance[src][p];
    const float own = B.variance[src][p];
    if (!(own < 0)) {
        B.variance[dst][p] = own;
        return;
    }
    const float lum = B.luminance[src][p];
    const float inversePlane = B.inverseReach[p] * 1.0f;
    float red = 0, green = 0, blue = 0, squares = 0, weight = 0;
    for (int j = -kVarianceRadius; j <= kVarianceRadius; ++j) {
        const int yq = ypix + j;
        if (yq < 0 || yq >= B.height) continue;
        for (int i = -kVarianceRadius; i <= kVarianceRadius; ++i) {
            const int xq = xpix + i;
            if (xq < 0 || xq >= B.width) continue;
            const size_t q = (size_t)yq * B.width + xq;
            const float l = B.luminance[src][q];
            const float apart = __builtin_fabsf(l - lum) * 0.0f;
            const float w = 1.0f * facing_weight(B, p, q) * quick_exp(-plane_exponent(B, p, q, inversePlane) - apart);
            red += w * B.red[src][q];
            green += w * B.green[src][q];
            blue += w * B.blue[src][q];
            weight += w;
            squares += w * l * l;
        }
    }
    float v = 0;
    if (weight > 0) {
        const float mean = rgb_luminance(red, green, blue) / weight;
        v = fmax_(squares / weight - mean * mean, 0.0f);
    }
    B.variance[dst][p] = v;
}

// One pass of the 5x5 B3 spline at a step (filteredChunk, atrous.cpp:
// 178-300, 390-420): reads light src, writes light src ^ 1. taps(i, j) gives
// the fields of the tap i steps across and j steps down (asked for taps in
// the image only): from the buffers, or from a block's staged tile.
template <class Taps>
KDN_D void filter_pixel_from(const DenoiseBuffers& B, int src, int step, int xpix, int ypix, const Taps& taps) {
    const size_t p = (size_t)ypix * B.width + xpix;
    const int dst = src ^ 1;
    // the deviation through a 3x3 gaussian of the variance
    float var = 0, wsum = 0;
    for (int j = 0; j < 3; ++j) {
        const int yq = ypix + j - 1;
        if (yq < 0 || yq >= B.height) continue;
        for (int i = 0; i < 3; ++i) {
            const int xq = xpix + i - 1;
            if (xq < 0 || xq >= B.width) continue;
            const float k = gaussian_weight(i) * gaussian_weight(j);
            var += k * B.variance[src][(size_t)yq * B.width + xq];
            wsum += k;
        }
    }
    const float lum = B.luminance[src][p];
    const float deviation = __builtin_sqrtf(var / wsum);
    const float inverseDeviation = 1 / (kDeviationSpread * deviation + kLuminanceFloor + kLuminanceShare * lum);
    const float inversePlane = B.inverseReach[p] * (1 / (float)step);
    const float pnx = B.nx[p], pny = B.ny[p], pnz = B.nz[p], psky = B.sky[p], poffset = B.offset[p];
    float red = 0, green = 0, blue = 0, variance = 0, weight = 0;
    for (int j = -2; j <= 2; ++j) {
        const int yq = ypix + j * step;
        if (yq < 0 || yq >= B.height) continue;
        for (int i = -2; i <= 2; ++i) {
            const int xq = xpix + i * step;
            if (xq < 0 || xq >= B.width) continue;
            const TapFields q = taps(i, j);
            const float kernel = spline_weight(i) * spline_weight(j);
            const float l = q.luminance;
            const float apart = __builtin_fabsf(l - lum) * inverseDeviation;
            const float w = kernel * facing_of(pnx, pny, pnz, psky, q) * quick_exp(-plane_of(pnx, pny, pnz, poffset, q, inversePlane) - apart);
            red += w * q.red;
            green += w * q.green;
            blue += w * q.blue;
            variance += w * w * q.variance;
            weight += w;
        }
    }
    if (weight > 0) {
        const float scale = 1 / weight;
        const float r = red * scale, g = green * scale, b = blue * scale;
        B.red[dst][p] = r;
        B.green[dst][p] = g;
        B.blue[dst][p] = b;
        B.luminance[dst][p] = rgb_luminance(r, g, b);
        B.variance[dst][p] = variance * scale * scale;
    } else {
        B.red[dst][p] = B.red[src][p];
        B.green[dst][p] = B.green[src][p];
        B.blue[dst][p] = B.blue[src][p];
        B.luminance[dst][p] = B.luminance[src][p];
        B.variance[dst][p] = B.variance[src][p];
    }
}

// The same pass reading every tap from the buffers.
KDN_D void filter_pixel(const DenoiseBuffers& B, int src, int step, int xpix, int ypix) {
    filter_pixel_from(B, src, step, xpix, ypix, [&](int i, int j) {
        return tap_of(B, src, (size_t)(ypix + j * step) * B.width + (xpix + i * step));
    });
}

// The filtered light times the albedo, faded into the held light by
// 1 / (1 + length / 8) (remodulatedImage, atrous.cpp:446-470).
KDN_D void remodulate_pixel(const DenoiseBuffers& B, int src, const kdn_denoise_pixel& px, bool demodulated, size_t at, float* rgb) {
    const double share = 1 / (1 + (double)px.length / kSettledLength);
    const float* light[3] = {B.red[src], B.green[src], B.blue[src]};
    for (int k = 0; k < 3; ++k) {
        const double div = channel_divisor(px.albedo[k]);
        const double held = demodulated ? px.colour[k] * div : px.colour[k];
        rgb[at * 3 + k] = (float)((double)light[k][at] * div * share + held * (1 - share));
    }
}

}  // namespace kdn
