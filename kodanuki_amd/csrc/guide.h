// The learned guide's per-lane arithmetic: the input encoding (a
// multiresolution hash grid of the position, real spherical harmonics of
// the outgoing direction, the normal, a frequency code of the hero
// wavelength), the mixture of von Mises-Fisher lobes the network's outputs
// decode to, its density, its sampling and the gradient of the training
// loss. Shared by the HIP kernels (which run the network's layers on MFMA)
// and the CPU build of the library.
//
// The reference has no learned guide; the guide joins its direction
// sampling where the Sampler's direction member is chosen
// (plugins/transport/paths/sampler.h:23,357) and, as there, is mixed with
// BSDF sampling by one-sample MIS so the estimate stays unbiased. Published
// methods restated: the hash encoding of Instant NGP (Mueller et al. 2022,
// its spatial hash primes), the vMF mixture of Neural Parametric Mixtures
// (Dong et al. 2023), the KL-divergence gradient estimate of Neural
// Importance Sampling (Mueller et al. 2019), vMF sampling by Jakob's
// numerically stable inversion (2012).
#pragma once

#include "integrator_dev.h"

namespace kdn {

constexpr int kGuideLevels = 16;          // hash grid levels
constexpr int kGuideFeatures = 2;         // features per level
constexpr uint32_t kGuideTable = 1u << 19;  // entries per level
constexpr int kGuideWidth = 64;           // inputs, hidden width and padded outputs
constexpr int kGuideLayers = 4;           // weight matrices: 3 hidden ReLU layers, then the outputs
constexpr int kGuideLobes = 8;
constexpr int kGuideOutputs = 5 * kGuideLobes + 1;  // logit, axis (3), log sharpness per lobe; the selection logit
constexpr int kGuideSelect = 5 * kGuideLobes;       // output of the selection logit
constexpr int kGuideDepths = 4;           // guided vertices: path depths 0 .. 3
constexpr uint32_t kGuideTrainPaths = 1u << 20;  // paths of a round that keep training records, at most (default)
constexpr float kGuideShare = 0.3f;       // the guide's share of directions (one-sample MIS; the most if learned)
constexpr float kGuideLogSharpMin = -3.0f;  // sharpness within [e^-3, e^6.2]
constexpr float kGuideLogSharpMax = 6.2f;
constexpr int kGuideMinRes = 16, kGuideMaxRes = 128;  // grid resolutions (the finest by default)
constexpr uint64_t kGuideGridParams = (uint64_t)kGuideLevels * kGuideTable * kGuideFeatures;
constexpr int kGuideMlpParams = kGuideLayers * kGuideWidth * kGuideWidth;
// input layout: grid features, SH of the outgoing direction, the normal,
// the wavelength code, a constant one (the first layer's bias), zeros
constexpr int kInGrid = 0, kInSh = 32, kInNormal = 48, kInWave = 51, kInOne = 59;

// A training record of one guided vertex (floats): point, outgoing
// direction, normal, hero wavelength, the direction taken, its density,
// the throughput after the vertex (4 lanes), the path's radiance when the
// next vertex begins (4 lanes), a state word (1 recorded, 3 snapshot
// taken), the BSDF's density of the direction and the throughput before
// the vertex (4 lanes).
constexpr int kRecPoint = 0, kRecOut = 3, kRecNormal = 6, kRecNm = 9, kRecIn = 10, kRecPdf = 13, kRecThr = 14,
              kRecSnap = 18, kRecState = 22, kRecBsdf = 23, kRecBefore = 24, kRecFloats = 28;
// A training sample (floats): point, outgoing direction, normal, hero
// wavelength, the direction taken, its weight (radiance over density),
// the BSDF's density of the direction (negative: train the lobes alone).
constexpr int kSamPoint = 0, kSamOut = 3, kSamNormal = 6, kSamNm = 9, kSamIn = 10, kSamWeight = 13, kSamBsdf = 14,
              kSamFloats = 16;

// The guide as the kernels read it.
struct GuideView {
    const float* grid;       // [level][entry][feature], FP32
    const uint16_t* mlp;     // BF16 weights [layer][in][out]
    uint32_t res[kGuideLevels];
    uint32_t dense;          // bit l: level l is indexed densely
    float lower[3], inv_size[3];
    float share_cap;         // the selection probability: share_cap * sigmoid(logit) if learned
    uint32_t learn_select;   // 0: the selection probability is share_cap itself
};

// ---- BF16 (round to nearest even; NaN kept quiet) ---------------------------
KDN_D uint16_t to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
KDN_D float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---- grid gradients in fixed point -------------------------------------------
// A grid gradient is summed as a 64-bit integer of 2^-32 units: integer sums
// are the same in any order, so a training step gives the same gradients
// however its atomics interleave. A contribution of magnitude 2^11 or more,
// or not finite, adds nothing and marks the step wild (it is skipped, as a
// step with a non-finite gradient is); so a step of at most 2^16 samples
// (eight corners a level each) keeps every sum below 2^31 in magnitude.
constexpr float kFixedUnit = 0x1p32f;
KDN_D uint64_t fixed_of(float v, bool* wild) {
    const float x = v * kFixedUnit;
    if (!(__builtin_fabsf(x) < 0x1p43f)) {
        *wild = true;
        return 0;
    }
    return (uint64_t)(int64_t)x;
}
KDN_D float float_of_fixed(uint64_t a) { return (float)((double)(int64_t)a * 0x1p-32); }

// ---- encoding -----------------------------------------------------------------
KDN_D uint32_t grid_index(const GuideView& g, int level, uint32_t x, uint32_t y, uint32_t z) {
    if (g.dense >> level & 1u) {
        const uint32_t side = g.res[level] + 1;
        return x + side * (y + side * z);
    }
    return (x * 1u ^ y * 2654435761u ^ z * 805459861u) & (kGuideTable - 1u);
}

// The eight corners of the point's cell at one level: entry indices and
// trilinear weights.
KDN_D void grid_corners(const GuideView& g, int level, const float p[3], uint32_t idx[8], float wt[8]) {
    const float n = (float)g.res[level];
    uint32_t c[3];
    float f[3];
    for (int k = 0; k < 3; ++k) {
        const float s = p[k] * n;
        float fl = __builtin_floorf(s);
        if (fl > n - 1) fl = n - 1;
        if (fl < 0) fl = 0;
        c[k] = (uint32_t)fl;
        f[k] = s - fl;
    }
    for (int corner = 0; corner < 8; ++corner) {
        const uint32_t dx = corner & 1, dy = corner >> 1 & 1, dz = corner >> 2 & 1;
        idx[corner] = grid_index(g, level, c[0] + dx, c[1] + dy, c[2] + dz);
        wt[corner] = (dx ? f[0] : 1 - f[0]) * (dy ? f[1] : 1 - f[1]) * (dz ? f[2] : 1 - f[2]);
    }
}

KDN_D void guide_position(const GuideView& g, const float x[3], float p[3]) {
    for (int k = 0; k < 3; ++k) {
        const float t = (x[k] - g.lower[k]) * g.inv_size[k];
        p[k] = t < 0 ? 0.0f : (t > 1 ? 1.0f : t);
    }
}

// Real spherical harmonics of degrees 0..3 of a unit direction.
KDN_D void sh16(float x, float y, float z, float* o) {
    const float xx = x * x, yy = y * y, zz = z * z;
    o[0] = 0.28209479177387814f;
    o[1] = -0.48860251190291987f * y;
    o[2] = 0.48860251190291987f * z;
    o[3] = -0.48860251190291987f * x;
    o[4] = 1.0925484305920792f * x * y;
    o[5] = -1.0925484305920792f * y * z;
    o[6] = 0.94617469575755997f * zz - 0.31539156525251999f;
    o[7] = -1.0925484305920792f * x * z;
    o[8] = 0.54627421529603959f * (xx - yy);
    o[9] = 0.59004358992664352f * y * (yy - 3 * xx);
    o[10] = 2.8906114426405538f * x * y * z;
    o[11] = 0.45704579946446572f * y * (1 - 5 * zz);
    o[12] = 0.3731763325901154f * z * (5 * zz - 3);
    o[13] = 0.45704579946446572f * x * (1 - 5 * zz);
    o[14] = 1.4453057213202769f * z * (xx - yy);
    o[15] = 0.59004358992664352f * x * (3 * yy - xx);
}

// The network's 64 inputs for a query.
KDN_D void guide_inputs(const GuideView& g, const float x[3], const float out[3], const float normal[3], float nm, float in[kGuideWidth]) {
    float p[3];
    guide_position(g, x, p);
    for (int level = 0; level < kGuideLevels; ++level) {
        uint32_t idx[8];
        float wt[8];
        grid_corners(g, level, p, idx, wt);
        float f0 = 0, f1 = 0;
        const float* table = g.grid + (size_t)level * kGuideTable * kGuideFeatures;
        for (int c = 0; c < 8; ++c) {
            f0 += wt[c] * table[(size_t)idx[c] * 2];
            f1 += wt[c] * table[(size_t)idx[c] * 2 + 1];
        }
        in[kInGrid + 2 * level] = f0;
        in[kInGrid + 2 * level + 1] = f1;
    }
    sh16(out[0], out[1], out[2], in + kInSh);
    for (int k = 0; k < 3; ++k) in[kInNormal + k] = normal[k];
    const float t = (nm - 380.0f) * (1.0f / 400.0f);
    float phase = kPi * t;
    for (int j = 0; j < 4; ++j) {
        in[kInWave + 2 * j] = sinf(phase);
        in[kInWave + 2 * j + 1] = cosf(phase);
        phase *= 2;
    }
    in[kInOne] = 1.0f;
    for (int k = kInOne + 1; k < kGuideWidth; ++k) in[k] = 0.0f;
}

// ---- the lobes ----------------------------------------------------------------
struct GuideLobes {
    float axis[kGuideLobes][3];
    float sharp[kGuideLobes];
    float share[kGuideLobes];
    float select;  // the probability of sampling the mixture rather than the BSDF
};

// The vMF normalisation sharpness / (2 pi (1 - e^(-2 sharpness))).
KDN_D float vmf_norm(float k) { return k / (2 * kPi * (1 - expf(-2 * k))); }

KDN_D float clamp_log_sharp(float c) { return c < kGuideLogSharpMin ? kGuideLogSharpMin : (c > kGuideLogSharpMax ? kGuideLogSharpMax : c); }

KDN_D float sigmoid(float x) { return 1 / (1 + expf(-x)); }

KDN_D void decode_lobes(const float* o, float cap, bool learned, GuideLobes& L) {
    float top = o[0];
    for (int k = 1; k < kGuideLobes; ++k) top = fmax_(top, o[5 * k]);
    float sum = 0;
    for (int k = 0; k < kGuideLobes; ++k) {
        L.share[k] = expf(o[5 * k] - top);
        sum += L.share[k];
        const float ax = o[5 * k + 1], ay = o[5 * k + 2], az = o[5 * k + 3];
        const float len = sqrtf(ax * ax + ay * ay + az * az);
        if (len > 1e-6f) {
            L.axis[k][0] = ax / len;
            L.axis[k][1] = ay / len;
            L.axis[k][2] = az / len;
        } else {
            L.axis[k][0] = 0;
            L.axis[k][1] = 0;
            L.axis[k][2] = 1;
        }
        L.sharp[k] = expf(clamp_log_sharp(o[5 * k + 4]));
    }
    for (int k = 0; k < kGuideLobes; ++k) L.share[k] /= sum;
    L.select = learned ? cap * sigmoid(o[kGuideSelect]) : cap;
}

KDN_D float mixture_pdf(const GuideLobes& L, V3 w) {
    float q = 0;
    for (int k = 0; k < kGuideLobes; ++k) {
        const float c = L.axis[k][0] * w.x + L.axis[k][1] * w.y + L.axis[k][2] * w.z;
        q += L.share[k] * vmf_norm(L.sharp[k]) * expf(L.sharp[k] * (c - 1));
    }
    return q;
}

// A direction of the mixture: the lobe by uLobe, the direction by (u, v).
KDN_D V3 mixture_sample(const GuideLobes& L, float uLobe, float u, float v) {
    int k = 0;
    float acc = L.share[0];
    while (k + 1 < kGuideLobes && uLobe >= acc) acc += L.share[++k];
    const float s = L.sharp[k];
    const float w = 1 + logf(u + (1 - u) * expf(-2 * s)) / s;
    const float r = sqrtf(fmax_(0.0f, 1 - w * w));
    const float phi = 2 * kPi * v;
    const Basis B = basis_around(v3(L.axis[k][0], L.axis[k][1], L.axis[k][2]));
    return to_world(B, v3(r * f_cos(phi), r * f_sin(phi), w));
}

// The training loss of one sample and its gradient with respect to the
// network's outputs (zero where the sample carries nothing or its density
// underflows). With the BSDF's density pB of the direction (bsdf >= 0)
// the loss is -weight * log p of the one-sample MIS density
// p = a q + (1 - a) pB, which trains the lobes and the selection a
// together (Neural Parametric Mixtures); without it, -weight * log q.
KDN_D float lobe_loss_gradient(const float* o, const float in[3], float weight, float bsdf, float cap, bool learned, float* g) {
    for (int c = 0; c < kGuideOutputs; ++c) g[c] = 0;
    if (!(weight > 0)) return 0.0f;
    GuideLobes L;
    decode_lobes(o, cap, learned, L);
    float v[kGuideLobes], dots[kGuideLobes];
    float q = 0;
    for (int k = 0; k < kGuideLobes; ++k) {
        dots[k] = L.axis[k][0] * in[0] + L.axis[k][1] * in[1] + L.axis[k][2] * in[2];
        v[k] = vmf_norm(L.sharp[k]) * expf(L.sharp[k] * (dots[k] - 1));
        q += L.share[k] * v[k];
    }
    if (!(q > 1e-30f)) return 0.0f;
    float loss;
    if (bsdf >= 0) {
        const float a = L.select;
        const float p = a * q + (1 - a) * bsdf;
        if (!(p > 1e-30f)) return 0.0f;
        if (learned) {
            const float s = sigmoid(o[kGuideSelect]);
            g[kGuideSelect] = -weight * (q - bsdf) / p * cap * s * (1 - s);
        }
        loss = -weight * logf(p);
        weight *= a * q / p;  // the lobes' share of the gradient
    } else {
        loss = -weight * logf(q);
    }
    for (int k = 0; k < kGuideLobes; ++k) {
        const float gamma = L.share[k] * v[k] / q;
        g[5 * k] = -weight * (gamma - L.share[k]);
        const float ax = o[5 * k + 1], ay = o[5 * k + 2], az = o[5 * k + 3];
        const float len = sqrtf(ax * ax + ay * ay + az * az);
        if (len > 1e-6f) {
            const float s = -weight * gamma * L.sharp[k] / len;
            for (int a = 0; a < 3; ++a) g[5 * k + 1 + a] = s * (in[a] - dots[k] * L.axis[k][a]);
        }
        const float c = o[5 * k + 4];
        if (c > kGuideLogSharpMin && c < kGuideLogSharpMax) {
            const float s = L.sharp[k];
            const float dlogv = 1 / s - 2 / (expf(2 * s) - 1) + (dots[k] - 1);
            g[5 * k + 4] = -weight * gamma * dlogv * s;
        }
    }
    return loss;
}

}  // namespace kdn
