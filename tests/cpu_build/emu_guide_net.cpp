// The learned guide's network launches for the bounds-checking CPU build:
// the device kernels (guide_net.hip) run their layers on MFMA across a
// wavefront, which this build's one-thread-at-a-time launches cannot; here
// each sample runs the same per-lane arithmetic (guide.h) with the layers
// as loops, rounding to BF16 where the device does (inputs, activations,
// gradients into each layer) and accumulating in FP32.
#include "guide_net.h"

#include <cstring>
#include <vector>

namespace kdn {
namespace {

float weight(const GuideView& g, int layer, int in, int out) { return from_bf16(g.mlp[(layer * 64 + in) * 64 + out]); }

// forward: H[0] the BF16 inputs, H[1..3] the BF16 ReLU activations, O the
// FP32 outputs
void forward(const GuideView& g, const float in[kGuideWidth], uint16_t H[kGuideLayers][kGuideWidth], float O[kGuideWidth]) {
    for (int k = 0; k < kGuideWidth; ++k) H[0][k] = to_bf16(in[k]);
    for (int layer = 0; layer < kGuideLayers; ++layer) {
        for (int o = 0; o < kGuideWidth; ++o) {
            float acc = 0;
            for (int k = 0; k < kGuideWidth; ++k) acc += from_bf16(H[layer][k]) * weight(g, layer, k, o);
            if (layer + 1 < kGuideLayers) {
                H[layer + 1][o] = to_bf16(acc > 0 ? acc : 0.0f);
            } else {
                O[o] = acc;
            }
        }
    }
}

void decode(const GuideView& g, const float O[kGuideWidth], GuideLobes& L) { decode_lobes(O, g.share_cap, g.learn_select != 0, L); }

}  // namespace

bool launchGuideInferQueries(const GuideView& g, const float* queries, uint32_t n, float* lobes, float* select, hipStream_t) {
    for (uint32_t item = 0; item < n; ++item) {
        const float* e = queries + (size_t)item * 10;
        float in[kGuideWidth], O[kGuideWidth];
        uint16_t H[kGuideLayers][kGuideWidth];
        guide_inputs(g, e, e + 3, e + 6, e[9], in);
        forward(g, in, H, O);
        GuideLobes L;
        decode(g, O, L);
        if (select) select[item] = L.select;
        float* o = lobes + (size_t)item * 5 * kGuideLobes;
        for (int k = 0; k < kGuideLobes; ++k) {
            o[5 * k] = L.axis[k][0];
            o[5 * k + 1] = L.axis[k][1];
            o[5 * k + 2] = L.axis[k][2];
            o[5 * k + 3] = L.sharp[k];
            o[5 * k + 4] = L.share[k];
        }
    }
    return true;
}

bool launchGuideInferPaths(const GuideView& g, const PassParams& P, uint32_t queue, uint32_t n, float* lobes, hipStream_t) {
    for (uint32_t item = 0; item < n; ++item) {
        const uint32_t i = P.s.queue[queue][item];
        KDN_CHECK(i < P.paths);
        float x[3], out[3], normal[3], nm;
        if (!guided_vertex(P, i, x, out, normal, &nm)) continue;
        float in[kGuideWidth], O[kGuideWidth];
        uint16_t H[kGuideLayers][kGuideWidth];
        guide_inputs(g, x, out, normal, nm, in);
        forward(g, in, H, O);
        GuideLobes L;
        decode(g, O, L);
        store_lobes(lobes, P.s.capacity, i, L);
    }
    return true;
}

uint32_t guideTrainPartials(uint32_t count) { return (count + 63) / 64; }  // one partial sum a tile here

bool launchGuideTrain(const GuideView& g, const float* samples, uint32_t n, uint32_t first, uint32_t count, const float* scale,
                      float* partial, uint64_t* gridGrad, float* loss, uint32_t* flag, hipStream_t) {
    const uint32_t tiles = (count + 63) / 64;
    for (uint32_t tile = 0; tile < tiles; ++tile) {
        float* P = partial + (size_t)tile * kGuideMlpParams;
        std::memset(P, 0, sizeof(float) * kGuideMlpParams);
        for (uint32_t lane = 0; lane < 64; ++lane) {
            const uint32_t b = tile * 64 + lane;
            if (b >= count || n == 0) continue;
            const float* smp = samples + (size_t)((first + b) % n) * kSamFloats;
            float in[kGuideWidth], O[kGuideWidth];
            uint16_t H[kGuideLayers][kGuideWidth];
            guide_inputs(g, smp + kSamPoint, smp + kSamOut, smp + kSamNormal, smp[kSamNm], in);
            forward(g, in, H, O);
            float grad[kGuideOutputs];
            const float l = lobe_loss_gradient(O, smp + kSamIn, smp[kSamWeight] * *scale, smp[kSamBsdf], g.share_cap, g.learn_select != 0, grad);
            if (l != 0) *loss += l;
            uint16_t G[kGuideWidth];
            for (int c = 0; c < kGuideWidth; ++c) G[c] = c < kGuideOutputs ? to_bf16(grad[c]) : (uint16_t)0;
            float dIn[kGuideWidth];
            for (int layer = kGuideLayers - 1; layer >= 0; --layer) {
                float* W = P + layer * 4096;
                for (int k = 0; k < kGuideWidth; ++k)
                    for (int o = 0; o < kGuideWidth; ++o) W[k * 64 + o] += from_bf16(H[layer][k]) * from_bf16(G[o]);
                float dH[kGuideWidth];
                for (int k = 0; k < kGuideWidth; ++k) {
                    float acc = 0;
                    for (int o = 0; o < kGuideWidth; ++o) acc += from_bf16(G[o]) * weight(g, layer, k, o);
                    dH[k] = acc;
                }
                if (layer > 0) {
                    for (int k = 0; k < kGuideWidth; ++k) {
                        const uint16_t h = H[layer][k];
                        G[k] = h != 0 && (h & 0x8000u) == 0 ? to_bf16(dH[k]) : (uint16_t)0;
                    }
                } else {
                    std::memcpy(dIn, dH, sizeof dIn);
                }
            }
            float p[3];
            guide_position(g, smp + kSamPoint, p);
            for (int level = 0; level < kGuideLevels; ++level) {
                uint32_t idx[8];
                float wt[8];
                grid_corners(g, level, p, idx, wt);
                uint64_t* table = gridGrad + (size_t)level * kGuideTable * kGuideFeatures;
                bool wild = false;
                for (int c = 0; c < 8; ++c) {
                    KDN_CHECK(idx[c] < kGuideTable);
                    table[(size_t)idx[c] * 2] += fixed_of(wt[c] * dIn[kInGrid + 2 * level], &wild);
                    table[(size_t)idx[c] * 2 + 1] += fixed_of(wt[c] * dIn[kInGrid + 2 * level + 1], &wild);
                }
                if (wild) *flag |= 1u;
            }
        }
    }
    return true;
}

}  // namespace kdn
