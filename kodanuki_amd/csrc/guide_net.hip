// The learned guide's network on MFMA: inference (queries or the paths of
// a queue → vMF lobes) and training (forward, loss gradient, backward:
// per-tile weight gradients and grid gradients).
//
// A wavefront takes a tile of 64 samples. The layers are 64x64 BF16
// matrices multiplied on v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
// a tile's layer is 4x4 output blocks of 16x16, two MFMAs deep. Weights sit
// in LDS (both orientations for training), activations of the tile in LDS
// as BF16 rows, outputs as FP32.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ctx.h"
#include "guide_net.h"

namespace kdn {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kLd = 72;         // LDS row stride of a 64-wide BF16 matrix (elements)
constexpr int kOutLd = 65;      // LDS row stride of FP32 outputs
constexpr int kInferWaves = 4;  // waves per inference block
constexpr uint32_t kTrainBlocks = 256;  // training blocks (one a CU: each holds 142 KB of LDS)

// acc = A · B for a 64x64 tile: A row-major (rows = samples, lda), B given
// transposed (Bt[n][k], ldb) so that a lane's eight consecutive k are
// contiguous. MFMA 16x16x32: lane l holds A[row l%16][k 8(l/16) .. +7],
// B[k 8(l/16) .. +7][col l%16] and D[rows 4(l/16) .. +3][col l%16].
__device__ __forceinline__ void wave_gemm(const uint16_t* A, int lda, const uint16_t* Bt, int ldb, f32x4 acc[16]) {
    const int l = threadIdx.x & 63;
    const int r = l & 15, kq = (l >> 4) * 8;
    for (int t = 0; t < 16; ++t) acc[t] = f32x4{0, 0, 0, 0};
    for (int ks = 0; ks < 64; ks += 32) {
        bf16x8 a[4], b[4];
        for (int rb = 0; rb < 4; ++rb) a[rb] = *(const bf16x8*)(A + (rb * 16 + r) * lda + ks + kq);
        for (int cb = 0; cb < 4; ++cb) b[cb] = *(const bf16x8*)(Bt + (cb * 16 + r) * ldb + ks + kq);
        for (int rb = 0; rb < 4; ++rb)
            for (int cb = 0; cb < 4; ++cb) acc[rb * 4 + cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rb], b[cb], acc[rb * 4 + cb], 0, 0, 0);
    }
}

// acc = A^T · B for a 64x64 tile where both are stored row-major with the
// reduced dimension (the samples) as rows: A[k][m], B[k][n]. Gathers the
// lane's eight k of a column element by element.
__device__ __forceinline__ void wave_gemm_tn(const uint16_t* A, int lda, const uint16_t* B, int ldb, f32x4 acc[16]) {
    const int l = threadIdx.x & 63;
    const int r = l & 15, kq = (l >> 4) * 8;
    for (int t = 0; t < 16; ++t) acc[t] = f32x4{0, 0, 0, 0};
    for (int ks = 0; ks < 64; ks += 32) {
        bf16x8 a[4], b[4];
        for (int rb = 0; rb < 4; ++rb) {
            uint16_t e[8];
            for (int j = 0; j < 8; ++j) e[j] = A[(ks + kq + j) * lda + rb * 16 + r];
            __builtin_memcpy(&a[rb], e, sizeof e);
        }
        for (int cb = 0; cb < 4; ++cb) {
            uint16_t e[8];
            for (int j = 0; j < 8; ++j) e[j] = B[(ks + kq + j) * ldb + cb * 16 + r];
            __builtin_memcpy(&b[cb], e, sizeof e);
        }
        for (int rb = 0; rb < 4; ++rb)
            for (int cb = 0; cb < 4; ++cb) acc[rb * 4 + cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rb], b[cb], acc[rb * 4 + cb], 0, 0, 0);
    }
}

// The tile's D elements: (row, col) of element j of block t.
__device__ __forceinline__ int d_row(int t, int j) { return (t >> 2) * 16 + ((threadIdx.x & 63) >> 4) * 4 + j; }
__device__ __forceinline__ int d_col(int t) { return (t & 3) * 16 + (threadIdx.x & 15); }

// ReLU(acc) as BF16 rows.
__device__ __forceinline__ void store_relu(const f32x4 acc[16], uint16_t* H) {
    for (int t = 0; t < 16; ++t)
        for (int j = 0; j < 4; ++j) {
            const float v = acc[t][j];
            H[d_row(t, j) * kLd + d_col(t)] = to_bf16(v > 0 ? v : 0.0f);
        }
}

// acc masked by the ReLU of the forward activation H (the gradient of the
// pre-activation), as BF16 rows.
__device__ __forceinline__ void store_masked(const f32x4 acc[16], const uint16_t* H, uint16_t* G) {
    for (int t = 0; t < 16; ++t)
        for (int j = 0; j < 4; ++j) {
            const int row = d_row(t, j), col = d_col(t);
            G[row * kLd + col] = H[row * kLd + col] != 0 && (H[row * kLd + col] & 0x8000u) == 0 ? to_bf16(acc[t][j]) : (uint16_t)0;
        }
}

__device__ __forceinline__ void store_f32(const f32x4 acc[16], float* O) {
    for (int t = 0; t < 16; ++t)
        for (int j = 0; j < 4; ++j) O[d_row(t, j) * kOutLd + d_col(t)] = acc[t][j];
}

// The lobes' outputs alone (columns below kGuideOutputs), rows kInferOutLd
// apart: a wave's outputs fit where its activations were.
constexpr int kInferOutLd = kGuideOutputs;  // odd: rows start in different banks
static_assert(64 * kInferOutLd * 4 >= 64 * kLd * 2, "a wave's outputs cover its activations' space");
__device__ __forceinline__ void store_outputs(const f32x4 acc[16], float* O) {
    for (int t = 0; t < 16; ++t) {
        const int col = d_col(t);
        if (col >= kGuideOutputs) continue;
        for (int j = 0; j < 4; ++j) O[d_row(t, j) * kInferOutLd + col] = acc[t][j];
    }
}

// Adds (v0, v1) to base[2 idx], base[2 idx + 1] (fixed point, fixed_of)
// for the active lanes of a wavefront (all 64 lanes call it). Lanes with
// the same entry are summed first (a butterfly in a fixed order), a leader
// adding the sum, while the first remaining lane's entry is shared by
// another; the rest add their own. The global atomics are executed far
// from the CU: the nearby samples of a tile share the coarse levels'
// entries, and one atomic a group instead of one a lane is what keeps
// training from waiting on them. A wild contribution raises *flag.
__device__ __forceinline__ void wave_scatter_add(uint64_t* base, uint32_t idx, float v0, float v1, bool active, uint32_t* flag) {
    bool wild = false;
    const int lane = threadIdx.x & 63;
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t a = __shfl(idx, leader);
        const uint64_t same = __ballot(active && idx == a) & pending;
        if (__popcll(same) < 2) break;
        const bool in = (same >> lane) & 1;
        float s0 = in ? v0 : 0.0f, s1 = in ? v1 : 0.0f;
        for (int off = 32; off > 0; off >>= 1) {
            s0 += __shfl_xor(s0, off);
            s1 += __shfl_xor(s1, off);
        }
        if (lane == leader) {
            atomicAdd((unsigned long long*)&base[(size_t)a * 2], (unsigned long long)fixed_of(s0, &wild));
            atomicAdd((unsigned long long*)&base[(size_t)a * 2 + 1], (unsigned long long)fixed_of(s1, &wild));
        }
        pending &= ~same;
    }
    if ((pending >> lane) & 1) {
        atomicAdd((unsigned long long*)&base[(size_t)idx * 2], (unsigned long long)fixed_of(v0, &wild));
        atomicAdd((unsigned long long*)&base[(size_t)idx * 2 + 1], (unsigned long long)fixed_of(v1, &wild));
    }
    if (wild) atomicOr(flag, 1u);
}

// The layers' BF16 weights into LDS, eight at a time: Wt[layer][out][in]
// for the forward products, and (Wn not null) Wn[layer][in][out] for the
// backward ones.
__device__ __forceinline__ void load_weights(const uint16_t* mlp, uint16_t (*Wt)[64 * kLd], uint16_t (*Wn)[64 * kLd], int threads) {
    const uint4* src = (const uint4*)mlp;
    for (int c = threadIdx.x; c < kGuideMlpParams / 8; c += threads) {
        const uint4 v = src[c];
        const int layer = c >> 9, in = (c >> 3) & 63, out0 = (c & 7) * 8;
        uint16_t e[8];
        __builtin_memcpy(e, &v, sizeof e);
        for (int j = 0; j < 8; ++j) Wt[layer][(out0 + j) * kLd + in] = e[j];
        if (Wn) *(uint4*)&Wn[layer][in * kLd + out0] = v;
    }
}

// the queries of a host-given array
struct ArrayQueries {
    const float* q;  // kdn_guide_query: point, out, normal, nm
    float* lobes;    // per query kGuideLobes x (axis, sharpness, share)
    float* select;   // per query the selection probability, or null
    uint32_t n;
    __device__ bool fetch(const GuideView& g, uint32_t item, float in[kGuideWidth]) const {
        if (item >= n) return false;
        const float* e = q + (size_t)item * 10;
        guide_inputs(g, e, e + 3, e + 6, e[9], in);
        return true;
    }
    __device__ void write(uint32_t item, const GuideLobes& L) const {
        float* o = lobes + (size_t)item * 5 * kGuideLobes;
        for (int k = 0; k < kGuideLobes; ++k) {
            o[5 * k] = L.axis[k][0];
            o[5 * k + 1] = L.axis[k][1];
            o[5 * k + 2] = L.axis[k][2];
            o[5 * k + 3] = L.sharp[k];
            o[5 * k + 4] = L.share[k];
        }
        if (select) select[item] = L.select;
    }
};

// the guided vertices of a queue of paths
struct PathQueries {
    PassParams P;
    uint32_t queue, n;
    float* lobes;  // lane-major, kGuideOutputs per path
    __device__ bool fetch(const GuideView& g, uint32_t item, float in[kGuideWidth]) const {
        if (item >= n) return false;
        const uint32_t i = P.s.queue[queue][item];
        float x[3], out[3], normal[3], nm;
        if (!guided_vertex(P, i, x, out, normal, &nm)) return false;
        guide_inputs(g, x, out, normal, nm, in);
        return true;
    }
    __device__ void write(uint32_t item, const GuideLobes& L) const {
        const uint32_t i = P.s.queue[queue][item];
        store_lobes(lobes, P.s.capacity, i, L);
    }
};

// Inference: 4 waves a block, persistent over tiles of 64 queries; the
// weights in LDS, each wave's activations (BF16 rows) and then its outputs
// in one buffer of its own: 78.8 KB a block, two blocks to a CU.
template <class Q>
__global__ __launch_bounds__(64 * kInferWaves) void k_guide_infer(GuideView g, Q src) {
    __shared__ __attribute__((aligned(16))) uint16_t Wt[kGuideLayers][64 * kLd];
    __shared__ __attribute__((aligned(16))) float buf[kInferWaves][64 * kInferOutLd];
    load_weights(g.mlp, Wt, nullptr, 64 * kInferWaves);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* A = (uint16_t*)buf[wave];
    float* O = buf[wave];
    const uint32_t tiles = (src.n + 63) / 64;
    for (uint32_t base = blockIdx.x * kInferWaves; base < tiles; base += gridDim.x * kInferWaves) {
        const uint32_t item = (base + wave) * 64 + lane;
        float in[kGuideWidth];
        const bool live = src.fetch(g, item, in);
        for (int k = 0; k < kGuideWidth; ++k) A[lane * kLd + k] = live ? to_bf16(in[k]) : (uint16_t)0;
        __syncthreads();
        f32x4 acc[16];
        for (int layer = 0; layer < kGuideLayers - 1; ++layer) {
            wave_gemm(A, kLd, Wt[layer], kLd, acc);
            __syncthreads();
            store_relu(acc, A);
            __syncthreads();
        }
        wave_gemm(A, kLd, Wt[kGuideLayers - 1], kLd, acc);
        __syncthreads();  // the last layer has read the activations the outputs replace
        store_outputs(acc, O);
        __syncthreads();
        if (live) {
            float o[kGuideOutputs];
            for (int c = 0; c < kGuideOutputs; ++c) o[c] = O[lane * kInferOutLd + c];
            GuideLobes L;
            decode_lobes(o, g.share_cap, g.learn_select != 0, L);
            src.write(item, L);
        }
        __syncthreads();
    }
}

// Training, one wave a block, persistent: each block takes tiles of 64
// samples in turn (forward, the loss gradient of each sample, its weight
// times *scale, backward) and sums its tiles' weight gradients into
// partial[block][layer][in][out]; grid gradients are added to grid_grad
// in fixed point (a wild one raises *flag) and each tile's loss to *loss.
__global__ __launch_bounds__(64) void k_guide_train(GuideView g, const float* samples, uint32_t n, uint32_t first, uint32_t count,
                                                    const float* scale, float* partial, uint64_t* grid_grad, float* loss,
                                                    uint32_t* flag) {
    __shared__ __attribute__((aligned(16))) uint16_t Wt[kGuideLayers][64 * kLd];  // [out][in]
    __shared__ __attribute__((aligned(16))) uint16_t Wn[kGuideLayers][64 * kLd];  // [in][out]
    __shared__ __attribute__((aligned(16))) uint16_t H[kGuideLayers][64 * kLd];   // inputs, h1, h2, h3
    __shared__ __attribute__((aligned(16))) uint16_t G[2][64 * kLd];              // gradients, ping-pong
    __shared__ float O[64 * kOutLd];
    load_weights(g.mlp, Wt, Wn, 64);
    const int lane = threadIdx.x & 63;
    float* P = partial + (size_t)blockIdx.x * kGuideMlpParams;
    const uint32_t tiles = (count + 63) / 64;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const bool firstTile = tile == blockIdx.x;
        const uint32_t b = tile * 64 + lane;
        const bool live = b < count && n > 0;
        const float* smp = samples + (size_t)(live ? (first + b) % n : 0) * kSamFloats;
        {
            float in[kGuideWidth];
            if (live) guide_inputs(g, smp + kSamPoint, smp + kSamOut, smp + kSamNormal, smp[kSamNm], in);
            for (int k = 0; k < kGuideWidth; ++k) H[0][lane * kLd + k] = live ? to_bf16(in[k]) : (uint16_t)0;
        }
        __syncthreads();
        f32x4 acc[16];
        for (int layer = 0; layer < kGuideLayers - 1; ++layer) {
            wave_gemm(H[layer], kLd, Wt[layer], kLd, acc);
            store_relu(acc, H[layer + 1]);
            __syncthreads();
        }
        wave_gemm(H[kGuideLayers - 1], kLd, Wt[kGuideLayers - 1], kLd, acc);
        store_f32(acc, O);
        __syncthreads();
        // the loss gradient of each output
        {
            float o[kGuideOutputs], grad[kGuideOutputs];
            for (int c = 0; c < kGuideOutputs; ++c) o[c] = O[lane * kOutLd + c];
            const float w = live ? smp[kSamWeight] * *scale : 0.0f;
            const float l = lobe_loss_gradient(o, smp + kSamIn, w, live ? smp[kSamBsdf] : -1.0f, g.share_cap, g.learn_select != 0, grad);
            for (int c = 0; c < kGuideWidth; ++c) G[0][lane * kLd + c] = c < kGuideOutputs ? to_bf16(grad[c]) : (uint16_t)0;
            if (l != 0) atomicAdd(loss, l);
        }
        __syncthreads();
        // backward, last layer first: dW = H^T dZ, dH = dZ W^T masked by ReLU
        int cur = 0;
        for (int layer = kGuideLayers - 1; layer >= 0; --layer) {
            wave_gemm_tn(H[layer], kLd, G[cur], kLd, acc);
            float* PW = P + layer * 4096;
            if (!firstTile) {
                // the block's sums so far, all loads issued before any store
                float before[64];
                for (int t = 0; t < 16; ++t)
                    for (int j = 0; j < 4; ++j) before[t * 4 + j] = PW[d_row(t, j) * 64 + d_col(t)];
                for (int t = 0; t < 16; ++t)
                    for (int j = 0; j < 4; ++j) acc[t][j] += before[t * 4 + j];
            }
            for (int t = 0; t < 16; ++t)
                for (int j = 0; j < 4; ++j) PW[d_row(t, j) * 64 + d_col(t)] = acc[t][j];
            wave_gemm(G[cur], kLd, Wn[layer], kLd, acc);
            if (layer > 0) {
                store_masked(acc, H[layer], G[cur ^ 1]);
            } else {
                store_f32(acc, O);
            }
            __syncthreads();
            cur ^= 1;
        }
        // the grid's share: each level's two features to its eight corners
        {
            float p[3] = {0, 0, 0};
            if (live) guide_position(g, smp + kSamPoint, p);
            for (int level = 0; level < kGuideLevels; ++level) {
                uint32_t idx[8];
                float wt[8];
                grid_corners(g, level, p, idx, wt);
                const float d0 = O[lane * kOutLd + kInGrid + 2 * level], d1 = O[lane * kOutLd + kInGrid + 2 * level + 1];
                const bool active = live && (d0 != 0 || d1 != 0);
                uint64_t* table = grid_grad + (size_t)level * kGuideTable * kGuideFeatures;
                for (int c = 0; c < 8; ++c) wave_scatter_add(table, idx[c], wt[c] * d0, wt[c] * d1, active, flag);
            }
        }
        __syncthreads();
    }
}

}  // namespace

bool launchGuideInferQueries(const GuideView& g, const float* queries, uint32_t n, float* lobes, float* select, hipStream_t st) {
    if (n == 0) return true;
    const uint32_t tiles = (n + 63) / 64;
    const uint32_t blocks = std::min<uint32_t>((tiles + kInferWaves - 1) / kInferWaves, 4096);
    hipLaunchKernelGGL(k_guide_infer<ArrayQueries>, dim3(blocks), dim3(64 * kInferWaves), 0, st, g, ArrayQueries{queries, lobes, select, n});
    return hipGetLastError() == hipSuccess;
}

bool launchGuideInferPaths(const GuideView& g, const PassParams& P, uint32_t queue, uint32_t n, float* lobes, hipStream_t st) {
    if (n == 0) return true;
    const uint32_t tiles = (n + 63) / 64;
    const uint32_t blocks = std::min<uint32_t>((tiles + kInferWaves - 1) / kInferWaves, 4096);
    hipLaunchKernelGGL(k_guide_infer<PathQueries>, dim3(blocks), dim3(64 * kInferWaves), 0, st, g, PathQueries{P, queue, n, lobes});
    return hipGetLastError() == hipSuccess;
}

uint32_t guideTrainPartials(uint32_t count) { return std::min<uint32_t>((count + 63) / 64, kTrainBlocks); }

bool launchGuideTrain(const GuideView& g, const float* samples, uint32_t n, uint32_t first, uint32_t count, const float* scale,
                      float* partial, uint64_t* gridGrad, float* loss, uint32_t* flag, hipStream_t st) {
    if (count == 0) return true;
    hipLaunchKernelGGL(k_guide_train, dim3(guideTrainPartials(count)), dim3(64), 0, st, g, samples, n, first, count, scale, partial,
                       gridGrad, loss, flag);
    return hipGetLastError() == hipSuccess;
}

}  // namespace kdn
