// The à-trous denoiser on the device: one kernel a stage, one thread a
// pixel (prepare, variance estimate, 4 filter passes, remodulate), and the
// G-buffer pass that turns a context's accumulated image into the
// denoiser's history. The per-pixel arithmetic is denoise_dev.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "ctx.h"
#include "denoise_dev.h"
#include "integrator_kernels.h"

namespace kdn {
namespace {

constexpr int kDenoiseBlock = 256;

__global__ void k_dn_prepare(DenoiseBuffers B, DenoiseLensDev L, const kdn_denoise_pixel* pixels, uint32_t demodulated) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= (uint32_t)(B.width * B.height)) return;
    prepare_pixel(B, L, pixels[at], (int)(at % B.width), (int)(at / B.width), demodulated != 0);
}

__global__ void k_dn_estimate(DenoiseBuffers B, int src) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= (uint32_t)(B.width * B.height)) return;
    estimate_pixel(B, src, (int)(at % B.width), (int)(at / B.width));
}

// A filter pass over LDS-staged tiles. At step s the taps of pixel
// (a + s u, b + s v) are the pixels (a + s (u + i), b + s (v + j)), |i|, |j| <=
// 2: the lattice of its residue (a, b) mod s. A block takes 16x16 lattice
// points of one residue and stages the 20x20 lattice points their taps
// reach (12 fields each, 19.2 KB) once, so each tap's guides and light are
// read from HBM once a block rather than once a pixel that reads it; every
// step's tile is the same size. Blocks: s^2 residues x tilesX x tilesY.
constexpr int kFilterTile = 16;
constexpr int kFilterWindow = kFilterTile + 4;
constexpr int kTapFields = 12;

__global__ void __launch_bounds__(kFilterTile * kFilterTile) k_dn_filter(DenoiseBuffers B, int src, int step, int tilesX, int tilesY) {
    __shared__ float window[kTapFields][kFilterWindow * kFilterWindow];
    const int tiles = tilesX * tilesY;
    const int residue = (int)blockIdx.x / tiles, tile = (int)blockIdx.x % tiles;
    const int ax = residue % step, ay = residue / step;
    const int u0 = (tile % tilesX) * kFilterTile, v0 = (tile / tilesX) * kFilterTile;
    for (int k = (int)threadIdx.x; k < kFilterWindow * kFilterWindow; k += kFilterTile * kFilterTile) {
        const int xq = ax + step * (u0 - 2 + k % kFilterWindow), yq = ay + step * (v0 - 2 + k / kFilterWindow);
        if (xq < 0 || xq >= B.width || yq < 0 || yq >= B.height) continue;  // outside the image: never read
        const TapFields q = tap_of(B, src, (size_t)yq * B.width + xq);
        const float f[kTapFields] = {q.x, q.y, q.z, q.nx, q.ny, q.nz, q.sky, q.red, q.green, q.blue, q.luminance, q.variance};
        for (int n = 0; n < kTapFields; ++n) window[n][k] = f[n];
    }
    __syncthreads();
    const int tx = (int)threadIdx.x % kFilterTile, ty = (int)threadIdx.x / kFilterTile;
    const int xpix = ax + step * (u0 + tx), ypix = ay + step * (v0 + ty);
    if (xpix >= B.width || ypix >= B.height) return;
    filter_pixel_from(B, src, step, xpix, ypix, [&](int i, int j) {
        const int k = (ty + 2 + j) * kFilterWindow + (tx + 2 + i);
        KDN_CHECK(k >= 0 && k < kFilterWindow * kFilterWindow);
        return TapFields{window[0][k], window[1][k], window[2][k], window[3][k],  window[4][k],  window[5][k],
                         window[6][k], window[7][k], window[8][k], window[9][k], window[10][k], window[11][k]};
    });
}
KDN_SYNCS_BLOCK(k_dn_filter)

__global__ void k_dn_remodulate(DenoiseBuffers B, int src, const kdn_denoise_pixel* pixels, uint32_t demodulated, float* rgb) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= (uint32_t)(B.width * B.height)) return;
    remodulate_pixel(B, src, pixels[at], demodulated != 0, at, rgb);
}

// primary rays through the pixel centres of a lens
__global__ void k_dn_rays(DenoiseLensDev L, int width, int height, float ex, float ey, float ez, kdn_ray* rays) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= (uint32_t)(width * height)) return;
    const int x = (int)(at % width), y = (int)(at / width);
    double h[3];
    for (int k = 0; k < 3; ++k) h[k] = L.corner[k] + L.across[k] * ((double)x + L.cx) + L.down[k] * ((double)y + L.cy);
    const double len = __builtin_sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    kdn_ray r;
    r.origin[0] = ex;
    r.origin[1] = ey;
    r.origin[2] = ez;
    for (int k = 0; k < 3; ++k) r.direction[k] = (float)(h[k] / len);
    r.far_ = kInf;
    r.inside = 0;
    r.within = KDN_NO_MATERIAL;
    rays[at] = r;
}

// the history of a pixel from its FP64 sums and its primary hit: the mean,
// the luminance's second moment (the channels' spreads taken as fully
// correlated: the sums hold squares per channel, not of the luminance),
// the samples, and distance, facing normal, leaf and albedo of the hit
__global__ void k_dn_history(const double* sum, const double* sq, const uint32_t* count, const kdn_hit* hits, const kdn_ray* rays,
                             const uint32_t* leafMaterial, const MaterialDev* materials, uint32_t nMaterials, uint32_t n,
                             kdn_denoise_pixel* out) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= n) return;
    kdn_denoise_pixel px;
    const double k = count[at] > 0 ? (double)count[at] : 1.0;
    double spread = 0;
    const double w[3] = {0.2126, 0.7152, 0.0722};
    for (int c = 0; c < 3; ++c) {
        const double mean = sum[3 * (size_t)at + c] / k;
        px.colour[c] = mean;
        spread += w[c] * __builtin_sqrt(fmax_d(sq[3 * (size_t)at + c] / k - mean * mean, 0.0));
    }
    const double lum = luminance_of(px.colour[0], px.colour[1], px.colour[2]);
    px.moment = lum * lum + spread * spread;
    px.length = (float)count[at];
    px.pad = 0;
    const kdn_hit h = hits[at];
    if (h.kind != KDN_HIT_SURFACE) {
        px.distance = kInf;
        px.leaf = KDN_NO_LEAF;
        for (int c = 0; c < 3; ++c) px.normal[c] = px.albedo[c] = 0;
    } else {
        px.distance = h.distance;
        px.leaf = h.leaf;
        const kdn_ray& r = rays[at];
        double nv[3] = {h.normal[0], h.normal[1], h.normal[2]};
        if (nv[0] * r.direction[0] + nv[1] * r.direction[1] + nv[2] * r.direction[2] > 0) {
            for (int c = 0; c < 3; ++c) nv[c] = -nv[c];
        }
        const double nl = __builtin_sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        for (int c = 0; c < 3; ++c) px.normal[c] = nl > 0 ? nv[c] / nl : 0.0;
        const uint32_t mi = h.leaf != KDN_NO_LEAF ? leafMaterial[h.leaf] : nMaterials - 1;
        const MaterialDev& m = materials[mi < nMaterials ? mi : nMaterials - 1];
        for (int c = 0; c < 3; ++c) px.albedo[c] = m.kind == KDN_GLASS ? 1.0 : (double)m.albedo[c];
    }
    out[at] = px;
}

DenoiseLensDev lensOf(const kdn_denoise_frame& f) {
    DenoiseLensDev L;
    const double sideways = f.reach * f.aspect;
    for (int k = 0; k < 3; ++k) {
        L.corner[k] = f.forward[k] - f.right[k] * sideways + f.up[k] * f.reach;
        L.across[k] = f.right[k] * (2 * sideways / (double)f.width);
        L.down[k] = f.up[k] * (-2 * f.reach / (double)f.height);
    }
    L.cx = f.centre_x;
    L.cy = f.centre_y;
    return L;
}

// The filter's working memory for the entry point that allocates nothing:
// its 19 fields of up to kStaticPixels pixels (76 MiB), held by every GPU
// the library runs on until the process ends. Host code reaches it only
// through the address hipGetSymbolAddress gives on each device
// (denoiseStaticScratch); every call on a GPU shares that GPU's array.
constexpr size_t kStaticPixels = size_t(1) << 20;
constexpr size_t kDenoiseFields = 19;
static __device__ __align__(256) float scratch[kStaticPixels * kDenoiseFields];

// The filter on device buffers: pixels in, rgb out, its 19 float fields a
// pixel in `base` (n floats each, one after another).
bool denoiseFields(Ctx& c, const kdn_denoise_frame& f, const kdn_denoise_pixel* pixels, float* rgb, hipStream_t st, float* base) {
    const size_t n = (size_t)f.width * f.height;
    DenoiseBuffers B;
    B.width = f.width;
    B.height = f.height;
    float** fields[] = {&B.x, &B.y, &B.z, &B.nx, &B.ny, &B.nz, &B.offset, &B.inverseReach, &B.sky, &B.red[0], &B.green[0], &B.blue[0],
                        &B.luminance[0], &B.variance[0], &B.red[1], &B.green[1], &B.blue[1], &B.luminance[1], &B.variance[1]};
    for (int k = 0; k < 19; ++k) *fields[k] = base + (size_t)k * n;
    const DenoiseLensDev L = lensOf(f);
    const dim3 grid((uint32_t)((n + kDenoiseBlock - 1) / kDenoiseBlock)), block(kDenoiseBlock);
    hipLaunchKernelGGL(k_dn_prepare, grid, block, 0, st, B, L, pixels, f.demodulated);
    hipLaunchKernelGGL(k_dn_estimate, grid, block, 0, st, B, 0);
    int src = 1;
    for (int k = 0; k < kDenoisePasses; ++k) {
        const int step = 1 << k;
        const int tilesX = ((f.width + step - 1) / step + kFilterTile - 1) / kFilterTile;
        const int tilesY = ((f.height + step - 1) / step + kFilterTile - 1) / kFilterTile;
        hipLaunchKernelGGL(k_dn_filter, dim3((uint32_t)(step * step * tilesX * tilesY)), dim3(kFilterTile * kFilterTile), 0, st, B, src,
                           step, tilesX, tilesY);
        src ^= 1;
    }
    hipLaunchKernelGGL(k_dn_remodulate, grid, block, 0, st, B, src, pixels, f.demodulated, rgb);
    KDN_HIP(hipGetLastError());
    return true;
}

bool checkFrame(Ctx& c, const kdn_denoise_frame* f) {
    if (!f || f->width <= 0 || f->height <= 0 || f->width > 16384 || f->height > 16384) {
        c.error = "kdn_denoise: frame size out of 1..16384";
        return false;
    }
    return true;
}

}  // namespace

// The filter with its fields in the context's scratch (grown on demand).
bool denoiseDevice(Ctx& c, const kdn_denoise_frame& f, const kdn_denoise_pixel* pixels, float* rgb, hipStream_t st) {
    const size_t n = (size_t)f.width * f.height;
    if (!ensure(c.scratch[4], n * kDenoiseFields * sizeof(float) + 256, c.error, false)) return false;
    return denoiseFields(c, f, pixels, rgb, st, (float*)c.scratch[4].ptr);
}

// The device address of this GPU's static working memory, looked up once
// per context with its device current (kdn_open calls it).
bool denoiseStaticScratch(Ctx& c) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
        c.error = std::string("hipGetSymbolAddress of the denoiser's static working memory: ") + hipGetErrorString(e);
        return false;
    }
    c.denoise_static = (float*)ptr;
    return true;
}

}  // namespace kdn

using namespace kdn;

extern "C" {

int kdn_denoise_device(kdn_ctx* ctx, const kdn_denoise_frame* frame, const void* pixels, void* rgb, void* stream) {
    if (!ctx || !pixels || !rgb) return KDN_E_ARGUMENT;
    Ctx& c = ctx->c;
    if (!checkFrame(c, frame)) return KDN_E_ARGUMENT;
    hipSetDevice(c.device);
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    return denoiseDevice(c, *frame, (const kdn_denoise_pixel*)pixels, (float*)rgb, st) ? KDN_OK : KDN_E_DEVICE;
}

int kdn_denoise_device_static(kdn_ctx* ctx, const kdn_denoise_frame* frame, const void* pixels, void* rgb, void* stream) {
    if (!ctx || !pixels || !rgb) return KDN_E_ARGUMENT;
    Ctx& c = ctx->c;
    if (!checkFrame(c, frame)) return KDN_E_ARGUMENT;
    if ((size_t)frame->width * frame->height > kStaticPixels) {
        c.error = "kdn_denoise_device_static: frame of " + std::to_string((size_t)frame->width * frame->height) +
                  " pixels passes the static working memory's " + std::to_string(kStaticPixels) + " (kdn_denoise_device has no limit)";
        return KDN_E_ARGUMENT;
    }
    if (c.denoise_static == nullptr) {
        c.error = "kdn_denoise_device_static: the static working memory was not found at kdn_open";
        return KDN_E_STATE;
    }
    hipSetDevice(c.device);
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    return denoiseFields(c, *frame, (const kdn_denoise_pixel*)pixels, (float*)rgb, st, c.denoise_static) ? KDN_OK : KDN_E_DEVICE;
}

int kdn_denoise(kdn_ctx* ctx, const kdn_denoise_frame* frame, const kdn_denoise_pixel* pixels, float* rgb) {
    if (!ctx || !pixels || !rgb) return KDN_E_ARGUMENT;
    Ctx& c = ctx->c;
    if (!checkFrame(c, frame)) return KDN_E_ARGUMENT;
    hipSetDevice(c.device);
    const size_t n = (size_t)frame->width * frame->height;
    if (!ensure(c.scratch[5], n * (sizeof(kdn_denoise_pixel) + 3 * sizeof(float)) + 256, c.error, false)) return KDN_E_MEMORY;
    kdn_denoise_pixel* dp = (kdn_denoise_pixel*)c.scratch[5].ptr;
    float* drgb = (float*)(dp + n);
    if (hipMemcpyAsync(dp, pixels, n * sizeof(kdn_denoise_pixel), hipMemcpyHostToDevice, c.stream) != hipSuccess ||
        !denoiseDevice(c, *frame, dp, drgb, c.stream) ||
        hipMemcpyAsync(rgb, drgb, n * 3 * sizeof(float), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess) {
        if (c.error.empty()) c.error = "kdn_denoise: device call failed";
        return KDN_E_DEVICE;
    }
    return KDN_OK;
}

int kdn_denoise_accumulation(kdn_ctx* ctx, float* rgb, kdn_denoise_stats* stats) {
    if (!ctx || !rgb) return KDN_E_ARGUMENT;
    Ctx& c = ctx->c;
    if (!c.accum || !c.tree.built || !c.integrator_ready) {
        c.error = "kdn_denoise_accumulation: no accumulated image (kdn_render_pass first)";
        return KDN_E_STATE;
    }
    hipSetDevice(c.device);
    hipStream_t st = c.stream;
    const int W = c.width, H = c.height;
    const size_t n = (size_t)W * H;
    // the lens of the scene's camera (core/camera.cpp:9-25)
    const kdn_camera& cam = c.scene.camera;
    double fwd[3], rt[3], upv[3], len = 0;
    for (int k = 0; k < 3; ++k) {
        fwd[k] = (double)cam.target[k] - cam.eye[k];
        len += fwd[k] * fwd[k];
    }
    len = std::sqrt(len);
    for (int k = 0; k < 3; ++k) fwd[k] /= len;
    rt[0] = fwd[1] * cam.up[2] - fwd[2] * cam.up[1];
    rt[1] = fwd[2] * cam.up[0] - fwd[0] * cam.up[2];
    rt[2] = fwd[0] * cam.up[1] - fwd[1] * cam.up[0];
    len = std::sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
    for (int k = 0; k < 3; ++k) rt[k] /= len;
    upv[0] = rt[1] * fwd[2] - rt[2] * fwd[1];
    upv[1] = rt[2] * fwd[0] - rt[0] * fwd[2];
    upv[2] = rt[0] * fwd[1] - rt[1] * fwd[0];
    kdn_denoise_frame f{};
    f.width = W;
    f.height = H;
    for (int k = 0; k < 3; ++k) {
        f.forward[k] = fwd[k];
        f.right[k] = rt[k];
        f.up[k] = upv[k];
    }
    f.reach = std::tan((double)cam.fov / 2);
    f.aspect = (double)W / H;
    f.centre_x = f.centre_y = 0.5;
    f.demodulated = 0;
    const size_t bytes = n * (sizeof(kdn_ray) + sizeof(kdn_hit) + sizeof(kdn_denoise_pixel) + 3 * sizeof(float)) + 1024;
    if (!ensure(c.scratch[5], bytes, c.error, false)) return KDN_E_MEMORY;
    char* at = (char*)c.scratch[5].ptr;
    kdn_denoise_pixel* px = (kdn_denoise_pixel*)at;
    at += (n * sizeof(kdn_denoise_pixel) + 255) & ~size_t(255);
    kdn_ray* rays = (kdn_ray*)at;
    at += (n * sizeof(kdn_ray) + 255) & ~size_t(255);
    kdn_hit* hits = (kdn_hit*)at;
    at += (n * sizeof(kdn_hit) + 255) & ~size_t(255);
    float* drgb = (float*)at;
    const DenoiseLensDev L = lensOf(f);
    const dim3 grid((uint32_t)((n + kDenoiseBlock - 1) / kDenoiseBlock)), block(kDenoiseBlock);
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    if (stats) {
        for (hipEvent_t& e : ev) {
            if (hipEventCreate(&e) != hipSuccess) return KDN_E_DEVICE;
        }
        hipEventRecord(ev[0], st);
    }
    hipLaunchKernelGGL(k_dn_rays, grid, block, 0, st, L, W, H, cam.eye[0], cam.eye[1], cam.eye[2], rays);
    if (!extendRays(c, rays, (uint32_t)n, hits, st)) return KDN_E_DEVICE;
    hipLaunchKernelGGL(k_dn_history, grid, block, 0, st, (const double*)c.accum, (const double*)c.accum_sq, (const uint32_t*)c.accum_n,
                       (const kdn_hit*)hits, (const kdn_ray*)rays, (const uint32_t*)c.scene.leaf_material,
                       (const MaterialDev*)c.materials_dev, c.n_materials_dev, (uint32_t)n, px);
    if (stats) hipEventRecord(ev[1], st);
    const bool filtered = denoiseDevice(c, f, px, drgb, st);
    if (stats) {
        hipEventRecord(ev[2], st);
        hipEventSynchronize(ev[2]);
        float a = 0, b = 0;
        hipEventElapsedTime(&a, ev[0], ev[1]);
        hipEventElapsedTime(&b, ev[1], ev[2]);
        stats->gbuffer_ms = a;
        stats->filter_ms = b;
        for (hipEvent_t e : ev) hipEventDestroy(e);
    }
    if (!filtered ||
        hipMemcpyAsync(rgb, drgb, n * 3 * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        if (c.error.empty()) c.error = "kdn_denoise_accumulation: device call failed";
        return KDN_E_DEVICE;
    }
    return KDN_OK;
}

}  // extern "C"
