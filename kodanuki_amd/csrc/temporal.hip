// The temporal member on the device: a viewer's frames drawn with a
// G-buffer at factor 1 (every pixel drawn each frame through its centre),
// the history demodulated and resident in HBM between frames, as the
// reference's temporalFrame decides it (plugins/reconstruction/temporal/
// temporal.cpp:162-1029 with demodulatedFrame true).
//
// One lane per pixel in 16x16 blocks. The block first stages its pixels'
// demodulated sample lights with a one-pixel halo in LDS (18x18 FP64
// triples, image edges clamped as the reference clamps them), so the 3x3
// neighbourhood each lane's clamp reads (rowNeighbourhoods,
// temporal.cpp:935-955) is nine LDS reads instead of nine L2 gathers; the
// history a moving camera reprojects is read bilinearly from the history
// before (four 96-byte pixels a lane). Everything is FP64 in the
// reference's order of operations. Per pixel the kernel reads the frame's
// FP64 sums and count (28 B), its record and G-buffer surfaces (64 B) and
// up to four history pixels, and writes one history pixel and the shown
// RGB (108 B): it is bound by HBM.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ctx.h"
#include "integrator_dev.h"

namespace kdn {

namespace {

constexpr int kTile = 16;
constexpr int kHalo = kTile + 2;
constexpr double kMovingCap = 32, kStillCap = 256, kYoung = 8;  // temporal.cpp:16-24
constexpr double kMovingSpread = 1.25, kStillSpread = 6;
constexpr double kDepthShare = 0.03, kNormalCosine = 0.9, kAlbedoFloor = 1e-3;

struct D3 {
    double x, y, z;
};
__host__ __device__ inline D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__host__ __device__ inline D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline D3 operator*(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline D3 times(D3 a, D3 b) { return D3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KDN_D double len(D3 a) { return __builtin_sqrt(dot(a, a)); }
KDN_D double luminance(D3 c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }
KDN_D int clamped(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// TemporalCamera of a lens at the grid's size (temporal.cpp:307-322)
struct CameraDev {
    D3 eye, forward, corner, across, down, across_inverse, down_inverse;
    double width, height;
};
CameraDev cameraOf(const kdn_lens& l, int width, int height) {
    const D3 right = d3(l.right[0], l.right[1], l.right[2]), up = d3(l.up[0], l.up[1], l.up[2]);
    const double sideways = l.reach * l.aspect;
    CameraDev c;
    c.eye = d3(l.eye[0], l.eye[1], l.eye[2]);
    c.forward = d3(l.forward[0], l.forward[1], l.forward[2]);
    c.across = right * (2 * sideways / (double)width);
    c.down = up * (-2 * l.reach / (double)height);
    c.corner = c.forward - right * sideways + up * l.reach;
    c.across_inverse = c.across * (1 / dot(c.across, c.across));
    c.down_inverse = c.down * (1 / dot(c.down, c.down));
    c.width = (double)width;
    c.height = (double)height;
    return c;
}
KDN_D D3 heading(const CameraDev& c, double x, double y) { return c.corner + c.across * x + c.down * y; }

struct TemporalArgs {
    int width, height;
    CameraDev cam, prev;
    uint32_t moved, recounted, kept;
    // the frame's lights: parts (3 doubles a pixel), or the FP64 sums of a
    // pass and its counts
    const double* parts;
    const double* sum;
    const uint32_t* count;
    const kdn_surface* record;
    const kdn_surface* gbuffer;
    const kdn_denoise_pixel* before;
    kdn_denoise_pixel* out;
    float* image;
};

KDN_D D3 divisor(const float a[3]) {
    return d3(a[0] < kAlbedoFloor ? 1.0 : (double)a[0], a[1] < kAlbedoFloor ? 1.0 : (double)a[1], a[2] < kAlbedoFloor ? 1.0 : (double)a[2]);
}
KDN_D D3 divisor_d(const double a[3]) {
    return d3(a[0] < kAlbedoFloor ? 1.0 : a[0], a[1] < kAlbedoFloor ? 1.0 : a[1], a[2] < kAlbedoFloor ? 1.0 : a[2]);
}

// sampleLight: the pixel's light this frame (frameLight: its sum times the
// reciprocal of its samples) over its sample's albedo
KDN_D D3 sample_light(const TemporalArgs& A, size_t p) {
    D3 light;
    if (A.parts) {
        light = d3(A.parts[3 * p], A.parts[3 * p + 1], A.parts[3 * p + 2]);
    } else {
        const uint32_t n = A.count[p];
        const double k = n > 0 ? 1.0 / (double)n : 0.0;
        light = d3(A.sum[3 * p] * k, A.sum[3 * p + 1] * k, A.sum[3 * p + 2] * k);
    }
    const D3 d = divisor(A.record[p].albedo);
    return d3(light.x / d.x, light.y / d.y, light.z / d.z);
}

KDN_D bool sky(const kdn_surface& s) { return s.leaf == KDN_NO_LEAF || !(fabs(s.distance) < INFINITY); }
KDN_D uint32_t shown_leaf(const kdn_surface& s) { return sky(s) ? KDN_NO_LEAF : s.leaf; }

struct Plane {
    uint32_t leaf;
    D3 point, normal;
    double reach;
};
KDN_D Plane sky_plane() { return Plane{KDN_NO_LEAF, d3(0, 0, 0), d3(0, 0, 0), 0}; }
// planeThrough (temporal.cpp:358-368)
KDN_D Plane plane_through(const kdn_surface& s, D3 eye, D3 h) {
    if (sky(s)) return sky_plane();
    const double d = (double)s.distance;
    return Plane{s.leaf, eye + h * (d / len(h)), d3(s.normal[0], s.normal[1], s.normal[2]), kDepthShare * d};
}
KDN_D bool same_leaf(uint32_t a, uint32_t b, bool recounted) { return recounted ? ((a == KDN_NO_LEAF) == (b == KDN_NO_LEAF)) : a == b; }
// onPlane (temporal.cpp:397-408)
KDN_D bool on_plane(const Plane& plane, const Plane& other, bool recounted) {
    if (!same_leaf(other.leaf, plane.leaf, recounted)) return false;
    if (plane.leaf == KDN_NO_LEAF) return true;
    if (dot(other.normal, plane.normal) < kNormalCosine) return false;
    return fabs(dot(other.point - plane.point, plane.normal)) <= plane.reach;
}

struct History {
    bool found;
    D3 colour;
    double moment;
    float length;
};

// the frame before's pixel (x, y) on its plane (lastPlane, temporal.cpp:378-391)
KDN_D Plane last_plane(const TemporalArgs& A, int x, int y) {
    const kdn_denoise_pixel& last = A.before[(size_t)y * A.width + x];
    if (last.leaf == KDN_NO_LEAF || !(fabs(last.distance) < INFINITY)) return sky_plane();
    const D3 h = heading(A.prev, (double)x + 0.5, (double)y + 0.5);
    return Plane{last.leaf, A.prev.eye + h * ((double)last.distance / len(h)), d3(last.normal[0], last.normal[1], last.normal[2]), 0};
}

// surfacedHistory and fetchedHistory (temporal.cpp:495-528, 584-603): the
// G-buffer's point moved into the camera before and read bilinearly from
// the pixels there on its plane
KDN_D History surfaced_history(const TemporalArgs& A, int x, int y) {
    History none{false, d3(0, 0, 0), 0, 0};
    const D3 h = heading(A.cam, (double)x + 0.5, (double)y + 0.5);
    const Plane plane = plane_through(A.gbuffer[(size_t)y * A.width + x], A.cam.eye, h);
    const D3 along = plane.leaf == KDN_NO_LEAF ? h : plane.point - A.prev.eye;
    // placeBefore (temporal.cpp:419-436)
    const double ahead = dot(along, A.prev.forward);
    if (!(ahead > 0)) return none;
    const double scale = 1 / ahead;
    const double px = dot(along, A.prev.across_inverse) * scale + 0.5 * A.prev.width;
    const double py = dot(along, A.prev.down_inverse) * scale + 0.5 * A.prev.height;
    if (!(px >= 0 && px < A.prev.width && py >= 0 && py < A.prev.height)) return none;
    const double ax = px - 0.5, ay = py - 0.5;
    const double lx = floor(ax), ly = floor(ay);
    const double tx = ax - lx, ty = ay - ly;
    D3 colour = d3(0, 0, 0);
    double moment = 0, weight = 0;
    float length = 0;
    for (int j = 0; j < 2; ++j) {
        for (int i = 0; i < 2; ++i) {
            const int cx = (int)lx + i, cy = (int)ly + j;
            const double w = (i == 0 ? 1 - tx : tx) * (j == 0 ? 1 - ty : ty);
            if (!(w > 0 && cx >= 0 && cx < A.width && cy >= 0 && cy < A.height)) continue;
            const kdn_denoise_pixel& last = A.before[(size_t)cy * A.width + cx];
            if (!(last.length > 0) || !on_plane(plane, last_plane(A, cx, cy), A.recounted != 0)) continue;
            colour = colour + d3(last.colour[0], last.colour[1], last.colour[2]) * w;
            moment += last.moment * w;
            length += (float)((double)last.length * w);
            weight += w;
        }
    }
    if (!(weight > 0)) return none;
    return History{true, colour * (1 / weight), moment / weight, (float)((double)length / weight)};
}

// jointFilled (temporal.cpp:711-765) at factor 1: the pixel's own sample
// where its first surface lies on the G-buffer's plane, else the mean of
// the 3x3 samples on it, else the bilinear fill (the sample itself)
KDN_D D3 joint_filled(const TemporalArgs& A, int x, int y, const D3 (*L)[kHalo], int lx, int ly) {
    const D3 h = heading(A.cam, (double)x + 0.5, (double)y + 0.5);
    const Plane plane = plane_through(A.gbuffer[(size_t)y * A.width + x], A.cam.eye, h);
    const Plane own = plane_through(A.record[(size_t)y * A.width + x], A.cam.eye, h);
    if (on_plane(plane, own, false)) return L[ly][lx] * 1.0 * (1 / 1.0);
    D3 sum = d3(0, 0, 0);
    double weight = 0;
    for (int dy = -1; dy <= 1; ++dy) {
        for (int dx = -1; dx <= 1; ++dx) {
            const int sx = clamped(x + dx, 0, A.width - 1), sy = clamped(y + dy, 0, A.height - 1);
            const D3 hs = heading(A.cam, (double)sx + 0.5, (double)sy + 0.5);
            if (!on_plane(plane, plane_through(A.record[(size_t)sy * A.width + sx], A.cam.eye, hs), false)) continue;
            sum = sum + L[ly + dy][lx + dx] * 1.0;
            weight += 1;
        }
    }
    if (weight > 0) return sum * (1 / weight);
    // filledColour: corners (x, y), (x + 1, y), ... with weights 1, 0, 0, 0
    const int xr = (x + 1 < A.width ? x + 1 : A.width - 1) - x, yb = (y + 1 < A.height ? y + 1 : A.height - 1) - y;
    const D3 top = L[ly][lx] * (1 - 0.0) + L[ly][lx + xr] * 0.0;
    const D3 bottom = L[ly + yb][lx] * (1 - 0.0) + L[ly + yb][lx + xr] * 0.0;
    return top * (1 - 0.0) + bottom * 0.0;
}

__global__ void __launch_bounds__(kTile * kTile) k_temporal(TemporalArgs A) {
    __shared__ D3 L[kHalo][kHalo];
    const int x0 = (int)blockIdx.x * kTile, y0 = (int)blockIdx.y * kTile;
    // stage the block's demodulated sample lights with a clamped halo
    for (int k = (int)threadIdx.x; k < kHalo * kHalo; k += kTile * kTile) {
        const int hx = k % kHalo, hy = k / kHalo;
        const int gx = clamped(x0 + hx - 1, 0, A.width - 1), gy = clamped(y0 + hy - 1, 0, A.height - 1);
        L[hy][hx] = sample_light(A, (size_t)gy * A.width + gx);
    }
    __syncthreads();
    const int tx = (int)threadIdx.x % kTile, ty = (int)threadIdx.x / kTile;
    const int x = x0 + tx, y = y0 + ty;
    if (x >= A.width || y >= A.height) return;
    const int lx = tx + 1, ly = ty + 1;
    const size_t p = (size_t)y * A.width + x;
    // the 3x3 neighbourhood: columns (middle, above, below), then the
    // middle column, the one left and the one right (rowNeighbourhoods)
    D3 sums[3], squares[3];
    for (int c = 0; c < 3; ++c) {
        const int cx = lx + (c == 0 ? 0 : (c == 1 ? -1 : 1));
        const D3 m = L[ly][cx], u = L[ly - 1][cx], b = L[ly + 1][cx];
        sums[c] = m + u + b;
        squares[c] = times(m, m) + times(u, u) + times(b, b);
    }
    const D3 sum = sums[0] + sums[1] + sums[2], sq = squares[0] + squares[1] + squares[2];
    const D3 mean = sum * (1.0 / 9);
    const D3 m9 = d3(sum.x / 9, sum.y / 9, sum.z / 9);
    const D3 deviation = d3(__builtin_sqrt(fmax(sq.x / 9 - m9.x * m9.x, 0.0)), __builtin_sqrt(fmax(sq.y / 9 - m9.y * m9.y, 0.0)),
                            __builtin_sqrt(fmax(sq.z / 9 - m9.z * m9.z, 0.0)));

    const kdn_surface g = A.gbuffer[p];
    const bool aimed = shown_leaf(A.record[p]) == shown_leaf(g);
    const bool moved = A.moved != 0;
    // historyFor (temporal.cpp:605-621)
    History hist{false, d3(0, 0, 0), 0, 0};
    if (A.kept) {
        if (!moved) {
            const kdn_denoise_pixel& last = A.before[p];
            if (last.length > 0) hist = History{true, d3(last.colour[0], last.colour[1], last.colour[2]), last.moment, last.length};
        } else {
            hist = surfaced_history(A, x, y);
        }
    }
    const D3 sample = L[ly][lx];
    D3 colour;
    double moment;
    float length;
    if (hist.found) {
        // heldBlend (temporal.cpp:846-866)
        const float cap = moved ? (float)kMovingCap : (float)kStillCap;
        length = fminf(hist.length, cap);
        D3 held = hist.colour;
        if (moved || hist.length < (float)kYoung) {
            const double spread = moved ? kMovingSpread : kStillSpread;
            const D3 reach = deviation * spread;
            held = d3(fmin(fmax(held.x, mean.x - reach.x), mean.x + reach.x), fmin(fmax(held.y, mean.y - reach.y), mean.y + reach.y),
                      fmin(fmax(held.z, mean.z - reach.z), mean.z + reach.z));
        }
        if (aimed) {
            const double lum = luminance(sample);
            const double alpha = fmax(1.0 / ((double)fminf(hist.length, cap) + 1), 1.0 / (double)cap);
            colour = held * (1 - alpha) + sample * alpha;
            moment = hist.moment * (1 - alpha) + lum * lum * alpha;
            length = fminf(length + 1, cap);
        } else {
            colour = held;
            moment = hist.moment;
        }
    } else {
        // freshBlend (temporal.cpp:796-804)
        colour = aimed ? sample : joint_filled(A, x, y, L, lx, ly);
        const double lum = luminance(colour);
        moment = lum * lum;
        length = aimed ? 1.0f : 0.0f;
    }
    kdn_denoise_pixel out;
    out.colour[0] = colour.x;
    out.colour[1] = colour.y;
    out.colour[2] = colour.z;
    out.moment = moment;
    out.length = length;
    out.distance = g.distance;
    for (int c = 0; c < 3; ++c) {
        out.normal[c] = (double)g.normal[c];
        out.albedo[c] = (double)g.albedo[c];
    }
    out.leaf = g.leaf;
    out.pad = 0;
    A.out[p] = out;
    const D3 shown = times(colour, divisor_d(out.albedo));
    A.image[3 * p] = (float)shown.x;
    A.image[3 * p + 1] = (float)shown.y;
    A.image[3 * p + 2] = (float)shown.z;
}
KDN_SYNCS_BLOCK(k_temporal)

// the G-buffer: a primary ray through every pixel's centre
__global__ void k_tm_rays(CameraDev cam, int width, int height, kdn_ray* rays) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= (uint32_t)width * (uint32_t)height) return;
    const D3 h = heading(cam, (double)(at % (uint32_t)width) + 0.5, (double)(at / (uint32_t)width) + 0.5);
    const double l = len(h);
    kdn_ray r;
    r.origin[0] = (float)cam.eye.x;
    r.origin[1] = (float)cam.eye.y;
    r.origin[2] = (float)cam.eye.z;
    r.direction[0] = (float)(h.x / l);
    r.direction[1] = (float)(h.y / l);
    r.direction[2] = (float)(h.z / l);
    r.far_ = kInf;
    r.inside = 0;
    r.within = KDN_NO_MATERIAL;
    rays[at] = r;
}

__global__ void k_tm_surfaces(const kdn_hit* hits, const kdn_ray* rays, const uint32_t* leafMaterial, const MaterialDev* materials,
                              uint32_t nMaterials, uint32_t n, kdn_surface* out) {
    const uint32_t at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at < n) out[at] = surface_met(hits[at], rays[at], leafMaterial, materials, nMaterials);
}

bool launchTemporal(Ctx& c, const TemporalArgs& A, hipStream_t st) {
    const dim3 grid((uint32_t)((A.width + kTile - 1) / kTile), (uint32_t)((A.height + kTile - 1) / kTile));
    hipLaunchKernelGGL(k_temporal, grid, dim3(kTile * kTile), 0, st, A);
    KDN_HIP(hipGetLastError());
    return true;
}

}  // namespace

// the lens of the resident scene's camera at an image's size (lensOf,
// core/camera.cpp:9-25)
kdn_lens lensOfCamera(const kdn_camera& cam, int width, int height) {
    kdn_lens l{};
    double fwd[3], len = 0;
    for (int k = 0; k < 3; ++k) {
        fwd[k] = (double)cam.target[k] - cam.eye[k];
        len += fwd[k] * fwd[k];
    }
    len = std::sqrt(len);
    for (int k = 0; k < 3; ++k) {
        l.eye[k] = cam.eye[k];
        l.forward[k] = fwd[k] / len;
    }
    double rt[3] = {l.forward[1] * cam.up[2] - l.forward[2] * cam.up[1], l.forward[2] * cam.up[0] - l.forward[0] * cam.up[2],
                    l.forward[0] * cam.up[1] - l.forward[1] * cam.up[0]};
    const double rl = std::sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
    for (int k = 0; k < 3; ++k) l.right[k] = rt[k] / rl;
    l.up[0] = l.right[1] * l.forward[2] - l.right[2] * l.forward[1];
    l.up[1] = l.right[2] * l.forward[0] - l.right[0] * l.forward[2];
    l.up[2] = l.right[0] * l.forward[1] - l.right[1] * l.forward[0];
    l.reach = std::tan((double)cam.fov / 2);
    l.aspect = (double)width / height;
    return l;
}

bool temporalStep(Ctx& c, int width, int height, const kdn_lens& lens, const kdn_lens& before, const kdn_temporal_desc& d,
                  const double* parts, const kdn_surface* record, const kdn_surface* gbuffer, const kdn_denoise_pixel* beforePixels,
                  int beforeWidth, int beforeHeight, kdn_denoise_pixel* pixels, float* shown) {
    const size_t n = (size_t)width * (size_t)height;
    const size_t nb = beforePixels ? (size_t)beforeWidth * (size_t)beforeHeight : 0;
    const size_t o0 = 0, o1 = o0 + ((n * 3 * sizeof(double) + 255) & ~size_t(255));
    const size_t o2 = o1 + ((n * sizeof(kdn_surface) + 255) & ~size_t(255)), o3 = o2 + ((n * sizeof(kdn_surface) + 255) & ~size_t(255));
    const size_t o4 = o3 + ((nb * sizeof(kdn_denoise_pixel) + 255) & ~size_t(255));
    const size_t o5 = o4 + ((n * sizeof(kdn_denoise_pixel) + 255) & ~size_t(255));
    const size_t bytes = o5 + n * 3 * sizeof(float);
    if (!ensure(c.scratch[6], bytes, c.error, false)) return false;
    char* base = (char*)c.scratch[6].ptr;
    hipStream_t st = c.stream;
    KDN_HIP(hipMemcpyAsync(base + o0, parts, n * 3 * sizeof(double), hipMemcpyHostToDevice, st));
    KDN_HIP(hipMemcpyAsync(base + o1, record, n * sizeof(kdn_surface), hipMemcpyHostToDevice, st));
    KDN_HIP(hipMemcpyAsync(base + o2, gbuffer, n * sizeof(kdn_surface), hipMemcpyHostToDevice, st));
    if (nb) KDN_HIP(hipMemcpyAsync(base + o3, beforePixels, nb * sizeof(kdn_denoise_pixel), hipMemcpyHostToDevice, st));
    TemporalArgs A{};
    A.width = width;
    A.height = height;
    A.cam = cameraOf(lens, width, height);
    A.prev = cameraOf(before, width, height);
    A.moved = d.moved;
    A.recounted = d.recounted;
    A.kept = !d.restarted && beforePixels && beforeWidth == width && beforeHeight == height && width > 0;
    A.parts = (const double*)(base + o0);
    A.record = (const kdn_surface*)(base + o1);
    A.gbuffer = (const kdn_surface*)(base + o2);
    A.before = (const kdn_denoise_pixel*)(base + o3);
    A.out = (kdn_denoise_pixel*)(base + o4);
    A.image = (float*)(base + o5);
    if (!launchTemporal(c, A, st)) return false;
    KDN_HIP(hipMemcpyAsync(pixels, base + o4, n * sizeof(kdn_denoise_pixel), hipMemcpyDeviceToHost, st));
    KDN_HIP(hipMemcpyAsync(shown, base + o5, n * 3 * sizeof(float), hipMemcpyDeviceToHost, st));
    KDN_HIP(hipStreamSynchronize(st));
    return true;
}

bool temporalFrame(Ctx& c, const kdn_temporal_desc& d, float* shown, float* denoised, kdn_temporal_stats* stats) {
    if (!c.accum || !c.firsts || c.firsts_width != c.width || c.firsts_height != c.height) {
        c.error = "kdn_temporal_frame: no frame record (a pass with flags bit 9 first)";
        return false;
    }
    const int W = c.width, H = c.height;
    const size_t n = (size_t)W * (size_t)H;
    hipStream_t st = c.stream;
    const kdn_lens lens = lensOfCamera(c.scene.camera, W, H);
    const bool kept = !d.restarted && c.tm_width == W && c.tm_height == H;
    if (!ensure(c.tm_history[0], n * sizeof(kdn_denoise_pixel), c.error, false) ||
        !ensure(c.tm_history[1], n * sizeof(kdn_denoise_pixel), c.error, false) ||
        !ensure(c.tm_gbuffer, n * sizeof(kdn_surface), c.error, false))
        return false;
    // scratch: the G-buffer's rays and hits, then the shown and denoised images
    const size_t o1 = (n * sizeof(kdn_ray) + 255) & ~size_t(255), o2 = o1 + ((n * sizeof(kdn_hit) + 255) & ~size_t(255));
    const size_t o3 = o2 + ((n * 3 * sizeof(float) + 255) & ~size_t(255));
    if (!ensure(c.scratch[5], o3 + n * 3 * sizeof(float), c.error, false)) return false;
    char* base = (char*)c.scratch[5].ptr;
    kdn_ray* rays = (kdn_ray*)base;
    kdn_hit* hits = (kdn_hit*)(base + o1);
    float* image = (float*)(base + o2);
    float* filtered = (float*)(base + o3);
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    struct Events {
        hipEvent_t* e;
        ~Events() {
            for (int k = 0; k < 4; ++k)
                if (e[k]) hipEventDestroy(e[k]);
        }
    } guard{ev};
    if (stats) {
        for (hipEvent_t& e : ev) KDN_HIP(hipEventCreate(&e));
        KDN_HIP(hipEventRecord(ev[0], st));
    }
    const CameraDev cam = cameraOf(lens, W, H);
    const dim3 grid((uint32_t)((n + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_tm_rays, grid, block, 0, st, cam, W, H, rays);
    KDN_HIP(hipGetLastError());
    if (!extendRays(c, rays, (uint32_t)n, hits, st)) return false;
    hipLaunchKernelGGL(k_tm_surfaces, grid, block, 0, st, (const kdn_hit*)hits, (const kdn_ray*)rays, (const uint32_t*)c.scene.leaf_material,
                       (const MaterialDev*)c.materials_dev, c.n_materials_dev, (uint32_t)n, (kdn_surface*)c.tm_gbuffer.ptr);
    KDN_HIP(hipGetLastError());
    if (stats) KDN_HIP(hipEventRecord(ev[1], st));
    const int next = 1 - c.tm_current;
    TemporalArgs A{};
    A.width = W;
    A.height = H;
    A.cam = cam;
    A.prev = kept ? cameraOf(c.tm_lens, W, H) : cam;
    A.moved = d.moved;
    A.recounted = d.recounted;
    A.kept = kept ? 1u : 0u;
    A.sum = c.accum;
    A.count = c.accum_n;
    A.record = c.firsts;
    A.gbuffer = (const kdn_surface*)c.tm_gbuffer.ptr;
    A.before = (const kdn_denoise_pixel*)c.tm_history[c.tm_current].ptr;
    A.out = (kdn_denoise_pixel*)c.tm_history[next].ptr;
    A.image = image;
    if (!launchTemporal(c, A, st)) return false;
    if (stats) KDN_HIP(hipEventRecord(ev[2], st));
    if (d.denoise) {
        kdn_denoise_frame f{};
        f.width = W;
        f.height = H;
        for (int k = 0; k < 3; ++k) {
            f.forward[k] = lens.forward[k];
            f.right[k] = lens.right[k];
            f.up[k] = lens.up[k];
        }
        f.reach = lens.reach;
        f.aspect = lens.aspect;
        f.centre_x = f.centre_y = 0.5;
        f.demodulated = 1;
        if (!denoiseDevice(c, f, A.out, filtered, st)) return false;
    }
    if (stats) KDN_HIP(hipEventRecord(ev[3], st));
    if (shown) KDN_HIP(hipMemcpyAsync(shown, image, n * 3 * sizeof(float), hipMemcpyDeviceToHost, st));
    if (denoised && d.denoise) KDN_HIP(hipMemcpyAsync(denoised, filtered, n * 3 * sizeof(float), hipMemcpyDeviceToHost, st));
    KDN_HIP(hipStreamSynchronize(st));
    if (stats) {
        float a = 0, b = 0, e = 0;
        hipEventElapsedTime(&a, ev[0], ev[1]);
        hipEventElapsedTime(&b, ev[1], ev[2]);
        hipEventElapsedTime(&e, ev[2], ev[3]);
        stats->gbuffer_ms = a;
        stats->temporal_ms = b;
        stats->filter_ms = e;
    }
    c.tm_current = next;
    c.tm_width = W;
    c.tm_height = H;
    c.tm_lens = lens;
    return true;
}

bool temporalHistory(Ctx& c, kdn_denoise_pixel* pixels) {
    if (c.tm_width <= 0) {
        c.error = "kdn_temporal_history: no frame yet";
        return false;
    }
    const size_t n = (size_t)c.tm_width * (size_t)c.tm_height;
    KDN_HIP(hipMemcpyAsync(pixels, c.tm_history[c.tm_current].ptr, n * sizeof(kdn_denoise_pixel), hipMemcpyDeviceToHost, c.stream));
    KDN_HIP(hipStreamSynchronize(c.stream));
    return true;
}

}  // namespace kdn
