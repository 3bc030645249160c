// Images on the host: a linear RGB frame graded to bytes and written as a
// binary PPM, or written as a little-endian PFM (the reference's
// shared/image.cpp:16-96 and the sink/file plugin's outputs,
// plugins/sink/file/plugin.cpp:30-36).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "kodanuki.h"

namespace kdn {

// the thread's last error without a context (abi.cpp)
int setGlobalError(const std::string& m, int code);

namespace {

// The filmic curve and its ceiling past which the display saturates
// (filmicCurve, shared/image.cpp:25-30).
constexpr double kFilmicCeiling = 2.51 / 2.43;
constexpr double kFilmicSaturated = 16.0;
constexpr double kByteLevels = 255.0;

double filmic(double v) {
    if (v >= kFilmicSaturated) return kFilmicCeiling;
    return v * (2.51 * v + 0.03) / (v * (2.43 * v + 0.59) + 0.14);
}

double share(double v) {
    if (!(v > 0)) return 0;
    return v < 1 ? v : 1;
}

// displayed and encodedByte (shared/image.cpp:39-53)
uint8_t byteOf(double value, float exposure, int32_t curve) {
    double shown = 0;
    if (value > 0) {
        const double exposed = value * (double)exposure;
        const double mapped = curve == KDN_CURVE_FILMIC ? filmic(exposed) : exposed;
        shown = std::pow(share(mapped), 1 / 2.2);
    }
    return (uint8_t)std::lround(shown * kByteLevels);
}

bool checked(const float* rgb, int32_t width, int32_t height, const char* what) {
    if (!rgb || width <= 0 || height <= 0) {
        setGlobalError(std::string(what) + ": an image needs pixels and a positive size", KDN_E_ARGUMENT);
        return false;
    }
    return true;
}

int cannotWrite(const char* path) { return setGlobalError(std::string(path) + ":1:1: cannot write", KDN_E_ARGUMENT); }

}  // namespace
}  // namespace kdn

extern "C" {

int kdn_image_encode(const float* rgb, int32_t width, int32_t height, float exposure, int32_t curve, uint8_t* bytes) {
    if (!kdn::checked(rgb, width, height, "kdn_image_encode") || !bytes) return KDN_E_ARGUMENT;
    if (curve != KDN_CURVE_FILMIC && curve != KDN_CURVE_CLAMPED) return kdn::setGlobalError("kdn_image_encode: a curve is filmic or clamped", KDN_E_ARGUMENT);
    const size_t n = (size_t)width * height * 3;
    for (size_t i = 0; i < n; ++i) bytes[i] = kdn::byteOf((double)rgb[i], exposure, curve);
    return KDN_OK;
}

int kdn_image_write_ppm(const float* rgb, int32_t width, int32_t height, float exposure, int32_t curve, const char* path) {
    if (!path) return kdn::setGlobalError("kdn_image_write_ppm: null path", KDN_E_ARGUMENT);
    std::vector<uint8_t> bytes((size_t)width * height * 3 > 0 ? (size_t)width * height * 3 : 0);
    const int rc = kdn_image_encode(rgb, width, height, exposure, curve, bytes.data());
    if (rc) return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f) return kdn::cannotWrite(path);
    const bool ok = std::fprintf(f, "P6\n%d %d\n255\n", width, height) > 0 && std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
    if (std::fclose(f) != 0 || !ok) return kdn::cannotWrite(path);
    return KDN_OK;
}

int kdn_image_write_pfm(const float* rgb, int32_t width, int32_t height, const char* path) {
    if (!kdn::checked(rgb, width, height, "kdn_image_write_pfm")) return KDN_E_ARGUMENT;
    if (!path) return kdn::setGlobalError("kdn_image_write_pfm: null path", KDN_E_ARGUMENT);
    FILE* f = std::fopen(path, "wb");
    if (!f) return kdn::cannotWrite(path);
    // little-endian (a negative scale), rows bottom first; this library's
    // hosts are little-endian, so the floats are written as they lie
    bool ok = std::fprintf(f, "PF\n%d %d\n-1\n", width, height) > 0;
    const size_t row = (size_t)width * 3;
    for (int32_t y = height - 1; ok && y >= 0; --y) ok = std::fwrite(rgb + (size_t)y * row, sizeof(float), row, f) == row;
    if (std::fclose(f) != 0 || !ok) return kdn::cannotWrite(path);
    return KDN_OK;
}

}  // extern "C"
