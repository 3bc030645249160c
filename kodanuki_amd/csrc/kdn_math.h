// Device arithmetic of the tape: FP32 point values, FP32 intervals with
// outward rounding, FP32 duals. Compiled by hipcc for gfx950 only, with
// -ffp-contract=off so every operation rounds once as written.
//
// Opcode semantics follow the reference number types
// (source/sdf/numbers.cpp:107-188 point, :198-430 Interval, :433-590 Dual).
// The reference computes intervals in FP64 with round-to-nearest; here an
// interval endpoint is rounded outward exactly (directed rounding from the
// sign of the fma residual), so an interval encloses the real value of the
// program with the scene's double constants. Transcendentals (sin, cos,
// atan2, pow) are computed in double by the polynomials below, then rounded
// to float (to nearest for points, outward with a 1e-13 relative slack for
// intervals), so that any IEEE double machine gives the same bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kodanuki.h"

#define KDN_D __device__ __forceinline__
// Marks a kernel that synchronises its block (__syncthreads); the device
// build needs no mark (the CPU build of the library runs such kernels'
// blocks with their threads at once).
#ifndef KDN_SYNCS_BLOCK
#define KDN_SYNCS_BLOCK(kernel)
#endif

namespace kdn {

constexpr float kFltMin = 1.17549435e-38f;
constexpr float kInf = __builtin_huge_valf();
constexpr double kPiD = 3.14159265358979323846;

KDN_D uint32_t fbits(float x) { return __float_as_uint(x); }
KDN_D float bitsf(uint32_t b) { return __uint_as_float(b); }
KDN_D bool isnanf_(float x) { return x != x; }
KDN_D float fmin_(float a, float b) { return b < a ? b : a; }  // std::min
KDN_D float fmax_(float a, float b) { return a < b ? b : a; }  // std::max

// The next float above x; values within FLT_MIN of zero go to FLT_MIN so
// that no subnormal is made.
KDN_D float up(float x) {
    if (isnanf_(x) || x == kInf) return x;
    if (x < kFltMin && x >= -kFltMin) return kFltMin;
    const uint32_t b = fbits(x);
    return bitsf(x > 0 ? b + 1 : b - 1);
}
KDN_D float down(float x) { return -up(-x); }

KDN_D bool tiny(float p) { return !(p != 0 && (p > 1e-30f || p < -1e-30f)) || p == kInf || p == -kInf || isnanf_(p); }

// ---- directed rounding of + - * / sqrt -----------------------------------
KDN_D float add_rd(float a, float b) {
    const float s = a + b;
    if (tiny(s) || a == kInf || a == -kInf || b == kInf || b == -kInf) {
        if (a == -kInf || b == -kInf) return s;  // an infinite operand is exact
        if (a == kInf || b == kInf) return s;
        return s == 0 && (a == 0 || b == 0) ? s : down(s);
    }
    const float bv = s - a, av = s - bv;
    const float err = (a - av) + (b - bv);
    return err < 0 ? down(s) : s;
}
KDN_D float add_ru(float a, float b) { return -add_rd(-a, -b); }
KDN_D float sub_rd(float a, float b) { return add_rd(a, -b); }
KDN_D float sub_ru(float a, float b) { return add_ru(a, -b); }

KDN_D float mul_rd(float a, float b) {
    if (a == 0 || b == 0) return 0.0f;
    const float p = a * b;
    if (tiny(p)) {
        if ((a == kInf || a == -kInf || b == kInf || b == -kInf)) return p;
        return down(p);
    }
    const float r = __builtin_fmaf(a, b, -p);
    return r < 0 ? down(p) : p;
}
KDN_D float mul_ru(float a, float b) { return -mul_rd(-a, b); }

KDN_D float div_rd(float a, float b) {
    if (a == 0) return 0.0f;
    const float q = a / b;
    if (tiny(q) || b == kInf || b == -kInf || a == kInf || a == -kInf) {
        if (q == kInf || q == -kInf) return (a == kInf || a == -kInf) ? q : down(q);
        return down(q);
    }
    const float r = __builtin_fmaf(q, b, -a);  // q*b - a = (q - a/b) * b
    const bool qAbove = (r > 0) == (b > 0);
    return (r != 0 && qAbove) ? down(q) : q;
}
KDN_D float div_ru(float a, float b) { return -div_rd(-a, b); }

KDN_D float sqrt_rd(float a) {
    if (a <= 0) return 0.0f;
    const float s = __builtin_sqrtf(a);
    if (s == kInf) return s;
    const float r = __builtin_fmaf(s, s, -a);
    return r > 0 ? down(s) : s;
}
KDN_D float sqrt_ru(float a) {
    if (a <= 0) return 0.0f;
    const float s = __builtin_sqrtf(a);
    if (s == kInf) return s;
    const float r = __builtin_fmaf(s, s, -a);
    return r < 0 ? up(s) : s;
}

// ---- double kernels for transcendentals ------------------------------------
// sin and cos: reduce by pi/2 in three parts (Cody-Waite), Taylor
// polynomials to degree 17/16 on |r| <= pi/4 (truncation below 2e-17).
KDN_D double poly_sin(double r) {
    const double r2 = r * r;
    double s = -1.0 / 355687428096000.0;                 // -1/17!
    s = s * r2 + 1.0 / 1307674368000.0;                  // 1/15!
    s = s * r2 - 1.0 / 6227020800.0;                     // -1/13!
    s = s * r2 + 1.0 / 39916800.0;                       // 1/11!
    s = s * r2 - 1.0 / 362880.0;                         // -1/9!
    s = s * r2 + 1.0 / 5040.0;                           // 1/7!
    s = s * r2 - 1.0 / 120.0;                            // -1/5!
    s = s * r2 + 1.0 / 6.0;                              // 1/3!
    return r - r * r2 * s;
}
KDN_D double poly_cos(double r) {
    const double r2 = r * r;
    double c = 1.0 / 20922789888000.0;                   // 1/16!
    c = c * r2 - 1.0 / 87178291200.0;                    // -1/14!
    c = c * r2 + 1.0 / 479001600.0;                      // 1/12!
    c = c * r2 - 1.0 / 3628800.0;                        // -1/10!
    c = c * r2 + 1.0 / 40320.0;                          // 1/8!
    c = c * r2 - 1.0 / 720.0;                            // -1/6!
    c = c * r2 + 1.0 / 24.0;                             // 1/4!
    c = c * r2 - 0.5;                                    // -1/2!
    return 1.0 + r2 * c;
}
// quadrant q in 0..3 and reduced r
KDN_D double reduced(double x, int* q) {
    const double k = __builtin_rint(x * 0.63661977236758134308);  // 2/pi
    const double p1 = 1.57079632673412561417e+00;  // the leading 33 bits of pi/2
    const double p2 = 6.07710050650619224932e-11;  // the rest of pi/2
    const double r = (x - k * p1) - k * p2;
    const long long ki = (long long)k;
    *q = (int)(ki & 3);
    return r;
}
KDN_D double d_sin(double x) {
    int q;
    const double r = reduced(x, &q);
    const double v = (q & 1) ? poly_cos(r) : poly_sin(r);
    return (q & 2) ? -v : v;
}
KDN_D double d_cos(double x) {
    int q;
    const double r = reduced(x, &q);
    const double v = (q & 1) ? poly_sin(r) : poly_cos(r);
    return ((q + 1) & 2) ? -v : v;
}
// atan on [0, 1] by halving the argument twice (atan x = 2 atan(x / (1 +
// sqrt(1 + x^2)))) then the Taylor series to degree 23.
KDN_D double d_atan01(double x) {
    double y = x;
    for (int k = 0; k < 2; ++k) y = y / (1.0 + __builtin_sqrt(1.0 + y * y));
    const double y2 = y * y;
    double s = -1.0 / 23.0;
    s = s * y2 + 1.0 / 21.0;
    s = s * y2 - 1.0 / 19.0;
    s = s * y2 + 1.0 / 17.0;
    s = s * y2 - 1.0 / 15.0;
    s = s * y2 + 1.0 / 13.0;
    s = s * y2 - 1.0 / 11.0;
    s = s * y2 + 1.0 / 9.0;
    s = s * y2 - 1.0 / 7.0;
    s = s * y2 + 1.0 / 5.0;
    s = s * y2 - 1.0 / 3.0;
    return 4.0 * (y + y * y2 * s);
}
KDN_D double d_atan2(double y, double x) {
    if (y == 0 && x == 0) return 0.0;
    const double ay = y < 0 ? -y : y, ax = x < 0 ? -x : x;
    double a = ay <= ax ? d_atan01(ay / ax) : kPiD / 2 - d_atan01(ax / ay);
    if (x < 0) a = kPiD - a;
    return y < 0 ? -a : a;
}
// log by frexp-style reduction to [sqrt(1/2), sqrt(2)) and the atanh series
// of s = (m-1)/(m+1); exp by k ln2 reduction and Taylor to degree 20.
KDN_D double d_log(double x) {
    if (!(x > 0)) return x == 0 ? -__builtin_huge_val() : __builtin_nan("");
    if (x == __builtin_huge_val()) return x;
    int e = 0;
    double m = __builtin_frexp(x, &e);  // m in [0.5, 1)
    if (m < 0.70710678118654752440) { m *= 2; e -= 1; }
    const double s = (m - 1) / (m + 1), s2 = s * s;
    double t = 1.0 / 25.0;
    for (int k = 23; k >= 1; k -= 2) t = t * s2 + 1.0 / k;
    return 2 * s * t + e * 0.69314718055994530942;
}
KDN_D double d_exp(double x) {
    if (x > 709.0) return __builtin_huge_val();
    if (x < -745.0) return 0.0;
    const double k = __builtin_rint(x * 1.44269504088896340736);
    const double r = (x - k * 0.693147180369123816490) - k * 1.90821492927058770002e-10;
    double h = 1.0;
    for (int n = 20; n >= 1; --n) h = 1.0 + h * r / n;
    return __builtin_ldexp(h, (int)k);
}
KDN_D double d_pow(double a, double b) {
    if (b == 0) return 1.0;
    if (a == 0) return b > 0 ? 0.0 : __builtin_huge_val();
    if (a < 0) {
        if (b != __builtin_floor(b)) return __builtin_nan("");
        const double m = d_exp(b * d_log(-a));
        const double half = b * 0.5;
        return half == __builtin_floor(half) ? m : -m;
    }
    return d_exp(b * d_log(a));
}

// float from double, to nearest / toward -inf / toward +inf
KDN_D float f_rn(double v) { return (float)v; }
KDN_D float f_rd(double v) {
    const float f = (float)v;
    return ((double)f > v) ? down(f) : f;
}
KDN_D float f_ru(double v) {
    const float f = (float)v;
    return ((double)f < v) ? up(f) : f;
}
KDN_D double slack(double v) { return 1e-13 * (v < 0 ? -v : v) + 1e-300; }

// ---- points -----------------------------------------------------------------
KDN_D float p_sign(float a) { return a > 0 ? 1.0f : (a < 0 ? -1.0f : 0.0f); }
KDN_D float p_smin(float a, float b, float r) {
    const float blend = fmax_(r - __builtin_fabsf(a - b), 0.0f) / r;
    return fmin_(a, b) - blend * blend * r * 0.25f;
}
KDN_D float p_pos2(float a) {
    const float p = fmax_(a, 0.0f);
    return p * p;
}
KDN_D float p_truth(bool b) { return b ? 1.0f : 0.0f; }

// ---- intervals -----------------------------------------------------------------
struct Iv {
    float lo, hi;
};

KDN_D Iv iv(float lo, float hi) { return Iv{lo, hi}; }
KDN_D bool i_point(Iv a) { return a.lo == a.hi; }
KDN_D bool i_is(Iv a, float v) { return a.lo == v && a.hi == v; }
KDN_D bool i_zero_in(Iv a) { return a.lo <= 0 && a.hi >= 0; }
KDN_D bool i_excludes_zero(Iv a) { return a.lo > 0 || a.hi < 0; }
KDN_D bool i_is_zero(Iv a) { return a.lo == 0 && a.hi == 0; }

KDN_D Iv i_const(const kdn_step& s) {
    if (s.flags == 1) return iv(s.value, up(s.value));
    if (s.flags == 2) return iv(down(s.value), s.value);
    return iv(s.value, s.value);
}
KDN_D Iv i_add(Iv a, Iv b) { return iv(add_rd(a.lo, b.lo), add_ru(a.hi, b.hi)); }
KDN_D Iv i_sub(Iv a, Iv b) { return iv(sub_rd(a.lo, b.hi), sub_ru(a.hi, b.lo)); }
KDN_D Iv i_neg(Iv a) { return iv(-a.hi, -a.lo); }
KDN_D Iv i_mul(Iv a, Iv b) {
    const float l = fmin_(fmin_(fmin_(mul_rd(a.lo, b.lo), mul_rd(a.lo, b.hi)), mul_rd(a.hi, b.lo)), mul_rd(a.hi, b.hi));
    const float h = fmax_(fmax_(fmax_(mul_ru(a.lo, b.lo), mul_ru(a.lo, b.hi)), mul_ru(a.hi, b.lo)), mul_ru(a.hi, b.hi));
    return iv(l, h);
}
KDN_D Iv i_div(Iv a, Iv b) {
    if (i_zero_in(b)) return iv(-kInf, kInf);
    const float l = fmin_(fmin_(fmin_(div_rd(a.lo, b.lo), div_rd(a.lo, b.hi)), div_rd(a.hi, b.lo)), div_rd(a.hi, b.hi));
    const float h = fmax_(fmax_(fmax_(div_ru(a.lo, b.lo), div_ru(a.lo, b.hi)), div_ru(a.hi, b.lo)), div_ru(a.hi, b.hi));
    return iv(l, h);
}
KDN_D Iv i_sqrt(Iv a) { return iv(sqrt_rd(fmax_(a.lo, 0.0f)), sqrt_ru(fmax_(a.hi, 0.0f))); }
KDN_D Iv i_abs(Iv a) {
    if (a.lo >= 0) return a;
    if (a.hi <= 0) return i_neg(a);
    return iv(0.0f, fmax_(-a.lo, a.hi));
}
KDN_D Iv i_floor(Iv a) { return iv(__builtin_floorf(a.lo), __builtin_floorf(a.hi)); }
KDN_D Iv i_sign(Iv a) { return iv(p_sign(a.lo), p_sign(a.hi)); }
// Whether [lo, hi] reaches offset + 2 pi k for some whole k, decided in
// double; the slack only ever turns a no into a yes.
KDN_D bool reaches_d(double lo, double hi, double offset) {
    const double twoPi = 2 * kPiD;
    const double k = __builtin_ceil((lo - offset) / twoPi - 1e-9);
    return offset + twoPi * k <= hi + 1e-9 * (1 + __builtin_fabs(hi));
}
// The sine over [lo, hi] (numbers.cpp:289-300) with its extremes.
KDN_D Iv i_sin_d(double lo, double hi) {
    if (hi - lo >= 2 * kPiD - 1e-9 || __builtin_fabs(lo) > 1e6 || __builtin_fabs(hi) > 1e6) return iv(-1.0f, 1.0f);
    const double sl = d_sin(lo), sh = d_sin(hi);
    const double mn = sl < sh ? sl : sh, mx = sl < sh ? sh : sl;
    float l = f_rd(mn - slack(mn) - 1e-15), h = f_ru(mx + slack(mx) + 1e-15);
    if (reaches_d(lo, hi, kPiD / 2)) h = 1.0f;
    if (reaches_d(lo, hi, -kPiD / 2)) l = -1.0f;
    return iv(fmax_(l, -1.0f), fmin_(h, 1.0f));
}
KDN_D Iv i_sin(Iv a) { return i_sin_d((double)a.lo, (double)a.hi); }
// cos a = sin(a + pi/2) (numbers.cpp:302-307), the shift made in double;
// the sine's slack covers its rounding.
KDN_D Iv i_cos(Iv a) { return i_sin_d((double)a.lo + kPiD / 2, (double)a.hi + kPiD / 2); }
KDN_D Iv i_atan2(Iv y, Iv x) {
    const bool nearOrigin = i_zero_in(x) && i_zero_in(y);
    const bool acrossCut = x.lo < 0 && i_zero_in(y);
    if (nearOrigin || acrossCut) return iv(-3.14159274f, 3.14159274f);  // [-up(pi), up(pi)]
    double v[4] = {d_atan2(y.lo, x.lo), d_atan2(y.lo, x.hi), d_atan2(y.hi, x.lo), d_atan2(y.hi, x.hi)};
    double mn = v[0], mx = v[0];
    for (int k = 1; k < 4; ++k) { mn = v[k] < mn ? v[k] : mn; mx = mx < v[k] ? v[k] : mx; }
    return iv(f_rd(mn - slack(mn) - 1e-15), f_ru(mx + slack(mx) + 1e-15));
}
KDN_D Iv i_pow_whole(Iv a, double n);
KDN_D Iv i_pow(Iv a, Iv b) {
    if (b.lo == b.hi && b.lo == __builtin_floorf(b.lo)) return i_pow_whole(a, (double)b.lo);
    if (a.lo <= 0) return iv(-kInf, kInf);
    double v[4] = {d_pow(a.lo, b.lo), d_pow(a.lo, b.hi), d_pow(a.hi, b.lo), d_pow(a.hi, b.hi)};
    double mn = v[0], mx = v[0];
    for (int k = 1; k < 4; ++k) { mn = v[k] < mn ? v[k] : mn; mx = mx < v[k] ? v[k] : mx; }
    return iv(f_rd(mn - slack(mn)), f_ru(mx + slack(mx)));
}
KDN_D Iv i_pow_whole(Iv a, double n) {
    // whole powers as the reference wholePower (numbers.cpp:340-356)
    Iv base = a;
    double m = n < 0 ? -n : n;
    const double half = m * 0.5;
    const bool even = half == __builtin_floor(half);
    if (even) base = i_abs(a);
    Iv r;
    if (m == 0) {
        r = iv(1.0f, 1.0f);
    } else {
        const double pl = d_pow(base.lo, m), ph = d_pow(base.hi, m);
        r = iv(f_rd(pl - slack(pl)), f_ru(ph + slack(ph)));
    }
    if (n < 0) return i_div(iv(1.0f, 1.0f), r);
    return r;
}
KDN_D Iv i_min(Iv a, Iv b) { return iv(fmin_(a.lo, b.lo), fmin_(a.hi, b.hi)); }
KDN_D Iv i_max(Iv a, Iv b) { return iv(fmax_(a.lo, b.lo), fmax_(a.hi, b.hi)); }
KDN_D Iv i_decided(bool yes, bool no) { return iv(yes ? 1.0f : 0.0f, no ? 0.0f : 1.0f); }
KDN_D Iv i_compare(uint8_t op, Iv a, Iv b) {
    const bool apart = a.hi < b.lo || b.hi < a.lo;
    const bool same = a.lo == a.hi && b.lo == b.hi && a.lo == b.lo;
    switch (op) {
        case KDN_LESS: return i_decided(a.hi < b.lo, a.lo >= b.hi);
        case KDN_LESS_EQUAL: return i_decided(a.hi <= b.lo, a.lo > b.hi);
        case KDN_EQUAL: return i_decided(same, apart);
        case KDN_NOT_EQUAL: return i_decided(apart, same);
        case KDN_AND: return i_decided(i_excludes_zero(a) && i_excludes_zero(b), i_is_zero(a) || i_is_zero(b));
        case KDN_OR: return i_decided(i_excludes_zero(a) || i_excludes_zero(b), i_is_zero(a) && i_is_zero(b));
        default: {  // Not: a == 0
            const Iv z = iv(0.0f, 0.0f);
            const bool ap = a.hi < z.lo || z.hi < a.lo;
            const bool sm = a.lo == a.hi && a.lo == 0;
            return i_decided(sm, ap);
        }
    }
}
KDN_D Iv i_choose(Iv c, Iv a, Iv b) {
    if (i_excludes_zero(c)) return a;
    if (i_is_zero(c)) return b;
    return iv(fmin_(a.lo, b.lo), fmax_(a.hi, b.hi));
}
// Bounds on |a - b| for floats a, b.
KDN_D void absdiff_bounds(float a, float b, float* lo, float* hi) {
    const float x = sub_rd(a, b), y = sub_ru(a, b);
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    *lo = (x <= 0 && y >= 0) ? 0.0f : fmin_(ax, ay);
    *hi = fmax_(ax, ay);
}
// The blend term max(r - |a-b|, 0)^2 / r^2 * r / 4 bounded above and below.
KDN_D float blend_term_ru(float a, float b, float r) {
    float dl, dh;
    absdiff_bounds(a, b, &dl, &dh);
    const float g = fmax_(sub_ru(r, dl), 0.0f);
    const float q = div_ru(g, r);
    return mul_ru(mul_ru(mul_ru(q, q), r), 0.25f);
}
KDN_D float blend_term_rd(float a, float b, float r) {
    float dl, dh;
    absdiff_bounds(a, b, &dl, &dh);
    const float g = fmax_(sub_rd(r, dh), 0.0f);
    const float q = div_rd(g, r);
    return mul_rd(mul_rd(mul_rd(q, q), r), 0.25f);
}
// The smooth minimum at the ends with r's lower end, as the reference
// (numbers.cpp:409-417): smin rises in a and in b, so its lower end is at
// the lower ends and its upper end at the upper ends; each end is bounded
// outward. A radius that is not positive gives the reference's fallback.
KDN_D Iv i_smin(Iv a, Iv b, Iv r) {
    const float rl = r.lo;
    if (!(rl > 0)) return iv(fmin_(a.lo, b.lo), fmin_(a.hi, b.hi));
    float lower = sub_rd(fmin_(a.lo, b.lo), blend_term_ru(a.lo, b.lo, rl));
    float upper = sub_ru(fmin_(a.hi, b.hi), blend_term_rd(a.hi, b.hi, rl));
    if (isnanf_(lower)) lower = sub_rd(fmin_(a.lo, b.lo), div_ru(rl, 4.0f));
    if (isnanf_(upper)) upper = fmin_(a.hi, b.hi);
    return iv(lower, upper);
}
KDN_D Iv i_pos2(Iv a) {
    const float l = fmax_(a.lo, 0.0f), h = fmax_(a.hi, 0.0f);
    return iv(mul_rd(l, l), mul_ru(h, h));
}

// ---- duals (value and gradient, FP32) ------------------------------------------------
struct Dl {
    float v, g[3];
};
KDN_D Dl dl(float v) { return Dl{v, {0, 0, 0}}; }
KDN_D Dl chain1(const Dl& a, float f, float d1) { return Dl{f, {d1 * a.g[0], d1 * a.g[1], d1 * a.g[2]}}; }
KDN_D Dl chain2(const Dl& a, const Dl& b, float f, float fa, float fb) {
    return Dl{f, {fa * a.g[0] + fb * b.g[0], fa * a.g[1] + fb * b.g[1], fa * a.g[2] + fb * b.g[2]}};
}

}  // namespace kdn
