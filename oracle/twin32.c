/*
 * FP32 twin of the GPU octree build and extend kernels — TEST
 * INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The arithmetic restates the FP32 semantics of the device path:
 *  - points: IEEE single precision, one rounding per operation, no
 *    contraction; sqrt and division correctly rounded;
 *  - intervals: the reference Interval operations (source/sdf/numbers.cpp:
 *    198-430) with every endpoint rounded outward exactly, the direction
 *    read from the sign of the fma residual (or of the TwoSum error);
 *  - transcendentals: double polynomials (Cody-Waite sin/cos, halved-argument
 *    atan series, atanh-series log, Taylor exp) rounded to float.
 * The cut rewrite and the octree growth restate plugins/geometry/tree/cut.cpp
 * and grow.cpp/build.cpp with the split rule "Surface and level < depth".
 * Build with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    float lo, hi;
} Iv;

static const float kFltMin = 1.17549435e-38f;
#define KINF ((float)INFINITY)
static const double kPiD = 3.14159265358979323846;

static uint32_t fbits(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    return b;
}
static float bitsf(uint32_t b) {
    float x;
    memcpy(&x, &b, 4);
    return x;
}
static int isnanf_(float x) { return x != x; }
static float fmin_(float a, float b) { return b < a ? b : a; }
static float fmax_(float a, float b) { return a < b ? b : a; }

static float up(float x) {
    if (isnanf_(x) || x == KINF) return x;
    if (x < kFltMin && x >= -kFltMin) return kFltMin;
    const uint32_t b = fbits(x);
    return bitsf(x > 0 ? b + 1 : b - 1);
}
static float down(float x) { return -up(-x); }
static int tiny(float p) { return !(p != 0 && (p > 1e-30f || p < -1e-30f)) || p == KINF || p == -KINF || isnanf_(p); }

static float add_rd(float a, float b) {
    const float s = a + b;
    if (tiny(s) || a == KINF || a == -KINF || b == KINF || b == -KINF) {
        if (a == -KINF || b == -KINF) return s;
        if (a == KINF || b == KINF) return s;
        return s == 0 && (a == 0 || b == 0) ? s : down(s);
    }
    const float bv = s - a, av = s - bv;
    const float err = (a - av) + (b - bv);
    return err < 0 ? down(s) : s;
}
static float add_ru(float a, float b) { return -add_rd(-a, -b); }
static float sub_rd(float a, float b) { return add_rd(a, -b); }
static float sub_ru(float a, float b) { return add_ru(a, -b); }
static float mul_rd(float a, float b) {
    if (a == 0 || b == 0) return 0.0f;
    const float p = a * b;
    if (tiny(p)) {
        if (a == KINF || a == -KINF || b == KINF || b == -KINF) return p;
        return down(p);
    }
    const float r = fmaf(a, b, -p);
    return r < 0 ? down(p) : p;
}
static float mul_ru(float a, float b) { return -mul_rd(-a, b); }
static float div_rd(float a, float b) {
    if (a == 0) return 0.0f;
    const float q = a / b;
    if (tiny(q) || b == KINF || b == -KINF || a == KINF || a == -KINF) {
        if (q == KINF || q == -KINF) return (a == KINF || a == -KINF) ? q : down(q);
        return down(q);
    }
    const float r = fmaf(q, b, -a);
    const int qAbove = (r > 0) == (b > 0);
    return (r != 0 && qAbove) ? down(q) : q;
}
static float div_ru(float a, float b) { return -div_rd(-a, b); }
static float sqrt_rd(float a) {
    if (a <= 0) return 0.0f;
    const float s = sqrtf(a);
    if (s == KINF) return s;
    const float r = fmaf(s, s, -a);
    return r > 0 ? down(s) : s;
}
static float sqrt_ru(float a) {
    if (a <= 0) return 0.0f;
    const float s = sqrtf(a);
    if (s == KINF) return s;
    const float r = fmaf(s, s, -a);
    return r < 0 ? up(s) : s;
}

/* ---- double kernels ---- */
static double poly_sin(double r) {
    const double r2 = r * r;
    double s = -1.0 / 355687428096000.0;
    s = s * r2 + 1.0 / 1307674368000.0;
    s = s * r2 - 1.0 / 6227020800.0;
    s = s * r2 + 1.0 / 39916800.0;
    s = s * r2 - 1.0 / 362880.0;
    s = s * r2 + 1.0 / 5040.0;
    s = s * r2 - 1.0 / 120.0;
    s = s * r2 + 1.0 / 6.0;
    return r - r * r2 * s;
}
static double poly_cos(double r) {
    const double r2 = r * r;
    double c = 1.0 / 20922789888000.0;
    c = c * r2 - 1.0 / 87178291200.0;
    c = c * r2 + 1.0 / 479001600.0;
    c = c * r2 - 1.0 / 3628800.0;
    c = c * r2 + 1.0 / 40320.0;
    c = c * r2 - 1.0 / 720.0;
    c = c * r2 + 1.0 / 24.0;
    c = c * r2 - 0.5;
    return 1.0 + r2 * c;
}
static double reduced(double x, int* q) {
    const double k = rint(x * 0.63661977236758134308);
    const double p1 = 1.57079632673412561417e+00;
    const double p2 = 6.07710050650619224932e-11;
    const double r = (x - k * p1) - k * p2;
    const long long ki = (long long)k;
    *q = (int)(ki & 3);
    return r;
}
static double d_sin(double x) {
    int q;
    const double r = reduced(x, &q);
    const double v = (q & 1) ? poly_cos(r) : poly_sin(r);
    return (q & 2) ? -v : v;
}
static double d_cos(double x) {
    int q;
    const double r = reduced(x, &q);
    const double v = (q & 1) ? poly_sin(r) : poly_cos(r);
    return ((q + 1) & 2) ? -v : v;
}
static double d_atan01(double x) {
    double y = x;
    for (int k = 0; k < 2; ++k) y = y / (1.0 + sqrt(1.0 + y * y));
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
static double d_atan2(double y, double x) {
    if (y == 0 && x == 0) return 0.0;
    const double ay = y < 0 ? -y : y, ax = x < 0 ? -x : x;
    double a = ay <= ax ? d_atan01(ay / ax) : kPiD / 2 - d_atan01(ax / ay);
    if (x < 0) a = kPiD - a;
    return y < 0 ? -a : a;
}
static double d_log(double x) {
    if (!(x > 0)) return x == 0 ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    int e = 0;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m *= 2;
        e -= 1;
    }
    const double s = (m - 1) / (m + 1), s2 = s * s;
    double t = 1.0 / 25.0;
    for (int k = 23; k >= 1; k -= 2) t = t * s2 + 1.0 / k;
    return 2 * s * t + e * 0.69314718055994530942;
}
static double d_exp(double x) {
    if (x > 709.0) return INFINITY;
    if (x < -745.0) return 0.0;
    const double k = rint(x * 1.44269504088896340736);
    const double r = (x - k * 0.693147180369123816490) - k * 1.90821492927058770002e-10;
    double h = 1.0;
    for (int n = 20; n >= 1; --n) h = 1.0 + h * r / n;
    return ldexp(h, (int)k);
}
static double d_pow(double a, double b) {
    if (b == 0) return 1.0;
    if (a == 0) return b > 0 ? 0.0 : INFINITY;
    if (a < 0) {
        if (b != floor(b)) return NAN;
        const double m = d_exp(b * d_log(-a));
        const double half = b * 0.5;
        return half == floor(half) ? m : -m;
    }
    return d_exp(b * d_log(a));
}
static float f_rn(double v) { return (float)v; }
static float f_rd(double v) {
    const float f = (float)v;
    return ((double)f > v) ? down(f) : f;
}
static float f_ru(double v) {
    const float f = (float)v;
    return ((double)f < v) ? up(f) : f;
}
static double slack(double v) { return 1e-13 * (v < 0 ? -v : v) + 1e-300; }

/* ---- points ---- */
static float p_sign(float a) { return a > 0 ? 1.0f : (a < 0 ? -1.0f : 0.0f); }
static float p_smin(float a, float b, float r) {
    const float blend = fmax_(r - fabsf(a - b), 0.0f) / r;
    return fmin_(a, b) - blend * blend * r * 0.25f;
}
static float p_pos2(float a) {
    const float p = fmax_(a, 0.0f);
    return p * p;
}
static float p_truth(int b) { return b ? 1.0f : 0.0f; }

/* ---- intervals ---- */
static Iv iv(float lo, float hi) {
    Iv r = {lo, hi};
    return r;
}
static int i_point(Iv a) { return a.lo == a.hi; }
static int i_is(Iv a, float v) { return a.lo == v && a.hi == v; }
static int i_zero_in(Iv a) { return a.lo <= 0 && a.hi >= 0; }
static int i_excludes_zero(Iv a) { return a.lo > 0 || a.hi < 0; }
static int i_is_zero(Iv a) { return a.lo == 0 && a.hi == 0; }
static Iv i_const(const kdn_step* s) {
    if (s->flags == 1) return iv(s->value, up(s->value));
    if (s->flags == 2) return iv(down(s->value), s->value);
    return iv(s->value, s->value);
}
static Iv i_add(Iv a, Iv b) { return iv(add_rd(a.lo, b.lo), add_ru(a.hi, b.hi)); }
static Iv i_sub(Iv a, Iv b) { return iv(sub_rd(a.lo, b.hi), sub_ru(a.hi, b.lo)); }
static Iv i_neg(Iv a) { return iv(-a.hi, -a.lo); }
static Iv i_mul(Iv a, Iv b) {
    const float l = fmin_(fmin_(fmin_(mul_rd(a.lo, b.lo), mul_rd(a.lo, b.hi)), mul_rd(a.hi, b.lo)), mul_rd(a.hi, b.hi));
    const float h = fmax_(fmax_(fmax_(mul_ru(a.lo, b.lo), mul_ru(a.lo, b.hi)), mul_ru(a.hi, b.lo)), mul_ru(a.hi, b.hi));
    return iv(l, h);
}
static Iv i_div(Iv a, Iv b) {
    if (i_zero_in(b)) return iv(-KINF, KINF);
    const float l = fmin_(fmin_(fmin_(div_rd(a.lo, b.lo), div_rd(a.lo, b.hi)), div_rd(a.hi, b.lo)), div_rd(a.hi, b.hi));
    const float h = fmax_(fmax_(fmax_(div_ru(a.lo, b.lo), div_ru(a.lo, b.hi)), div_ru(a.hi, b.lo)), div_ru(a.hi, b.hi));
    return iv(l, h);
}
static Iv i_sqrt(Iv a) { return iv(sqrt_rd(fmax_(a.lo, 0.0f)), sqrt_ru(fmax_(a.hi, 0.0f))); }
static Iv i_abs(Iv a) {
    if (a.lo >= 0) return a;
    if (a.hi <= 0) return i_neg(a);
    return iv(0.0f, fmax_(-a.lo, a.hi));
}
static Iv i_floor(Iv a) { return iv(floorf(a.lo), floorf(a.hi)); }
static Iv i_sign(Iv a) { return iv(p_sign(a.lo), p_sign(a.hi)); }
static int reaches_d(double lo, double hi, double offset) {
    const double twoPi = 2 * kPiD;
    const double k = ceil((lo - offset) / twoPi - 1e-9);
    return offset + twoPi * k <= hi + 1e-9 * (1 + fabs(hi));
}
static Iv i_sin_d(double lo, double hi) {
    if (hi - lo >= 2 * kPiD - 1e-9 || fabs(lo) > 1e6 || fabs(hi) > 1e6) return iv(-1.0f, 1.0f);
    const double sl = d_sin(lo), sh = d_sin(hi);
    const double mn = sl < sh ? sl : sh, mx = sl < sh ? sh : sl;
    float l = f_rd(mn - slack(mn) - 1e-15), h = f_ru(mx + slack(mx) + 1e-15);
    if (reaches_d(lo, hi, kPiD / 2)) h = 1.0f;
    if (reaches_d(lo, hi, -kPiD / 2)) l = -1.0f;
    return iv(fmax_(l, -1.0f), fmin_(h, 1.0f));
}
static Iv i_sin(Iv a) { return i_sin_d((double)a.lo, (double)a.hi); }
static Iv i_cos(Iv a) { return i_sin_d((double)a.lo + kPiD / 2, (double)a.hi + kPiD / 2); }
static Iv i_atan2(Iv y, Iv x) {
    const int nearOrigin = i_zero_in(x) && i_zero_in(y);
    const int acrossCut = x.lo < 0 && i_zero_in(y);
    if (nearOrigin || acrossCut) return iv(-3.14159274f, 3.14159274f);
    double v[4] = {d_atan2(y.lo, x.lo), d_atan2(y.lo, x.hi), d_atan2(y.hi, x.lo), d_atan2(y.hi, x.hi)};
    double mn = v[0], mx = v[0];
    for (int k = 1; k < 4; ++k) {
        mn = v[k] < mn ? v[k] : mn;
        mx = mx < v[k] ? v[k] : mx;
    }
    return iv(f_rd(mn - slack(mn) - 1e-15), f_ru(mx + slack(mx) + 1e-15));
}
static Iv i_pow_whole(Iv a, double n) {
    Iv base = a;
    double m = n < 0 ? -n : n;
    const double half = m * 0.5;
    const int even = half == floor(half);
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
static Iv i_pow(Iv a, Iv b) {
    if (b.lo == b.hi && b.lo == floorf(b.lo)) return i_pow_whole(a, (double)b.lo);
    if (a.lo <= 0) return iv(-KINF, KINF);
    double v[4] = {d_pow(a.lo, b.lo), d_pow(a.lo, b.hi), d_pow(a.hi, b.lo), d_pow(a.hi, b.hi)};
    double mn = v[0], mx = v[0];
    for (int k = 1; k < 4; ++k) {
        mn = v[k] < mn ? v[k] : mn;
        mx = mx < v[k] ? v[k] : mx;
    }
    return iv(f_rd(mn - slack(mn)), f_ru(mx + slack(mx)));
}
static Iv i_min(Iv a, Iv b) { return iv(fmin_(a.lo, b.lo), fmin_(a.hi, b.hi)); }
static Iv i_max(Iv a, Iv b) { return iv(fmax_(a.lo, b.lo), fmax_(a.hi, b.hi)); }
static Iv i_decided(int yes, int no) { return iv(yes ? 1.0f : 0.0f, no ? 0.0f : 1.0f); }
static Iv i_compare(uint8_t op, Iv a, Iv b) {
    const int apart = a.hi < b.lo || b.hi < a.lo;
    const int same = a.lo == a.hi && b.lo == b.hi && a.lo == b.lo;
    switch (op) {
        case KDN_LESS: return i_decided(a.hi < b.lo, a.lo >= b.hi);
        case KDN_LESS_EQUAL: return i_decided(a.hi <= b.lo, a.lo > b.hi);
        case KDN_EQUAL: return i_decided(same, apart);
        case KDN_NOT_EQUAL: return i_decided(apart, same);
        case KDN_AND: return i_decided(i_excludes_zero(a) && i_excludes_zero(b), i_is_zero(a) || i_is_zero(b));
        case KDN_OR: return i_decided(i_excludes_zero(a) || i_excludes_zero(b), i_is_zero(a) && i_is_zero(b));
        default: {
            const int ap = a.hi < 0 || 0 < a.lo;
            const int sm = a.lo == a.hi && a.lo == 0;
            return i_decided(sm, ap);
        }
    }
}
static Iv i_choose(Iv c, Iv a, Iv b) {
    if (i_excludes_zero(c)) return a;
    if (i_is_zero(c)) return b;
    return iv(fmin_(a.lo, b.lo), fmax_(a.hi, b.hi));
}
static void absdiff_bounds(float a, float b, float* lo, float* hi) {
    const float x = sub_rd(a, b), y = sub_ru(a, b);
    const float ax = fabsf(x), ay = fabsf(y);
    *lo = (x <= 0 && y >= 0) ? 0.0f : fmin_(ax, ay);
    *hi = fmax_(ax, ay);
}
static float blend_term_ru(float a, float b, float r) {
    float dl, dh;
    absdiff_bounds(a, b, &dl, &dh);
    const float g = fmax_(sub_ru(r, dl), 0.0f);
    const float q = div_ru(g, r);
    return mul_ru(mul_ru(mul_ru(q, q), r), 0.25f);
}
static float blend_term_rd(float a, float b, float r) {
    float dl, dh;
    absdiff_bounds(a, b, &dl, &dh);
    const float g = fmax_(sub_rd(r, dh), 0.0f);
    const float q = div_rd(g, r);
    return mul_rd(mul_rd(mul_rd(q, q), r), 0.25f);
}
static Iv i_smin(Iv a, Iv b, Iv r) {
    const float rl = r.lo;
    if (!(rl > 0)) return iv(fmin_(a.lo, b.lo), fmin_(a.hi, b.hi));
    float lower = sub_rd(fmin_(a.lo, b.lo), blend_term_ru(a.lo, b.lo, rl));
    float upper = sub_ru(fmin_(a.hi, b.hi), blend_term_rd(a.hi, b.hi, rl));
    if (isnanf_(lower)) lower = sub_rd(fmin_(a.lo, b.lo), div_ru(rl, 4.0f));
    if (isnanf_(upper)) upper = fmin_(a.hi, b.hi);
    return iv(lower, upper);
}
static Iv i_pos2(Iv a) {
    const float l = fmax_(a.lo, 0.0f), h = fmax_(a.hi, 0.0f);
    return iv(mul_rd(l, l), mul_ru(h, h));
}

/* The interval of one step, registers strided by `stride`. */
static Iv interval_step(const kdn_step* s, const Iv* r, const Iv box[3]) {
    switch (s->op) {
        case KDN_CONSTANT:
        case KDN_ARGUMENT: return i_const(s);
        case KDN_POINT_X: return box[0];
        case KDN_POINT_Y: return box[1];
        case KDN_POINT_Z: return box[2];
        case KDN_NORMAL_X:
        case KDN_NORMAL_Y:
        case KDN_NORMAL_Z: return iv(0.0f, 0.0f);
        case KDN_ADD: return i_add(r[s->a], r[s->b]);
        case KDN_SUBTRACT: return i_sub(r[s->a], r[s->b]);
        case KDN_MULTIPLY: return i_mul(r[s->a], r[s->b]);
        case KDN_DIVIDE: return i_div(r[s->a], r[s->b]);
        case KDN_NEGATE: return i_neg(r[s->a]);
        case KDN_SQUARE_ROOT: return i_sqrt(r[s->a]);
        case KDN_ABSOLUTE: return i_abs(r[s->a]);
        case KDN_FLOOR: return i_floor(r[s->a]);
        case KDN_SINE: return i_sin(r[s->a]);
        case KDN_COSINE: return i_cos(r[s->a]);
        case KDN_SIGN: return i_sign(r[s->a]);
        case KDN_ARC_TANGENT: return i_atan2(r[s->a], r[s->b]);
        case KDN_POWER: return i_pow(r[s->a], r[s->b]);
        case KDN_MINIMUM: return i_min(r[s->a], r[s->b]);
        case KDN_MAXIMUM: return i_max(r[s->a], r[s->b]);
        case KDN_SMOOTH_MINIMUM: return i_smin(r[s->a], r[s->b], r[s->c]);
        case KDN_POSITIVE_SQUARE: return i_pos2(r[s->a]);
        case KDN_SELECT: return i_choose(r[s->a], r[s->b], r[s->c]);
        default: return i_compare(s->op, r[s->a], r[s->b]);
    }
}

static float p_compare(uint8_t op, float a, float b) {
    switch (op) {
        case KDN_LESS: return p_truth(a < b);
        case KDN_LESS_EQUAL: return p_truth(a <= b);
        case KDN_EQUAL: return p_truth(a == b);
        case KDN_NOT_EQUAL: return p_truth(a != b);
        case KDN_AND: return p_truth(a != 0 && b != 0);
        case KDN_OR: return p_truth(a != 0 || b != 0);
        default: return p_truth(a == 0);
    }
}

static float point_step(const kdn_step* s, const float* r, float px, float py, float pz) {
    switch (s->op) {
        case KDN_CONSTANT:
        case KDN_ARGUMENT: return s->value;
        case KDN_POINT_X: return px;
        case KDN_POINT_Y: return py;
        case KDN_POINT_Z: return pz;
        case KDN_NORMAL_X:
        case KDN_NORMAL_Y:
        case KDN_NORMAL_Z: return 0.0f;
        case KDN_ADD: return r[s->a] + r[s->b];
        case KDN_SUBTRACT: return r[s->a] - r[s->b];
        case KDN_MULTIPLY: return r[s->a] * r[s->b];
        case KDN_DIVIDE: return r[s->a] / r[s->b];
        case KDN_NEGATE: return -r[s->a];
        case KDN_SQUARE_ROOT: return sqrtf(r[s->a]);
        case KDN_ABSOLUTE: return fabsf(r[s->a]);
        case KDN_FLOOR: return floorf(r[s->a]);
        case KDN_SINE: return f_rn(d_sin((double)r[s->a]));
        case KDN_COSINE: return f_rn(d_cos((double)r[s->a]));
        case KDN_SIGN: return p_sign(r[s->a]);
        case KDN_ARC_TANGENT: return f_rn(d_atan2((double)r[s->a], (double)r[s->b]));
        case KDN_POWER: return f_rn(d_pow((double)r[s->a], (double)r[s->b]));
        case KDN_MINIMUM: return fmin_(r[s->a], r[s->b]);
        case KDN_MAXIMUM: return fmax_(r[s->a], r[s->b]);
        case KDN_SMOOTH_MINIMUM: return p_smin(r[s->a], r[s->b], r[s->c]);
        case KDN_POSITIVE_SQUARE: return p_pos2(r[s->a]);
        case KDN_SELECT: return r[s->a] != 0 ? r[s->b] : r[s->c];
        default: return p_compare(s->op, r[s->a], r[s->b]);
    }
}

/* duals */
typedef struct {
    float v, g[3];
} Dl;
static Dl dl(float v) {
    Dl d = {v, {0, 0, 0}};
    return d;
}
static Dl chain1(Dl a, float f, float d1) {
    Dl o = {f, {d1 * a.g[0], d1 * a.g[1], d1 * a.g[2]}};
    return o;
}
static Dl chain2(Dl a, Dl b, float f, float fa, float fb) {
    Dl o = {f, {fa * a.g[0] + fb * b.g[0], fa * a.g[1] + fb * b.g[1], fa * a.g[2] + fb * b.g[2]}};
    return o;
}
static float blend_slope(float d, float r) {
    const float v = (r - d) / (2 * r);
    return v < 0 ? 0.0f : (v > 1 ? 1.0f : v);
}
static Dl dual_step(const kdn_step* s, const Dl* r, float px, float py, float pz) {
    switch (s->op) {
        case KDN_CONSTANT:
        case KDN_ARGUMENT: return dl(s->value);
        case KDN_POINT_X: { Dl d = {px, {1, 0, 0}}; return d; }
        case KDN_POINT_Y: { Dl d = {py, {0, 1, 0}}; return d; }
        case KDN_POINT_Z: { Dl d = {pz, {0, 0, 1}}; return d; }
        case KDN_NORMAL_X:
        case KDN_NORMAL_Y:
        case KDN_NORMAL_Z: return dl(0.0f);
        case KDN_ADD: return chain2(r[s->a], r[s->b], r[s->a].v + r[s->b].v, 1, 1);
        case KDN_SUBTRACT: return chain2(r[s->a], r[s->b], r[s->a].v - r[s->b].v, 1, -1);
        case KDN_MULTIPLY: return chain2(r[s->a], r[s->b], r[s->a].v * r[s->b].v, r[s->b].v, r[s->a].v);
        case KDN_DIVIDE: {
            const float b = r[s->b].v;
            return chain2(r[s->a], r[s->b], r[s->a].v / b, 1 / b, -r[s->a].v / (b * b));
        }
        case KDN_NEGATE: return chain1(r[s->a], -r[s->a].v, -1);
        case KDN_SQUARE_ROOT: {
            const float f = sqrtf(r[s->a].v);
            return chain1(r[s->a], f, f > 0 ? 0.5f / f : 0.0f);
        }
        case KDN_ABSOLUTE: return r[s->a].v < 0 ? chain1(r[s->a], -r[s->a].v, -1) : r[s->a];
        case KDN_FLOOR: return dl(floorf(r[s->a].v));
        case KDN_SINE: {
            const double x = r[s->a].v;
            return chain1(r[s->a], f_rn(d_sin(x)), f_rn(d_cos(x)));
        }
        case KDN_COSINE: {
            const double x = r[s->a].v;
            return chain1(r[s->a], f_rn(d_cos(x)), -f_rn(d_sin(x)));
        }
        case KDN_SIGN: return dl(p_sign(r[s->a].v));
        case KDN_ARC_TANGENT: {
            const float y = r[s->a].v, x = r[s->b].v, q = x * x + y * y;
            return chain2(r[s->a], r[s->b], f_rn(d_atan2(y, x)), x / q, -y / q);
        }
        case KDN_POWER: {
            const double a = r[s->a].v, b = r[s->b].v;
            const float f = f_rn(d_pow(a, b));
            const float fa = f_rn(b * d_pow(a, b - 1));
            const float fb = a > 0 ? f * f_rn(d_log(a)) : 0.0f;
            return chain2(r[s->a], r[s->b], f, fa, fb);
        }
        case KDN_MINIMUM: return r[s->a].v <= r[s->b].v ? r[s->a] : r[s->b];
        case KDN_MAXIMUM: return r[s->a].v >= r[s->b].v ? r[s->a] : r[s->b];
        case KDN_SMOOTH_MINIMUM: {
            const float fa = blend_slope(r[s->a].v - r[s->b].v, r[s->c].v);
            return chain2(r[s->a], r[s->b], p_smin(r[s->a].v, r[s->b].v, r[s->c].v), fa, 1 - fa);
        }
        case KDN_POSITIVE_SQUARE: return chain1(r[s->a], p_pos2(r[s->a].v), 2 * fmax_(r[s->a].v, 0.0f));
        case KDN_SELECT: return r[s->a].v != 0 ? r[s->b] : r[s->c];
        default: return dl(p_compare(s->op, r[s->a].v, r[s->b].v));
    }
}

/* ---- public evaluators ---- */
void or_interval32(const kdn_step* steps, uint32_t count, const float box6[6], float* lo, float* hi) {
    Iv box[3] = {{box6[0], box6[1]}, {box6[2], box6[3]}, {box6[4], box6[5]}};
    Iv* r = (Iv*)malloc(sizeof(Iv) * (count ? count : 1));
    for (uint32_t i = 0; i < count; ++i) {
        r[i] = interval_step(&steps[i], r, box);
        lo[i] = r[i].lo;
        hi[i] = r[i].hi;
    }
    free(r);
}

float or_eval32(const kdn_step* steps, uint32_t count, uint32_t result, const float p[3], float* regs) {
    for (uint32_t i = 0; i < count; ++i) regs[i] = point_step(&steps[i], regs, p[0], p[1], p[2]);
    return regs[result];
}

/* ---- the cut rewrite (cut.cpp:60-253) ---- */
#define NO_OPERAND 0xFFFFFFFFu
static int operand_count(uint8_t op) {
    switch (op) {
        case KDN_CONSTANT: case KDN_ARGUMENT: case KDN_POINT_X: case KDN_POINT_Y: case KDN_POINT_Z:
        case KDN_NORMAL_X: case KDN_NORMAL_Y: case KDN_NORMAL_Z: return 0;
        case KDN_NEGATE: case KDN_SQUARE_ROOT: case KDN_ABSOLUTE: case KDN_FLOOR: case KDN_SINE:
        case KDN_COSINE: case KDN_SIGN: case KDN_NOT: return 1;
        case KDN_SELECT: case KDN_SMOOTH_MINIMUM: return 3;
        default: return 2;
    }
}
static int is_literal(uint8_t op) { return op == KDN_CONSTANT || op == KDN_ARGUMENT; }
static uint32_t separated(const kdn_step* s, const Iv* v, int lowerWins) {
    const Iv a = v[s->a], b = v[s->b];
    if (a.hi <= b.lo) return lowerWins ? s->a : s->b;
    if (b.hi <= a.lo) return lowerWins ? s->b : s->a;
    return NO_OPERAND;
}
static uint32_t outside_blend(const kdn_step* s, const Iv* v) {
    const float radius = v[s->c].lo;
    const Iv d = i_sub(v[s->a], v[s->b]);
    if (d.hi <= -radius) return s->a;
    if (d.lo >= radius) return s->b;
    return NO_OPERAND;
}
static uint32_t neutral_operand(const kdn_step* s, const Iv* v, float neutral) {
    const int commutes = s->op == KDN_ADD || s->op == KDN_MULTIPLY;
    if (i_is(v[s->b], neutral)) return s->a;
    if (commutes && i_is(v[s->a], neutral)) return s->b;
    return NO_OPERAND;
}
static uint32_t shortcut(const kdn_step* s, const Iv* v) {
    switch (s->op) {
        case KDN_ADD:
        case KDN_SUBTRACT: return neutral_operand(s, v, 0.0f);
        case KDN_MULTIPLY:
        case KDN_DIVIDE: return neutral_operand(s, v, 1.0f);
        case KDN_MINIMUM: return separated(s, v, 1);
        case KDN_MAXIMUM: return separated(s, v, 0);
        case KDN_SMOOTH_MINIMUM: return outside_blend(s, v);
        case KDN_ABSOLUTE: return v[s->a].lo >= 0 ? s->a : NO_OPERAND;
        case KDN_SELECT: {
            const Iv c = v[s->a];
            if (c.lo > 0 || c.hi < 0) return s->b;
            if (c.lo == 0 && c.hi == 0) return s->c;
            return NO_OPERAND;
        }
        default: return NO_OPERAND;
    }
}
static uint32_t negated_operand(const kdn_step* s, const Iv* v) {
    const int product = s->op == KDN_MULTIPLY;
    if (s->op == KDN_ABSOLUTE && v[s->a].hi <= 0) return s->a;
    if (s->op == KDN_SUBTRACT && i_is(v[s->a], 0.0f)) return s->b;
    if (product && i_is(v[s->a], -1.0f)) return s->b;
    if (product && i_is(v[s->b], -1.0f)) return s->a;
    return NO_OPERAND;
}
static kdn_step resolved_step(kdn_step s, const Iv* v, uint32_t index) {
    const Iv self = v[index];
    if (i_point(self) && !is_literal(s.op)) {
        kdn_step k;
        memset(&k, 0, sizeof(k));
        k.op = KDN_CONSTANT;
        k.value = self.lo;
        k.leaf = s.leaf;
        return k;
    }
    if (s.op == KDN_POSITIVE_SQUARE && v[s.a].lo >= 0) {
        s.op = KDN_MULTIPLY;
        s.b = s.a;
        return s;
    }
    const uint32_t n = negated_operand(&s, v);
    if (n != NO_OPERAND) {
        s.op = KDN_NEGATE;
        s.a = (uint16_t)n;
        s.b = 0;
        s.c = 0;
    }
    return s;
}
static kdn_step unsquared(const kdn_step* steps, kdn_step s) {
    if (s.op != KDN_SQUARE_ROOT) return s;
    const kdn_step inner = steps[s.a];
    if (inner.op == KDN_POSITIVE_SQUARE) {
        kdn_step q = inner;
        q.op = KDN_MAXIMUM;
        return q;
    }
    if (inner.op != KDN_MULTIPLY || inner.a != inner.b) return s;
    s.op = KDN_ABSOLUTE;
    s.a = inner.a;
    return s;
}
static int is_zero_step(const kdn_step* steps, uint32_t i) {
    return steps[i].op == KDN_CONSTANT && steps[i].value == 0.0f && steps[i].flags == 0;
}
static kdn_step positive_squared(const kdn_step* steps, kdn_step s) {
    if (s.op != KDN_MULTIPLY || s.a != s.b) return s;
    const kdn_step inner = steps[s.a];
    if (inner.op != KDN_MAXIMUM || !(is_zero_step(steps, inner.a) || is_zero_step(steps, inner.b))) return s;
    const int zeroFirst = is_zero_step(steps, inner.a);
    s.op = KDN_POSITIVE_SQUARE;
    s.a = zeroFirst ? inner.b : inner.a;
    s.b = zeroFirst ? inner.a : inner.b;
    return s;
}
static int same_step(const kdn_step* x, const kdn_step* y) {
    return x->op == y->op && x->flags == y->flags && x->a == y->a && x->b == y->b && x->c == y->c &&
           fbits(x->value) == fbits(y->value) && x->leaf == y->leaf;
}

/* One cut of a program over a box. */
typedef struct {
    float lo, hi;
    int status; /* 0 Empty, 1 Filled, 2 Surface */
    int same;
    uint32_t length, result;
    kdn_step* steps; /* compacted program when !same */
} Cut32;

static Cut32 cut32(const kdn_step* prog, uint32_t L, uint32_t res, const Iv box[3], Iv* IV, kdn_step* ST, uint32_t* AL) {
    Cut32 c;
    memset(&c, 0, sizeof(c));
    for (uint32_t i = 0; i < L; ++i) IV[i] = interval_step(&prog[i], IV, box);
    const Iv range = IV[res];
    c.lo = range.lo;
    c.hi = range.hi;
    c.result = res;
    if (range.lo > 0) {
        c.status = 0;
        return c;
    }
    c.status = range.hi < 0 ? 1 : 2;
    int changed = 0;
    for (uint32_t i = 0; i < L; ++i) {
        const kdn_step orig = prog[i];
        if (is_literal(orig.op)) {
            AL[i] = i;
            kdn_step lit = orig;
            lit.b = 0;
            lit.c = 0;
            ST[i] = lit;
            changed = changed || !same_step(&lit, &orig);
            continue;
        }
        kdn_step s = orig;
        const int count = operand_count(orig.op);
        s.a = count > 0 ? (uint16_t)AL[orig.a] : orig.a;
        s.b = count > 1 ? (uint16_t)AL[orig.b] : 0;
        s.c = count > 2 ? (uint16_t)AL[orig.c] : 0;
        s = positive_squared(ST, unsquared(ST, s));
        const uint32_t other = i_point(IV[i]) ? NO_OPERAND : shortcut(&s, IV);
        AL[i] = other != NO_OPERAND ? other : i;
        const kdn_step resolved = resolved_step(s, IV, i);
        ST[i] = resolved;
        changed = changed || !same_step(&resolved, &orig);
    }
    const uint32_t result = AL[res];
    for (uint32_t i = 0; i < L; ++i) AL[i] = 0;
    AL[result] = 1;
    uint32_t live = 0;
    for (uint32_t i = L; i-- > 0;) {
        if (!AL[i]) continue;
        ++live;
        const int count = operand_count(ST[i].op);
        if (count > 0) AL[ST[i].a] = 1;
        if (count > 1) AL[ST[i].b] = 1;
        if (count > 2) AL[ST[i].c] = 1;
    }
    c.same = (!changed && live == L && result == res);
    c.length = live;
    if (c.same) return c;
    c.steps = (kdn_step*)malloc(sizeof(kdn_step) * (live + 1));
    uint32_t kept = 0;
    for (uint32_t i = 0; i < L; ++i) {
        if (!AL[i]) continue;
        kdn_step s = ST[i];
        const int count = operand_count(s.op);
        s.a = count > 0 ? (uint16_t)AL[s.a] : s.a;
        s.b = count > 1 ? (uint16_t)AL[s.b] : 0;
        s.c = count > 2 ? (uint16_t)AL[s.c] : 0;
        c.steps[kept] = s;
        AL[i] = kept++;
    }
    c.result = AL[result];
    return c;
}

/* ---- the octree, level by level ---- */
static void* grow(void* p, uint64_t* cap, uint64_t need, size_t item) {
    if (need <= *cap) return p;
    uint64_t n = *cap ? *cap : 1024;
    while (n < need) n *= 2;
    p = realloc(p, n * item);
    *cap = n;
    return p;
}

static uint32_t header_length(const kdn_step* h) { return (uint32_t)h->a | ((uint32_t)h->b << 16); }

static uint64_t push_program(or_octree* t, const kdn_step* steps, uint32_t count, uint32_t result, uint32_t leaf) {
    t->pool = (kdn_step*)grow(t->pool, &t->cap_pool, t->n_pool + count + 1, sizeof(kdn_step));
    const uint64_t at = t->n_pool;
    kdn_step h;
    memset(&h, 0, sizeof(h));
    h.op = 255;
    h.a = (uint16_t)(count & 0xFFFF);
    h.b = (uint16_t)(count >> 16);
    h.c = (uint16_t)result;
    h.leaf = leaf;
    /* the header's value: the FP32 operations of one point evaluation
     * (constants and coordinates free, smooth minimum 10, positive square
     * 2, transcendentals 20, every other step 1) */
    uint32_t flops = 0;
    for (uint32_t i = 0; i < count; ++i) {
        switch (steps[i].op) {
            case KDN_CONSTANT: case KDN_ARGUMENT: case KDN_POINT_X: case KDN_POINT_Y: case KDN_POINT_Z:
            case KDN_NORMAL_X: case KDN_NORMAL_Y: case KDN_NORMAL_Z: break;
            case KDN_SMOOTH_MINIMUM: flops += 10; break;
            case KDN_POSITIVE_SQUARE: flops += 2; break;
            case KDN_SINE: case KDN_COSINE: case KDN_ARC_TANGENT: case KDN_POWER: flops += 20; break;
            default: flops += 1;
        }
    }
    h.value = (float)flops;
    t->pool[at] = h;
    memcpy(t->pool + at + 1, steps, sizeof(kdn_step) * count);
    t->n_pool += count + 1;
    return at;
}

static void box_of(const float lower[3], const float size[3], uint32_t level, uint32_t x, uint32_t y, uint32_t z, Iv box[3]) {
    const uint32_t c[3] = {x, y, z};
    for (int k = 0; k < 3; ++k) {
        const float cell = ldexpf(size[k], -(int)level);
        const float lo = lower[k] + (float)c[k] * cell;
        box[k] = iv(lo, lo + cell);
    }
}

/* Cut every entry of one parent over one child box; returns the child. The
 * parent's entries are [first, first + count) of t->entries. Programs are
 * appended in the order the GPU's prefix sums place them: by parent entry,
 * then octant... which is item order (entry-major, octant-minor) over the
 * whole level. Here a level is cut item by item in that order. */
typedef struct {
    Cut32 cut;
    uint32_t object, program;
} ItemT;

int or_build_octree32(const kdn_step* steps, const kdn_object* objects, uint32_t object_count, const float lower[3],
                      const float size[3], int depth, or_octree* t) {
    memset(t, 0, sizeof(*t));
    /* the object programs at the front of the pool, in order */
    t->entries = (kdn_entry_record*)grow(NULL, &t->cap_entries, object_count + 1, sizeof(kdn_entry_record));
    for (uint32_t i = 0; i < object_count; ++i) {
        const uint64_t at = push_program(t, steps + objects[i].first, objects[i].count, objects[i].result, i);
        t->entries[i].object = i;
        t->entries[i].program = (uint32_t)at;
    }
    t->n_entries = object_count;
    uint32_t maxL = 1;
    for (uint32_t i = 0; i < object_count; ++i) maxL = objects[i].count > maxL ? objects[i].count : maxL;
    Iv* IV = (Iv*)malloc(sizeof(Iv) * maxL);
    kdn_step* ST = (kdn_step*)malloc(sizeof(kdn_step) * maxL);
    uint32_t* AL = (uint32_t*)malloc(sizeof(uint32_t) * maxL);

    /* parents of the current level: (node index or root marker, entry range) */
    uint64_t levelFirst = 0, levelCount = 0;
    for (int level = -1; level < depth; ++level) {
        /* active parents */
        uint64_t nParents = 0;
        uint32_t* parents = NULL;  /* node indices; for level -1 none */
        uint32_t octants = level < 0 ? 1 : 8;
        if (level < 0) {
            nParents = 1;
        } else {
            parents = (uint32_t*)malloc(sizeof(uint32_t) * (levelCount + 1));
            for (uint64_t i = 0; i < levelCount; ++i) {
                if (t->nodes[levelFirst + i].kind == 2) parents[nParents++] = (uint32_t)(levelFirst + i);
            }
            if (nParents == 0) {
                free(parents);
                break;
            }
        }
        const uint64_t childBase = level < 0 ? 0 : t->n_nodes;
        /* branches */
        for (uint64_t p = 0; p < nParents && level >= 0; ++p) {
            t->nodes[parents[p]].kind = 3;
            t->nodes[parents[p]].first_child = (uint32_t)(childBase + 8 * p);
        }
        /* the level's items in order: parent, entry, octant */
        uint64_t nItems = 0;
        for (uint64_t p = 0; p < nParents; ++p) nItems += (level < 0 ? object_count : t->nodes[parents[p]].entry_count) * octants;
        ItemT* items = (ItemT*)calloc(nItems + 1, sizeof(ItemT));
        uint64_t w = 0;
        for (uint64_t p = 0; p < nParents; ++p) {
            const uint32_t ef = level < 0 ? 0 : t->nodes[parents[p]].entry_first;
            const uint32_t ec = level < 0 ? object_count : t->nodes[parents[p]].entry_count;
            for (uint32_t j = 0; j < ec; ++j) {
                const kdn_entry_record e = t->entries[ef + j];
                const kdn_step* h = &t->pool[e.program];
                for (uint32_t o = 0; o < octants; ++o, ++w) {
                    Iv box[3];
                    if (level < 0) {
                        box_of(lower, size, 0, 0, 0, 0, box);
                    } else {
                        const kdn_node_record* pn = &t->nodes[parents[p]];
                        box_of(lower, size, (uint32_t)level + 1, 2 * pn->x + (o & 1), 2 * pn->y + ((o >> 1) & 1),
                               2 * pn->z + ((o >> 2) & 1), box);
                    }
                    items[w].cut = cut32(h + 1, header_length(h), h->c, box, IV, ST, AL);
                    items[w].object = e.object;
                    items[w].program = e.program;
                }
            }
        }
        /* children: the cell decision (cut.cpp:446-476) */
        const uint64_t nChildren = nParents * octants;
        t->nodes = (kdn_node_record*)grow(t->nodes, &t->cap_nodes, childBase + nChildren, sizeof(kdn_node_record));
        uint32_t* keep = (uint32_t*)calloc(nItems + 1, sizeof(uint32_t));
        uint32_t* kind = (uint32_t*)calloc(nChildren + 1, sizeof(uint32_t));
        uint32_t* kept = (uint32_t*)calloc(nChildren + 1, sizeof(uint32_t));
        uint64_t itemBase = 0;
        for (uint64_t p = 0; p < nParents; ++p) {
            const uint32_t ec = level < 0 ? object_count : t->nodes[parents[p]].entry_count;
            for (uint32_t o = 0; o < octants; ++o) {
                const uint64_t ch = p * octants + o;
                int filled = -1;
                for (uint32_t j = 0; j < ec; ++j) {
                    if (items[itemBase + (uint64_t)j * octants + o].cut.status == 1) {
                        filled = (int)j;
                        break;
                    }
                }
                if (filled >= 0) {
                    keep[itemBase + (uint64_t)filled * octants + o] = 1;
                    kept[ch] = 1;
                    kind[ch] = 1;
                    continue;
                }
                float lowest = KINF;
                for (uint32_t j = 0; j < ec; ++j) lowest = fmin_(lowest, items[itemBase + (uint64_t)j * octants + o].cut.hi);
                uint32_t k = 0;
                for (uint32_t j = 0; j < ec; ++j) {
                    const Cut32* c = &items[itemBase + (uint64_t)j * octants + o].cut;
                    if (c->status == 2 && !(c->lo > lowest)) keep[itemBase + (uint64_t)j * octants + o] = ++k;
                }
                kept[ch] = k;
                kind[ch] = k == 0 ? 0 : 2;
            }
            itemBase += (uint64_t)ec * octants;
        }
        /* child entry offsets (prefix over children) and program offsets
         * (prefix over items) */
        uint64_t* childOffsets = (uint64_t*)calloc(nChildren + 1, sizeof(uint64_t));
        uint64_t run = 0;
        for (uint64_t ch = 0; ch < nChildren; ++ch) {
            childOffsets[ch] = run;
            run += kept[ch];
        }
        const uint64_t entryBase = t->n_entries;
        t->entries = (kdn_entry_record*)grow(t->entries, &t->cap_entries, entryBase + run + 1, sizeof(kdn_entry_record));
        t->n_entries = entryBase + run;
        /* programs in item order */
        itemBase = 0;
        w = 0;
        for (uint64_t p = 0; p < nParents; ++p) {
            const uint32_t ec = level < 0 ? object_count : t->nodes[parents[p]].entry_count;
            for (uint32_t j = 0; j < ec; ++j) {
                for (uint32_t o = 0; o < octants; ++o, ++w) {
                    Cut32* c = &items[w].cut;
                    if (keep[w]) {
                        const uint64_t ch = p * octants + o;
                        const uint64_t slot = entryBase + childOffsets[ch] + keep[w] - 1;
                        uint32_t prog = items[w].program;
                        if (!c->same) prog = (uint32_t)push_program(t, c->steps, c->length, c->result, items[w].object);
                        t->entries[slot].object = items[w].object;
                        t->entries[slot].program = prog;
                    }
                    free(c->steps);
                    c->steps = NULL;
                }
            }
        }
        /* child nodes */
        for (uint64_t ch = 0; ch < nChildren; ++ch) {
            kdn_node_record* n = &t->nodes[childBase + ch];
            n->level = (uint32_t)(level + 1);
            if (level < 0) {
                n->x = n->y = n->z = 0;
            } else {
                const kdn_node_record* pn = &t->nodes[parents[ch / 8]];
                const uint32_t o = (uint32_t)(ch % 8);
                n->x = 2 * pn->x + (o & 1);
                n->y = 2 * pn->y + ((o >> 1) & 1);
                n->z = 2 * pn->z + ((o >> 2) & 1);
            }
            n->kind = kind[ch];
            n->first_child = 0;
            n->entry_first = (uint32_t)(entryBase + childOffsets[ch]);
            n->entry_count = kept[ch];
        }
        t->n_nodes = childBase + nChildren;
        levelFirst = childBase;
        levelCount = nChildren;
        free(items);
        free(keep);
        free(kind);
        free(kept);
        free(childOffsets);
        free(parents);
    }
    free(IV);
    free(ST);
    free(AL);
    return 0;
}

void or_free_octree(or_octree* t) {
    free(t->nodes);
    free(t->entries);
    free(t->pool);
    memset(t, 0, sizeof(*t));
}

/* ---- the march (twin of trace_dev.h) ---- */
#define MAX_ITER 4096
static const float kTolAbs = 1e-5f, kTolRel = 2e-6f, kTie = 1e-6f, kGraze = 0.02f;
static const float kNewtonFacing = 0.01f, kNewtonReach = 100.0f;

typedef struct {
    const or_octree* t;
    const uint32_t* leafMaterial;
    float lower[3], size[3];
    int depth;
    Dl* buf;
} View;

/* The point step over the dual buffer's values. */
static float point_step_dl(const kdn_step* s, const Dl* r, float px, float py, float pz) {
    float ra = r[s->a].v, rb = r[s->b].v, rc = r[s->c].v;
    switch (s->op) {
        case KDN_CONSTANT:
        case KDN_ARGUMENT: return s->value;
        case KDN_POINT_X: return px;
        case KDN_POINT_Y: return py;
        case KDN_POINT_Z: return pz;
        case KDN_NORMAL_X:
        case KDN_NORMAL_Y:
        case KDN_NORMAL_Z: return 0.0f;
        case KDN_ADD: return ra + rb;
        case KDN_SUBTRACT: return ra - rb;
        case KDN_MULTIPLY: return ra * rb;
        case KDN_DIVIDE: return ra / rb;
        case KDN_NEGATE: return -ra;
        case KDN_SQUARE_ROOT: return sqrtf(ra);
        case KDN_ABSOLUTE: return fabsf(ra);
        case KDN_FLOOR: return floorf(ra);
        case KDN_SINE: return f_rn(d_sin((double)ra));
        case KDN_COSINE: return f_rn(d_cos((double)ra));
        case KDN_SIGN: return p_sign(ra);
        case KDN_ARC_TANGENT: return f_rn(d_atan2((double)ra, (double)rb));
        case KDN_POWER: return f_rn(d_pow((double)ra, (double)rb));
        case KDN_MINIMUM: return fmin_(ra, rb);
        case KDN_MAXIMUM: return fmax_(ra, rb);
        case KDN_SMOOTH_MINIMUM: return p_smin(ra, rb, rc);
        case KDN_POSITIVE_SQUARE: return p_pos2(ra);
        case KDN_SELECT: return ra != 0 ? rb : rc;
        default: return p_compare(s->op, ra, rb);
    }
}

static float eval_prog(const View* v, uint32_t at, float px, float py, float pz) {
    const kdn_step* h = &v->t->pool[at];
    const uint32_t L = header_length(h);
    for (uint32_t i = 0; i < L; ++i) v->buf[i].v = point_step_dl(&h[1 + i], v->buf, px, py, pz);
    return v->buf[h->c].v;
}

static Dl eval_dual_prog(const View* v, uint32_t at, float px, float py, float pz) {
    const kdn_step* h = &v->t->pool[at];
    const uint32_t L = header_length(h);
    for (uint32_t i = 0; i < L; ++i) v->buf[i] = dual_step(&h[1 + i], v->buf, px, py, pz);
    return v->buf[h->c];
}

typedef struct {
    float least, nearest;
    uint32_t arg;
} Field;

static Field leaf_field(const View* v, const kdn_node_record* n, float px, float py, float pz) {
    Field f = {KINF, KINF, 0};
    for (uint32_t j = 0; j < n->entry_count; ++j) {
        const float x = eval_prog(v, v->t->entries[n->entry_first + j].program, px, py, pz);
        if (x < f.least) {
            f.least = x;
            f.arg = j;
        }
        f.nearest = fmin_(f.nearest, fabsf(x));
    }
    return f;
}

static uint32_t chosen(const kdn_step* s, const Dl* r) {
    const float a = r[s->a].v, b = r[s->b].v;
    switch (s->op) {
        case KDN_MINIMUM:
        case KDN_SMOOTH_MINIMUM: return a <= b ? s->a : s->b;
        case KDN_MAXIMUM: return a >= b ? s->a : s->b;
        case KDN_SELECT: return a != 0 ? s->b : s->c;
        default: return s->a;
    }
}
static int is_join(uint8_t op) {
    return op == KDN_MINIMUM || op == KDN_SMOOTH_MINIMUM || op == KDN_MAXIMUM || op == KDN_NEGATE || op == KDN_SELECT;
}
static uint32_t leaf_from(const kdn_step* prog, const Dl* r, uint32_t at) {
    while (is_join(prog[at].op)) at = chosen(&prog[at], r);
    return prog[at].leaf;
}
static void source_of(const kdn_step* prog, uint32_t result, const Dl* r, uint32_t* leaf, uint32_t* other, float* share) {
    *other = KDN_NO_LEAF;
    *share = 0.0f;
    uint32_t at = result;
    while (is_join(prog[at].op)) {
        const kdn_step s = prog[at];
        const uint32_t next = chosen(&s, r);
        float sh = 0.0f;
        if (s.op == KDN_SMOOTH_MINIMUM) {
            const float radius = r[s.c].v;
            const float apart = fabsf(r[s.a].v - r[s.b].v);
            sh = (radius > 0 && apart < radius) ? 0.5f * (radius - apart) / radius : 0.0f;
        }
        if (sh > 0 && *other == KDN_NO_LEAF) {
            *other = leaf_from(prog, r, next == s.a ? s.b : s.a);
            *share = sh;
        }
        at = next;
    }
    *leaf = prog[at].leaf;
}

static void faced_normal(const float g[3], const float d[3], int inside, float n[3]) {
    const float len = sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (!(len > 0)) {
        n[0] = -d[0];
        n[1] = -d[1];
        n[2] = -d[2];
        return;
    }
    float u[3] = {g[0] / len, g[1] / len, g[2] / len};
    const float along = u[0] * d[0] + u[1] * d[1] + u[2] * d[2];
    if (inside) {
        const float s = along < 0 ? -1.0f : 1.0f;
        n[0] = s * u[0];
        n[1] = s * u[1];
        n[2] = s * u[2];
        return;
    }
    if (along > 0) {
        float w[3] = {u[0] - d[0] * (along + kGraze), u[1] - d[1] * (along + kGraze), u[2] - d[2] * (along + kGraze)};
        const float l = sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        n[0] = w[0] / l;
        n[1] = w[1] / l;
        n[2] = w[2] / l;
        return;
    }
    n[0] = u[0];
    n[1] = u[1];
    n[2] = u[2];
}

static void land(const View* v, const kdn_node_record* n, uint32_t index, uint32_t winner, const float o[3], const float d[3],
                 float t, int inside, kdn_hit* hit) {
    const float px = o[0] + d[0] * t, py = o[1] + d[1] * t, pz = o[2] + d[2] * t;
    const uint32_t at = v->t->entries[n->entry_first + winner].program;
    const Dl value = eval_dual_prog(v, at, px, py, pz);
    /* one Newton step along the ray onto the surface, bounded */
    const float slope = value.g[0] * d[0] + value.g[1] * d[1] + value.g[2] * d[2];
    const float glen = sqrtf(value.g[0] * value.g[0] + value.g[1] * value.g[1] + value.g[2] * value.g[2]);
    const float bound = kNewtonReach * (kTolAbs + kTolRel * t);
    float tr = t;
    if (fabsf(slope) > kNewtonFacing * glen) {
        const float dt = -value.v / slope;
        tr = t + fmax_(-bound, fmin_(bound, dt));
    }
    hit->kind = KDN_HIT_SURFACE;
    hit->distance = tr;
    hit->point[0] = o[0] + d[0] * tr;
    hit->point[1] = o[1] + d[1] * tr;
    hit->point[2] = o[2] + d[2] * tr;
    faced_normal(value.g, d, inside, hit->normal);
    source_of(&v->t->pool[at + 1], v->t->pool[at].c, v->buf, &hit->leaf, &hit->other, &hit->share);
    hit->cell = index;
}

static float vmax3(float a, float b, float c) { return fmax_(fmax_(a, b), c); }

static uint32_t winner_at(const View* v, const kdn_node_record* n, const float o[3], const float d[3], float t, float inset,
                          Field f) {
    if (n->entry_count == 1) return 0;
    const float px = o[0] + d[0] * t, py = o[1] + d[1] * t, pz = o[2] + d[2] * t;
    const float band = kTie * (1 + vmax3(fabsf(px), fabsf(py), fabsf(pz)));
    const float qx = o[0] + d[0] * (t + inset), qy = o[1] + d[1] * (t + inset), qz = o[2] + d[2] * (t + inset);
    uint32_t winner = f.arg;
    float winnerOn = eval_prog(v, v->t->entries[n->entry_first + f.arg].program, qx, qy, qz);
    for (uint32_t j = 0; j < n->entry_count; ++j) {
        const uint32_t at = v->t->entries[n->entry_first + j].program;
        const float x = eval_prog(v, at, px, py, pz);
        if (!(x - f.least <= band)) continue;
        const float on = eval_prog(v, at, qx, qy, qz);
        if (on < winnerOn) {
            winner = j;
            winnerOn = on;
        }
    }
    return winner;
}

static void missed(kdn_hit* hit) {
    memset(hit, 0, sizeof(*hit));
    hit->kind = KDN_HIT_MISSED;
    hit->distance = KINF;
    hit->leaf = KDN_NO_LEAF;
    hit->other = KDN_NO_LEAF;
    hit->cell = KDN_NO_LEAF;
}

static float slope_of(float part) {
    const float m = 1e-12f;
    return 1.0f / (fabsf(part) < m ? (part < 0 ? -m : m) : part);
}

static void trace_one(const View* v, const kdn_ray* ray, kdn_hit* hit) {
    const float o[3] = {ray->origin[0], ray->origin[1], ray->origin[2]};
    const float d[3] = {ray->direction[0], ray->direction[1], ray->direction[2]};
    const int inside = ray->inside != 0;
    float t0 = 0.0f, t1 = ray->far_;
    float inv[3];
    for (int k = 0; k < 3; ++k) {
        inv[k] = slope_of(d[k]);
        const float a = (v->lower[k] - o[k]) * inv[k];
        const float b = (v->lower[k] + v->size[k] - o[k]) * inv[k];
        t0 = fmax_(t0, fmin_(a, b));
        t1 = fmin_(t1, fmax_(a, b));
    }
    missed(hit);
    if (!(t0 <= t1)) return;
    const int D = v->depth;
    const uint32_t N = 1u << D;
    float finest[3];
    uint32_t ci[3];
    for (int k = 0; k < 3; ++k) {
        finest[k] = ldexpf(v->size[k], -D);
        const float p = o[k] + d[k] * t0;
        const float q = floorf((p - v->lower[k]) / finest[k]);
        ci[k] = q < 0 ? 0u : (q >= (float)N ? N - 1 : (uint32_t)q);
    }
    float t = t0;
    int iterations = 0;
    while (iterations < MAX_ITER) {
        uint32_t index = 0;
        kdn_node_record n = v->t->nodes[0];
        int lvl = 0;
        while (n.kind == 3) {
            ++lvl;
            const int sh = D - lvl;
            const uint32_t oct = ((ci[0] >> sh) & 1u) | (((ci[1] >> sh) & 1u) << 1) | (((ci[2] >> sh) & 1u) << 2);
            index = n.first_child + oct;
            n = v->t->nodes[index];
        }
        const uint32_t nc[3] = {n.x, n.y, n.z};
        float tExit = KINF;
        int axis = 0;
        for (int k = 0; k < 3; ++k) {
            const float cell = ldexpf(v->size[k], -lvl);
            const float lo = v->lower[k] + (float)nc[k] * cell;
            const float hi = lo + cell;
            const float bound = d[k] > 0 ? hi : lo;
            const float te = d[k] == 0 ? KINF : (bound - o[k]) * inv[k];
            if (te < tExit) {
                tExit = te;
                axis = k;
            }
        }
        const float tEnd = fmin_(tExit, t1);
        if (n.kind == 2) {
            while (t <= tEnd && iterations < MAX_ITER) {
                ++iterations;
                const float px = o[0] + d[0] * t, py = o[1] + d[1] * t, pz = o[2] + d[2] * t;
                const Field f = leaf_field(v, &n, px, py, pz);
                const float tol = kTolAbs + kTolRel * t;
                const float inset = 2 * tol;
                float step;
                if (!inside) {
                    if (f.least < tol) {
                        const uint32_t w = winner_at(v, &n, o, d, t, inset, f);
                        land(v, &n, index, w, o, d, t, 0, hit);
                        return;
                    }
                    step = f.least;
                } else {
                    if (f.nearest < tol) {
                        const float qx = o[0] + d[0] * (t + inset), qy = o[1] + d[1] * (t + inset), qz = o[2] + d[2] * (t + inset);
                        const Field past = leaf_field(v, &n, qx, qy, qz);
                        if (past.least > 0) {
                            const Field here = leaf_field(v, &n, px, py, pz);
                            const uint32_t w = winner_at(v, &n, o, d, t, inset, here);
                            land(v, &n, index, w, o, d, t, 1, hit);
                            return;
                        }
                        kdn_hit other;
                        land(v, &n, index, past.arg, o, d, t + inset, 1, &other);
                        /* the crossing's hit keeps the landing's Newton step
                         * onto the solid past it (the device's rule) */
                        if (other.leaf != KDN_NO_LEAF && v->leafMaterial[other.leaf] != ray->within) {
                            *hit = other;
                            return;
                        }
                        step = inset;
                    } else {
                        step = f.nearest;
                    }
                }
                /* a step stops at the leaf's exit: the programs of the
                 * objects dropped from the leaf bound nothing beyond it */
                const float next = t + step;
                if (next >= tExit) {
                    t = tExit;
                    break;
                }
                t = next;
            }
        } else if (n.kind == 1 && !inside) {
            hit->kind = KDN_HIT_SOLID;
            hit->distance = t;
            hit->point[0] = o[0] + d[0] * t;
            hit->point[1] = o[1] + d[1] * t;
            hit->point[2] = o[2] + d[2] * t;
            hit->normal[0] = hit->normal[1] = hit->normal[2] = 0;
            hit->normal[axis] = d[axis] > 0 ? -1.0f : 1.0f;
            hit->cell = index;
            /* the solid's leaf: the source of its first program's value */
            if (n.entry_count > 0) {
                const uint32_t at = v->t->entries[n.entry_first].program;
                (void)eval_prog(v, at, hit->point[0], hit->point[1], hit->point[2]);
                float sh;
                source_of(&v->t->pool[at + 1], v->t->pool[at].c, v->buf, &hit->leaf, &hit->other, &sh);
                hit->share = sh;
            }
            return;
        }
        /* a ray from inside crosses Empty leaves (the reference's inside march
         * crosses open space, sphere_trace.cpp:189-201) */
        ++iterations;
        if (tExit >= t1) return;
        t = fmax_(t, tExit);
        const int sh = D - lvl;
        uint32_t next[3];
        for (int k = 0; k < 3; ++k) {
            const uint32_t first = nc[k] << sh, last = ((nc[k] + 1) << sh) - 1;
            if (k == axis) {
                if (d[k] > 0) {
                    next[k] = last + 1;
                } else {
                    if (first == 0) return;
                    next[k] = first - 1;
                }
            } else {
                const float p = o[k] + d[k] * tExit;
                const float q = floorf((p - v->lower[k]) / finest[k]);
                next[k] = q < (float)first ? first : (q > (float)last ? last : (uint32_t)q);
            }
        }
        if (next[axis] >= N) return;
        ci[0] = next[0];
        ci[1] = next[1];
        ci[2] = next[2];
    }
}

void or_trace32(const or_octree* tree, const float lower[3], const float size[3], int depth, const uint32_t* leaf_material,
                const kdn_ray* rays, uint32_t count, kdn_hit* hits) {
    uint32_t maxL = 1;
    for (uint64_t e = 0; e < tree->n_entries; ++e) {
        const uint32_t L = header_length(&tree->pool[tree->entries[e].program]);
        maxL = L > maxL ? L : maxL;
    }
    View v;
    v.t = tree;
    v.leafMaterial = leaf_material;
    for (int k = 0; k < 3; ++k) {
        v.lower[k] = lower[k];
        v.size[k] = size[k];
    }
    v.depth = depth;
    v.buf = (Dl*)malloc(sizeof(Dl) * maxL);
    for (uint32_t i = 0; i < count; ++i) trace_one(&v, &rays[i], &hits[i]);
    free(v.buf);
}
