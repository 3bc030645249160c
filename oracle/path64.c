/*
 * The plain spectral path tracer in FP64 — TEST INFRASTRUCTURE ONLY (see
 * oracle.h).
 *
 * It restates the reference transport for the plain configuration
 * (plugins/transport/paths): the counter-based numbers (trace.cpp:47-74,
 * path.cpp:15-46, math/random.cpp), four wavelength lanes by the
 * CIE-importance cuts (math/colour.cpp:134-257), RGB to spectrum
 * (colour.cpp:88-103, spectrum.cpp:49-67), matte, GGX metal with energy
 * compensation (material.cpp:416-660, microfacet.cpp, energy.cpp:45-127)
 * and exact dispersive glass (material.cpp:766-1325, dispersion.cpp),
 * quad lamps met and aimed at (light.cpp), the gradient sky
 * (light.cpp:582-591), power-heuristic MIS and roulette from depth 4
 * (trace.cpp:1020-1074). Geometry is the FP64 sphere trace of ref64.c.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define LANES 4
#define TABLE 65
#define CUTS 256
#define VIEWS 32
static const double PI = 3.14159265358979323846;

/* ---- colour tables (colour.cpp:105-211) ---- */
typedef struct {
    double x[TABLE], y[TABLE], z[TABLE], white[3], cuts[CUTS + 1];
} Matching;

static double skewed(double nm, double peak, double below, double above) {
    const double w = nm < peak ? below : above, o = (nm - peak) / w;
    return exp(-0.5 * o * o);
}
static void linear_of(double x, double y, double z, double rgb[3]) {
    rgb[0] = 3.2406 * x - 1.5372 * y - 0.4986 * z;
    rgb[1] = -0.9689 * x + 1.8758 * y + 0.0415 * z;
    rgb[2] = 0.0557 * x - 0.2040 * y + 1.0570 * z;
}
static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

static Matching g_match;
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

static void build_matching(void) {
    Matching* m = &g_match;
    double sx = 0, sy = 0, sz = 0;
    for (int s = 0; s < TABLE; ++s) {
        const double nm = 380.0 + 5.0 * s;
        m->x[s] = 1.056 * skewed(nm, 599.8, 37.9, 31.0) + 0.362 * skewed(nm, 442.0, 16.0, 26.7) - 0.065 * skewed(nm, 501.1, 20.4, 26.2);
        m->y[s] = 0.821 * skewed(nm, 568.8, 46.9, 40.5) + 0.286 * skewed(nm, 530.9, 16.3, 31.1);
        m->z[s] = 1.217 * skewed(nm, 437.0, 11.8, 36.0) + 0.681 * skewed(nm, 459.0, 26.0, 13.8);
        const double w = (s == 0 || s == TABLE - 1 ? 0.5 : 1.0) * 5.0;
        sx += m->x[s] * w;
        sy += m->y[s] * w;
        sz += m->z[s] * w;
    }
    linear_of(sx, sy, sz, m->white);
    double seen[TABLE];
    for (int s = 0; s < TABLE; ++s) {
        double rgb[3];
        linear_of(m->x[s], m->y[s], m->z[s], rgb);
        seen[s] = fabs(rgb[0] / m->white[0]) + fabs(rgb[1] / m->white[1]) + fabs(rgb[2] / m->white[2]);
    }
    double total = 0, shares[TABLE - 1], before[TABLE] = {0};
    for (int s = 0; s < TABLE - 1; ++s) total += seen[s] + seen[s + 1];
    for (int s = 0; s < TABLE - 1; ++s) shares[s] = 0.7 * (seen[s] + seen[s + 1]) / total + 0.3 / (TABLE - 1);
    for (int s = 0; s < TABLE - 1; ++s) before[s + 1] = before[s] + shares[s];
    int slice = 0;
    for (int i = 0; i <= CUTS; ++i) {
        const double u = (double)i / CUTS * before[TABLE - 1];
        while (slice < TABLE - 2 && before[slice + 1] <= u) ++slice;
        m->cuts[i] = 380.0 + 5.0 * (slice + clampd((u - before[slice]) / shares[slice], 0.0, 1.0));
    }
    m->cuts[0] = 380.0;
    m->cuts[CUTS] = 700.0;
}
static const Matching* matching(void) {
    pthread_once(&g_once, build_matching);
    return &g_match;
}

void or_sampled_wavelengths(double u, double nm[4], double density[4]) {
    const Matching* m = matching();
    for (int l = 0; l < LANES; ++l) {
        const double raw = u + 0.25 * l, turned = raw >= 1 ? raw - 1 : raw;
        const double place = turned * CUTS;
        const int index = (int)place < CUTS - 1 ? (int)place : CUTS - 1;
        const double width = m->cuts[index + 1] - m->cuts[index];
        const double v = m->cuts[index] + (place - index) * width;
        nm[l] = v < nextafter(700.0, 380.0) ? v : nextafter(700.0, 380.0);
        density[l] = 1 / (CUTS * width);
    }
}

static double table_at(const double* curve, double nm) {
    const double slot = (nm - 380.0) / 5.0;
    int below = (int)slot;
    below = below < 0 ? 0 : (below > TABLE - 2 ? TABLE - 2 : below);
    const double blend = clampd(slot - below, 0.0, 1.0);
    return curve[below] + blend * (curve[below + 1] - curve[below]);
}

void or_colour_of(const double nm[4], const double density[4], const double s[4], double rgb[3]) {
    const Matching* m = matching();
    double x = 0, y = 0, z = 0;
    for (int l = 0; l < LANES; ++l) {
        if (density[l] <= 0) continue;
        const double w = s[l] / density[l];
        x += table_at(m->x, nm[l]) * w;
        y += table_at(m->y, nm[l]) * w;
        z += table_at(m->z, nm[l]) * w;
    }
    linear_of(x / LANES, y / LANES, z / LANES, rgb);
    for (int k = 0; k < 3; ++k) rgb[k] /= m->white[k];
}

static double smooth_rise(double nm, double lo, double hi) {
    const double b = clampd((nm - lo) / (hi - lo), 0.0, 1.0);
    return b * b * (3 - 2 * b);
}
void or_spectrum_of(const double rgb[3], const double nm[4], double out[4]) {
    const int least = (rgb[0] <= rgb[1] && rgb[0] <= rgb[2]) ? 0 : ((rgb[1] <= rgb[0] && rgb[1] <= rgb[2]) ? 1 : 2);
    const int first = least == 0 ? 1 : 0, second = least == 2 ? 1 : 2;
    const int most = rgb[first] <= rgb[second] ? second : first, mid = 3 - least - most;
    const double base = rgb[least], low = rgb[mid] - base, high = rgb[most] - rgb[mid];
    for (int l = 0; l < LANES; ++l) {
        const double r = smooth_rise(nm[l], 575.0, 605.0), b = 1 - smooth_rise(nm[l], 445.0, 530.0), g = (1 - b) * (1 - r);
        const double basis[3] = {r, g, b};
        out[l] = base + low * (1 - basis[least]) + high * basis[most];
    }
}

/* ---- vectors ---- */
typedef struct {
    double x, y, z;
} V;
static V vv(double x, double y, double z) {
    V v = {x, y, z};
    return v;
}
static V add(V a, V b) { return vv(a.x + b.x, a.y + b.y, a.z + b.z); }
static V sub(V a, V b) { return vv(a.x - b.x, a.y - b.y, a.z - b.z); }
static V mul(V a, double s) { return vv(a.x * s, a.y * s, a.z * s); }
static double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V cross(V a, V b) { return vv(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static double len(V a) { return sqrt(dot(a, a)); }
static V norm(V a) { return mul(a, 1 / len(a)); }
static V mirror(V out, V n) { return sub(mul(n, 2 * dot(out, n)), out); }
typedef struct {
    V t, b, n;
} B;
static B basis(V n) {
    const double sign = n.z >= 0 ? 1 : -1, a = -1 / (sign + n.z), bb = n.x * n.y * a;
    B o = {vv(1 + sign * n.x * n.x * a, sign * bb, -sign * n.x), vv(bb, sign + n.y * n.y * a, -n.y), n};
    return o;
}
static V world(B b, V v) { return add(add(mul(b.t, v.x), mul(b.b, v.y)), mul(b.n, v.z)); }
static V local(B b, V v) { return vv(dot(v, b.t), dot(v, b.b), dot(v, b.n)); }
static V cosine_hemisphere(B b, double u, double v) {
    const double r = sqrt(u), a = 2 * PI * v, h = sqrt(u < 1 ? 1 - u : 0);
    return world(b, vv(r * cos(a), r * sin(a), h));
}

/* ---- materials ---- */
typedef struct {
    uint32_t kind, exact;
    double albedo[3], width, share, esc[VIEWS], esc_avg, index, spread, glow[3];
} Mat;

static double smith_vis(double lv, double ar, double w) {
    const double sq = w * w, f = ar * sqrt(sq + (1 - sq) * lv * lv), s = lv * sqrt(sq + (1 - sq) * ar * ar);
    return f + s > 0 ? 0.5 / (f + s) : 0;
}
static double smith_mask(double c, double w) {
    const double sq = w * w, u = c + sqrt(sq + (1 - sq) * c * c);
    return u > 0 ? 2 * c / u : 0;
}
static double facet_d(double c, double w) {
    if (c <= 0) return 0;
    const double sq = w * w, lo = c * c * (sq - 1) + 1;
    return sq / (PI * lo * lo);
}
static void escape_of(double width, double* table, double* avg) {
    *avg = 0;
    for (int view = 0; view < VIEWS; ++view) {
        const double leaving = (view + 0.5) / VIEWS;
        const V out = vv(sqrt(1 - leaving * leaving > 0 ? 1 - leaving * leaving : 0), 0, leaving);
        double total = 0;
        for (int ring = 0; ring < 24; ++ring) {
            for (int spoke = 0; spoke < 12; ++spoke) {
                const double along = (ring + 0.5) / 24;
                const double tilt = sqrt((1 - along) / (1 + (width * width - 1) * along));
                const double spread = sqrt(1 - tilt * tilt > 0 ? 1 - tilt * tilt : 0), ang = 2 * PI * (spoke + 0.5) / 12;
                const V f = vv(spread * cos(ang), spread * sin(ang), tilt);
                const double onto = dot(out, f);
                if (onto <= 0) continue;
                const double inz = f.z * 2 * onto - out.z;
                if (inz <= 0) continue;
                total += 4 * smith_vis(out.z, inz, width) * inz * onto / f.z;
            }
        }
        table[view] = total / (24 * 12);
        *avg += 2 * table[view] * leaving / VIEWS;
    }
}
/* escapeOf (energy.cpp:100-127), exported for the pins. */
void or_escape_of(double width, double table[32], double* average) { escape_of(width, table, average); }
static double escape_at(const Mat* m, double c) {
    const double place = clampd(c, 0, 1) * VIEWS - 0.5;
    int lo = (int)floor(place);
    lo = lo < 0 ? 0 : (lo > VIEWS - 1 ? VIEWS - 1 : lo);
    const int hi = lo + 1 < VIEWS - 1 ? lo + 1 : VIEWS - 1;
    const double s = clampd(place - lo, 0, 1);
    return m->esc[lo] * (1 - s) + m->esc[hi] * s;
}
static V halfway(V o, V i, V n) {
    const V t = add(o, i);
    const double s = len(t);
    return s > 1e-12 ? mul(t, 1 / s) : n;
}
static void rough_value(const Mat* m, const double nm[4], V n, V out, V in, double f[4]) {
    const double lv = dot(n, out), ar = dot(n, in);
    for (int l = 0; l < LANES; ++l) f[l] = 0;
    if (lv <= 0 || ar <= 0) return;
    if (m->kind != KDN_METAL) {
        const double a[3] = {m->albedo[0] / PI, m->albedo[1] / PI, m->albedo[2] / PI};
        or_spectrum_of(a, nm, f);
        return;
    }
    const V h = halfway(out, in, n);
    const double onto = dot(out, h) > 0 ? dot(out, h) : 0;
    const double facets = facet_d(dot(n, h), m->width) * smith_vis(lv, ar, m->width);
    const double rest = pow(clampd(1 - onto, 0, 1), 5.0);
    double sch[3], avg[3], face[4], flat[4];
    for (int k = 0; k < 3; ++k) {
        sch[k] = m->albedo[k] + (1 - m->albedo[k]) * rest;
        avg[k] = m->albedo[k] + (1 - m->albedo[k]) * (1.0 / 21);
    }
    or_spectrum_of(sch, nm, face);
    or_spectrum_of(avg, nm, flat);
    const double room = 1 - m->esc_avg;
    const double lost = room <= 1e-6 ? 0 : (1 - escape_at(m, lv)) * (1 - escape_at(m, ar)) / (PI * room);
    for (int l = 0; l < LANES; ++l) {
        const double twice = flat[l] * flat[l] * m->esc_avg / (1 - flat[l] * (1 - m->esc_avg));
        f[l] = face[l] * facets + twice * lost;
    }
}
static double rough_density(const Mat* m, V n, V out, V in) {
    const double lv = dot(n, out), ar = dot(n, in);
    if (lv <= 0 || ar <= 0) return 0;
    if (m->kind != KDN_METAL) return ar / PI;
    const V h = halfway(out, in, n);
    const double facets = dot(out, h) > 0 ? facet_d(dot(n, h), m->width) * smith_mask(lv, m->width) / (4 * lv) : 0;
    return m->share * facets + (1 - m->share) * ar / PI;
}
static V visible_facet(V n, V out, double w, double u, double v) {
    const B b = basis(n);
    const V l = local(b, out);
    const V s = norm(vv(w * l.x, w * l.y, l.z));
    const double flat = s.x * s.x + s.y * s.y;
    const V first = flat > 1e-12 ? mul(vv(-s.y, s.x, 0), 1 / sqrt(flat)) : vv(1, 0, 0);
    const V second = cross(s, first);
    const double reach = sqrt(u), ang = 2 * PI * v, across = reach * cos(ang), lean = 0.5 * (1 + s.z);
    const double along = (1 - lean) * sqrt(1 - across * across > 0 ? 1 - across * across : 0) + lean * reach * sin(ang);
    const double hh = 1 - across * across - along * along;
    const V r = add(add(mul(first, across), mul(second, along)), mul(s, sqrt(hh > 0 ? hh : 0)));
    return world(b, norm(vv(w * r.x, w * r.y, r.z > 1e-12 ? r.z : 1e-12)));
}
static double dielectric(double c, double ratio) {
    const double left = 1 - ratio * ratio * (1 - c * c);
    if (left <= 0) return 1;
    const double on = sqrt(left), pe = (ratio * c - on) / (ratio * c + on), pa = (ratio * on - c) / (ratio * on + c);
    return 0.5 * (pe * pe + pa * pa);
}

typedef struct {
    int found, through, exact, collapse;
    V in;
    double weight[4], density;
} Sc;

static Sc glass_scatter(const Mat* m, const double nm[4], V n, V out, int inside, double uside) {
    Sc s;
    memset(&s, 0, sizeof(s));
    s.exact = 1;
    const double lv = dot(n, out);
    if (lv <= 0) return s;
    const int parted = m->spread != 0;
    double tr[4], be[4], R[4];
    for (int l = 0; l < LANES; ++l) {
        const double um = nm[l] * 1e-3, idx = parted ? m->index + m->spread / (um * um) : m->index;
        tr[l] = inside ? idx : 1;
        be[l] = inside ? 1 : idx;
        R[l] = dielectric(lv, tr[l] / be[l]);
    }
    if (!parted) R[1] = R[2] = R[3] = R[0];
    const double chance = parted ? 0.25 * (R[0] + R[1] + R[2] + R[3]) : R[0];
    const double ratio = tr[0] / be[0], left = 1 - ratio * ratio * (1 - lv * lv);
    const int passed = left > 0;
    s.found = 1;
    if (uside < chance || !passed) {
        s.in = mirror(out, n);
        const double c = passed ? chance : 1;
        for (int l = 0; l < LANES; ++l) s.weight[l] = R[l] / c;
        return s;
    }
    s.in = sub(mul(sub(mul(n, lv), out), ratio), mul(n, sqrt(left)));
    s.through = 1;
    for (int l = 0; l < LANES; ++l) s.weight[l] = (1 - R[l]) / (1 - chance);
    s.collapse = parted && ratio != 1;
    return s;
}
static Sc rough_scatter(const Mat* m, const double nm[4], V n, V out, double ulobe, double u, double v) {
    Sc s;
    memset(&s, 0, sizeof(s));
    if (dot(n, out) <= 0) return s;
    V in;
    if (m->kind == KDN_METAL && ulobe < m->share) {
        const V f = visible_facet(n, out, m->width, u, v);
        if (dot(out, f) <= 0) return s;
        in = mirror(out, f);
    } else {
        in = cosine_hemisphere(basis(n), u, v);
    }
    const double ar = dot(n, in);
    if (ar <= 0) return s;
    const double d = rough_density(m, n, out, in);
    if (!(d > 0)) return s;
    double f[4];
    rough_value(m, nm, n, out, in, f);
    s.found = 1;
    s.in = in;
    s.density = d;
    for (int l = 0; l < LANES; ++l) s.weight[l] = f[l] * ar / d;
    return s;
}

/* ---- lamps ---- */
typedef struct {
    V at, along, across, facing;
    double radiance[3], area, pick, cdf;
} Lamp;

static int meet_quad(const Lamp* L, V o, V d, double tmax, double* t, double* cosine) {
    const double toward = -dot(L->facing, d);
    if (!(toward > 0)) return 0;
    const double tt = dot(sub(L->at, o), L->facing) / dot(d, L->facing);
    if (!(tt > 0) || !(tt < tmax)) return 0;
    const V p = sub(add(o, mul(d, tt)), L->at);
    const double aa = dot(L->along, L->along), ac = dot(L->along, L->across), cc = dot(L->across, L->across);
    const double pa = dot(p, L->along), pc = dot(p, L->across), det = aa * cc - ac * ac;
    const double sa = (pa * cc - pc * ac) / det, sc = (pc * aa - pa * ac) / det;
    if (sa < 0 || sa > 1 || sc < 0 || sc > 1) return 0;
    *t = tt;
    *cosine = toward;
    return 1;
}
static double power_weight(double a, double b) { return a * a + b * b > 0 ? a * a / (a * a + b * b) : 0; }

/* ---- the Sobol member (plugins/sampler/sobol/sobol.cpp:8-156) ----
 * 28 coordinates, each its own dimension of Joe and Kuo's sequence over
 * 16 bits; coordinates from 6 on take the dimensions again on a second
 * index shuffle; place and value scrambled by Laine-Karras hashes seeded by
 * the upper half of the stream. */
static uint32_t stir32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7FEB352Du;
    v ^= v >> 15;
    v *= 0x846CA68Bu;
    return v ^ (v >> 16);
}
static uint32_t rev32(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 32; ++b) r |= ((v >> b) & 1u) << (31 - b);
    return r;
}
static uint32_t rev16(uint32_t v) { return rev32(v & 0xFFFFu) >> 16; }
static uint32_t lk_scramble(uint32_t v, uint32_t seed) {
    v += seed;
    v ^= v * 0x6C50B47Cu;
    v ^= v * 0xB82F1E52u;
    v ^= v * 0xC7AFE638u;
    return v ^ (v * 0x8D22F6E6u);
}
static const uint32_t kSobolRules[28][9] = {
    {0, 0, 1, 1, 1, 1, 1, 1, 1},      {1, 0, 1, 0, 0, 0, 0, 0, 0},      {2, 1, 1, 3, 0, 0, 0, 0, 0},
    {3, 1, 1, 3, 1, 0, 0, 0, 0},      {3, 2, 1, 1, 1, 0, 0, 0, 0},      {4, 1, 1, 1, 3, 3, 0, 0, 0},
    {4, 4, 1, 3, 5, 13, 0, 0, 0},     {5, 2, 1, 1, 5, 5, 17, 0, 0},     {5, 4, 1, 1, 5, 5, 5, 0, 0},
    {5, 7, 1, 1, 7, 11, 19, 0, 0},    {5, 11, 1, 1, 5, 1, 1, 0, 0},     {5, 13, 1, 1, 1, 3, 11, 0, 0},
    {5, 14, 1, 3, 5, 5, 31, 0, 0},    {6, 1, 1, 3, 3, 9, 7, 49, 0},     {6, 13, 1, 1, 1, 15, 21, 21, 0},
    {6, 16, 1, 3, 1, 13, 27, 49, 0},  {6, 19, 1, 1, 1, 15, 7, 5, 0},    {6, 22, 1, 3, 1, 15, 13, 25, 0},
    {6, 25, 1, 1, 5, 5, 19, 61, 0},   {7, 1, 1, 3, 7, 11, 23, 15, 103}, {7, 4, 1, 3, 7, 13, 13, 15, 69},
    {7, 7, 1, 1, 3, 13, 7, 35, 63},   {7, 8, 1, 3, 5, 9, 1, 25, 53},    {7, 14, 1, 3, 1, 13, 9, 35, 107},
    {7, 19, 1, 3, 1, 5, 27, 61, 31},  {7, 21, 1, 1, 5, 11, 19, 41, 61}, {7, 28, 1, 3, 5, 3, 3, 13, 69},
    {7, 31, 1, 1, 7, 13, 1, 19, 1}};
static uint32_t g_sobol_columns[28][16];
static pthread_once_t g_sobol_once = PTHREAD_ONCE_INIT;
static void build_sobol(void) {
    for (int d = 0; d < 28; ++d) {
        const uint32_t degree = kSobolRules[d][0], poly = kSobolRules[d][1];
        uint32_t m[16];
        for (uint32_t b = 0; b < 16; ++b) {
            if (degree == 0 || b < degree) {
                m[b] = kSobolRules[d][2 + (degree == 0 ? 0 : b)];
            } else {
                uint32_t n = m[b - degree] ^ (m[b - degree] << degree);
                for (uint32_t k = 1; k < degree; ++k)
                    if ((poly >> (degree - 1 - k)) & 1u) n ^= m[b - k] << k;
                m[b] = n;
            }
            g_sobol_columns[d][b] = rev32(m[b] << (31 - b));
        }
    }
}
double or_sobol_number(uint64_t stream, uint64_t index, uint64_t coordinate) {
    pthread_once(&g_sobol_once, build_sobol);
    const uint32_t scramble = (uint32_t)(stream >> 32), c = (uint32_t)coordinate;
    const uint32_t place = rev16(lk_scramble(rev16((uint32_t)index), stir32(scramble + (c < 6 ? 0u : 0x68E31DA4u))));
    const uint32_t seed = stir32(scramble + c * 0xA136AAADu + 0x2545F491u);
    const uint32_t* col = g_sobol_columns[c < 6 ? c : c - 6];
    uint32_t value = 0;
    for (int b = 0; b < 16; ++b)
        if ((place >> b) & 1u) value ^= col[b];
    return (double)rev32(lk_scramble(value, seed)) * 0x1p-32;
}

/* ---- numbers (path.cpp:22-37, sampler.cpp:15-30, random.cpp:19-21) ---- */
static double sequence_number(int sobol, uint64_t stream, uint32_t sample, uint64_t dim) {
    return sobol && dim < 28 ? or_sobol_number(stream, sample, dim) : or_hashed_number(stream, sample, dim);
}
#define number(stream, sample, event, draw) \
    sequence_number(J->sobol, (stream), (sample), (draw) < 4 ? (uint64_t)(draw) : (uint64_t)(4 + (event) * 8 + (draw) - 4))
enum { PIXEL_X = 0, PIXEL_Y = 1, LIGHT_U = 4, LIGHT_V = 5, DIR_U = 6, DIR_V = 7, LOBE = 8, SIDE = 9, SURVIVAL = 10, PICK = 11 };

typedef struct {
    const or_render_scene* scene;
    Mat* mats;
    uint32_t nmat;
    Lamp* lamps;
    uint32_t nlamps;
    int width, height, x0, y0, x1, y1;
    uint32_t first, spp, surfaces, min_depth;
    uint64_t seed_bits;
    int sobol;
    V eye, fwd, right, up;
    double reach, aspect;
    double* rgb;
    int next_row;
    pthread_mutex_t lock;
} Job;

static void trace_one(const or_render_scene* sc, V o, V d, double far_, int inside, uint32_t within, kdn_hit* hit, double* t) {
    kdn_ray r;
    r.origin[0] = (float)o.x;
    r.origin[1] = (float)o.y;
    r.origin[2] = (float)o.z;
    r.direction[0] = (float)d.x;
    r.direction[1] = (float)d.y;
    r.direction[2] = (float)d.z;
    r.far_ = (float)far_;
    r.inside = inside;
    r.within = within;
    or_sphere_trace64(&sc->geometry, &r, 1, hit, t);
}

static void sample_path(const Job* J, int x, int y, uint32_t sample, double rgb[3]) {
    const or_render_scene* sc = J->scene;
    const uint64_t stream = J->seed_bits ^ or_mixed_bits(((uint64_t)(uint32_t)x << 32) | (uint64_t)(uint32_t)y);
    const double start = or_hashed_number(stream, ~0ull, 0);
    const double turned = (double)sample * 0.6180339887498949 + start;
    double nm[4], dens[4];
    or_sampled_wavelengths(turned - floor(turned), nm, dens);
    const double fx = x + number(stream, sample, 0, PIXEL_X), fy = y + number(stream, sample, 0, PIXEL_Y);
    const double u = 2 * fx / J->width - 1, v = 1 - 2 * fy / J->height;
    V d = norm(add(add(J->fwd, mul(J->right, u * J->reach * J->aspect)), mul(J->up, v * J->reach)));
    V o = J->eye;
    double thr[4] = {1, 1, 1, 1}, rad[4] = {0, 0, 0, 0};
    int inside = 0, dispersed = 0, exactBefore = 1;
    uint32_t within = KDN_NO_MATERIAL;
    double pdfBefore = 0;
    for (int depth = 0; depth < (int)J->surfaces; ++depth) {
        kdn_hit hit;
        double t;
        trace_one(sc, o, d, INFINITY, inside, within, &hit, &t);
        const double tHit = hit.kind == KDN_HIT_MISSED ? INFINITY : t;
        if (!inside) {
            for (uint32_t j = 0; j < J->nlamps; ++j) {
                double tl, cl;
                if (!meet_quad(&J->lamps[j], o, d, tHit, &tl, &cl)) continue;
                double Le[4];
                or_spectrum_of(J->lamps[j].radiance, nm, Le);
                const double pdfL = J->lamps[j].pick * tl * tl / (cl * J->lamps[j].area);
                const double w = exactBefore ? 1 : power_weight(pdfBefore, pdfL);
                for (int l = 0; l < LANES; ++l) rad[l] += thr[l] * Le[l] * w;
            }
        }
        if (hit.kind == KDN_HIT_MISSED) {
            if (!inside) {
                const double share = sqrt(fabs(d.z));
                const float* far_ = d.z >= 0 ? sc->sky_zenith : sc->sky_ground;
                double c[3], sky[4];
                for (int k = 0; k < 3; ++k) c[k] = sc->sky_horizon[k] * (1 - share) + far_[k] * share;
                or_spectrum_of(c, nm, sky);
                for (int l = 0; l < LANES; ++l) rad[l] += thr[l] * sky[l];
            }
            break;
        }
        if (hit.kind != KDN_HIT_SURFACE) break;
        uint32_t mi = hit.leaf != KDN_NO_LEAF ? sc->geometry.leaf_material[hit.leaf] : J->nmat - 1;
        if (mi >= J->nmat) mi = J->nmat - 1;
        const Mat* m = &J->mats[mi];
        const V xp = add(o, mul(d, t));
        const V out = mul(d, -1);
        V n = vv(hit.normal[0], hit.normal[1], hit.normal[2]);
        n = norm(n);
        if (dot(n, out) < 0) n = mul(n, -1);
        if (m->glow[0] > 0 || m->glow[1] > 0 || m->glow[2] > 0) {
            double g[4];
            or_spectrum_of(m->glow, nm, g);
            for (int l = 0; l < LANES; ++l) rad[l] += thr[l] * g[l];
        }
        const double mx = fmax(fabs(xp.x), fmax(fabs(xp.y), fabs(xp.z)));
        const double lift = 1e-5 * (1 + mx);
        const int exact = m->kind == KDN_GLASS && m->exact;
        if (!exact && J->nlamps > 0) {
            const double up = number(stream, sample, depth, PICK);
            uint32_t j = 0;
            while (j + 1 < J->nlamps && up >= J->lamps[j].cdf) ++j;
            const Lamp* L = &J->lamps[j];
            const double lu = number(stream, sample, depth, LIGHT_U), lv = number(stream, sample, depth, LIGHT_V);
            const V yp = add(add(L->at, mul(L->along, lu)), mul(L->across, lv));
            const V toward = sub(yp, xp);
            const double dist = len(toward);
            const V in = mul(toward, 1 / dist);
            const double cl = -dot(L->facing, in), cs = dot(n, in);
            if (cl > 0 && cs > 0 && dist > 2 * lift) {
                double f[4], Le[4], c[4], any = 0;
                rough_value(m, nm, n, out, in, f);
                or_spectrum_of(L->radiance, nm, Le);
                const double pdfL = L->pick * dist * dist / (cl * L->area);
                const double w = power_weight(pdfL, rough_density(m, n, out, in));
                for (int l = 0; l < LANES; ++l) {
                    c[l] = thr[l] * f[l] * cs * Le[l] / pdfL * w;
                    any += c[l];
                }
                if (any > 0) {
                    kdn_hit sh;
                    double st;
                    trace_one(sc, add(xp, mul(n, lift)), in, dist - lift, 0, KDN_NO_MATERIAL, &sh, &st);
                    if (sh.kind == KDN_HIT_MISSED || !(st < dist - lift)) {
                        for (int l = 0; l < LANES; ++l) rad[l] += c[l];
                    }
                }
            }
        }
        if (depth + 1 >= (int)J->surfaces) break;
        const double ulobe = number(stream, sample, depth, LOBE), uside = number(stream, sample, depth, SIDE);
        const double du = number(stream, sample, depth, DIR_U), dv = number(stream, sample, depth, DIR_V);
        const Sc s = m->kind == KDN_GLASS ? glass_scatter(m, nm, n, out, inside, uside) : rough_scatter(m, nm, n, out, ulobe, du, dv);
        if (!s.found) break;
        const double side = dot(s.in, n);
        if (s.through ? side >= 0 : side <= 0) break;
        double carried[4];
        for (int l = 0; l < LANES; ++l) carried[l] = thr[l] * s.weight[l];
        if (s.collapse && !dispersed) {
            carried[0] *= LANES;
            carried[1] = carried[2] = carried[3] = 0;
            dispersed = 1;
        }
        if (depth >= (int)J->min_depth) {
            const double survival = fmin(0.95, fmax(fmax(carried[0], carried[1]), fmax(carried[2], carried[3])));
            if (!(number(stream, sample, depth, SURVIVAL) < survival)) break;
            for (int l = 0; l < LANES; ++l) carried[l] /= survival;
        }
        for (int l = 0; l < LANES; ++l) thr[l] = carried[l];
        o = add(xp, mul(n, s.through ? -lift : lift));
        d = s.in;
        if (s.through) {
            inside = !inside;
            within = inside ? mi : KDN_NO_MATERIAL;
        }
        pdfBefore = s.exact ? 0 : s.density;
        exactBefore = s.exact;
    }
    or_colour_of(nm, dens, rad, rgb);
}

static void* worker(void* arg) {
    Job* J = (Job*)arg;
    const int w = J->x1 - J->x0;
    for (;;) {
        pthread_mutex_lock(&J->lock);
        const int row = J->next_row++;
        pthread_mutex_unlock(&J->lock);
        if (row >= J->y1) break;
        for (int x = J->x0; x < J->x1; ++x) {
            double* px = J->rgb + 3 * ((size_t)(row - J->y0) * w + (x - J->x0));
            for (uint32_t k = 0; k < J->spp; ++k) {
                double c[3];
                sample_path(J, x, row, J->first + k, c);
                px[0] += c[0];
                px[1] += c[1];
                px[2] += c[2];
            }
        }
    }
    return NULL;
}

void or_render64(const or_render_scene* scene, int width, int height, int x0, int y0, int x1, int y1, uint32_t first, uint32_t spp,
                 uint32_t seed, uint32_t surfaces, uint32_t min_depth, double* rgb, int threads) {
    or_render64_sequenced(scene, width, height, x0, y0, x1, y1, first, spp, seed, surfaces, min_depth, 0, rgb, threads);
}

void or_render64_sequenced(const or_render_scene* scene, int width, int height, int x0, int y0, int x1, int y1, uint32_t first,
                           uint32_t spp, uint32_t seed, uint32_t surfaces, uint32_t min_depth, int sobol, double* rgb, int threads) {
    Job J;
    memset(&J, 0, sizeof(J));
    J.sobol = sobol;
    J.scene = scene;
    J.nmat = scene->material_count;
    J.mats = (Mat*)calloc(J.nmat ? J.nmat : 1, sizeof(Mat));
    for (uint32_t i = 0; i < J.nmat; ++i) {
        const kdn_material* s = &scene->materials[i];
        Mat* m = &J.mats[i];
        m->kind = s->kind;
        for (int k = 0; k < 3; ++k) {
            m->albedo[k] = s->albedo[k];
            m->glow[k] = s->glow[k] * s->strength;
        }
        const double r = clampd(s->kind == KDN_MATTE ? 1.0 : s->roughness, 0, 1);
        m->width = fmax(r * r, 1e-3);
        if (s->kind == KDN_METAL) {
            escape_of(m->width, m->esc, &m->esc_avg);
        } else {
            for (int v = 0; v < VIEWS; ++v) m->esc[v] = 1;
            m->esc_avg = 1;
        }
        m->share = clampd(m->esc_avg, 0.25, 1.0);
        m->exact = s->kind == KDN_GLASS && s->roughness <= 0;
        m->index = s->refraction;
        m->spread = s->dispersion;
    }
    J.lamps = (Lamp*)calloc(scene->light_count ? scene->light_count : 1, sizeof(Lamp));
    double total = 0;
    for (uint32_t i = 0; i < scene->light_count; ++i) {
        const kdn_light* l = &scene->lights[i];
        if (l->kind != KDN_LIGHT_QUAD) continue;
        Lamp* L = &J.lamps[J.nlamps];
        L->at = vv(l->at[0], l->at[1], l->at[2]);
        L->along = vv(l->along[0], l->along[1], l->along[2]);
        L->across = vv(l->across[0], l->across[1], l->across[2]);
        const V f = cross(L->along, L->across);
        L->area = len(f);
        if (!(L->area > 1e-12)) continue;
        L->facing = mul(f, 1 / L->area);
        for (int k = 0; k < 3; ++k) L->radiance[k] = l->power[k] / L->area;
        L->pick = fmax((l->power[0] + l->power[1] + l->power[2]) / 3.0, 0) * PI;
        total += L->pick;
        J.nlamps++;
    }
    double cdf = 0;
    for (uint32_t i = 0; i < J.nlamps; ++i) {
        J.lamps[i].pick = total > 0 ? J.lamps[i].pick / total : 1.0 / J.nlamps;
        cdf += J.lamps[i].pick;
        J.lamps[i].cdf = cdf;
    }
    if (J.nlamps) J.lamps[J.nlamps - 1].cdf = 1;
    J.width = width;
    J.height = height;
    J.x0 = x0;
    J.y0 = y0;
    J.x1 = x1;
    J.y1 = y1;
    J.first = first;
    J.spp = spp;
    J.surfaces = surfaces;
    J.min_depth = min_depth;
    J.seed_bits = or_mixed_bits(seed);
    const kdn_camera* c = &scene->camera;
    const V eye = vv(c->eye[0], c->eye[1], c->eye[2]), target = vv(c->target[0], c->target[1], c->target[2]);
    J.eye = eye;
    J.fwd = norm(sub(target, eye));
    J.right = norm(cross(J.fwd, vv(c->up[0], c->up[1], c->up[2])));
    J.up = cross(J.right, J.fwd);
    J.reach = tan((double)c->fov / 2);
    J.aspect = (double)width / height;
    J.rgb = rgb;
    J.next_row = y0;
    pthread_mutex_init(&J.lock, NULL);
    if (threads < 1) threads = 1;
    pthread_t* th = (pthread_t*)calloc(threads, sizeof(pthread_t));
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, worker, &J);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&J.lock);
    free(J.mats);
    free(J.lamps);
}
