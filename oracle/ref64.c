/*
 * FP64 restatement of the reference hot path — TEST INFRASTRUCTURE ONLY
 * (see oracle.h). Each function names the reference code it follows.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static const double kPi = 3.14159265358979323846;

static double dmin(double a, double b) { return b < a ? b : a; } /* std::min */
static double dmax(double a, double b) { return a < b ? b : a; } /* std::max */

/* ---- point evaluation: runStep (source/sdf/program.cpp:61-98) over the
 * double operations of source/sdf/numbers.cpp:107-188 ---- */
static double d_sign(double a) { return a > 0 ? 1.0 : (a < 0 ? -1.0 : 0.0); }
static double d_truth(int b) { return b ? 1.0 : 0.0; }
static double d_smin(double a, double b, double r) {
    const double blend = dmax(r - fabs(a - b), 0.0) / r;
    return dmin(a, b) - blend * blend * r * 0.25;
}
static double d_pos2(double a) {
    const double p = dmax(a, 0.0);
    return p * p;
}
static double d_compare(uint8_t op, double a, double b) {
    switch (op) {
        case KDN_LESS: return d_truth(a < b);
        case KDN_LESS_EQUAL: return d_truth(a <= b);
        case KDN_EQUAL: return d_truth(a == b);
        case KDN_NOT_EQUAL: return d_truth(a != b);
        case KDN_AND: return d_truth(a != 0 && b != 0);
        case KDN_OR: return d_truth(a != 0 || b != 0);
        default: return d_truth(a == 0);
    }
}

double or_eval64(const kdn_step* steps, const double* values, uint32_t count, uint32_t result, const double p[3], double* r) {
    for (uint32_t i = 0; i < count; ++i) {
        const kdn_step* s = &steps[i];
        double v;
        switch (s->op) {
            case KDN_CONSTANT:
            case KDN_ARGUMENT: v = values[i]; break;
            case KDN_POINT_X: v = p[0]; break;
            case KDN_POINT_Y: v = p[1]; break;
            case KDN_POINT_Z: v = p[2]; break;
            case KDN_NORMAL_X:
            case KDN_NORMAL_Y:
            case KDN_NORMAL_Z: v = 0; break;
            case KDN_ADD: v = r[s->a] + r[s->b]; break;
            case KDN_SUBTRACT: v = r[s->a] - r[s->b]; break;
            case KDN_MULTIPLY: v = r[s->a] * r[s->b]; break;
            case KDN_DIVIDE: v = r[s->a] / r[s->b]; break;
            case KDN_NEGATE: v = -r[s->a]; break;
            case KDN_SQUARE_ROOT: v = sqrt(r[s->a]); break;
            case KDN_ABSOLUTE: v = fabs(r[s->a]); break;
            case KDN_FLOOR: v = floor(r[s->a]); break;
            case KDN_SINE: v = sin(r[s->a]); break;
            case KDN_COSINE: v = cos(r[s->a]); break;
            case KDN_SIGN: v = d_sign(r[s->a]); break;
            case KDN_ARC_TANGENT: v = atan2(r[s->a], r[s->b]); break;
            case KDN_POWER: v = pow(r[s->a], r[s->b]); break;
            case KDN_MINIMUM: v = dmin(r[s->a], r[s->b]); break;
            case KDN_MAXIMUM: v = dmax(r[s->a], r[s->b]); break;
            case KDN_SMOOTH_MINIMUM: v = d_smin(r[s->a], r[s->b], r[s->c]); break;
            case KDN_POSITIVE_SQUARE: v = d_pos2(r[s->a]); break;
            case KDN_SELECT: v = r[s->a] != 0 ? r[s->b] : r[s->c]; break;
            default: v = d_compare(s->op, r[s->a], r[s->b]); break;
        }
        r[i] = v;
    }
    return r[result];
}

/* ---- intervals: source/sdf/numbers.cpp:198-430, round to nearest ---- */
typedef struct {
    double lo, hi;
} I;
static I mk(double lo, double hi) {
    I r = {lo, hi};
    return r;
}
static I hull4(double a, double b, double c, double d) {
    /* std::min / std::max over an initializer list, in order */
    double lo = a, hi = a;
    if (b < lo) lo = b;
    if (c < lo) lo = c;
    if (d < lo) lo = d;
    if (hi < b) hi = b;
    if (hi < c) hi = c;
    if (hi < d) hi = d;
    return mk(lo, hi);
}
static int zero_in(I a) { return a.lo <= 0 && a.hi >= 0; }
static int excludes_zero(I a) { return a.lo > 0 || a.hi < 0; }
static int is_zero(I a) { return a.lo == 0 && a.hi == 0; }
static I i_mul(I a, I b) { return hull4(a.lo * b.lo, a.lo * b.hi, a.hi * b.lo, a.hi * b.hi); }
static I i_abs(I a) {
    if (a.lo >= 0) return a;
    if (a.hi <= 0) return mk(-a.hi, -a.lo);
    return mk(0, dmax(-a.lo, a.hi));
}
static int reaches(I a, double offset) {
    const double k = ceil((a.lo - offset) / (2 * kPi));
    return offset + 2 * kPi * k <= a.hi;
}
static I i_sin(I a) {
    if (a.hi - a.lo >= 2 * kPi) return mk(-1, 1);
    const double sl = sin(a.lo), sh = sin(a.hi);
    I out = mk(dmin(sl, sh), dmax(sl, sh));
    out.hi = reaches(a, kPi / 2) ? 1 : out.hi;
    out.lo = reaches(a, -kPi / 2) ? -1 : out.lo;
    return out;
}
static I i_div(I a, I b) {
    if (zero_in(b)) return mk(-INFINITY, INFINITY);
    return i_mul(a, mk(1 / b.hi, 1 / b.lo));
}
static I whole_power(I a, double n) {
    if (n >= 0 && fmod(n, 2.0) == 0) {
        const I m = i_abs(a);
        return mk(pow(m.lo, n), pow(m.hi, n));
    }
    if (n > 0) return mk(pow(a.lo, n), pow(a.hi, n));
    return i_div(mk(1, 1), whole_power(a, -n));
}
static I decided(int yes, int no) { return mk(yes ? 1.0 : 0.0, no ? 0.0 : 1.0); }
static double settled(double v, double fallback) { return isnan(v) ? fallback : v; }

void or_interval64(const kdn_step* steps, const double* values, uint32_t count, const double box[6], double* lo, double* hi) {
    I* r = (I*)malloc(sizeof(I) * (count ? count : 1));
    for (uint32_t i = 0; i < count; ++i) {
        const kdn_step* s = &steps[i];
        const I A = s->a < i ? r[s->a] : mk(0, 0);
        const I B = s->b < i ? r[s->b] : mk(0, 0);
        const I Cc = s->c < i ? r[s->c] : mk(0, 0);
        I v;
        switch (s->op) {
            case KDN_CONSTANT:
            case KDN_ARGUMENT: v = mk(values[i], values[i]); break;
            case KDN_POINT_X: v = mk(box[0], box[1]); break;
            case KDN_POINT_Y: v = mk(box[2], box[3]); break;
            case KDN_POINT_Z: v = mk(box[4], box[5]); break;
            case KDN_NORMAL_X:
            case KDN_NORMAL_Y:
            case KDN_NORMAL_Z: v = mk(0, 0); break;
            case KDN_ADD: v = mk(A.lo + B.lo, A.hi + B.hi); break;
            case KDN_SUBTRACT: v = mk(A.lo - B.hi, A.hi - B.lo); break;
            case KDN_MULTIPLY: v = i_mul(A, B); break;
            case KDN_DIVIDE: v = i_div(A, B); break;
            case KDN_NEGATE: v = mk(-A.hi, -A.lo); break;
            case KDN_SQUARE_ROOT: v = mk(sqrt(dmax(A.lo, 0.0)), sqrt(dmax(A.hi, 0.0))); break;
            case KDN_ABSOLUTE: v = i_abs(A); break;
            case KDN_FLOOR: v = mk(floor(A.lo), floor(A.hi)); break;
            case KDN_SINE: v = i_sin(A); break;
            case KDN_COSINE: v = i_sin(mk(A.lo + kPi / 2, A.hi + kPi / 2)); break;
            case KDN_SIGN: v = mk(d_sign(A.lo), d_sign(A.hi)); break;
            case KDN_ARC_TANGENT: {
                const int nearOrigin = zero_in(B) && zero_in(A);
                const int acrossCut = B.lo < 0 && zero_in(A);
                v = (nearOrigin || acrossCut) ? mk(-kPi, kPi)
                                              : hull4(atan2(A.lo, B.lo), atan2(A.lo, B.hi), atan2(A.hi, B.lo), atan2(A.hi, B.hi));
                break;
            }
            case KDN_POWER:
                if (B.lo == B.hi && B.lo == floor(B.lo)) v = whole_power(A, B.lo);
                else if (A.lo <= 0) v = mk(-INFINITY, INFINITY);
                else v = hull4(pow(A.lo, B.lo), pow(A.lo, B.hi), pow(A.hi, B.lo), pow(A.hi, B.hi));
                break;
            case KDN_MINIMUM: v = mk(dmin(A.lo, B.lo), dmin(A.hi, B.hi)); break;
            case KDN_MAXIMUM: v = mk(dmax(A.lo, B.lo), dmax(A.hi, B.hi)); break;
            case KDN_SMOOTH_MINIMUM: {
                const double l = d_smin(A.lo, B.lo, Cc.lo), h = d_smin(A.hi, B.hi, Cc.lo);
                v = mk(settled(l, dmin(A.lo, B.lo) - Cc.lo / 4), settled(h, dmin(A.hi, B.hi)));
                break;
            }
            case KDN_POSITIVE_SQUARE: v = mk(d_pos2(A.lo), d_pos2(A.hi)); break;
            case KDN_SELECT:
                v = excludes_zero(A) ? B : (is_zero(A) ? Cc : mk(dmin(B.lo, Cc.lo), dmax(B.hi, Cc.hi)));
                break;
            default: {
                const int apart = A.hi < B.lo || B.hi < A.lo;
                const int same = A.lo == A.hi && B.lo == B.hi && A.lo == B.lo;
                switch (s->op) {
                    case KDN_LESS: v = decided(A.hi < B.lo, A.lo >= B.hi); break;
                    case KDN_LESS_EQUAL: v = decided(A.hi <= B.lo, A.lo > B.hi); break;
                    case KDN_EQUAL: v = decided(same, apart); break;
                    case KDN_NOT_EQUAL: v = decided(apart, same); break;
                    case KDN_AND: v = decided(excludes_zero(A) && excludes_zero(B), is_zero(A) || is_zero(B)); break;
                    case KDN_OR: v = decided(excludes_zero(A) || excludes_zero(B), is_zero(A) && is_zero(B)); break;
                    default: {
                        const int ap = A.hi < 0 || 0 < A.lo;
                        const int sm = A.lo == A.hi && A.lo == 0;
                        v = decided(sm, ap);
                    }
                }
            }
        }
        r[i] = v;
        lo[i] = v.lo;
        hi[i] = v.hi;
    }
    free(r);
}

/* ---- duals: source/sdf/numbers.cpp:433-590 ---- */
typedef struct {
    double v, g[3];
} D;
static D cd(double v) {
    D d = {v, {0, 0, 0}};
    return d;
}
static D ch1(D a, double f, double d1) {
    D o = {f, {d1 * a.g[0], d1 * a.g[1], d1 * a.g[2]}};
    return o;
}
static D ch2(D a, D b, double f, double fa, double fb) {
    D o = {f, {fa * a.g[0] + fb * b.g[0], fa * a.g[1] + fb * b.g[1], fa * a.g[2] + fb * b.g[2]}};
    return o;
}
static double clamp01(double v) { return v < 0 ? 0 : (v > 1 ? 1 : v); }

double or_dual64(const kdn_step* steps, const double* values, uint32_t count, uint32_t result, const double p[3], double g[3],
                 double* regs4) {
    D* r = (D*)regs4;
    for (uint32_t i = 0; i < count; ++i) {
        const kdn_step* s = &steps[i];
        D v;
        switch (s->op) {
            case KDN_CONSTANT:
            case KDN_ARGUMENT: v = cd(values[i]); break;
            case KDN_POINT_X: v = cd(p[0]); v.g[0] = 1; break;
            case KDN_POINT_Y: v = cd(p[1]); v.g[1] = 1; break;
            case KDN_POINT_Z: v = cd(p[2]); v.g[2] = 1; break;
            case KDN_NORMAL_X:
            case KDN_NORMAL_Y:
            case KDN_NORMAL_Z: v = cd(0); break;
            case KDN_ADD: v = ch2(r[s->a], r[s->b], r[s->a].v + r[s->b].v, 1, 1); break;
            case KDN_SUBTRACT: v = ch2(r[s->a], r[s->b], r[s->a].v - r[s->b].v, 1, -1); break;
            case KDN_MULTIPLY: v = ch2(r[s->a], r[s->b], r[s->a].v * r[s->b].v, r[s->b].v, r[s->a].v); break;
            case KDN_DIVIDE: {
                const double b = r[s->b].v;
                v = ch2(r[s->a], r[s->b], r[s->a].v / b, 1 / b, -r[s->a].v / (b * b));
                break;
            }
            case KDN_NEGATE: v = ch1(r[s->a], -r[s->a].v, -1); break;
            case KDN_SQUARE_ROOT: {
                const double f = sqrt(r[s->a].v);
                v = ch1(r[s->a], f, f > 0 ? 0.5 / f : 0);
                break;
            }
            case KDN_ABSOLUTE: v = r[s->a].v < 0 ? ch1(r[s->a], -r[s->a].v, -1) : r[s->a]; break;
            case KDN_FLOOR: v = cd(floor(r[s->a].v)); break;
            case KDN_SINE: v = ch1(r[s->a], sin(r[s->a].v), cos(r[s->a].v)); break;
            case KDN_COSINE: v = ch1(r[s->a], cos(r[s->a].v), -sin(r[s->a].v)); break;
            case KDN_SIGN: v = cd(d_sign(r[s->a].v)); break;
            case KDN_ARC_TANGENT: {
                const double y = r[s->a].v, x = r[s->b].v, q = x * x + y * y;
                v = ch2(r[s->a], r[s->b], atan2(y, x), x / q, -y / q);
                break;
            }
            case KDN_POWER: {
                const double a = r[s->a].v, b = r[s->b].v, f = pow(a, b);
                v = ch2(r[s->a], r[s->b], f, b * pow(a, b - 1), f * (a > 0 ? log(a) : 0));
                break;
            }
            case KDN_MINIMUM: v = r[s->a].v <= r[s->b].v ? r[s->a] : r[s->b]; break;
            case KDN_MAXIMUM: v = r[s->a].v >= r[s->b].v ? r[s->a] : r[s->b]; break;
            case KDN_SMOOTH_MINIMUM: {
                const double rr = r[s->c].v;
                const double fa = clamp01((rr - (r[s->a].v - r[s->b].v)) / (2 * rr));
                v = ch2(r[s->a], r[s->b], d_smin(r[s->a].v, r[s->b].v, rr), fa, 1 - fa);
                break;
            }
            case KDN_POSITIVE_SQUARE: v = ch1(r[s->a], d_pos2(r[s->a].v), 2 * dmax(r[s->a].v, 0.0)); break;
            case KDN_SELECT: v = r[s->a].v != 0 ? r[s->b] : r[s->c]; break;
            default: v = cd(d_compare(s->op, r[s->a].v, r[s->b].v)); break;
        }
        r[i] = v;
    }
    g[0] = r[result].g[0];
    g[1] = r[result].g[1];
    g[2] = r[result].g[2];
    return r[result].v;
}

/* ---- the field straight from the scene tree (source/sdf/scene.cpp:36-75)
 * with the standard.sdf primitives in closed form (scenes/standard.sdf) ---- */
static double len3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }
static double len2(double x, double y) { return sqrt(x * x + y * y); }

static double primitive(const kdn_leaf* leaf, const double q[3]) {
    const double* a = leaf->args;
    switch (leaf->primitive) {
        case KDN_PRIM_SPHERE: return len3(q[0], q[1], q[2]) - a[0];
        case KDN_PRIM_BOX: {
            double m[3], mx = -INFINITY;
            for (int k = 0; k < 3; ++k) {
                m[k] = fabs(q[k]) - a[k] / 2 + a[3];
                mx = dmax(mx, m[k]);
            }
            return len3(dmax(m[0], 0), dmax(m[1], 0), dmax(m[2], 0)) + dmin(mx, 0.0) - a[3];
        }
        case KDN_PRIM_TORUS: return len2(len2(q[0], q[1]) - a[0], q[2]) - a[1];
        case KDN_PRIM_PLANE: return q[0] * a[0] + q[1] * a[1] + q[2] * a[2] - a[3];
        case KDN_PRIM_CYLINDER: {
            const double dx = fabs(len2(q[0], q[1])) - a[1], dz = fabs(q[2]) - a[0] / 2;
            return dmin(dmax(dx, dz), 0.0) + len2(dmax(dx, 0), dmax(dz, 0));
        }
        case KDN_PRIM_CAPSULE: {
            const double pa[3] = {q[0] - a[0], q[1] - a[1], q[2] - a[2]}, ba[3] = {a[3] - a[0], a[4] - a[1], a[5] - a[2]};
            double t = (pa[0] * ba[0] + pa[1] * ba[1] + pa[2] * ba[2]) / (ba[0] * ba[0] + ba[1] * ba[1] + ba[2] * ba[2]);
            t = t < 0 ? 0 : (t > 1 ? 1 : t);
            return len3(pa[0] - ba[0] * t, pa[1] - ba[1] * t, pa[2] - ba[2] * t) - a[6];
        }
        case KDN_PRIM_ELLIPSOID:
            return (len3(q[0] / a[0], q[1] / a[1], q[2] / a[2]) - 1) * dmin(a[0], dmin(a[1], a[2]));
    }
    return 1e30;
}

static double leaf_distance(const kdn_leaf* leaf, const double p[3]) {
    const double d[3] = {p[0] - leaf->origin[0], p[1] - leaf->origin[1], p[2] - leaf->origin[2]};
    double local[3];
    for (int i = 0; i < 3; ++i) local[i] = leaf->linear[3 * i] * d[0] + leaf->linear[3 * i + 1] * d[1] + leaf->linear[3 * i + 2] * d[2];
    return primitive(leaf, local) * leaf->stretch;
}

static double smooth_min(double a, double b, double radius) { return radius <= 0 ? dmin(a, b) : d_smin(a, b, radius); }

static double join_pair(const kdn_node* n, double a, double b) {
    switch (n->join) {
        case 1: return smooth_min(a, b, n->radius);
        case 2: return -smooth_min(-a, b, n->radius);
        case 3: return -smooth_min(-a, -b, n->radius);
        default: return a;
    }
}

static double node_distance(const kdn_leaf* leaves, const kdn_node* nodes, const uint32_t* children, uint32_t index, const double p[3]) {
    const kdn_node* n = &nodes[index];
    if (n->join == 0) return leaf_distance(&leaves[n->leaf], p);
    double d = node_distance(leaves, nodes, children, children[n->first_child], p);
    for (uint32_t i = 1; i < n->child_count; ++i) d = join_pair(n, d, node_distance(leaves, nodes, children, children[n->first_child + i], p));
    return d;
}

double or_scene_distance(const kdn_leaf* leaves, const kdn_node* nodes, const uint32_t* children, const uint32_t* tops,
                         uint32_t top_count, const double p[3]) {
    double d = 1e30;
    for (uint32_t i = 0; i < top_count; ++i) d = dmin(d, node_distance(leaves, nodes, children, tops[i], p));
    return d;
}

/* ---- the FP64 sphere trace (plugins/geometry/sphere_trace/sphere_trace.cpp) ---- */
#define SPHERE_TOLERANCE 1e-6
#define SPHERE_INSET (2 * SPHERE_TOLERANCE)
#define SPHERE_TIE 1e-12
#define SPHERE_GRAZE 0.02
#define SPHERE_STEPS (1 << 16)

typedef struct {
    const or_scene64* s;
    const kdn_ray* ray;
    double o[3], d[3];
    double* vals;
    double* regs;
} March;

static void point_at(const March* m, double t, double p[3]) {
    for (int k = 0; k < 3; ++k) p[k] = m->o[k] + m->d[k] * t;
}

static void values_at(const March* m, double t, double* out) {
    double p[3];
    point_at(m, t, p);
    for (uint32_t j = 0; j < m->s->object_count; ++j) {
        const kdn_object* ob = &m->s->objects[j];
        out[j] = or_eval64(m->s->steps + ob->first, m->s->values + ob->first, ob->count, ob->result, p, m->regs);
    }
}

static uint32_t least_of(const double* v, uint32_t n) {
    uint32_t least = 0;
    for (uint32_t k = 0; k < n; ++k) least = v[k] < v[least] ? k : least;
    return least;
}

/* sphereField + keptSourceOf at p */
static void field_at(const March* m, uint32_t object, const double p[3], double g[3], uint32_t* leaf, uint32_t* other, double* share) {
    const kdn_object* ob = &m->s->objects[object];
    const kdn_step* prog = m->s->steps + ob->first;
    D* r = (D*)m->regs;
    or_dual64(prog, m->s->values + ob->first, ob->count, ob->result, p, g, m->regs);
    /* laneSourceOf (program.cpp:482-502) over the dual values */
    *other = KDN_NO_LEAF;
    *share = 0;
    uint32_t at = ob->result;
    for (;;) {
        const kdn_step* s = &prog[at];
        const int join = s->op == KDN_MINIMUM || s->op == KDN_SMOOTH_MINIMUM || s->op == KDN_MAXIMUM || s->op == KDN_NEGATE ||
                         s->op == KDN_SELECT;
        if (!join) break;
        const double a = r[s->a].v, b = r[s->b].v;
        uint32_t next;
        switch (s->op) {
            case KDN_MINIMUM:
            case KDN_SMOOTH_MINIMUM: next = a <= b ? s->a : s->b; break;
            case KDN_MAXIMUM: next = a >= b ? s->a : s->b; break;
            case KDN_SELECT: next = a != 0 ? s->b : s->c; break;
            default: next = s->a;
        }
        double sh = 0;
        if (s->op == KDN_SMOOTH_MINIMUM) {
            const double radius = r[s->c].v, apart = fabs(a - b);
            sh = radius > 0 && apart < radius ? 0.5 * (radius - apart) / radius : 0;
        }
        if (sh > 0 && *other == KDN_NO_LEAF) {
            uint32_t o2 = next == s->a ? s->b : s->a;
            for (;;) {
                const kdn_step* q = &prog[o2];
                const int j2 = q->op == KDN_MINIMUM || q->op == KDN_SMOOTH_MINIMUM || q->op == KDN_MAXIMUM || q->op == KDN_NEGATE ||
                               q->op == KDN_SELECT;
                if (!j2) break;
                const double qa = r[q->a].v, qb = r[q->b].v;
                switch (q->op) {
                    case KDN_MINIMUM:
                    case KDN_SMOOTH_MINIMUM: o2 = qa <= qb ? q->a : q->b; break;
                    case KDN_MAXIMUM: o2 = qa >= qb ? q->a : q->b; break;
                    case KDN_SELECT: o2 = qa != 0 ? q->b : q->c; break;
                    default: o2 = q->a;
                }
            }
            *other = prog[o2].leaf;
            *share = sh;
        }
        at = next;
    }
    *leaf = prog[at].leaf;
}

static void normal_of(const kdn_ray* ray, const double g[3], const double d[3], double n[3]) {
    const double len = len3(g[0], g[1], g[2]);
    if (!(len > 0)) {
        n[0] = -d[0];
        n[1] = -d[1];
        n[2] = -d[2];
        return;
    }
    double u[3] = {g[0] / len, g[1] / len, g[2] / len};
    const double along = u[0] * d[0] + u[1] * d[1] + u[2] * d[2];
    if (ray->inside) {
        const double s = along < 0 ? -1 : 1;
        for (int k = 0; k < 3; ++k) n[k] = s * u[k];
        return;
    }
    if (along > 0) {
        double w[3];
        for (int k = 0; k < 3; ++k) w[k] = u[k] - d[k] * (along + SPHERE_GRAZE);
        const double l = len3(w[0], w[1], w[2]);
        for (int k = 0; k < 3; ++k) n[k] = w[k] / l;
        return;
    }
    for (int k = 0; k < 3; ++k) n[k] = u[k];
}

static void hit_at(const March* m, double t, const double g[3], uint32_t leaf, uint32_t other, double share, kdn_hit* hit, double* t64) {
    double p[3], n[3];
    point_at(m, t, p);
    normal_of(m->ray, g, m->d, n);
    hit->kind = KDN_HIT_SURFACE;
    hit->distance = (float)t;
    for (int k = 0; k < 3; ++k) {
        hit->point[k] = (float)p[k];
        hit->normal[k] = (float)n[k];
    }
    hit->leaf = leaf;
    hit->other = other;
    hit->share = (float)share;
    hit->cell = KDN_NO_LEAF;
    *t64 = t;
}

static void landed(const March* m, const double* values, double t, kdn_hit* hit, double* t64) {
    const uint32_t n = m->s->object_count;
    double p[3];
    point_at(m, t, p);
    const double band = SPHERE_TIE * (1 + dmax(dmax(fabs(p[0]), fabs(p[1])), fabs(p[2])));
    const uint32_t first = least_of(values, n);
    double* on = (double*)malloc(sizeof(double) * n);
    values_at(m, t + SPHERE_INSET, on);
    uint32_t winner = first;
    for (uint32_t k = 0; k < n; ++k) winner = values[k] - values[first] <= band && on[k] < on[winner] ? k : winner;
    free(on);
    double g[3], share;
    uint32_t leaf, other;
    field_at(m, winner, p, g, &leaf, &other, &share);
    hit_at(m, t, g, leaf, other, share, hit, t64);
}

static void trace64(const or_scene64* s, const kdn_ray* ray, kdn_hit* hit, double* t64, double* vals, double* regs) {
    March m;
    m.s = s;
    m.ray = ray;
    for (int k = 0; k < 3; ++k) {
        m.o[k] = ray->origin[k];
        m.d[k] = ray->direction[k];
    }
    m.vals = vals;
    m.regs = regs;
    memset(hit, 0, sizeof(*hit));
    hit->kind = KDN_HIT_MISSED;
    hit->distance = INFINITY;
    hit->leaf = hit->other = hit->cell = KDN_NO_LEAF;
    *t64 = INFINITY;
    /* sphereSpan (sphere_trace.cpp:55-69) */
    double lower = 0, upper = ray->far_;
    for (int k = 0; k < 3; ++k) {
        const double part = m.d[k];
        const double slope = 1 / (fabs(part) < 1e-12 ? copysign(1e-12, part) : part);
        const double a = (s->lower[k] - m.o[k]) * slope, b = (s->lower[k] + s->size[k] - m.o[k]) * slope;
        lower = dmax(lower, dmin(a, b));
        upper = dmin(upper, dmax(a, b));
    }
    if (s->object_count == 0 || upper < lower) return;
    const uint32_t n = s->object_count;
    double t = lower;
    if (!ray->inside) {
        for (int step = 0; step < SPHERE_STEPS && t <= upper; ++step) {
            values_at(&m, t, vals);
            const double least = vals[least_of(vals, n)];
            if (least < SPHERE_TOLERANCE) {
                landed(&m, vals, t, hit, t64);
                return;
            }
            t += least;
        }
        return;
    }
    double* past = (double*)malloc(sizeof(double) * n);
    for (int step = 0; step < SPHERE_STEPS && t <= upper; ++step) {
        values_at(&m, t, vals);
        double nearest = fabs(vals[0]);
        for (uint32_t k = 0; k < n; ++k) nearest = dmin(nearest, fabs(vals[k]));
        if (nearest < SPHERE_TOLERANCE) {
            values_at(&m, t + SPHERE_INSET, past);
            const uint32_t least = least_of(past, n);
            if (past[least] > 0) {
                landed(&m, vals, t, hit, t64);
                free(past);
                return;
            }
            double p[3], g[3], share;
            uint32_t leaf, other;
            point_at(&m, t + SPHERE_INSET, p);
            field_at(&m, least, p, g, &leaf, &other, &share);
            if (leaf != KDN_NO_LEAF && s->leaf_material[leaf] != ray->within) {
                hit_at(&m, t, g, leaf, other, share, hit, t64);
                free(past);
                return;
            }
        }
        t += nearest < SPHERE_TOLERANCE ? SPHERE_INSET : nearest;
    }
    free(past);
}

void or_sphere_trace64(const or_scene64* s, const kdn_ray* rays, uint32_t count, kdn_hit* hits, double* distances64) {
    uint32_t maxL = 1;
    for (uint32_t j = 0; j < s->object_count; ++j) maxL = s->objects[j].count > maxL ? s->objects[j].count : maxL;
    double* vals = (double*)malloc(sizeof(double) * (s->object_count + 1));
    double* regs = (double*)malloc(sizeof(double) * 4 * maxL);
    for (uint32_t i = 0; i < count; ++i) {
        double t64;
        trace64(s, &rays[i], &hits[i], &t64, vals, regs);
        if (distances64) distances64[i] = t64;
    }
    free(vals);
    free(regs);
}

/* ---- counter-based random numbers (source/math/random.cpp:8-21) ---- */
uint64_t or_mixed_bits(uint64_t value) {
    value += 0x9E3779B97F4A7C15ULL;
    value = (value ^ (value >> 30)) * 0xBF58476D1CE4E5B9ULL;
    value = (value ^ (value >> 27)) * 0x94D049BB133111EBULL;
    return value ^ (value >> 31);
}

double or_hashed_number(uint64_t stream, uint64_t index, uint64_t coordinate) {
    const uint64_t bits = or_mixed_bits(stream + or_mixed_bits(index + or_mixed_bits(coordinate)));
    return (double)(bits >> 11) * 0x1p-53;
}
