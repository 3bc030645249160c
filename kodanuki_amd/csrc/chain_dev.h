// The highlight survey's chain walk, in FP64: a point walked over one shiny
// leaf until the leaf's normal halves the directions to the start and to a
// lamp, then how the highlight spreads and moves there (walkedChain,
// plugins/sampler/highlight_guide/chain.h:76-97, chain.cpp:157-521).
//
// The leaf is its own program over world space (leafProgram,
// sdf/scene.cpp:284-288), evaluated with dual numbers (value and gradient)
// where the walk puts a point on the surface and with hyper-dual numbers
// (the six Hessian entries too) where it reads a vertex (the rules of
// sdf/numbers.cpp:433-823: a one-operand step by its first and second
// derivative, a two-operand step by its partials; min and max pick an
// operand, ties to the first).
//
// The walk, as the reference takes it: the seed put on the surface by two
// Newton steps along the gradient; then up to 16 damped Newton steps in the
// tangent plane on the mismatch between the unit normal and the unit half
// vector of the directions to the start and to the lamp (the lamp's middle,
// or its direction for a far lamp), its 2x2 Jacobian from the turns of the
// two legs and the leaf's Hessian; a step at most a tenth of the distance to
// the start, halved up to six times while the mismatch grows, each trial
// point put back on the surface. It settles with the mismatch under 1e-6,
// the start and the lamp on one side, the value under 1e-6, outside a
// sphere lamp. The spread: the facets' alpha^2 / 2 and the lamp's own
// spread through the lamp rows, carried back through the inverse Jacobian
// and seen from the start on the plane one unit along the axis (its
// across/up basis from basisAround); the motion: the start's sideways
// pushes carried the same way.
#pragma once

#include "kdn_math.h"
#include "tables.h"

namespace kdn {

constexpr int kLeafMostSteps = 64;  // the longest leaf program walked

struct D3 {
    double x, y, z;
};
KDN_D D3 d3(double x, double y, double z) { return D3{x, y, z}; }
KDN_D D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
KDN_D D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
KDN_D D3 operator*(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
KDN_D double dotd(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KDN_D double lengthd(D3 a) { return __builtin_sqrt(dotd(a, a)); }
KDN_D D3 crossd(D3 a, D3 b) { return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
KDN_D double compd(D3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
KDN_D bool finited(D3 a) { return __builtin_isfinite(a.x + a.y + a.z); }

// basisAround (math/vector.cpp:23-35): the frame of a unit normal
KDN_D void basis_of(D3 n, D3* t, D3* b) {
    const double side = __builtin_copysign(1.0, n.z);
    const double a = -1.0 / (side + n.z);
    const double c = n.x * n.y * a;
    *t = d3(1.0 + side * n.x * n.x * a, side * c, -side * n.x);
    *b = d3(c, side + n.y * n.y * a, -n.y);
}

// ---- dual and hyper-dual numbers --------------------------------------------------
// Num<true>: value, gradient and the Hessian's six entries (xx, xy, xz, yy,
// yz, zz); Num<false>: value and gradient (the projections onto the leaf
// read these alone, as leafSlope does, chain.cpp:140-151).
template <bool H>
struct Num {
    double v, g[3], h[H ? 6 : 1];
};
using HD = Num<true>;
using DL = Num<false>;
// a lane's registers for one leaf program, read either way
union LeafRegs {
    HD hd[kLeafMostSteps];
    DL dl[kLeafMostSteps];
};

KDN_D int hd_row(int k) { return k < 3 ? 0 : (k < 5 ? 1 : 2); }
KDN_D int hd_col(int k) { return k < 3 ? k : (k < 5 ? k - 2 : 2); }
template <bool H>
KDN_D Num<H> num_const(double v) {
    Num<H> o;
    o.v = v;
    for (int i = 0; i < 3; ++i) o.g[i] = 0;
    if constexpr (H)
        for (int k = 0; k < 6; ++k) o.h[k] = 0;
    return o;
}
template <bool H>
KDN_D Num<H> num_one(const Num<H>& a, double f, double d1, double d2) {
    Num<H> o = num_const<H>(f);
    for (int i = 0; i < 3; ++i) o.g[i] = d1 * a.g[i];
    if constexpr (H)
        for (int k = 0; k < 6; ++k) o.h[k] = d1 * a.h[k] + d2 * (a.g[hd_row(k)] * a.g[hd_col(k)]);
    return o;
}
template <bool H>
KDN_D Num<H> num_two(const Num<H>& a, const Num<H>& b, double f, double fa, double fb, double faa, double fab, double fbb) {
    Num<H> o = num_const<H>(f);
    for (int i = 0; i < 3; ++i) o.g[i] = fa * a.g[i] + fb * b.g[i];
    if constexpr (H) {
        for (int k = 0; k < 6; ++k) {
            const int i = hd_row(k), j = hd_col(k);
            const double ga = a.g[i] * a.g[j], gab = a.g[i] * b.g[j] + b.g[i] * a.g[j], gb = b.g[i] * b.g[j];
            o.h[k] = fa * a.h[k] + fb * b.h[k] + faa * ga + fab * gab + fbb * gb;
        }
    }
    return o;
}
KDN_D double truthd(bool t) { return t ? 1.0 : 0.0; }

// One step of a leaf program on (hyper-)dual registers (evaluate,
// sdf/program.cpp:60-92). Normal inputs are zero, as the walk's reads give.
template <bool H>
KDN_D Num<H> num_step(const LeafStepDev& s, const Num<H>* r, const double p[3]) {
    using N = Num<H>;
    const uint32_t op = s.op;
    if (op == KDN_CONSTANT) return num_const<H>(s.value);
    if (op >= KDN_POINT_X && op <= KDN_POINT_Z) {
        N o = num_const<H>(p[op - KDN_POINT_X]);
        o.g[op - KDN_POINT_X] = 1;
        return o;
    }
    if (op == KDN_ARGUMENT || (op >= KDN_NORMAL_X && op <= KDN_NORMAL_Z)) return num_const<H>(0);
    const N& a = r[s.a];
    switch (op) {
        case KDN_NEGATE: return num_one<H>(a, -a.v, -1, 0);
        case KDN_SQUARE_ROOT: {
            const double f = __builtin_sqrt(a.v);
            if (f <= 0) return num_const<H>(f);
            return num_one<H>(a, f, 0.5 / f, -0.25 / (a.v * f));
        }
        case KDN_ABSOLUTE: return a.v < 0 ? num_one<H>(a, -a.v, -1, 0) : a;
        case KDN_FLOOR: return num_const<H>(__builtin_floor(a.v));
        case KDN_SINE: return num_one<H>(a, sin(a.v), cos(a.v), -sin(a.v));
        case KDN_COSINE: return num_one<H>(a, cos(a.v), -sin(a.v), -cos(a.v));
        case KDN_SIGN: return num_const<H>(a.v > 0 ? 1.0 : (a.v < 0 ? -1.0 : 0.0));
        case KDN_NOT: return num_const<H>(truthd(a.v == 0));
        case KDN_POSITIVE_SQUARE: {
            const double q = a.v > 0 ? a.v : 0.0;
            return num_one<H>(a, q * q, 2 * q, a.v > 0 ? 2.0 : 0.0);
        }
        default: break;
    }
    const N& b = r[s.b];
    switch (op) {
        case KDN_ADD: return num_two<H>(a, b, a.v + b.v, 1, 1, 0, 0, 0);
        case KDN_SUBTRACT: return num_two<H>(a, b, a.v - b.v, 1, -1, 0, 0, 0);
        case KDN_MULTIPLY: return num_two<H>(a, b, a.v * b.v, b.v, a.v, 0, 1, 0);
        case KDN_DIVIDE: {
            const double v = b.v;
            return num_two<H>(a, b, a.v / v, 1 / v, -a.v / (v * v), 0, -1 / (v * v), 2 * a.v / (v * v * v));
        }
        case KDN_ARC_TANGENT: {
            const double rr = b.v * b.v + a.v * a.v, xy = b.v * a.v;
            return num_two<H>(a, b, atan2(a.v, b.v), b.v / rr, -a.v / rr, -2 * xy / (rr * rr), (a.v * a.v - b.v * b.v) / (rr * rr),
                              2 * xy / (rr * rr));
        }
        case KDN_POWER: {
            const double v = a.v, n = b.v, f = pow(v, n), l = v > 0 ? log(v) : 0.0, below = pow(v, n - 1);
            return num_two<H>(a, b, f, n * below, f * l, n * (n - 1) * pow(v, n - 2), below * (1 + n * l), f * l * l);
        }
        case KDN_MINIMUM: return a.v <= b.v ? a : b;
        case KDN_MAXIMUM: return a.v >= b.v ? a : b;
        case KDN_SMOOTH_MINIMUM: {
            const double rad = r[s.c].v, d = a.v - b.v;
            const double blend = (rad - __builtin_fabs(d) > 0 ? rad - __builtin_fabs(d) : 0.0) / rad;
            const double f = (a.v <= b.v ? a.v : b.v) - blend * blend * rad * 0.25;
            double fa = (rad - d) / (2 * rad);
            fa = fa < 0 ? 0.0 : (fa > 1 ? 1.0 : fa);
            const double k = __builtin_fabs(d) < rad ? 1 / (2 * rad) : 0.0;
            return num_two<H>(a, b, f, fa, 1 - fa, -k, k, -k);
        }
        case KDN_SELECT: return a.v != 0 ? b : r[s.c];
        case KDN_LESS: return num_const<H>(truthd(a.v < b.v));
        case KDN_LESS_EQUAL: return num_const<H>(truthd(a.v <= b.v));
        case KDN_EQUAL: return num_const<H>(truthd(a.v == b.v));
        case KDN_NOT_EQUAL: return num_const<H>(truthd(a.v != b.v));
        case KDN_AND: return num_const<H>(truthd(a.v != 0 && b.v != 0));
        case KDN_OR: return num_const<H>(truthd(a.v != 0 || b.v != 0));
        default: return num_const<H>(0);
    }
}

// The leaf's value and gradient (and, with H, its Hessian) at p.
template <bool H>
KDN_D Num<H> leaf_read(const LeafStepDev* steps, const LeafProgDev& prog, D3 at, Num<H>* regs) {
    const double p[3] = {at.x, at.y, at.z};
    for (uint32_t k = 0; k < prog.count; ++k) regs[k] = num_step<H>(steps[prog.first + k], regs, p);
    return regs[prog.result];
}

// ---- 3x3 and 2x2 matrices ---------------------------------------------------------
struct M3 {
    double m[3][3];
};
struct M2 {
    double m[2][2];
};
// (I - u u^T) scale
KDN_D M3 across_of(D3 u, double scale) {
    M3 o;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) o.m[r][c] = ((r == c ? 1.0 : 0.0) - compd(u, r) * compd(u, c)) * scale;
    return o;
}
KDN_D M3 zero3() { return across_of(d3(0, 0, 0), 0); }
KDN_D M3 eye3() { return across_of(d3(0, 0, 0), 1); }
KDN_D M3 outer3(D3 v) {
    M3 o;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) o.m[r][c] = compd(v, r) * compd(v, c);
    return o;
}
KDN_D M3 mul3(const M3& a, const M3& b) {
    M3 o = zero3();
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            for (int i = 0; i < 3; ++i) o.m[r][c] += a.m[r][i] * b.m[i][c];
    return o;
}
// left a + b
KDN_D M3 sum3(const M3& a, double left, const M3& b) {
    M3 o;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) o.m[r][c] = left * a.m[r][c] + b.m[r][c];
    return o;
}
KDN_D M3 scale3(const M3& a, double f) { return sum3(a, f, zero3()); }
// the row vector v^T m
KDN_D D3 row_times(D3 v, const M3& m) {
    return d3(v.x * m.m[0][0] + v.y * m.m[1][0] + v.z * m.m[2][0], v.x * m.m[0][1] + v.y * m.m[1][1] + v.z * m.m[2][1],
              v.x * m.m[0][2] + v.y * m.m[1][2] + v.z * m.m[2][2]);
}

// ---- the walk ----------------------------------------------------------------------
struct ChainLamp {
    bool far_;
    D3 place;   // the middle of a near lamp, the unit direction of a far one
    uint32_t kind;
    D3 at;
    double radius;
    M3 spreadNear;  // a quad's spread (it does not turn with the view)
    double cosine;
};

KDN_D ChainLamp chain_lamp(const LampDev& L) {
    ChainLamp o;
    o.kind = L.kind;
    o.at = d3(L.at[0], L.at[1], L.at[2]);
    o.radius = L.radius;
    o.cosine = L.cosine;
    const D3 along = d3(L.along[0], L.along[1], L.along[2]), across = d3(L.across[0], L.across[1], L.across[2]);
    o.far_ = L.kind == KDN_LIGHT_DIRECTIONAL || L.kind == KDN_LIGHT_SUN || L.kind == KDN_LIGHT_MOON;
    if (o.far_) {
        const D3 d = d3(L.direction[0], L.direction[1], L.direction[2]);
        o.place = d * (1 / lengthd(d));
    } else if (L.kind == KDN_LIGHT_QUAD) {
        o.place = o.at + (along + across) * 0.5;
    } else {
        o.place = o.at;
    }
    o.spreadNear = sum3(outer3(along), 1.0 / 12, scale3(outer3(across), 1.0 / 12));
    return o;
}

// lampSpreadOf (chain.cpp:207-221)
KDN_D M3 lamp_spread(const ChainLamp& L, D3 towards) {
    double c = L.cosine < -1 ? -1.0 : (L.cosine > 1 ? 1.0 : L.cosine);
    const double angle = 2 * acos(c);
    if (L.kind == KDN_LIGHT_SPHERE) return across_of(towards, 0.25 * L.radius * L.radius);
    if (L.kind == KDN_LIGHT_QUAD) return L.spreadNear;
    if (L.kind == KDN_LIGHT_SUN || L.kind == KDN_LIGHT_MOON) return across_of(towards, angle * angle / 16);
    return zero3();
}

struct ChainAt {
    D3 point, normal, t0, t1;
    double value, strength;
    M3 hess;
    // the bend there
    double res[2];
    M2 slope;
    D3 start[2], lampRow[2];
    bool sided;
    double error;
};

KDN_D void chain_bend(D3 from, const ChainLamp& L, ChainAt& h) {
    const D3 toStart = from - h.point;
    const double ds = lengthd(toStart);
    const D3 uBefore = toStart * (1 / ds);
    const M3 turnBefore = across_of(uBefore, 1 / ds);
    D3 uAfter;
    M3 turnAfter;
    if (L.far_) {
        uAfter = L.place;
        turnAfter = zero3();
    } else {
        const D3 toLamp = L.place - h.point;
        const double dl = lengthd(toLamp);
        uAfter = toLamp * (1 / dl);
        turnAfter = across_of(uAfter, 1 / dl);
    }
    const M3 lampTurn = L.far_ ? eye3() : turnAfter;
    h.sided = dotd(uBefore, h.normal) * dotd(uAfter, h.normal) > 0;
    const D3 raw = uBefore + uAfter;
    const double side = dotd(raw, h.normal) < 0 ? -1.0 : 1.0;
    const D3 half = raw * side;
    const double hl = lengthd(half);
    const D3 uHalf = half * (1 / hl);
    const M3 turnHalf = across_of(uHalf, 1 / hl);
    const M3 turnNormal = mul3(across_of(h.normal, 1 / h.strength), h.hess);
    const D3 off = uHalf - h.normal;
    const M3 byBefore = scale3(mul3(turnHalf, turnBefore), side);
    const M3 byAfter = scale3(mul3(turnHalf, turnAfter), side);
    const M3 byHere = sum3(sum3(byBefore, 1, byAfter), -1, scale3(turnNormal, -1));
    const M3 byLamp = scale3(mul3(turnHalf, lampTurn), side);
    const D3 t[2] = {h.t0, h.t1};
    h.res[0] = dotd(h.t0, off);
    h.res[1] = dotd(h.t1, off);
    for (int r = 0; r < 2; ++r) {
        const D3 left = row_times(t[r], byHere);
        for (int c = 0; c < 2; ++c) h.slope.m[r][c] = dotd(left, t[c]);
        h.start[r] = row_times(t[r], byBefore);
        h.lampRow[r] = row_times(t[r], byLamp);
    }
    h.error = __builtin_sqrt(h.res[0] * h.res[0] + h.res[1] * h.res[1]);
}

// chainVertexAt / settledVertex (chain.cpp:157-184): the point put on the
// surface by two Newton steps, then read there
KDN_D ChainAt chain_vertex(const LeafStepDev* steps, const LeafProgDev& prog, D3 start, LeafRegs& regs) {
    D3 p = start;
    for (int k = 0; k < 2; ++k) {
        const DL r = leaf_read<false>(steps, prog, p, regs.dl);
        const D3 g = d3(r.g[0], r.g[1], r.g[2]);
        const double steep = dotd(g, g);
        if (!(steep > 0)) break;
        p = p - g * (r.v / steep);
    }
    const HD r = leaf_read<true>(steps, prog, p, regs.hd);
    ChainAt h;
    h.point = p;
    h.value = r.v;
    const D3 g = d3(r.g[0], r.g[1], r.g[2]);
    h.strength = lengthd(g);
    h.normal = g * (1 / h.strength);
    basis_of(h.normal, &h.t0, &h.t1);
    h.hess.m[0][0] = r.h[0];
    h.hess.m[0][1] = h.hess.m[1][0] = r.h[1];
    h.hess.m[0][2] = h.hess.m[2][0] = r.h[2];
    h.hess.m[1][1] = r.h[3];
    h.hess.m[1][2] = h.hess.m[2][1] = r.h[4];
    h.hess.m[2][2] = r.h[5];
    return h;
}

// solvedPair (chain.cpp:298-318): a 2x2 system by partial pivoting
KDN_D bool solve2(const M2& m, double r0, double r1, double* x0, double* x1) {
    double largest = __builtin_fabs(m.m[0][0]);
    largest = __builtin_fmax(largest, __builtin_fabs(m.m[0][1]));
    largest = __builtin_fmax(largest, __builtin_fabs(m.m[1][0]));
    largest = __builtin_fmax(largest, __builtin_fabs(m.m[1][1]));
    const bool swapped = __builtin_fabs(m.m[1][0]) > __builtin_fabs(m.m[0][0]);
    const double t0 = swapped ? m.m[1][0] : m.m[0][0], t1 = swapped ? m.m[1][1] : m.m[0][1];
    const double b0 = swapped ? m.m[0][0] : m.m[1][0], b1 = swapped ? m.m[0][1] : m.m[1][1];
    const double first = swapped ? r1 : r0, second = swapped ? r0 : r1;
    if (!(__builtin_fabs(t0) > 1e-12 * largest)) return false;
    const double factor = b0 / t0, pivot = b1 - factor * t1;
    if (!(__builtin_fabs(pivot) > 1e-12 * largest)) return false;
    const double y = (second - factor * first) / pivot;
    *x0 = (first - t1 * y) / t0;
    *x1 = y;
    return true;
}

struct ChainOut {
    bool found;
    D3 point, axis, across, up;
    double spread[3];
    D3 motion[2];
};

KDN_D ChainOut chain_unsettled() {
    ChainOut o;
    o.found = false;
    o.point = o.axis = o.across = o.up = d3(0, 0, 0);
    o.spread[0] = o.spread[1] = o.spread[2] = 0;
    o.motion[0] = o.motion[1] = d3(0, 0, 0);
    return o;
}

// walkedChain (chain.cpp:507-519): from a start with its two sideways
// directions, over a leaf of facet width `width`, toward lamp L, from seed
KDN_D ChainOut walked_chain(const LeafStepDev* steps, const LeafProgDev& prog, D3 from, D3 side0, D3 side1, double width,
                            const ChainLamp& L, D3 seed, LeafRegs& regs) {
    if (prog.count == 0 || prog.count > (uint32_t)kLeafMostSteps) return chain_unsettled();
    ChainAt now = chain_vertex(steps, prog, seed, regs);
    chain_bend(from, L, now);
    for (int iteration = 0; iteration < 16 && now.error > 1e-10; ++iteration) {
        double s0, s1;
        if (!solve2(now.slope, -now.res[0], -now.res[1], &s0, &s1)) break;
        // cappedStep: at most a tenth of the way to the start
        const double room = 0.1 * lengthd(from - now.point), moved = hypot(s0, s1);
        const double scale = moved > room ? room / moved : 1.0;
        s0 *= scale;
        s1 *= scale;
        bool taken = false;
        ChainAt next;
        for (int halving = 0; halving <= 6; ++halving) {
            next = chain_vertex(steps, prog, now.point + now.t0 * s0 + now.t1 * s1, regs);
            chain_bend(from, L, next);
            if (next.error <= now.error) {
                taken = true;
                break;
            }
            s0 *= 0.5;
            s1 *= 0.5;
        }
        if (!taken) break;
        now = next;
    }
    // chainSettledAt (chain.cpp:430-433)
    const bool outsideLamp = L.kind != KDN_LIGHT_SPHERE || lengthd(now.point - L.at) > L.radius;
    if (!(now.error < 1e-6 && now.sided && __builtin_fabs(now.value) < 1e-6 && finited(now.point) && outsideLamp))
        return chain_unsettled();
    // invertedPair: the inverse Jacobian
    M2 inv;
    double a0, a1, b0, b1;
    if (!solve2(now.slope, 1, 0, &a0, &a1) || !solve2(now.slope, 0, 1, &b0, &b1)) return chain_unsettled();
    inv.m[0][0] = a0;
    inv.m[0][1] = b0;
    inv.m[1][0] = a1;
    inv.m[1][1] = b1;
    // chainNoise: the facets' and the lamp's spread on the tangent plane
    const D3 towards = L.far_ ? L.place : (L.place - now.point) * (1 / lengthd(L.place - now.point));
    const M3 lampSpread = lamp_spread(L, towards);
    const double facet = 0.5 * width * width;
    M2 noise;
    noise.m[0][0] = facet;
    noise.m[0][1] = 0;
    noise.m[1][0] = 0;
    noise.m[1][1] = facet;
    for (int r = 0; r < 2; ++r) {
        const D3 row = row_times(now.lampRow[r], lampSpread);
        for (int c = 0; c < 2; ++c) noise.m[r][c] += dotd(row, now.lampRow[c]);
    }
    // sandwiched(inverse, noise): inverse noise inverse^T
    M2 moved;
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) {
            double s = 0;
            for (int p = 0; p < 2; ++p)
                for (int q = 0; q < 2; ++q) s += inv.m[r][p] * noise.m[p][q] * inv.m[c][q];
            moved.m[r][c] = s;
        }
    ChainOut o;
    o.found = true;
    o.point = now.point;
    const D3 toFirst = now.point - from;
    const double distance = lengthd(toFirst);
    o.axis = toFirst * (1 / distance);
    basis_of(o.axis, &o.across, &o.up);
    const D3 plane[2] = {o.across, o.up};
    const D3 tan[2] = {now.t0, now.t1};
    M2 seen;
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) seen.m[r][c] = dotd(plane[r], tan[c]) / distance;
    M2 spread;
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) {
            double s = 0;
            for (int p = 0; p < 2; ++p)
                for (int q = 0; q < 2; ++q) s += seen.m[r][p] * moved.m[p][q] * seen.m[c][q];
            spread.m[r][c] = s;
        }
    o.spread[0] = spread.m[0][0];
    o.spread[1] = spread.m[0][1];
    o.spread[2] = spread.m[1][1];
    // pushedMotion: the start's sideways pushes through the inverse
    const D3 push[2] = {side0, side1};
    for (int k = 0; k < 2; ++k) {
        double change[2] = {0, 0};
        for (int r = 0; r < 2; ++r)
            for (int p = 0; p < 2; ++p) change[r] -= inv.m[r][p] * dotd(now.start[p], push[k]);
        o.motion[k] = now.t0 * change[0] + now.t1 * change[1];
    }
    o.found = finited(o.axis) && finited(o.motion[0]) && finited(o.motion[1]) &&
              __builtin_isfinite(o.spread[0] + o.spread[1] + o.spread[2]);
    return o;
}

}  // namespace kdn
