// Tape interpreters on the device: interval, point and dual evaluation of
// one step, the cut rewrite of one step, and the source walk.
//
// Registers are addressed through a strided view so that the eight
// octants of one parent program, which run the same steps, sit side by
// side in memory (index i of octant o at base + i * stride + o).
#pragma once

#include "kdn_math.h"
#include "layout.h"

namespace kdn {

template <class T>
struct Strided {
    T* base;
    uint32_t stride;
    KDN_D T& operator[](uint32_t i) const { return base[(size_t)i * stride]; }
};

// ---- intervals (source/sdf/program.cpp:61-98 over numbers.cpp Interval) ----
template <class R>
KDN_D Iv interval_step(const kdn_step& s, const R& r, const Iv box[3]) {
    switch (s.op) {
        case KDN_CONSTANT: return i_const(s);
        case KDN_ARGUMENT: return i_const(s);
        case KDN_POINT_X: return box[0];
        case KDN_POINT_Y: return box[1];
        case KDN_POINT_Z: return box[2];
        case KDN_NORMAL_X:
        case KDN_NORMAL_Y:
        case KDN_NORMAL_Z: return iv(0.0f, 0.0f);
        case KDN_ADD: return i_add(r[s.a], r[s.b]);
        case KDN_SUBTRACT: return i_sub(r[s.a], r[s.b]);
        case KDN_MULTIPLY: return i_mul(r[s.a], r[s.b]);
        case KDN_DIVIDE: return i_div(r[s.a], r[s.b]);
        case KDN_NEGATE: return i_neg(r[s.a]);
        case KDN_SQUARE_ROOT: return i_sqrt(r[s.a]);
        case KDN_ABSOLUTE: return i_abs(r[s.a]);
        case KDN_FLOOR: return i_floor(r[s.a]);
        case KDN_SINE: return i_sin(r[s.a]);
        case KDN_COSINE: return i_cos(r[s.a]);
        case KDN_SIGN: return i_sign(r[s.a]);
        case KDN_ARC_TANGENT: return i_atan2(r[s.a], r[s.b]);
        case KDN_POWER: return i_pow(r[s.a], r[s.b]);
        case KDN_MINIMUM: return i_min(r[s.a], r[s.b]);
        case KDN_MAXIMUM: return i_max(r[s.a], r[s.b]);
        case KDN_SMOOTH_MINIMUM: return i_smin(r[s.a], r[s.b], r[s.c]);
        case KDN_POSITIVE_SQUARE: return i_pos2(r[s.a]);
        case KDN_SELECT: return i_choose(r[s.a], r[s.b], r[s.c]);
        default: return i_compare(s.op, r[s.a], r[s.b]);
    }
}

// ---- points (numbers.cpp:107-188) -------------------------------------------
KDN_D float p_compare(uint8_t op, float a, float b) {
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

template <class R>
KDN_D float point_step(const kdn_step& s, const R& r, float px, float py, float pz) {
    switch (s.op) {
        case KDN_CONSTANT:
        case KDN_ARGUMENT: return s.value;
        case KDN_POINT_X: return px;
        case KDN_POINT_Y: return py;
        case KDN_POINT_Z: return pz;
        case KDN_NORMAL_X:
        case KDN_NORMAL_Y:
        case KDN_NORMAL_Z: return 0.0f;
        case KDN_ADD: return r[s.a] + r[s.b];
        case KDN_SUBTRACT: return r[s.a] - r[s.b];
        case KDN_MULTIPLY: return r[s.a] * r[s.b];
        case KDN_DIVIDE: return r[s.a] / r[s.b];
        case KDN_NEGATE: return -r[s.a];
        case KDN_SQUARE_ROOT: return __builtin_sqrtf(r[s.a]);
        case KDN_ABSOLUTE: return __builtin_fabsf(r[s.a]);
        case KDN_FLOOR: return __builtin_floorf(r[s.a]);
        case KDN_SINE: return f_rn(d_sin((double)r[s.a]));
        case KDN_COSINE: return f_rn(d_cos((double)r[s.a]));
        case KDN_SIGN: return p_sign(r[s.a]);
        case KDN_ARC_TANGENT: return f_rn(d_atan2((double)r[s.a], (double)r[s.b]));
        case KDN_POWER: return f_rn(d_pow((double)r[s.a], (double)r[s.b]));
        case KDN_MINIMUM: return fmin_(r[s.a], r[s.b]);
        case KDN_MAXIMUM: return fmax_(r[s.a], r[s.b]);
        case KDN_SMOOTH_MINIMUM: return p_smin(r[s.a], r[s.b], r[s.c]);
        case KDN_POSITIVE_SQUARE: return p_pos2(r[s.a]);
        case KDN_SELECT: return r[s.a] != 0 ? r[s.b] : r[s.c];
        default: return p_compare(s.op, r[s.a], r[s.b]);
    }
}

// ---- duals (numbers.cpp:433-590) ----------------------------------------------
KDN_D float blend_slope(float d, float r) {
    const float v = (r - d) / (2 * r);
    return v < 0 ? 0.0f : (v > 1 ? 1.0f : v);
}

template <class R>
KDN_D Dl dual_step(const kdn_step& s, const R& r, float px, float py, float pz) {
    switch (s.op) {
        case KDN_CONSTANT:
        case KDN_ARGUMENT: return dl(s.value);
        case KDN_POINT_X: return Dl{px, {1, 0, 0}};
        case KDN_POINT_Y: return Dl{py, {0, 1, 0}};
        case KDN_POINT_Z: return Dl{pz, {0, 0, 1}};
        case KDN_NORMAL_X:
        case KDN_NORMAL_Y:
        case KDN_NORMAL_Z: return dl(0.0f);
        case KDN_ADD: return chain2(r[s.a], r[s.b], r[s.a].v + r[s.b].v, 1, 1);
        case KDN_SUBTRACT: return chain2(r[s.a], r[s.b], r[s.a].v - r[s.b].v, 1, -1);
        case KDN_MULTIPLY: return chain2(r[s.a], r[s.b], r[s.a].v * r[s.b].v, r[s.b].v, r[s.a].v);
        case KDN_DIVIDE: {
            const float b = r[s.b].v;
            return chain2(r[s.a], r[s.b], r[s.a].v / b, 1 / b, -r[s.a].v / (b * b));
        }
        case KDN_NEGATE: return chain1(r[s.a], -r[s.a].v, -1);
        case KDN_SQUARE_ROOT: {
            const float f = __builtin_sqrtf(r[s.a].v);
            return chain1(r[s.a], f, f > 0 ? 0.5f / f : 0.0f);
        }
        case KDN_ABSOLUTE: return r[s.a].v < 0 ? chain1(r[s.a], -r[s.a].v, -1) : r[s.a];
        case KDN_FLOOR: return dl(__builtin_floorf(r[s.a].v));
        case KDN_SINE: {
            const double x = r[s.a].v;
            return chain1(r[s.a], f_rn(d_sin(x)), f_rn(d_cos(x)));
        }
        case KDN_COSINE: {
            const double x = r[s.a].v;
            return chain1(r[s.a], f_rn(d_cos(x)), -f_rn(d_sin(x)));
        }
        case KDN_SIGN: return dl(p_sign(r[s.a].v));
        case KDN_ARC_TANGENT: {
            const float y = r[s.a].v, x = r[s.b].v, q = x * x + y * y;
            return chain2(r[s.a], r[s.b], f_rn(d_atan2(y, x)), x / q, -y / q);
        }
        case KDN_POWER: {
            const double a = r[s.a].v, b = r[s.b].v;
            const float f = f_rn(d_pow(a, b));
            const float fa = f_rn(b * d_pow(a, b - 1));
            const float fb = a > 0 ? f * f_rn(d_log(a)) : 0.0f;
            return chain2(r[s.a], r[s.b], f, fa, fb);
        }
        case KDN_MINIMUM: return r[s.a].v <= r[s.b].v ? r[s.a] : r[s.b];
        case KDN_MAXIMUM: return r[s.a].v >= r[s.b].v ? r[s.a] : r[s.b];
        case KDN_SMOOTH_MINIMUM: {
            const float fa = blend_slope(r[s.a].v - r[s.b].v, r[s.c].v);
            return chain2(r[s.a], r[s.b], p_smin(r[s.a].v, r[s.b].v, r[s.c].v), fa, 1 - fa);
        }
        case KDN_POSITIVE_SQUARE: return chain1(r[s.a], p_pos2(r[s.a].v), 2 * fmax_(r[s.a].v, 0.0f));
        case KDN_SELECT: return r[s.a].v != 0 ? r[s.b] : r[s.c];
        default: return dl(p_compare(s.op, r[s.a].v, r[s.b].v));
    }
}

// ---- the cut rewrite of one step (plugins/geometry/tree/cut.cpp:60-253) --------
constexpr uint32_t kNoOperand = 0xFFFFFFFFu;

KDN_D int operand_count(uint8_t op) {
    switch (op) {
        case KDN_CONSTANT: case KDN_ARGUMENT: case KDN_POINT_X: case KDN_POINT_Y: case KDN_POINT_Z:
        case KDN_NORMAL_X: case KDN_NORMAL_Y: case KDN_NORMAL_Z: return 0;
        case KDN_NEGATE: case KDN_SQUARE_ROOT: case KDN_ABSOLUTE: case KDN_FLOOR: case KDN_SINE:
        case KDN_COSINE: case KDN_SIGN: case KDN_NOT: return 1;
        case KDN_SELECT: case KDN_SMOOTH_MINIMUM: return 3;
        default: return 2;
    }
}
KDN_D bool is_literal(uint8_t op) { return op == KDN_CONSTANT || op == KDN_ARGUMENT; }
// The word of the march's lane register file a coordinate or normal step
// reads (the point's x, y, z; a normal is zero at a point), 0 for others.
KDN_D uint32_t point_word_of(uint8_t op) {
    switch (op) {
        case KDN_POINT_X: return kPointWord;
        case KDN_POINT_Y: return kPointWord + 1;
        case KDN_POINT_Z: return kPointWord + 2;
        case KDN_NORMAL_X: case KDN_NORMAL_Y: case KDN_NORMAL_Z: return kPointWord + 3;
        default: return 0;
    }
}

// FP32 operations one point evaluation of a step takes: the algorithmic
// work the march's roofline counts (loads of constants and coordinates
// are free, a smooth minimum is ten operations, a transcendental twenty).
KDN_D uint32_t step_flops(uint8_t op) {
    switch (op) {
        case KDN_CONSTANT: case KDN_ARGUMENT: case KDN_POINT_X: case KDN_POINT_Y: case KDN_POINT_Z:
        case KDN_NORMAL_X: case KDN_NORMAL_Y: case KDN_NORMAL_Z: return 0;
        case KDN_SMOOTH_MINIMUM: return 10;
        case KDN_POSITIVE_SQUARE: return 2;
        case KDN_SINE: case KDN_COSINE: case KDN_ARC_TANGENT: case KDN_POWER: return 20;
        default: return 1;
    }
}

template <class IV>
KDN_D uint32_t separated(const kdn_step& s, const IV& iv_, bool lowerWins) {
    const Iv a = iv_[s.a], b = iv_[s.b];
    if (a.hi <= b.lo) return lowerWins ? s.a : s.b;
    if (b.hi <= a.lo) return lowerWins ? s.b : s.a;
    return kNoOperand;
}
template <class IV>
KDN_D uint32_t outside_blend(const kdn_step& s, const IV& iv_) {
    const float radius = iv_[s.c].lo;
    const Iv d = i_sub(iv_[s.a], iv_[s.b]);
    if (d.hi <= -radius) return s.a;
    if (d.lo >= radius) return s.b;
    return kNoOperand;
}
template <class IV>
KDN_D uint32_t neutral_operand(const kdn_step& s, const IV& iv_, float neutral) {
    const bool commutes = s.op == KDN_ADD || s.op == KDN_MULTIPLY;
    if (i_is(iv_[s.b], neutral)) return s.a;
    if (commutes && i_is(iv_[s.a], neutral)) return s.b;
    return kNoOperand;
}
template <class IV>
KDN_D uint32_t shortcut(const kdn_step& s, const IV& iv_) {
    switch (s.op) {
        case KDN_ADD:
        case KDN_SUBTRACT: return neutral_operand(s, iv_, 0.0f);
        case KDN_MULTIPLY:
        case KDN_DIVIDE: return neutral_operand(s, iv_, 1.0f);
        case KDN_MINIMUM: return separated(s, iv_, true);
        case KDN_MAXIMUM: return separated(s, iv_, false);
        case KDN_SMOOTH_MINIMUM: return outside_blend(s, iv_);
        case KDN_ABSOLUTE: return iv_[s.a].lo >= 0 ? s.a : kNoOperand;
        case KDN_SELECT: {
            const Iv c = iv_[s.a];
            if (c.lo > 0 || c.hi < 0) return s.b;
            if (c.lo == 0 && c.hi == 0) return s.c;
            return kNoOperand;
        }
        default: return kNoOperand;
    }
}
template <class IV>
KDN_D uint32_t negated_operand(const kdn_step& s, const IV& iv_) {
    const bool product = s.op == KDN_MULTIPLY;
    if (s.op == KDN_ABSOLUTE && iv_[s.a].hi <= 0) return s.a;
    if (s.op == KDN_SUBTRACT && i_is(iv_[s.a], 0.0f)) return s.b;
    if (product && i_is(iv_[s.a], -1.0f)) return s.b;
    if (product && i_is(iv_[s.b], -1.0f)) return s.a;
    return kNoOperand;
}
template <class IV>
KDN_D kdn_step resolved_step(const kdn_step& s, const IV& iv_, uint32_t index) {
    const Iv self = iv_[index];
    if (i_point(self) && !is_literal(s.op)) {
        kdn_step k;
        k.op = KDN_CONSTANT;
        k.flags = 0;
        k.a = k.b = k.c = 0;
        k.value = self.lo;
        k.leaf = s.leaf;
        return k;
    }
    if (s.op == KDN_POSITIVE_SQUARE && iv_[s.a].lo >= 0) {
        kdn_step q = s;
        q.op = KDN_MULTIPLY;
        q.b = s.a;
        return q;
    }
    const uint32_t n = negated_operand(s, iv_);
    if (n != kNoOperand) {
        kdn_step q = s;
        q.op = KDN_NEGATE;
        q.a = (uint16_t)n;
        q.b = 0;
        q.c = 0;
        return q;
    }
    return s;
}
template <class ST>
KDN_D kdn_step unsquared(const ST& steps, const kdn_step& s) {
    if (s.op != KDN_SQUARE_ROOT) return s;
    const kdn_step inner = steps[s.a];
    if (inner.op == KDN_POSITIVE_SQUARE) {
        kdn_step q = inner;
        q.op = KDN_MAXIMUM;
        return q;
    }
    if (inner.op != KDN_MULTIPLY || inner.a != inner.b) return s;
    kdn_step q = s;
    q.op = KDN_ABSOLUTE;
    q.a = inner.a;
    return q;
}
template <class ST>
KDN_D bool is_zero_step(const ST& steps, uint32_t i) {
    const kdn_step z = steps[i];
    return z.op == KDN_CONSTANT && z.value == 0.0f && z.flags == 0;
}
template <class ST>
KDN_D kdn_step positive_squared(const ST& steps, const kdn_step& s) {
    if (s.op != KDN_MULTIPLY || s.a != s.b) return s;
    const kdn_step inner = steps[s.a];
    if (inner.op != KDN_MAXIMUM || !(is_zero_step(steps, inner.a) || is_zero_step(steps, inner.b))) return s;
    const bool zeroFirst = is_zero_step(steps, inner.a);
    kdn_step q = s;
    q.op = KDN_POSITIVE_SQUARE;
    q.a = zeroFirst ? inner.b : inner.a;
    q.b = zeroFirst ? inner.a : inner.b;
    return q;
}
KDN_D bool same_step(const kdn_step& x, const kdn_step& y) {
    return x.op == y.op && x.flags == y.flags && x.a == y.a && x.b == y.b && x.c == y.c &&
           __float_as_uint(x.value) == __float_as_uint(y.value) && x.leaf == y.leaf;
}

// ---- the source walk (source/sdf/program.cpp:442-502) ---------------------------
template <class R>
KDN_D uint32_t chosen_operand(const kdn_step& s, const R& v) {
    const float a = v[s.a], b = v[s.b];
    switch (s.op) {
        case KDN_MINIMUM:
        case KDN_SMOOTH_MINIMUM: return a <= b ? s.a : s.b;
        case KDN_MAXIMUM: return a >= b ? s.a : s.b;
        case KDN_SELECT: return a != 0 ? s.b : s.c;
        default: return s.a;
    }
}
KDN_D bool is_join(uint8_t op) {
    return op == KDN_MINIMUM || op == KDN_SMOOTH_MINIMUM || op == KDN_MAXIMUM || op == KDN_NEGATE || op == KDN_SELECT;
}
template <class R>
KDN_D uint32_t leaf_from(const kdn_step* prog, const R& v, uint32_t at) {
    while (is_join(prog[at].op)) at = chosen_operand(prog[at], v);
    return prog[at].leaf;
}
template <class R>
KDN_D void source_of(const kdn_step* prog, uint32_t result, const R& v, uint32_t* leaf, uint32_t* other, float* share) {
    *other = KDN_NO_LEAF;
    *share = 0.0f;
    uint32_t at = result;
    while (is_join(prog[at].op)) {
        const kdn_step s = prog[at];
        const uint32_t next = chosen_operand(s, v);
        float sh = 0.0f;
        if (s.op == KDN_SMOOTH_MINIMUM) {
            const float radius = v[s.c];
            const float apart = __builtin_fabsf(v[s.a] - v[s.b]);
            sh = (radius > 0 && apart < radius) ? 0.5f * (radius - apart) / radius : 0.0f;
        }
        if (sh > 0 && *other == KDN_NO_LEAF) {
            *other = leaf_from(prog, v, next == s.a ? s.b : s.a);
            *share = sh;
        }
        at = next;
    }
    *leaf = prog[at].leaf;
}

}  // namespace kdn
