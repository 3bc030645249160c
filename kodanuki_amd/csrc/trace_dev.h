// Sphere tracing through the interval-cut octree, one ray per lane.
//
// The ray walks the octree's leaves front to back: it descends from the
// root by integer cell coordinates (exact, no point location in floats),
// skips Empty leaves to their exit, and inside a Surface leaf marches by
// the least value of the leaf's cut programs, which equal the whole
// programs inside the leaf (cut.cpp's invariant), so each step is a safe
// sphere-tracing step. Leaving a leaf it steps to the face neighbour
// across its exit face.
//
// Landing, tie-break, inside rays, normals and the hit leaf follow the
// reference sphere_trace plugin (plugins/geometry/sphere_trace/
// sphere_trace.cpp:103-228) with FP32 tolerances.
#pragma once

#include "layout.h"
#include "tape_dev.h"

namespace kdn {

constexpr int kMaxIterations = 4096;  // cells plus march steps of one ray
constexpr float kTolAbs = 1e-5f;      // landing tolerance: kTolAbs + kTolRel * t
constexpr float kTolRel = 2e-6f;
constexpr float kTie = 1e-6f;         // values within kTie (1 + max|p|) tie (sphere_trace.cpp:103-113)
constexpr float kGraze = 0.02f;       // sphere_trace.cpp:16
// The landing's Newton step: taken where the ray meets the surface at a
// cosine above kNewtonFacing, and no longer than kNewtonReach tolerances
// (a march stops at most one tolerance above the surface, which is
// 1/kNewtonFacing tolerances along such a ray).
constexpr float kNewtonFacing = 0.01f;
constexpr float kNewtonReach = 100.0f;

struct TreeView {
    const NodeDev* nodes;
    const EntryDev* entries;
    const kdn_step* pool;
    const SlotStep* spool;  // the pool's programs on slot registers, same offsets
    const SlotStep* mpool;  // the same with their literal steps folded away (the march's)
    const uint32_t* leafMaterial;
    float lower[3];
    float size[3];
    int depth;
    double* work;  // when set, the FP32 operations of every evaluation are added here
};

// A lane's slot registers: slot k at s[k * stride] (in LDS, one column per
// lane of the block, so a wave's lanes touch consecutive banks).
struct SlotRegs {
    float* s;
    uint32_t stride;
    KDN_D float& operator[](uint32_t i) const {
        KDN_CHECK(i < (uint32_t)kLaneWords);
        return s[i * stride];
    }
};

// A lane's ancestry in the octree: the node index at each level of the
// current leaf's path from the root (in LDS, one column per lane).
struct PathRegs {
    uint32_t* s;
    uint32_t stride;
    KDN_D uint32_t& operator[](uint32_t level) const {
        KDN_CHECK(level < (uint32_t)kPathLevels);
        return s[level * stride];
    }
};

// A lane's register files: the slot registers every march evaluation runs
// in and its octree ancestry.
struct LaneRegs {
    SlotRegs slots;
    PathRegs path;
};

// One step of a march program (its operands read from the lane's register
// file or its constant): the arithmetic most steps are (sum, difference,
// product, least, most, negation, magnitude, positive square, copy)
// computed without a branch and chosen by opcode, so that a wave's lanes,
// each in its own program, run the same instructions; square roots, smooth
// minima and the rare operations under branches a wave skips where none of
// its lanes takes them. The same operations in the same order as
// point_step, so the same bits.
KDN_D float march_step(uint32_t code, float value, const SlotRegs& r) {
    const uint32_t op = code & 63u;
    float a = r[(code >> 10) & 31u], b = r[(code >> 15) & 31u], c = r[(code >> 20) & 31u];
    a = (code & kImmA) ? value : a;
    b = (code & kImmB) ? value : b;
    c = (code & kImmC) ? value : c;
    const float sum = a + b, difference = a - b, product = a * b;
    const float pos = fmax_(a, 0.0f);
    float v = op == KDN_SUBTRACT ? difference : sum;
    v = op == KDN_MULTIPLY ? product : v;
    v = op == KDN_MINIMUM ? fmin_(a, b) : v;
    v = op == KDN_MAXIMUM ? fmax_(a, b) : v;
    v = op == KDN_NEGATE ? -a : v;
    v = op == KDN_ABSOLUTE ? __builtin_fabsf(a) : v;
    v = op == KDN_POSITIVE_SQUARE ? pos * pos : v;
    v = op == kMoveOp ? a : v;
    if (op == KDN_SQUARE_ROOT) {
        v = __builtin_sqrtf(a);
    } else if (op == KDN_SMOOTH_MINIMUM) {
        v = p_smin(a, b, c);
    } else if (op == KDN_DIVIDE || op == KDN_FLOOR || op == KDN_SINE || op == KDN_COSINE || op == KDN_SIGN ||
               op == KDN_ARC_TANGENT || op == KDN_POWER || op == KDN_SELECT || (op >= KDN_LESS && op <= KDN_NOT)) {
        const float w[3] = {a, b, c};
        kdn_step st;
        st.op = (uint8_t)op;
        st.flags = 0;
        st.a = 0;
        st.b = 1;
        st.c = 2;
        st.value = value;
        st.leaf = 0;
        v = point_step(st, w, 0.0f, 0.0f, 0.0f);
    }
    return v;
}

// A march program (the build's slot allocation with its literal steps
// folded away, same offsets as the SSA pool) at a point: the point goes
// into the lane's point words, then the steps run in batches of kSlotFetch
// fetched together (the loads of a batch are in flight at once, so a lane
// waits once per batch; the pools keep kPoolPad steps readable past their
// end for the last batch), each batch's steps unrolled.
KDN_D float eval_slots(const TreeView& tv, uint32_t at, const SlotRegs& r, float px, float py, float pz) {
    const SlotStep* p = tv.mpool + at;
    const SlotStep head = p[0];
    SlotStep q[kSlotFetch];
#pragma unroll
    for (int k = 0; k < kSlotFetch; ++k) q[k] = p[1 + k];
    r[kPointWord] = px;
    r[kPointWord + 1] = py;
    r[kPointWord + 2] = pz;
    r[kPointWord + 3] = 0.0f;
    if (tv.work) *tv.work += head.value;
    const uint32_t L = head.code & 0xFFFFFFu;
    for (uint32_t base = 0; base < L; base += kSlotFetch) {
        if (base > 0) {
#pragma unroll
            for (int k = 0; k < kSlotFetch; ++k) q[k] = p[1 + base + k];
        }
#pragma unroll
        for (int k = 0; k < kSlotFetch; ++k) {
            if (base + k < L) r[(q[k].code >> 6) & 15u] = march_step(q[k].code, q[k].value, r);
        }
    }
    return r[(head.code >> 24) & 15u];
}

struct LeafField {
    float least;      // least value over the leaf's programs
    float nearest;    // least |value|
    uint32_t arg;     // entry of the least value (first on ties)
};

KDN_D LeafField leaf_field(const TreeView& tv, const NodeDev& n, const SlotRegs& r, float px, float py, float pz) {
    LeafField f{kInf, kInf, 0};
    for (uint32_t j = 0; j < n.entry_count; ++j) {
        const float v = eval_slots(tv, tv.entries[n.entry_first + j].program, r, px, py, pz);
        if (v < f.least) {
            f.least = v;
            f.arg = j;
        }
        f.nearest = fmin_(f.nearest, __builtin_fabsf(v));
    }
    return f;
}

KDN_D float vmax3(float a, float b, float c) { return fmax_(fmax_(a, b), c); }

// Normal from the gradient, faced as sphereNormal (sphere_trace.cpp:133-149).
KDN_D void faced_normal(const float g[3], const float d[3], bool inside, float n[3]) {
    const float len = __builtin_sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (!(len > 0)) {
        n[0] = -d[0]; n[1] = -d[1]; n[2] = -d[2];
        return;
    }
    float u[3] = {g[0] / len, g[1] / len, g[2] / len};
    const float along = u[0] * d[0] + u[1] * d[1] + u[2] * d[2];
    if (inside) {
        const float s = along < 0 ? -1.0f : 1.0f;
        n[0] = s * u[0]; n[1] = s * u[1]; n[2] = s * u[2];
        return;
    }
    if (along > 0) {
        float w[3] = {u[0] - d[0] * (along + kGraze), u[1] - d[1] * (along + kGraze), u[2] - d[2] * (along + kGraze)};
        const float l = __builtin_sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        n[0] = w[0] / l; n[1] = w[1] / l; n[2] = w[2] / l;
        return;
    }
    n[0] = u[0]; n[1] = u[1]; n[2] = u[2];
}

// The source of a value (program.cpp:442-502): the leaf that decides it,
// and, where a smooth minimum blends two solids on its way, the other
// leaf and the share the blend gives it.
struct Source {
    uint32_t leaf, other;
    float share;
};

// A program on slot registers evaluated in duals at a point, each value's
// source carried forward with it. The reference walks back from the
// result through the joins (min, smooth-min, max, negate, select) to the
// deciding step, keeping the outermost smooth-min blend it passes; a
// value here takes its chosen operand's source at each join, the blend of
// a smooth-min in its band replacing what the operand carried, so the
// result's source is the walk's. With Dual false only the values and
// sources are made. Registers are private: kSlots duals and sources a
// lane, the most values a program keeps live.
template <bool Dual>
KDN_D Dl eval_sourced(const TreeView& tv, uint32_t at, float px, float py, float pz, Source* source) {
    Dl reg[kSlots];
    Source src[kSlots];
    struct View {
        Dl* r;
        KDN_D Dl& operator[](uint32_t i) const {
            KDN_CHECK(i < (uint32_t)kSlots);
            return r[i];
        }
    } const view{reg};
    struct Values {
        const Dl* r;
        KDN_D float operator[](uint32_t i) const {
            KDN_CHECK(i < (uint32_t)kSlots);
            return r[i].v;
        }
    } const values{reg};
    const SlotStep head = tv.spool[at];
    if (tv.work) *tv.work += (Dual ? 4 : 1) * head.value;
    const uint32_t L = head.code & 0xFFFFFFu;
    for (uint32_t i = 0; i < L; ++i) {
        const uint32_t code = tv.spool[at + 1 + i].code;
        kdn_step st;
        st.op = (uint8_t)(code & 63u);
        st.flags = 0;
        st.a = (uint16_t)((code >> 10) & 15u);
        st.b = (uint16_t)((code >> 14) & 15u);
        st.c = (uint16_t)((code >> 18) & 15u);
        st.value = tv.spool[at + 1 + i].value;
        st.leaf = 0;
        Dl v;
        if (Dual) {
            v = dual_step(st, view, px, py, pz);
        } else {
            v = dl(point_step(st, values, px, py, pz));
        }
        Source sv;
        if (is_join(st.op)) {
            const uint32_t next = chosen_operand(st, values);
            sv = src[next & 15u];
            if (st.op == KDN_SMOOTH_MINIMUM) {
                const float radius = values[st.c];
                const float apart = __builtin_fabsf(values[st.a] - values[st.b]);
                const float sh = (radius > 0 && apart < radius) ? 0.5f * (radius - apart) / radius : 0.0f;
                if (sh > 0) {
                    sv.other = src[(next == st.a ? st.b : st.a) & 15u].leaf;
                    sv.share = sh;
                }
            }
        } else {
            sv = Source{tv.pool[at + 1 + i].leaf, KDN_NO_LEAF, 0.0f};
        }
        const uint32_t dst = (code >> 6) & 15u;
        reg[dst] = v;
        src[dst] = sv;
    }
    const uint32_t result = (head.code >> 24) & 15u;
    *source = src[result];
    return reg[result];
}

// The hit at t from entry `winner` of leaf n. The march stops within its
// tolerance of the surface; one Newton step along the ray, t - v / (g.d),
// bounded and taken only where the ray is not grazing, brings the distance
// onto the surface.
KDN_D void land(const TreeView& tv, const NodeDev& n, uint32_t nodeIndex, uint32_t winner, const float o[3], const float d[3], float t,
                bool inside, kdn_hit& hit) {
    const float px = o[0] + d[0] * t, py = o[1] + d[1] * t, pz = o[2] + d[2] * t;
    const uint32_t at = tv.entries[n.entry_first + winner].program;
    Source src;
    const Dl value = eval_sourced<true>(tv, at, px, py, pz, &src);
    const float slope = value.g[0] * d[0] + value.g[1] * d[1] + value.g[2] * d[2];
    const float glen = __builtin_sqrtf(value.g[0] * value.g[0] + value.g[1] * value.g[1] + value.g[2] * value.g[2]);
    const float bound = kNewtonReach * (kTolAbs + kTolRel * t);
    float tr = t;
    if (__builtin_fabsf(slope) > kNewtonFacing * glen) {
        const float dt = -value.v / slope;
        tr = t + fmax_(-bound, fmin_(bound, dt));
    }
    hit.kind = KDN_HIT_SURFACE;
    hit.distance = tr;
    hit.point[0] = o[0] + d[0] * tr; hit.point[1] = o[1] + d[1] * tr; hit.point[2] = o[2] + d[2] * tr;
    faced_normal(value.g, d, inside, hit.normal);
    hit.leaf = src.leaf;
    hit.other = src.other;
    hit.share = src.share;
    hit.cell = nodeIndex;
}

// The winner of a landing: the least value, ties going to the least
// value a little further along (sphereWinner, sphere_trace.cpp:103-113).
KDN_D uint32_t winner_at(const TreeView& tv, const NodeDev& n, const SlotRegs& r, const float o[3], const float d[3], float t,
                         float inset, const LeafField& f) {
    if (n.entry_count == 1) return 0;
    const float px = o[0] + d[0] * t, py = o[1] + d[1] * t, pz = o[2] + d[2] * t;
    const float band = kTie * (1 + vmax3(__builtin_fabsf(px), __builtin_fabsf(py), __builtin_fabsf(pz)));
    const float qx = o[0] + d[0] * (t + inset), qy = o[1] + d[1] * (t + inset), qz = o[2] + d[2] * (t + inset);
    uint32_t winner = f.arg;
    float winnerOn = eval_slots(tv, tv.entries[n.entry_first + f.arg].program, r, qx, qy, qz);
    for (uint32_t j = 0; j < n.entry_count; ++j) {
        const uint32_t at = tv.entries[n.entry_first + j].program;
        const float v = eval_slots(tv, at, r, px, py, pz);
        if (!(v - f.least <= band)) continue;
        const float on = eval_slots(tv, at, r, qx, qy, qz);
        if (on < winnerOn) {
            winner = j;
            winnerOn = on;
        }
    }
    return winner;
}

KDN_D void missed(kdn_hit& hit) {
    hit.kind = KDN_HIT_MISSED;
    hit.distance = kInf;
    hit.point[0] = hit.point[1] = hit.point[2] = 0;
    hit.normal[0] = hit.normal[1] = hit.normal[2] = 0;
    hit.leaf = KDN_NO_LEAF;
    hit.other = KDN_NO_LEAF;
    hit.share = 0;
    hit.cell = KDN_NO_LEAF;
}

KDN_D float slope_of(float part) {
    const float m = 1e-12f;
    return 1.0f / (__builtin_fabsf(part) < m ? (part < 0 ? -m : m) : part);
}

// AnyHit: a shadow test, which only asks whether a surface lies on the ray
// before its far end: the march stops where it first comes within its
// tolerance of a surface (or enters a filled cell) and reports that point,
// with no tie-break, landing or source.
// (AnyHit is a run-time flag here: march_ray fixes it per call site, and a
// lane that marches its path's rays and its shadow rays by turns passes it
// per ray, so that a wave's lanes run one march between them.)
KDN_D void march_any(const TreeView& tv, const kdn_ray& ray, const LaneRegs& lr, kdn_hit& hit, int& iterations, const bool AnyHit) {
    const float o[3] = {ray.origin[0], ray.origin[1], ray.origin[2]};
    const float d[3] = {ray.direction[0], ray.direction[1], ray.direction[2]};
    const bool inside = ray.inside != 0;
    // the span inside the region (sphereSpan, sphere_trace.cpp:51-69)
    float t0 = 0.0f, t1 = ray.far_;
    float inv[3];
    for (int k = 0; k < 3; ++k) {
        inv[k] = slope_of(d[k]);
        const float a = (tv.lower[k] - o[k]) * inv[k];
        const float b = (tv.lower[k] + tv.size[k] - o[k]) * inv[k];
        t0 = fmax_(t0, fmin_(a, b));
        t1 = fmin_(t1, fmax_(a, b));
    }
    missed(hit);
    if (!(t0 <= t1)) return;
    const int D = tv.depth;
    const uint32_t N = 1u << D;
    float finest[3];
    uint32_t ci[3];
    for (int k = 0; k < 3; ++k) {
        finest[k] = __builtin_ldexpf(tv.size[k], -D);
        const float p = o[k] + d[k] * t0;
        const float q = __builtin_floorf((p - tv.lower[k]) / finest[k]);
        ci[k] = q < 0 ? 0u : (q >= (float)N ? N - 1 : (uint32_t)q);
    }
    const SlotRegs& r = lr.slots;
    const PathRegs& path = lr.path;
    float t = t0;
    // the deepest known ancestor of the next cell: the root at first, then
    // the lowest common ancestor of the cell left and the cell entered
    int lvl = 0;
    path[0] = 0;
    while (iterations < kMaxIterations) {
        // descend by integer coordinates from that ancestor
        uint32_t index = path[(uint32_t)lvl];
        NodeDev n = tv.nodes[index];
        while (n.kind == KindBranch) {
            ++lvl;
            const int sh = D - lvl;
            const uint32_t oct = ((ci[0] >> sh) & 1u) | (((ci[1] >> sh) & 1u) << 1) | (((ci[2] >> sh) & 1u) << 2);
            index = n.first_child + oct;
            path[(uint32_t)lvl] = index;
            n = tv.nodes[index];
        }
        const uint32_t nc[3] = {n.x, n.y, n.z};
        float lo[3], hi[3];
        float tExit = kInf;
        int axis = 0;
        for (int k = 0; k < 3; ++k) {
            const float cell = __builtin_ldexpf(tv.size[k], -lvl);
            lo[k] = tv.lower[k] + (float)nc[k] * cell;
            hi[k] = lo[k] + cell;
            const float bound = d[k] > 0 ? hi[k] : lo[k];
            const float te = d[k] == 0 ? kInf : (bound - o[k]) * inv[k];
            if (te < tExit) {
                tExit = te;
                axis = k;
            }
        }
        const float tEnd = fmin_(tExit, t1);
        if (n.kind == KindSurface) {
            while (t <= tEnd && iterations < kMaxIterations) {
                ++iterations;
                const float px = o[0] + d[0] * t, py = o[1] + d[1] * t, pz = o[2] + d[2] * t;
                const LeafField f = leaf_field(tv, n, r, px, py, pz);
                const float tol = kTolAbs + kTolRel * t;
                const float inset = 2 * tol;
                float step;
                if (!inside) {
                    if (f.least < tol) {
                        if (AnyHit) {
                            hit.kind = KDN_HIT_SURFACE;
                            hit.distance = t;
                            hit.cell = index;
                            return;
                        }
                        const uint32_t w = winner_at(tv, n, r, o, d, t, inset, f);
                        land(tv, n, index, w, o, d, t, false, hit);
                        return;
                    }
                    step = f.least;
                } else {
                    if (f.nearest < tol) {
                        const float qx = o[0] + d[0] * (t + inset), qy = o[1] + d[1] * (t + inset), qz = o[2] + d[2] * (t + inset);
                        const LeafField past = leaf_field(tv, n, r, qx, qy, qz);
                        if (past.least > 0) {
                            const LeafField here = leaf_field(tv, n, r, px, py, pz);
                            const uint32_t w = winner_at(tv, n, r, o, d, t, inset, here);
                            land(tv, n, index, w, o, d, t, true, hit);
                            return;
                        }
                        // another solid past the crossing: a hit where it is of
                        // another material (sphereCrossing, sphere_trace.cpp:176-187),
                        // its field taken past the crossing; the landing's Newton
                        // step from there puts the hit on that solid's surface
                        // (the reference reports t, within its FP64 tolerance of it)
                        kdn_hit other;
                        land(tv, n, index, past.arg, o, d, t + inset, true, other);
                        if (other.leaf != KDN_NO_LEAF && tv.leafMaterial[other.leaf] != ray.within) {
                            hit = other;
                            return;
                        }
                        step = inset;
                    } else {
                        step = f.nearest;
                    }
                }
                // The leaf's programs bound the distance to surfaces inside
                // it only (objects Empty here were dropped), so a step stops
                // at the leaf's exit.
                const float next = t + step;
                if (next >= tExit) {
                    t = tExit;
                    break;
                }
                t = next;
            }
        } else if (n.kind == KindFilled && !inside) {
            // a ray from outside enters a filled leaf: the side it meets
            hit.kind = KDN_HIT_SOLID;
            hit.distance = t;
            hit.point[0] = o[0] + d[0] * t; hit.point[1] = o[1] + d[1] * t; hit.point[2] = o[2] + d[2] * t;
            hit.normal[0] = hit.normal[1] = hit.normal[2] = 0;
            hit.normal[axis] = d[axis] > 0 ? -1.0f : 1.0f;
            hit.cell = index;
            // the solid's leaf: the source of its first program's value
            if (!AnyHit && n.entry_count > 0) {
                Source src;
                (void)eval_sourced<false>(tv, tv.entries[n.entry_first].program, hit.point[0], hit.point[1], hit.point[2], &src);
                hit.leaf = src.leaf;
                hit.other = src.other;
                hit.share = src.share;
            }
            return;
        }
        // (a ray from inside crosses Empty leaves as the reference's inside
        // march crosses open space, sphere_trace.cpp:189-201: it lands
        // where it next comes within its tolerance of a surface)
        ++iterations;
        if (tExit >= t1) return;
        t = fmax_(t, tExit);
        // the face neighbour across the exit face
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
                const float q = __builtin_floorf((p - tv.lower[k]) / finest[k]);
                const uint32_t qi = q < (float)first ? first : (q > (float)last ? last : (uint32_t)q);
                next[k] = qi;
            }
        }
        if (next[axis] >= N) return;
        // levels whose cell holds both the cell left and the one entered:
        // those whose shift passes every bit where their coordinates differ
        const uint32_t differ = (next[0] ^ ci[0]) | (next[1] ^ ci[1]) | (next[2] ^ ci[2]);
        const int shared = D - (32 - __builtin_clz(differ));
        lvl = shared < lvl ? shared : lvl;
        ci[0] = next[0]; ci[1] = next[1]; ci[2] = next[2];
    }
}

template <bool AnyHit>
KDN_D void march_ray(const TreeView& tv, const kdn_ray& ray, const LaneRegs& lr, kdn_hit& hit, int& iterations) {
    march_any(tv, ray, lr, hit, iterations, AnyHit);
}

// The extend step of one ray: its hit (missed when it leaves the region or
// its span). KDN_TRACE_ITERATIONS sees the cells and march steps the ray
// took (a hook of the CPU build; nothing on the device).
#ifndef KDN_TRACE_ITERATIONS
#define KDN_TRACE_ITERATIONS(n) ((void)0)
#endif
KDN_D void trace_ray(const TreeView& tv, const kdn_ray& ray, const LaneRegs& lr, kdn_hit& hit) {
    int iterations = 0;
    march_ray<false>(tv, ray, lr, hit, iterations);
    KDN_TRACE_ITERATIONS(iterations);
}

// The shadow test of a ray: its first approach within the tolerance of a
// surface (kind SURFACE, distance) or MISSED.
KDN_D void occlusion_ray(const TreeView& tv, const kdn_ray& ray, const LaneRegs& lr, kdn_hit& hit) {
    int iterations = 0;
    march_ray<true>(tv, ray, lr, hit, iterations);
    KDN_TRACE_ITERATIONS(iterations);
}

}  // namespace kdn
