// The interval-cut octree, built on the GPU level by level.
//
// Each level takes the Surface cells of the level above (the parents),
// and for every (parent entry, octant) pair — one thread each, the eight
// octants of one entry in adjacent lanes so they read the same steps —
// evaluates the entry's program in intervals over the child box, rewrites
// it (cut.cpp:117-253), and marks its live steps. A thread per child then
// takes the reference cell decision (cutCell, cut.cpp:446-476), prefix
// sums place the kept entries and their compacted programs, and a last
// pass writes the child nodes, the entries and the programs.
//
// Cell boxes are exact in FP32: the region is a power-of-two cube whose
// corner is a multiple of the finest cell, so every box end is
// lower + i * size / 2^level without rounding, and halving a box (the
// reference octantBox, cut.cpp:478-490) gives the same ends.
#include <hip/hip_runtime.h>

#include <chrono>

#include "ctx.h"
#include "tape_dev.h"

namespace kdn {

constexpr uint8_t kHeaderOp = 255;

struct ActEntry {
    uint32_t parent;   // index of the parent among the level's active cells
    uint32_t entry;    // index into the entry array
};

struct ItemOut {
    float lo, hi;      // the range of the program's result over the child box
    uint32_t status;   // 0 Empty, 1 Filled, 2 Surface
    uint32_t length;   // steps of the compacted cut program
    uint32_t same;     // 1 when the cut is the parent's program unchanged
    uint32_t result;   // the result index after the rewrite (old numbering)
};

KDN_D uint32_t header_length(const kdn_step& h) { return (uint32_t)h.a | ((uint32_t)h.b << 16); }

KDN_D void child_box(const float* lower, const float* size, uint32_t level, uint32_t x, uint32_t y, uint32_t z, Iv box[3]) {
    const uint32_t c[3] = {x, y, z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float cell = __builtin_ldexpf(size[k], -(int)level);
        const float lo = lower[k] + (float)c[k] * cell;
        box[k] = iv(lo, lo + cell);
    }
}

struct Region {
    float lower[3];
    float size[3];
};

// Active parents: the Surface nodes of the level, which become branches.
__global__ void k_flag_surface(const NodeDev* nodes, uint32_t n, uint32_t* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = nodes[i].kind == KindSurface ? 1u : 0u;
}

__global__ void k_mark_branches(NodeDev* nodes, uint32_t n, const uint32_t* flags, const uint32_t* rank, uint32_t childBase,
                                uint32_t* active, uint32_t* entryCounts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flags[i]) return;
    const uint32_t r = rank[i];
    active[r] = i;
    entryCounts[r] = nodes[i].entry_count;
    nodes[i].kind = KindBranch;
    nodes[i].first_child = childBase + 8 * r;
}

__global__ void k_fill_act(const NodeDev* nodes, const uint32_t* active, uint32_t nActive, const uint32_t* entryOffsets,
                           ActEntry* act, const EntryDev* entries, const kdn_step* pool, uint32_t* lengths) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nActive) return;
    const NodeDev node = nodes[active[p]];
    const uint32_t off = entryOffsets[p];
    for (uint32_t j = 0; j < node.entry_count; ++j) {
        const uint32_t e = node.entry_first + j;
        act[off + j] = ActEntry{p, e};
        lengths[off + j] = header_length(pool[entries[e].program]);
    }
}

// One thread per (entry, octant): intervals, rewrite, live marks.
__global__ void __launch_bounds__(256) k_cut_items(const ActEntry* act, uint32_t nItems, uint32_t octants, const NodeDev* nodes,
                                                   const uint32_t* active, const EntryDev* entries, const kdn_step* pool,
                                                   const uint32_t* scratchOffsets, Iv* ivScratch, kdn_step* stScratch,
                                                   uint32_t* alScratch, Region region, uint32_t childLevel, ItemOut* out) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nItems) return;
    const uint32_t a = w / octants, o = w % octants;
    const ActEntry ae = act[a];
    const EntryDev entry = entries[ae.entry];
    const kdn_step* prog = pool + entry.program + 1;
    const kdn_step header = pool[entry.program];
    const uint32_t L = header_length(header);
    const uint32_t res = header.c;

    Iv box[3];
    if (octants == 1) {
        child_box(region.lower, region.size, 0, 0, 0, 0, box);
    } else {
        const NodeDev parent = nodes[active[ae.parent]];
        child_box(region.lower, region.size, childLevel, 2 * parent.x + (o & 1), 2 * parent.y + ((o >> 1) & 1), 2 * parent.z + ((o >> 2) & 1), box);
    }

    const size_t base = (size_t)scratchOffsets[a] * octants + o;
    const Strided<Iv> IV{ivScratch + base, octants};
    const Strided<kdn_step> ST{stScratch + base, octants};
    const Strided<uint32_t> AL{alScratch + base, octants};

    for (uint32_t i = 0; i < L; ++i) IV[i] = interval_step(prog[i], IV, box);
    const Iv range = IV[res];
    ItemOut r;
    r.lo = range.lo;
    r.hi = range.hi;
    r.same = 0;
    r.length = 0;
    r.result = res;
    if (range.lo > 0) {  // occupancyOf: Empty
        r.status = 0;
        out[w] = r;
        return;
    }
    r.status = range.hi < 0 ? 1u : 2u;

    // rewriteFrom (cut.cpp:285-305) step by step
    bool changed = false;
    for (uint32_t i = 0; i < L; ++i) {
        const kdn_step orig = prog[i];
        if (is_literal(orig.op)) {
            AL[i] = i;
            kdn_step lit = orig;
            lit.b = 0;
            lit.c = 0;
            ST[i] = lit;
            changed = changed || !same_step(lit, orig);
            continue;
        }
        kdn_step s = orig;
        const int count = operand_count(orig.op);
        s.a = count > 0 ? (uint16_t)AL[orig.a] : orig.a;
        s.b = count > 1 ? (uint16_t)AL[orig.b] : 0;
        s.c = count > 2 ? (uint16_t)AL[orig.c] : 0;
        s = positive_squared(ST, unsquared(ST, s));
        const uint32_t other = i_point(IV[i]) ? kNoOperand : shortcut(s, IV);
        AL[i] = other != kNoOperand ? other : i;
        const kdn_step resolved = resolved_step(s, IV, i);
        ST[i] = resolved;
        changed = changed || !same_step(resolved, orig);
    }
    const uint32_t result = AL[res];
    r.result = result;
    // live marks (program.cpp:338-356), written over the aliases
    for (uint32_t i = 0; i < L; ++i) AL[i] = 0;
    AL[result] = 1;
    uint32_t live = 0;
    for (uint32_t i = L; i-- > 0;) {
        if (!AL[i]) continue;
        ++live;
        const kdn_step s = ST[i];
        const int count = operand_count(s.op);
        if (count > 0) AL[s.a] = 1;
        if (count > 1) AL[s.b] = 1;
        if (count > 2) AL[s.c] = 1;
    }
    r.same = (!changed && live == L && result == res) ? 1u : 0u;
    r.length = live;
    out[w] = r;
}

// One thread per child: the cell decision (cut.cpp:446-476).
__global__ void k_cut_children(uint32_t nChildren, uint32_t octants, const uint32_t* entryOffsets, const uint32_t* entryCounts,
                               const ItemOut* items, uint32_t* keep, uint32_t* keptCount, uint32_t* kind) {
    const uint32_t ch = blockIdx.x * blockDim.x + threadIdx.x;
    if (ch >= nChildren) return;
    const uint32_t p = ch / octants, o = ch % octants;
    const uint32_t first = entryOffsets[p], count = entryCounts[p];
    // the first Filled object fills the cell
    for (uint32_t j = 0; j < count; ++j) {
        const uint32_t w = (first + j) * octants + o;
        if (items[w].status == 1) {
            for (uint32_t k = 0; k < count; ++k) keep[(first + k) * octants + o] = 0;
            keep[w] = 1;
            keptCount[ch] = 1;
            kind[ch] = KindFilled;
            return;
        }
    }
    float lowest = kInf;
    for (uint32_t j = 0; j < count; ++j) lowest = fmin_(lowest, items[(first + j) * octants + o].hi);
    uint32_t kept = 0;
    for (uint32_t j = 0; j < count; ++j) {
        const uint32_t w = (first + j) * octants + o;
        const bool k = items[w].status == 2 && !(items[w].lo > lowest);
        keep[w] = k ? ++kept : 0;
    }
    keptCount[ch] = kept;
    kind[ch] = kept == 0 ? KindEmpty : KindSurface;
}

__global__ void k_program_sizes(uint32_t nItems, const uint32_t* keep, const ItemOut* items, uint32_t* sizes) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nItems) return;
    sizes[w] = (keep[w] && !items[w].same) ? items[w].length + 1 : 0;
}

// Write the kept entries and their compacted programs.
__global__ void __launch_bounds__(256) k_emit(const ActEntry* act, uint32_t nItems, uint32_t octants, const EntryDev* entries,
                                              const uint32_t* scratchOffsets, const kdn_step* stScratch, uint32_t* alScratch,
                                              const ItemOut* items, const uint32_t* keep, const uint32_t* poolOffsets,
                                              uint64_t poolBase, const uint32_t* childEntryOffsets, uint64_t entryBase,
                                              EntryDev* entriesOut, kdn_step* pool) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nItems || !keep[w]) return;
    const uint32_t a = w / octants, o = w % octants;
    const ActEntry ae = act[a];
    const EntryDev entry = entries[ae.entry];
    const ItemOut it = items[w];
    const uint32_t child = ae.parent * octants + o;
    const uint64_t slot = entryBase + childEntryOffsets[child] + keep[w] - 1;
    if (it.same) {
        entriesOut[slot] = entry;
        return;
    }
    const uint64_t at = poolBase + poolOffsets[w];
    entriesOut[slot] = EntryDev{entry.object, (uint32_t)at};
    const uint32_t L = (uint32_t)(pool[entry.program].a) | ((uint32_t)pool[entry.program].b << 16);
    const size_t base = (size_t)scratchOffsets[a] * octants + o;
    const Strided<kdn_step> ST{const_cast<kdn_step*>(stScratch) + base, octants};
    const Strided<uint32_t> MAP{alScratch + base, octants};
    // compactedProgram (program.cpp:358-380): marks become new indices
    uint32_t kept = 0, flops = 0;
    kdn_step* dst = pool + at + 1;
    for (uint32_t i = 0; i < L; ++i) {
        if (!MAP[i]) continue;
        kdn_step s = ST[i];
        const int count = operand_count(s.op);
        s.a = count > 0 ? (uint16_t)MAP[s.a] : s.a;
        s.b = count > 1 ? (uint16_t)MAP[s.b] : 0;
        s.c = count > 2 ? (uint16_t)MAP[s.c] : 0;
        dst[kept] = s;
        flops += step_flops(s.op);
        MAP[i] = kept++;
    }
    kdn_step h;
    h.op = kHeaderOp;
    h.flags = 0;
    h.a = (uint16_t)(kept & 0xFFFF);
    h.b = (uint16_t)(kept >> 16);
    h.c = (uint16_t)MAP[it.result];
    h.value = (float)flops;  // the program's point-evaluation work
    h.leaf = entry.object;
    pool[at] = h;
}

__global__ void k_write_children(uint32_t nChildren, uint32_t octants, const NodeDev* nodes, const uint32_t* active,
                                 const uint32_t* kind, const uint32_t* keptCount, const uint32_t* childEntryOffsets,
                                 uint64_t entryBase, uint32_t level, uint32_t cap, NodeDev* out) {
    const uint32_t ch = blockIdx.x * blockDim.x + threadIdx.x;
    if (ch >= nChildren) return;
    NodeDev n;
    n.level = level;
    if (octants == 1) {
        n.x = n.y = n.z = 0;
    } else {
        const NodeDev parent = nodes[active[ch / octants]];
        const uint32_t o = ch % octants;
        n.x = 2 * parent.x + (o & 1);
        n.y = 2 * parent.y + ((o >> 1) & 1);
        n.z = 2 * parent.z + ((o >> 2) & 1);
    }
    n.kind = kind[ch];
    n.first_child = 0;
    n.entry_first = (uint32_t)(entryBase + childEntryOffsets[ch]);
    n.entry_count = keptCount[ch];
    (void)cap;
    out[ch] = n;
}

__global__ void k_root_entries(uint32_t nObjects, const kdn_object* objects, uint64_t poolBase, const uint32_t* poolOffsets,
                               const kdn_step* steps, kdn_step* pool, EntryDev* entries, ActEntry* act, uint32_t* lengths) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nObjects) return;
    const kdn_object ob = objects[i];
    const uint64_t at = poolBase + poolOffsets[i];
    kdn_step h;
    h.op = kHeaderOp;
    h.flags = 0;
    h.a = (uint16_t)(ob.count & 0xFFFF);
    h.b = (uint16_t)(ob.count >> 16);
    h.c = (uint16_t)ob.result;
    uint32_t flops = 0;
    for (uint32_t k = 0; k < ob.count; ++k) {
        pool[at + 1 + k] = steps[ob.first + k];
        flops += step_flops(steps[ob.first + k].op);
    }
    h.value = (float)flops;
    h.leaf = i;
    pool[at] = h;
    entries[i] = EntryDev{i, (uint32_t)at};
    act[i] = ActEntry{0, i};
    lengths[i] = ob.count;
}

__global__ void k_object_sizes(uint32_t n, const kdn_object* objects, uint32_t* sizes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sizes[i] = objects[i].count + 1;
}

// Slot registers for every program an entry names (one lane per entry): a
// step's value takes the lowest free slot once the operands whose last use
// it is have given theirs back, so a program needs as many slots as it
// keeps values live at once. The result lives to the end. Writes the
// program as 8-byte slot steps at the same offsets of `spool`, and again,
// with its literal steps folded into the steps that read them, at the same
// offsets of `mpool` (the march's programs). A program that needs more than
// kSlots raises *overflow. Entries share programs (a cut that changes nothing keeps its
// parent's), so the first lane to claim a program's bit in `claimed`
// allocates it and the others leave it alone.
__global__ void k_slot_programs(const EntryDev* entries, uint32_t nEntries, const kdn_step* pool, SlotStep* spool, SlotStep* mpool,
                                uint16_t* lastUse, uint8_t* slotOf, uint32_t* claimed, uint32_t* overflow) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nEntries) return;
    const uint32_t at = entries[e].program;
    const uint32_t bit = 1u << (at & 31);
    if (atomicOr(&claimed[at >> 5], bit) & bit) return;
    const kdn_step h = pool[at];
    const uint32_t L = (uint32_t)h.a | ((uint32_t)h.b << 16);
    const kdn_step* prog = pool + at + 1;
    uint16_t* last = lastUse + at + 1;
    uint8_t* slot = slotOf + at + 1;
    constexpr uint16_t kUnread = 0xFFFF, kKept = 0xFFFE;
    for (uint32_t i = 0; i < L; ++i) last[i] = kUnread;
    for (uint32_t i = 0; i < L; ++i) {
        const kdn_step st = prog[i];
        const int n = operand_count(st.op);
        if (n > 0) last[st.a] = (uint16_t)i;
        if (n > 1) last[st.b] = (uint16_t)i;
        if (n > 2) last[st.c] = (uint16_t)i;
    }
    if (L > 0) last[h.c] = kKept;
    uint32_t freeSlots = kSlots >= 32 ? 0xFFFFFFFFu : ((1u << kSlots) - 1);
    SlotStep* out = spool + at + 1;
    for (uint32_t i = 0; i < L; ++i) {
        const kdn_step st = prog[i];
        const int n = operand_count(st.op);
        if (n > 0 && last[st.a] == i) freeSlots |= 1u << slot[st.a];
        if (n > 1 && last[st.b] == i) freeSlots |= 1u << slot[st.b];
        if (n > 2 && last[st.c] == i) freeSlots |= 1u << slot[st.c];
        if (freeSlots == 0) {
            atomicMax(overflow, L);
            return;
        }
        const uint32_t d = (uint32_t)__builtin_ctz(freeSlots);
        freeSlots &= ~(1u << d);
        slot[i] = (uint8_t)d;
        const uint32_t a = n > 0 ? slot[st.a] : 0, b = n > 1 ? slot[st.b] : 0, cc = n > 2 ? slot[st.c] : 0;
        out[i] = SlotStep{(uint32_t)st.op | (d << 6) | (a << 10) | (b << 14) | (cc << 18), st.value};
        if (last[i] == kUnread) freeSlots |= 1u << d;  // a value nothing reads
    }
    spool[at] = SlotStep{L | ((uint32_t)(L > 0 ? slot[h.c] : 0) << 24), h.value};

    // The march's program: a coordinate or normal step is read from the
    // lane's point words, a constant from the reading step's value (the
    // first constant operand of a step; a second one stays in its slot), so
    // only the steps that compute remain, on the same slots. A literal step
    // is kept (as kMoveOp) where it is the result or some reader keeps it in
    // its slot. (last[] is free again: it marks the kept literals.)
    for (uint32_t i = 0; i < L; ++i) last[i] = 0;
    for (uint32_t i = 0; i < L; ++i) {
        const kdn_step st = prog[i];
        const int n = operand_count(st.op);
        bool immTaken = false;
        for (int j = 0; j < n; ++j) {
            const uint32_t src = j == 0 ? st.a : (j == 1 ? st.b : st.c);
            const uint8_t sop = prog[src].op;
            if (!is_literal(sop)) continue;
            if (immTaken) last[src] = 1;
            immTaken = true;
        }
    }
    SlotStep* mout = mpool + at + 1;
    uint32_t M = 0;
    for (uint32_t i = 0; i < L; ++i) {
        const kdn_step st = prog[i];
        const bool kept = i == (uint32_t)h.c || last[i] != 0;
        uint32_t code = 0;
        float value = 0.0f;
        if (is_literal(st.op) || point_word_of(st.op) != 0) {
            if (!kept) continue;
            // a kept literal: a copy of its own operand word or constant
            code = kMoveOp;
            const uint32_t w = point_word_of(st.op);
            if (w) {
                code |= w << 10;
            } else {
                code |= kImmA;
                value = st.value;
            }
        } else {
            code = st.op;
            const int n = operand_count(st.op);
            bool immTaken = false;
            for (int j = 0; j < n; ++j) {
                const uint32_t src = j == 0 ? st.a : (j == 1 ? st.b : st.c);
                const uint8_t sop = prog[src].op;
                uint32_t w = point_word_of(sop);
                if (w == 0 && is_literal(sop) && !immTaken) {
                    immTaken = true;
                    code |= kImmA << j;
                    value = prog[src].value;
                } else if (w == 0) {
                    w = slot[src];
                }
                code |= w << (10 + 5 * j);
            }
        }
        code |= (uint32_t)slot[i] << 6;
        mout[M++] = SlotStep{code, value};
    }
    mpool[at] = SlotStep{M | ((uint32_t)(L > 0 ? slot[h.c] : 0) << 24), h.value};
}

static inline uint32_t blocksFor(uint64_t n, uint32_t t = 256) { return (uint32_t)((n + t - 1) / t); }

// Grow a device buffer, keeping its first `keep` bytes.
static bool growKeep(Ctx& c, DevBuf& b, size_t need, size_t keepBytes) {
    if (b.bytes >= need) return true;
    size_t bytes = b.bytes ? b.bytes : (1u << 20);
    while (bytes < need) bytes *= 2;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        c.error = "hipMalloc of " + std::to_string(bytes) + " bytes for the octree failed";
        return false;
    }
    if (b.ptr && keepBytes) KDN_HIP(hipMemcpyAsync(p, b.ptr, keepBytes, hipMemcpyDeviceToDevice, c.stream));
    KDN_HIP(hipStreamSynchronize(c.stream));
    if (b.ptr) KDN_HIP(hipFree(b.ptr));
    b.ptr = p;
    b.bytes = bytes;
    return true;
}

bool buildOctree(Ctx& c, const float lower[3], const float size[3], int depth, kdn_octree_stats* stats) {
    const auto start = std::chrono::steady_clock::now();
    OctreeDev& t = c.tree;
    t.built = false;
    Region region;
    for (int k = 0; k < 3; ++k) {
        t.lower[k] = region.lower[k] = lower[k];
        t.size[k] = region.size[k] = size[k];
    }
    t.depth = depth;
    t.level_first.clear();
    t.level_count.clear();
    t.n_nodes = t.n_entries = t.n_pool = 0;
    hipStream_t s = c.stream;
    const uint32_t nObj = c.scene.n_objects;
    if (nObj == 0) {
        c.error = "no scene on the device";
        return false;
    }
    DevBuf& tmp = c.scratch[0];
    DevBuf& bufA = c.scratch[1];  // act entries + lengths + offsets
    DevBuf& bufI = c.scratch[2];  // item outputs, keep, sizes, offsets
    DevBuf& bufS = c.scratch[3];  // interval / step / alias scratch
    DevBuf& bufC = c.scratch[4];  // per child / per active arrays

    // ---- the root: every object program, cut over the region box -------
    {
        if (!ensure(bufC, (size_t)nObj * 8 + 64, c.error, false)) return false;
        uint32_t* sizes = (uint32_t*)bufC.ptr;
        uint32_t* offs = sizes + nObj;
        hipLaunchKernelGGL(k_object_sizes, dim3(blocksFor(nObj)), dim3(256), 0, s, nObj, c.scene.objects, sizes);
        uint64_t total = 0;
        if (!scanU32(sizes, offs, nObj, &total, tmp, s, c.error)) return false;
        if (!growKeep(c, t.pool, (total + kPoolPad) * sizeof(kdn_step), 0)) return false;
        if (!growKeep(c, t.entries, (size_t)nObj * sizeof(EntryDev), 0)) return false;
        if (!ensure(bufA, (size_t)nObj * (sizeof(ActEntry) + 8) + 64, c.error, false)) return false;
        ActEntry* act = (ActEntry*)bufA.ptr;
        uint32_t* lengths = (uint32_t*)(act + nObj);
        hipLaunchKernelGGL(k_root_entries, dim3(blocksFor(nObj)), dim3(256), 0, s, nObj, c.scene.objects, (uint64_t)0, offs,
                           c.scene.steps, (kdn_step*)t.pool.ptr, (EntryDev*)t.entries.ptr, act, lengths);
        t.n_pool = total;
        t.n_entries = nObj;
    }

    // Level step: `nAct` act entries of `nParents` parents, octants per parent.
    // Active parent list `active` indexes `parentNodes`; the root uses a
    // virtual parent (octants = 1).
    auto levelStep = [&](uint32_t nParents, uint32_t octants, const NodeDev* parentNodes, const uint32_t* active,
                         const uint32_t* entryOffsets, const uint32_t* entryCounts, uint32_t nAct, ActEntry* act,
                         uint32_t* lengths, uint32_t childLevel, NodeDev* childOut) -> bool {
        const uint64_t nItems = (uint64_t)nAct * octants;
        const uint32_t nChildren = nParents * octants;
        // scratch offsets from program lengths
        uint32_t* scratchOffsets = lengths + nAct;
        uint64_t steps = 0;
        if (!scanU32(lengths, scratchOffsets, nAct, &steps, tmp, s, c.error)) return false;
        const size_t cells = (size_t)steps * octants;
        if (!ensure(bufS, cells * (sizeof(Iv) + sizeof(kdn_step) + 4) + 256, c.error, false)) return false;
        Iv* ivs = (Iv*)bufS.ptr;
        kdn_step* sts = (kdn_step*)(ivs + cells);
        uint32_t* als = (uint32_t*)(sts + cells);
        // per item arrays
        const size_t itemBytes = nItems * (sizeof(ItemOut) + 12) + 256;
        if (!ensure(bufI, itemBytes, c.error, false)) return false;
        ItemOut* items = (ItemOut*)bufI.ptr;
        uint32_t* keep = (uint32_t*)(items + nItems);
        uint32_t* sizes = keep + nItems;
        uint32_t* poolOffsets = sizes + nItems;
        hipLaunchKernelGGL(k_cut_items, dim3(blocksFor(nItems)), dim3(256), 0, s, act, (uint32_t)nItems, octants, parentNodes,
                           active, (const EntryDev*)t.entries.ptr, (const kdn_step*)t.pool.ptr, scratchOffsets, ivs, sts, als,
                           region, childLevel, items);
        // children
        if (!ensure(c.scratch[5], (size_t)nChildren * 12 + 64, c.error, false)) return false;
        uint32_t* keptCount = (uint32_t*)c.scratch[5].ptr;
        uint32_t* kind = keptCount + nChildren;
        uint32_t* childOffsets = kind + nChildren;
        hipLaunchKernelGGL(k_cut_children, dim3(blocksFor(nChildren)), dim3(256), 0, s, nChildren, octants, entryOffsets, entryCounts,
                           items, keep, keptCount, kind);
        uint64_t keptTotal = 0;
        if (!scanU32(keptCount, childOffsets, nChildren, &keptTotal, tmp, s, c.error)) return false;
        hipLaunchKernelGGL(k_program_sizes, dim3(blocksFor(nItems)), dim3(256), 0, s, (uint32_t)nItems, keep, items, sizes);
        uint64_t newSteps = 0;
        if (!scanU32(sizes, poolOffsets, (uint32_t)nItems, &newSteps, tmp, s, c.error)) return false;
        if (t.n_pool + newSteps > 0xFFFFFFFFull) {
            c.error = "the octree program pool passes 2^32 steps";
            return false;
        }
        if (!growKeep(c, t.pool, (t.n_pool + newSteps + kPoolPad) * sizeof(kdn_step), t.n_pool * sizeof(kdn_step))) return false;
        if (!growKeep(c, t.entries, (t.n_entries + keptTotal) * sizeof(EntryDev), t.n_entries * sizeof(EntryDev))) return false;
        hipLaunchKernelGGL(k_emit, dim3(blocksFor(nItems)), dim3(256), 0, s, act, (uint32_t)nItems, octants,
                           (const EntryDev*)t.entries.ptr, scratchOffsets, sts, als, items, keep, poolOffsets, t.n_pool,
                           childOffsets, t.n_entries, (EntryDev*)t.entries.ptr, (kdn_step*)t.pool.ptr);
        hipLaunchKernelGGL(k_write_children, dim3(blocksFor(nChildren)), dim3(256), 0, s, nChildren, octants, parentNodes, active,
                           kind, keptCount, childOffsets, t.n_entries, childLevel, (uint32_t)depth, childOut);
        t.n_pool += newSteps;
        t.n_entries += keptTotal;
        return true;
    };

    // root level
    {
        if (!growKeep(c, t.nodes, sizeof(NodeDev) * 64, 0)) return false;
        ActEntry* act = (ActEntry*)bufA.ptr;
        uint32_t* lengths = (uint32_t*)(act + nObj);
        // offsets/counts of the virtual parent
        uint32_t* pc = (uint32_t*)bufC.ptr;  // reuse: [0] offset 0, [1] count
        const uint32_t host[2] = {0, nObj};
        KDN_HIP(hipMemcpyAsync(pc, host, 8, hipMemcpyHostToDevice, s));
        // lengths already hold program lengths; grow bufA for scratch offsets
        if (!levelStep(1, 1, nullptr, nullptr, pc, pc + 1, nObj, act, lengths, 0, (NodeDev*)t.nodes.ptr)) return false;
        t.level_first.push_back(0);
        t.level_count.push_back(1);
        t.n_nodes = 1;
    }

    for (int level = 0; level < depth; ++level) {
        const uint32_t first = t.level_first[level], count = t.level_count[level];
        NodeDev* levelNodes = (NodeDev*)t.nodes.ptr + first;
        if (!ensure(bufC, (size_t)count * 20 + 64, c.error, false)) return false;
        uint32_t* flags = (uint32_t*)bufC.ptr;
        uint32_t* rank = flags + count;
        hipLaunchKernelGGL(k_flag_surface, dim3(blocksFor(count)), dim3(256), 0, s, levelNodes, count, flags);
        uint64_t nActive = 0;
        if (!scanU32(flags, rank, count, &nActive, tmp, s, c.error)) return false;
        if (nActive == 0) break;
        const uint64_t childBase = t.n_nodes;
        const uint64_t nChildren = nActive * 8;
        if (!growKeep(c, t.nodes, (childBase + nChildren) * sizeof(NodeDev), childBase * sizeof(NodeDev))) return false;
        levelNodes = (NodeDev*)t.nodes.ptr + first;
        uint32_t* active = rank + count;
        uint32_t* entryCounts = active + nActive;
        uint32_t* entryOffsets = entryCounts + nActive;
        hipLaunchKernelGGL(k_mark_branches, dim3(blocksFor(count)), dim3(256), 0, s, levelNodes, count, flags, rank,
                           (uint32_t)childBase, active, entryCounts);
        uint64_t nAct = 0;
        if (!scanU32(entryCounts, entryOffsets, (uint32_t)nActive, &nAct, tmp, s, c.error)) return false;
        if (!ensure(bufA, (size_t)nAct * (sizeof(ActEntry) + 8) + 64, c.error, false)) return false;
        ActEntry* act = (ActEntry*)bufA.ptr;
        uint32_t* lengths = (uint32_t*)(act + nAct);
        hipLaunchKernelGGL(k_fill_act, dim3(blocksFor(nActive)), dim3(256), 0, s, levelNodes, active, (uint32_t)nActive, entryOffsets,
                           act, (const EntryDev*)t.entries.ptr, (const kdn_step*)t.pool.ptr, lengths);
        NodeDev* childOut = (NodeDev*)t.nodes.ptr + childBase;
        if (!levelStep((uint32_t)nActive, 8, levelNodes, active, entryOffsets, entryCounts, (uint32_t)nAct, act, lengths,
                       (uint32_t)level + 1, childOut)) {
            return false;
        }
        t.level_first.push_back((uint32_t)childBase);
        t.level_count.push_back((uint32_t)nChildren);
        t.n_nodes = childBase + nChildren;
    }
    {
        // the march's programs on slot registers
        if (!growKeep(c, t.spool, (t.n_pool + kPoolPad) * sizeof(SlotStep), 0)) return false;
        if (!growKeep(c, t.mpool, (t.n_pool + kPoolPad) * sizeof(SlotStep), 0)) return false;
        const size_t claimWords = (t.n_pool + 31) / 32;
        if (!ensure(bufS, t.n_pool * 3 + 256, c.error, false)) return false;
        if (!ensure(bufC, claimWords * 4 + 64, c.error, false)) return false;
        uint16_t* lastUse = (uint16_t*)bufS.ptr;
        uint8_t* slotOf = (uint8_t*)(lastUse + t.n_pool);
        uint32_t* overflow = (uint32_t*)bufC.ptr;
        uint32_t* claimed = overflow + 16;
        KDN_HIP(hipMemsetAsync(bufC.ptr, 0, claimWords * 4 + 64, s));
        hipLaunchKernelGGL(k_slot_programs, dim3(blocksFor(t.n_entries)), dim3(256), 0, s, (const EntryDev*)t.entries.ptr,
                           (uint32_t)t.n_entries, (const kdn_step*)t.pool.ptr, (SlotStep*)t.spool.ptr, (SlotStep*)t.mpool.ptr, lastUse,
                           slotOf, claimed, overflow);
        uint32_t longest = 0;
        KDN_HIP(hipMemcpyAsync(&longest, overflow, 4, hipMemcpyDeviceToHost, s));
        KDN_HIP(hipStreamSynchronize(s));
        if (longest) {
            c.error = "a program of " + std::to_string(longest) + " steps keeps more than " + std::to_string(kSlots) +
                      " values live (the march's slot registers)";
            return false;
        }
    }
    KDN_HIP(hipStreamSynchronize(s));
    KDN_HIP(hipGetLastError());
    if (t.level_first_dev) KDN_HIP(hipFree(t.level_first_dev));
    KDN_HIP(hipMalloc(&t.level_first_dev, 4 * (t.level_first.size() + 1)));
    KDN_HIP(hipMemcpy(t.level_first_dev, t.level_first.data(), 4 * t.level_first.size(), hipMemcpyHostToDevice));
    t.built = true;

    if (stats) {
        *stats = kdn_octree_stats{};
        stats->levels = (uint32_t)t.level_first.size();
        std::vector<NodeDev> host(t.n_nodes);
        KDN_HIP(hipMemcpy(host.data(), t.nodes.ptr, t.n_nodes * sizeof(NodeDev), hipMemcpyDeviceToHost));
        for (size_t l = 0; l < t.level_first.size() && l < 16; ++l) {
            stats->nodes[l] = t.level_count[l];
            for (uint32_t i = 0; i < t.level_count[l]; ++i) {
                const NodeDev& n = host[t.level_first[l] + i];
                stats->surface[l] += (n.kind == KindSurface || n.kind == KindBranch) ? 1 : 0;
                stats->filled[l] += n.kind == KindFilled ? 1 : 0;
            }
        }
        stats->entries = (uint32_t)t.n_entries;
        stats->pool_steps = t.n_pool;
        stats->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
    }
    return true;
}

}  // namespace kdn
