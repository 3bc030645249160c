// Device data layouts shared by the kernels and the host: octree nodes and
// entries. Free of HIP headers so that a CPU build of the kernels can use
// them.
#pragma once

#include <cstdint>

// A bounds assertion that only the bounds-checking CPU build of the
// kernels defines; on the device it compiles to nothing.
#ifndef KDN_CHECK
#define KDN_CHECK(cond) ((void)0)
#endif

namespace kdn {

// An octree node as kept on the device, 32 bytes: the downloaded
// kdn_node_record layout.
struct NodeDev {
    uint32_t level, x, y, z, kind, first_child, entry_first, entry_count;
};

// Slot registers of the march's LDS register file per lane: the most
// values a program keeps live at once after the build's allocation (the
// build fails beyond it), and the trace kernels' block size: one wave, so
// that a block's LDS and wave slot free as soon as its own slowest ray
// ends rather than its block's (1.3x the samples per second of 256-lane
// blocks on the production scene at 1080p).
constexpr int kSlots = 16;
// The march's per-lane LDS register file: the kSlots slot registers, then
// the point being evaluated (x, y, z) and a zero, which the march's
// programs read as operands in place of their coordinate and normal steps.
constexpr int kLaneWords = kSlots + 4;
constexpr uint32_t kPointWord = kSlots;  // x, y, z at kPointWord + 0..2, zero at + 3
constexpr int kTraceBlock = 64;
// Waves per SIMD the trace kernels are compiled for: the march is bound by
// the latency of its dependent fetches, so more waves in flight hide more
// of it (the LDS of a lane, 128 B, allows five).
constexpr int kTraceWaves = 5;
// A pass finishes its paths in one launch once no more than one in
// kFinishShare of a chunk's paths are left: with 1, every path runs in the
// finishing launch from its camera ray on (its lanes pull paths and run a
// segment an iteration, so a wave waits only for its slowest segment; 1.2x
// the samples per second of a wavefront launch per bounce at 1080p). A
// learned-guide round keeps a launch per bounce over the guided depths.
constexpr uint32_t kFinishShare = 1;
// Queues of at least this many rays are sorted by direction octant and
// origin cell before their march launch (smaller ones are not worth a sort).
constexpr uint32_t kSortMin = 1u << 16;
// Octree levels a march keeps the ancestry of (depth caps 0..kMaxDepth):
// with the lane's kLaneWords slot words, 128 bytes of LDS a lane, which
// lets five one-wave blocks share each SIMD.
constexpr int kPathLevels = 12;
constexpr int kMaxDepth = kPathLevels - 1;

// A step of a program on slot registers, 8 bytes: code = op (bits 0-5) |
// destination slot (6-9) | operand slots a (10-13), b (14-17), c (18-21);
// value = the constant. A program's header has code = length (bits 0-23) |
// result slot (24-27) and value = its FP32 operation count.
//
// The march reads the same programs with their literal steps folded away
// (the march pool): code = op (bits 0-5) | destination slot (6-9) | operand
// words a (10-14), b (15-19), c (20-24) of the lane's register file (a slot,
// the point's x, y, z or the zero) | bits 25-27: operand a, b or c is the
// step's constant, `value` (a step takes at most one). A literal step stays
// only where it is the result or a second constant of its reader, as
// kMoveOp copying its operand a. The header is as above.
struct SlotStep {
    uint32_t code;
    float value;
};
constexpr uint32_t kMoveOp = 63;
constexpr uint32_t kImmA = 1u << 25, kImmB = 1u << 26, kImmC = 1u << 27;

// Steps fetched at once by an evaluation (independent loads in flight
// together), and the steps every program pool keeps readable past its
// last program so that a fetch may run over the end.
constexpr int kSlotFetch = 8;
constexpr int kPoolPad = 16;

enum NodeKind : uint32_t { KindEmpty = 0, KindFilled = 1, KindSurface = 2, KindBranch = 3 };

struct EntryDev {
    uint32_t object, program;
};

}  // namespace kdn
