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
// Octree levels a march keeps the ancestry of (depth caps 0..14).
constexpr int kPathLevels = 16;

// A step of a program on slot registers, 8 bytes: code = op (bits 0-5) |
// destination slot (6-9) | operand slots a (10-13), b (14-17), c (18-21);
// value = the constant. A program's header has code = length (bits 0-23) |
// result slot (24-27) and value = its FP32 operation count.
struct SlotStep {
    uint32_t code;
    float value;
};

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
