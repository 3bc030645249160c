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

// Registers of the trace's per-lane register file: the longest program a
// surface leaf may hold (the octree build fails beyond it).
constexpr int kMaxRegs = 384;
// Slot registers of the march's LDS register file per lane: the most
// values a program keeps live at once after the build's allocation (the
// build fails beyond it), and the trace kernels' block size.
constexpr int kSlots = 16;
constexpr int kTraceBlock = 256;
// Octree levels a march keeps the ancestry of (depth caps 0..14).
constexpr int kPathLevels = 16;

enum NodeKind : uint32_t { KindEmpty = 0, KindFilled = 1, KindSurface = 2, KindBranch = 3 };

struct EntryDev {
    uint32_t object, program;
};

}  // namespace kdn
