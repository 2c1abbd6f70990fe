// topology.h — seed-independent scaffold topology, scheduled for wave64 lanes.
//
// The reference uploads the raw topology (np, p0, p1, level order, slot map,
// rank) and labels level by level (common/src/cuda/label_gpu_common.cu:72-254).
// Here the host additionally turns the DAG into a *step schedule*: a sequence
// of steps of at most `lanes` nodes each, every node placed in a later step
// than all of its predecessors.  A workgroup of `lanes` threads labels one
// step per iteration with one barrier between steps.  Every node also gets a
// *location*: outputs live directly at their challenge rank in the persisted
// super-chunk, every other node in a recycled scratch slot.
//
// Packed record per scheduled position (4 x u32, 16 B):
//   x = destination location     y = pred0 location     z = pred1 location
//   w = (node_id << 2) | num_preds, or kEmptyRecord for an idle lane
// A location with bit 31 set is rank (loc & 0x7fffffff) in the persisted
// super-chunk; otherwise it is a scratch slot index.
//
// Records are laid out as n_steps + 1 rows of exactly `lanes` entries (short
// steps are padded with idle records {kDiscardSlot, kZeroSlot, kZeroSlot,
// kEmptyRecord}; the extra last row is all idle), so lane t of step s reads
// record s*lanes + t: the device needs no step table, prefetches row s+1
// unconditionally, and runs the same instruction stream in every lane.
//
// Near table, one u32 per record position: bits 0-15 for pred0, 16-31 for
// pred1, each the lane that wrote that predecessor if it was written in the
// step immediately before, else kFarPred.  The four-lanes-per-node kernel reads
// such predecessors from an on-chip copy of the previous step's labels and
// loads all others one step early.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace pose {

// Allocator whose resize() leaves new elements uninitialised: the schedule
// arrays (hundreds of MB at 2^17) are written in full by parallel passes, so a
// serial zero fill first would only cost time.
template <typename T>
struct NoInitAllocator : std::allocator<T> {
    template <typename U> struct rebind { using other = NoInitAllocator<U>; };
    NoInitAllocator() = default;
    template <typename U> NoInitAllocator(const NoInitAllocator<U>&) noexcept {}
    template <typename U> void construct(U* p) noexcept { ::new (static_cast<void*>(p)) U; }
    template <typename U, typename... A> void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};
using U32Buf = std::vector<uint32_t, NoInitAllocator<uint32_t>>;
using I32Buf = std::vector<int32_t, NoInitAllocator<int32_t>>;
using U8Buf = std::vector<uint8_t, NoInitAllocator<uint8_t>>;

constexpr uint32_t kPersistBit = 0x80000000u;
constexpr uint32_t kEmptyRecord = 3u;   // num_preds field 3: no node in this lane
// Reserved scratch slots: slot 0 always holds 32 zero bytes and stands for an
// absent predecessor (the message block is zero-filled there); slot 1 takes the
// discarded result of an idle lane.  Real labels use slots 2 and up.
constexpr uint32_t kZeroSlot = 0, kDiscardSlot = 1, kFirstSlot = 2;
constexpr uint32_t kFarPred = 0xFFFFu;   // near-table entry: not from the previous step

enum SchedMode : int {
    kSchedLevels = 0,   // ASAP dependency levels, each split into <= lanes steps
    kSchedList = 1,     // list schedule, longest remaining path first
};

// A step schedule of one node range of the scaffold.  Scratch slots are
// numbered within the range's own pool; outputs keep their absolute rank.
struct Schedule {
    uint32_t id_lo = 0, id_hi = 0;      // node ids [id_lo, id_hi)
    uint32_t rank_lo = 0, rank_hi = 0;  // its output ranks [rank_lo, rank_hi)
    uint32_t n_steps = 0;
    uint32_t pool = 0;                  // scratch slots (incl. reserved)
    uint32_t ring_steps = 0;
    std::vector<uint32_t> step_start;   // n_steps + 1 prefix offsets
    U32Buf rec;                         // 4 words per position, (n_steps+1)*lanes positions
    U32Buf near;                        // 1 word per position
};

struct HostTopology {
    uint32_t cb = 0;
    int n = 0;                      // graph parameter n
    uint64_t count = 0;             // scaffold nodes per super-chunk
    uint32_t lanes = 0;             // max nodes per step
    int mode = 0;
    uint32_t levels = 0;            // ASAP depth (critical path, in nodes)
    uint32_t n_steps = 0;
    uint32_t pool = 0;              // scratch slots per super-chunk (incl. reserved)
    uint32_t ring_steps = 0;        // lifetime bound of ring-page slots (steps)
    std::vector<uint32_t> step_start;   // n_steps + 1 prefix offsets
    U32Buf rec;                         // 4 words per position, (n_steps+1)*lanes positions
    U32Buf near;                        // 1 word per position (see the near table above)
    // The scaffold is two copies of one standalone graph with no edge between
    // them (ids [0, count/2) and [count/2, count)); part[p] schedules copy p
    // alone.  A copy's critical path is the scaffold's, with half the nodes,
    // so labeling the copies on two workgroups shortens a super-chunk's
    // latency when the device has idle compute units.
    bool has_parts = false;
    Schedule part[2];
    // Copy 1's schedule is copy 0's with node ids shifted by part1_id_shift
    // and output ranks by part1_rank_shift (mirrored copies whose ranks shift
    // by one constant: every power-of-two chunk_blocks).  Its device locations
    // are then copy 0's, so on the device the two copies share records and
    // near entries and differ only in block-0 CVs.
    bool part1_shifted = false;
    uint32_t part1_id_shift = 0, part1_rank_shift = 0;
    uint8_t desc[32] = {0};
    uint32_t key_src[8] = {0}, key_int[8] = {0};
    double build_seconds = 0;
};

// Graph parameter n = max(0, bitlen(m-1) - 1) (common/src/graph.c:562-570).
int graph_parameter_n(uint64_t m);
// 2 * S(n+1) (common/src/graph.c:33-54, 647-652).
uint64_t scaffold_node_count(uint64_t m);

// Raw topology in id order (the reference's pose_graph_edges,
// common/src/graph.c:1044-1100): np, sorted preds, challenge rank or -1.
int emit_topology(uint32_t cb, U8Buf& np, U32Buf& p0, U32Buf& p1, I32Buf& rank);

// Builds the whole-scaffold schedule and, with with_parts, the schedules of
// its two copies (built concurrently).
int build_host_topology(uint32_t cb, uint32_t lanes, int mode, HostTopology& out,
                        std::string* err, bool with_parts = false);

}  // namespace pose
