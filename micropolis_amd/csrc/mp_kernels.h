// Kernel launch wrappers of the REYES hot path (implemented in mp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mp_types.h"

namespace mp {

// Work-sharing state of the persistent split (reset by launch_reset).
struct ShareState {
    uint32_t head, tail;     // chunk queue: popped / reserved chunk counts
    uint32_t active;         // work-groups currently holding work
    uint32_t epoch;          // frame stamp of published chunks
};

struct SplitQueues {
    RangeRec* queue;         // private LIFO queues, stride entries each
    int32_t* queue_cnt;      // entries per private queue
    int64_t stride;
    RangeRec* shared_q;      // BALANCED: one shared LIFO queue
    int32_t* shared_cnt;
    int32_t n_groups;        // persistent work-groups
    int32_t stack_cap;       // LDS ring-stack entries per work-group
    RangeRec* chunks;        // spilled stack chunks, 64 ranges each
    uint32_t* chunk_count;   // ranges in each chunk
    uint32_t* chunk_ready;   // epoch stamp once a chunk is published
    uint32_t chunk_cap;      // 0 disables work sharing
    int32_t spill_at;        // spill 64 ranges when the stack exceeds this
    int32_t linger;          // idle groups wait (bounded) for spilled chunks
    ShareState* share;
    const uint32_t* pending; // if set and zero on entry, there is no work: exit
    int32_t spread;          // 1: narrow steps one sample per lane; 2: also two levels per
                             // step on the narrowest (speculative children)
    // step budget: a group that has run step_budget steps (0: no budget)
    // appends its queued and stacked ranges to `left` and exits; a second
    // launch deals them round-robin to all groups
    int32_t step_budget;
    RangeRec* left;
    uint32_t* left_cnt;
    uint32_t left_cap;
    // if set: the private queue of group g is strided[g], strided[g + G], ...
    // (G groups; *pending entries in all, at most left_cap)
    const RangeRec* strided;
    int32_t waves;           // kernel build: 2 or 3 waves per SIMD (registers vs groups)
};

struct CanonBuffers {
    RangeRec* draw;          // DRAW records in emission order
    RangeRec* bucket;        // grouped by global patch index
    RangeRec* canon;         // canonical order
    unsigned long long* keys;  // canonical keys of the bucketed records
    uint2* span;             // [offsets[g], offsets[g + 1]) of each bucketed record's patch
    uint32_t cap;            // capacity of the three lists
    uint32_t* counts;        // per global patch: DRAW count
    uint32_t* offsets;       // exclusive scan of counts (n_gpids + 1)
    uint32_t* cursor;        // scatter cursors
    uint32_t* part;          // block sums of the counts (many patches)
    uint32_t n_gpids;
    // radix-sort lists of buckets too large for the rank count: 2 x cap keys
    // and 2 x cap slot indices (ping-pong)
    unsigned long long* sort_keys;
    uint32_t* sort_vals;
    // forest split with in-group ranks (null where the path is not used):
    // group g's DRAW ranges at forest_out[g * kForestKeys ..], their ranks
    // among their patches' DRAW ranges at the same indices of rank, their
    // number at forest_total[g] (forest_groups groups)
    RangeRec* forest_out = nullptr;
    uint32_t* rank = nullptr;
    uint32_t* forest_total = nullptr;
    uint32_t forest_groups = 0;
};

// DRAW ranges one forest group can rank (its LDS key store and its output
// region); a group with more flags overflow bit 32.
constexpr int kForestKeys = 1536;

hipError_t launch_reset(FrameCounters* fc, ShareState* share, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, size_t n, unsigned long long v, hipStream_t s);
hipError_t launch_atomic_rate(unsigned long long* keys, uint32_t n, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);

// reset_fc / reset_share (may be null): the frame's counters are reset by the
// same launch (as launch_reset does)
hipError_t launch_init_queues(const DevObject* objs, int n_objs, uint32_t total_roots,
                              const SplitQueues& q, int balanced, FrameCounters* reset_fc,
                              ShareState* reset_share, hipStream_t s);
hipError_t launch_split_persistent(const DevObject* objs, const DevView* views,
                                   const DevParams& prm, const SplitQueues& q,
                                   const CanonBuffers& cb, FrameCounters* fc, int balanced,
                                   hipStream_t s);

// BREADTH / BOUNDED segment engine (one level or batch).
hipError_t launch_bound_segment(const DevObject* objs, const DevView* views,
                                const DevParams& prm, const RangeRec* in, RangeRec* pad,
                                uint32_t n, uint8_t* flags, uint32_t* blk_counts,
                                FrameCounters* fc, hipStream_t s);
hipError_t launch_scan_blocks(uint32_t* blk_counts, uint32_t n_blocks, uint32_t* totals,
                              hipStream_t s);
hipError_t launch_move_segment(const DevObject* objs, const RangeRec* pad, uint32_t n,
                               const uint8_t* flags, const uint32_t* blk_offsets,
                               RangeRec* draw_out, uint32_t draw_base, uint32_t draw_limit,
                               RangeRec* split_out, uint32_t* counts, hipStream_t s);
uint32_t segment_blocks(uint32_t n);

hipError_t launch_canonicalize(const DevObject* objs, DevView* views, int n_views,
                               const CanonBuffers& cb, FrameCounters* fc, int num_cus, int fused,
                               hipStream_t s);
// count == 0: every range of the frame (fc->total_ranges); otherwise the
// canonical ranges [first, first + count) with the static range schedule
hipError_t launch_raster(int patch_size, const RangeRec* canon, const DevObject* objs,
                         const DevView* views, const DevParams& prm, FrameCounters* fc,
                         float* quad_store, int grid, hipStream_t s, uint32_t first = 0, uint32_t count = 0);
// sets views[view].range_base (view >= 0) and, when fc is given, fc->total_ranges
hipError_t launch_set_view_base(DevView* views, int view, uint32_t base, FrameCounters* fc, uint32_t total,
                                hipStream_t s);
int raster_grid(int patch_size, int num_cus, const DevParams& prm);
// End of a frame: the frame's counters are copied to `out` (what the host
// and the later kernels read) and reset for the next frame together with
// the split's shared state (work-sharing chunks, the list of ranges left by a
// step budget, the BALANCED root count), so the next device frame starts
// directly with its split.  fc == nullptr: nothing is done.
struct FrameEnd {
    FrameCounters* fc = nullptr;
    FrameCounters* out = nullptr;
    ShareState* share = nullptr;
    uint32_t* left_cnt = nullptr;     // reset to 0
    int32_t* shared_cnt = nullptr;    // reset to shared_total
    int32_t shared_total = 0;
    uint32_t* counts = nullptr;       // per-patch DRAW counts cleared (n_counts), after a
    uint32_t n_counts = 0;            // ranked frame that wrote them with plain stores
    bool resolve_clears = false;      // set by launch_resolve: its threads clear them
};
hipError_t launch_resolve(const DevView* views, int n_views, int max_pixels,
                          const RangeRec* canon, const DevObject* objs, const float* quad_store,
                          int patch_size, int write_rgba32, const float* display_tab, hipStream_t s,
                          const FrameEnd& fe = FrameEnd());
// the frame-end step alone (frames that resolve nothing)
hipError_t launch_frame_end(const FrameEnd& fe, hipStream_t s);

hipError_t launch_init_roots(int obj, uint32_t n, RangeRec* out, hipStream_t s);

// debug tap: the diced grids of canonical ranges [first, first + n)
hipError_t launch_read_grids(int patch_size, const RangeRec* canon, const DevObject* objs, const DevView* views,
                             const DevParams& prm, uint32_t first, uint32_t n, float4* pos, int2* pxl, float* dep,
                             hipStream_t s);

// reverse pass of the differentiable extension
hipError_t launch_adjoint(int patch_size, const RangeRec* canon, const DevObject* objs,
                          const DevView* views, int n_views, int max_pixels, const DevParams& prm,
                          const FrameCounters* fc, const float* quad_store, const float* dl_dcolor,
                          uint32_t* visible, uint32_t* vis_list, uint32_t* vis_count, float* grad_cps,
                          float* grad_color, float* grad_amp, int grid, hipStream_t s);
int adjoint_grid(int patch_size, int num_cus);
int level_grid(int num_cus, int lanes_per_range, int per_cu);
hipError_t launch_split_levels(const DevObject* objs, int n_objs, const DevView* views, const DevParams& prm,
                               uint32_t total_roots, RangeRec* list_a, RangeRec* list_b, uint32_t list_cap,
                               uint32_t* level_cnt, int n_levels, const SplitQueues& q,
                               const CanonBuffers& cb, FrameCounters* fc, int grid, int lanes_per_range,
                               ShareState* reset_share, hipStream_t s);
// After a budgeted persistent split (q.step_budget > 0): finish the split
// without a budget, group g taking q.left[g], q.left[g + G], ...
hipError_t launch_split_rebalanced(const DevObject* objs, const DevView* views, const DevParams& prm,
                                   const SplitQueues& q, const CanonBuffers& cb, FrameCounters* fc,
                                   hipStream_t s);
hipError_t launch_xyz_to_cps(const float* xyz, float4* cps, size_t n, hipStream_t s);
int split_persistent_occupancy(size_t lds_bytes, int waves);   // waves: 2 or 3 (SplitQueues::waves)
// Forest split (LOCAL / BALANCED): group g splits the whole trees of the roots
// in its private queue (k_init_queues' non-shared layout) on an LDS stack of
// q.stack_cap entries (at least kForestMinStack).
// patch_types: bit 0 some object has Bezier patches, bit 1 some has Gregory
// patches (the kernel build without the other evaluation is launched)
// ranked: the groups also rank their DRAW ranges inside each patch
// (cb.rank: the range's global patch index << 16 | its rank) and write
// every root's DRAW count with a plain store, for
// launch_place_ranked; a group with more DRAW ranges than its LDS key store
// flags overflow bit 32 (the frame must then run unranked).  Unranked: the
// counts are added atomically (zero on entry) for launch_canonicalize.
hipError_t launch_split_forest(const DevObject* objs, const DevView* views, const DevParams& prm,
                               const SplitQueues& q, const CanonBuffers& cb, FrameCounters* fc, int patch_types,
                               int n_objs, int n_views, int ranked, hipStream_t s);
// whether a frame of n_gpids patches can take the ranked forest order
bool forest_rank_fits(uint32_t n_gpids);
// The canonical order of a ranked forest frame in one launch: every
// work-group scans the per-patch counts and places its share of the draw
// list at offset(patch) + rank (no bucket pass, no rank kernel).
hipError_t launch_place_ranked(DevView* views, int n_views, const CanonBuffers& cb,
                               FrameCounters* fc, int num_cus, hipStream_t s);
// whether a frame's object and view tables fit the forest split's LDS copy
bool split_forest_fits(int n_objs, int n_views);
int split_forest_occupancy(size_t lds_bytes);
size_t split_forest_lds(int stack_cap);
constexpr int kForestMinStack = 512;
size_t split_persistent_lds(int stack_cap);   // LDS bytes of one persistent split group

}  // namespace mp
