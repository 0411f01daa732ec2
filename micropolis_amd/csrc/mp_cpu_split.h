// CPU bound&split method: host depth-first split of every object of a frame.
//
// Replaces Reyes::BoundNSplitCLCPU (src/Reyes/BoundNSplitCLCPU.cpp:31-336):
// the same per-range predicate (eye-space 3x3 samples, Projection::bound
// screen size, split the longer eye-space side), evaluated on the host cores
// with a pool of threads over the patch roots.  The DRAW set does not depend
// on the thread count or order; the per-pass counts are those of the
// reference's single-threaded stack (batches of reyes_patches_per_pass).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <vector>

#include "mp_types.h"

namespace mp {

struct CpuSplitObject {
    const float* cps;        // host control points, 4 floats per point (w = 1)
    int type;                // MP_TYPE_BEZIER / MP_TYPE_GREGORY
    int n_patches;
    float mv[16];            // column-major modelview
    int width, height;       // viewport of the object's view
    float fx, fy, near_;     // Projection::f() and near plane
    float split_limit;
    uint32_t gpid_base;      // global patch index of patch 0
    uint32_t slot;           // object slot stored in RangeRec::obj_depth
};

struct CpuSplitFrame {
    std::vector<RangeRec> draws;        // all DRAW ranges, any order
    std::vector<uint32_t> counts;       // DRAW ranges per global patch index
    std::vector<int32_t> pass_counts;   // per object, batches of patches_per_pass
    int64_t bounds = 0;                 // ranges bounded
    int64_t max_stack = 0;              // largest height of the reference's stack
};

// Splits all objects.  n_threads <= 0 picks the host's core count (at most 16).
// out.draws receives every DRAW range in canonical order (object slot, patch,
// umin, vmin).
void cpu_split_frame(const std::vector<CpuSplitObject>& objs, uint32_t n_gpids,
                     float cull_ribbon, int max_split_depth, int patches_per_pass,
                     int n_threads, CpuSplitFrame& out);

// The same split, streamed: chunks of patches are split by n_threads worker
// threads while the calling thread hands the finished chunks, in canonical
// order (object slot, patch, umin, vmin), to `emit`.  out.draws stays empty;
// counts, pass counts, bounds and stack height are filled as by
// cpu_split_frame.
void cpu_split_stream(const std::vector<CpuSplitObject>& objs, uint32_t n_gpids, float cull_ribbon,
                      int max_split_depth, int patches_per_pass, int n_threads, CpuSplitFrame& out,
                      const std::function<void(const RangeRec*, size_t)>& emit);

// One range (exposed for tests of the library's host code): returns MP_CULL,
// MP_DRAW, MP_CPU_DROP, MP_HSPLIT (halve v) or MP_VSPLIT (halve u).
#define MP_CPU_DROP 8
int cpu_bound_range(const CpuSplitObject& o, const float* cp, int depth, float umin, float vmin,
                    float umax, float vmax, float cull_ribbon, int max_split_depth);

}  // namespace mp
