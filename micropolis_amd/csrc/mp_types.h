// Plain-old-data records shared by host and device code.
#pragma once

#include <stdint.h>

#define MP_CULL   0
#define MP_DRAW   1
#define MP_HSPLIT 2
#define MP_VSPLIT 6

#define MP_TYPE_BEZIER  0
#define MP_TYPE_GREGORY 1

#define MP_EMPTY_KEY 0xFFFFFFFFFFFFFFFFull

namespace mp {

// One node of the split tree: patch id, object slot (24 bits) | depth << 24,
// and the (u,v) range.  24 bytes.  (PatchRange, src/Reyes/PatchRange.h:9-27;
// LOCAL packs depth the same way, bound_n_split_local.cl:103-108.)
struct RangeRec {
    uint32_t pid;
    uint32_t obj_depth;
    float umin, vmin, umax, vmax;
};

// Camera / framebuffer of one view (Reyes::Projection + Framebuffer).
struct DevView {
    float proj[16];
    float screen[4];
    float f[2];
    float near_, far_, eps;
    int32_t width, height;
    int32_t grid_w, grid_h;      // 8x8 tiles per row / column
    int32_t vs_x, vs_y;          // width << subpixel_bits, height << subpixel_bits
    float clear_linear[4];
    unsigned long long* keys;    // tiled grid_w*grid_h*64 visibility keys
    uint8_t* rgba8;              // resolve outputs, row-major width*height
    float* depth;
    uint32_t* ids;
    float* rgba32;
    uint32_t first_gpid;         // global patch index of the view's first object
    uint32_t range_base;         // canonical index of its first range (device-written)
    uint32_t range_count;        // device-written
    uint32_t gpid_end;           // one past the global patch index of its last object
    uint32_t pixel_offset;       // first pixel of this view in multi-view buffers
    uint32_t clear_rgba8;        // display bytes of the clear colour (alpha 255)
    int32_t solo_obj;            // the view's only object, -1 when it has several or none
};

// One draw_patches call: patch buffer + modelview + colour.
struct DevObject {
    const float4* cps;           // 16 or 20 float4 per patch (w = 1)
    float mv[16];
    float color[4];
    float split_limit;
    float disp_amplitude;
    int32_t type;
    int32_t n_patches;
    int32_t view;
    uint32_t gpid_base;          // global patch index of patch 0
    uint32_t grad_offset;        // float offset of its control-point gradients
    int32_t draw_index;          // position in draw-call order
};

struct DevParams {
    int32_t patch_size;
    int32_t max_split_depth;
    float cull_ribbon;
    int32_t backface_culling;
    int32_t subpixel_bits;
    int32_t patches_per_pass;
    int32_t displacement;
    float disp_frequency;
    uint32_t disp_seed;
    int32_t disp_octaves;
    int32_t noise_wave_cells;    // hash wave-uniform lattice cells once per wave
    int32_t raster_one_wave;     // N = 16: one-wave raster work-groups
    int32_t raster_chunk;        // 0: static range stride; -1: dynamic, sized
                                 // from the range count; > 0: dynamic, fixed;
                                 // <= -2: static chunks of -raster_chunk ranges
};

// Device-side frame counters, copied back once per frame.
struct FrameCounters {
    uint32_t draw_count;         // DRAW records written to the draw list
    uint32_t overflow;           // draw list capacity exceeded
    unsigned long long bounds;   // bound() evaluations
    uint32_t slow_blocks;        // raster blocks on the wide-integer path
    uint32_t max_stack;          // largest LDS stack height seen
    uint32_t total_ranges;       // canonical ranges (written by the scan)
    uint32_t big_ranges;         // ranges of buckets ordered by the radix sort
    unsigned long long atomics;  // (micropolygon, pixel) coverage events = key atomics
    uint32_t raster_next;        // next canonical range to rasterise
};

// The counters every work-group of a large launch adds to (coverage samples,
// bound() evaluations, the stack maximum) are accumulated in kStripes
// stripes, one per 128-byte line, that follow the FrameCounters in the same
// allocation: atomics from thousands of work-groups to one address are
// performed one after another (4096 of them at a launch's end cost ~50 us),
// to 64 addresses in parallel.  The frame end folds the stripes into the
// counters.
constexpr int kStripes = 64;
constexpr int kStripeWords = 16;
constexpr size_t kFrameCountersBytes = sizeof(FrameCounters) + (size_t)kStripes * kStripeWords * 8;
enum { kStripeAtomics = 0, kStripeBounds = 1, kStripeMaxStack = 2, kStripeUsed = 3 };

}  // namespace mp
