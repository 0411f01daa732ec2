/*
 * micropolis_amd -- C ABI of the MI355X (gfx950) REYES hot path.
 *
 * This is the drop-in boundary for the OpenCL backend of the micropolis
 * renderer.  Each entry point names the reference interface it replaces.
 * All arguments are plain pointers and sizes; nothing throws across the ABI;
 * every call returns MP_OK (0) or a negative MP_E* code and records a
 * message retrievable with mp_last_error().
 *
 * Threading: a context is single-threaded (like Reyes::RendererCL, which is
 * driven by one host thread through one command queue).  All device work of
 * a context is ordered on one HIP stream; only mp_finish() and mp_read_*()
 * block the caller.
 */
#ifndef MICROPOLIS_AMD_H
#define MICROPOLIS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MP_OK           0
#define MP_E_INVALID   -1   /* bad argument / unsupported option          */
#define MP_E_HIP       -2   /* HIP runtime error                          */
#define MP_E_NOMEM     -3   /* device or host allocation failed           */
#define MP_E_STATE     -4   /* call out of order (e.g. draw before prepare) */
#define MP_E_NOTFOUND  -5   /* unknown patch handle                       */

/* bound_n_split_method (src/Reyes/config.xml:34-39; BALANCED from
 * kernels/bound_n_split_balanced.cl, which the reference does not wire) */
enum mp_bns_method {
    MP_BNS_CPU      = 0,
    MP_BNS_BOUNDED  = 1,
    MP_BNS_LOCAL    = 2,
    MP_BNS_BREADTH  = 3,
    MP_BNS_BALANCED = 4
};

/* PatchType (src/Reyes/PatchType.h) */
enum mp_patch_type {
    MP_BEZIER  = 0,   /* 16 control points per patch */
    MP_GREGORY = 1    /* 20 control points per patch */
};

/* Hot-path options of reyes.options / src/Reyes/config.xml:62-124.  Names
 * follow the reference option names. */
typedef struct mp_config {
    int32_t window_width, window_height;   /* window_size                     */
    float   clear_color[4];                /* clear_color                     */
    int32_t reyes_patch_size;              /* N: 8 or 16                      */
    float   bound_n_split_limit;           /* screen length limit (pixels)    */
    float   cull_ribbon;
    int32_t backface_culling;
    int32_t bound_n_split_method;          /* enum mp_bns_method              */
    int32_t bound_sample_rate;             /* must be 3 on the GPU path       */
    int32_t max_split_depth;               /* <= 23                           */
    int32_t reyes_patches_per_pass;        /* pass size (BREADTH truncation,  */
                                           /* BOUNDED batch, LOCAL out cap)   */
    int32_t framebuffer_tile_size;         /* must be 8                       */
    int32_t subpixel_bits;                 /* PXLCOORD_SHIFT                  */
    int32_t dice_group_width;              /* accepted, unused on gfx950      */
    int32_t bns_pipeline_length;           /* CPU method transfer ring        */
    int32_t local_bns_work_groups;         /* LOCAL/BALANCED work-groups;     */
                                           /* 0 = fill the device             */
    int32_t displacement;                  /* 0 off, 1 3D value noise         */
    float   displacement_amplitude;
    float   displacement_frequency;
    uint32_t displacement_seed;
    int32_t displacement_octaves;
    int32_t dummy_render;                  /* split only, no raster           */
    int32_t write_rgba32;                  /* resolve also writes the linear  */
                                           /* RGBA32F framebuffer             */
} mp_config;

/* Reyes::Projection constructor arguments (src/base/Projection.h:44) */
typedef struct mp_projection {
    float   fovy_deg;
    float   near_plane;
    int32_t width, height;
} mp_projection;

/* Statistics of the last finished frame (src/base/Statistics.cpp:160-218). */
typedef struct mp_stats {
    int64_t patches_per_frame;        /* DRAW ranges emitted by bound&split    */
    int64_t diced_ranges;             /* ranges rasterised (BREADTH truncates) */
    int64_t dropped_ranges;           /* BREADTH ranges beyond the pass size   */
    int64_t bounded_patches;          /* bound() evaluations                   */
    int64_t micropolygons;            /* reyes_patch_size^2 * diced_ranges     */
    int64_t pass_count;               /* render passes                         */
    int64_t max_patches;              /* largest in-flight split level/stack   */
    int64_t slow_blocks;              /* raster blocks on the wide-int path    */
    int64_t overflow_relaunches;      /* persistent split relaunches           */
    int64_t coverage_samples;         /* (micropolygon, pixel) samples written */
    double  backward_ms;              /* device time of the last mp_backward    */
    double  frame_ms;                 /* mp_prepare..mp_finish device time     */
    int64_t sorted_ranges;            /* ranges of patches with more than 4096 */
                                      /* DRAW ranges, put in canonical order   */
                                      /* by a radix sort                       */
} mp_stats;

typedef struct mp_context mp_context;

/* Fill *cfg with the values of the reference's reyes.options. */
int mp_config_defaults(mp_config* cfg);

/* Reyes::RendererCL::RendererCL (src/Reyes/RendererCL.cpp:21-106).
 * Returns NULL on failure; mp_create_error() then holds the reason. */
mp_context* mp_create(const mp_config* cfg);
const char* mp_create_error(void);

void mp_destroy(mp_context* ctx);

const char* mp_last_error(const mp_context* ctx);

/* Renderer::are_patches_loaded (src/Reyes/Renderer.h:36) */
int mp_are_patches_loaded(mp_context* ctx, uint64_t handle);

/* Renderer::load_patches / PatchIndex::load_patches
 * (src/Reyes/Renderer.h:37, src/Reyes/PatchIndex.cpp:57-100):
 * xyz holds n_points * 3 floats (16 points per Bezier, 20 per Gregory). */
int mp_load_patches(mp_context* ctx, uint64_t handle, const float* xyz,
                    size_t n_points, int32_t patch_type);

/* PatchIndex::delete_patches (src/Reyes/PatchIndex.cpp:103-108) */
int mp_unload_patches(mp_context* ctx, uint64_t handle);

/* Renderer::prepare (src/Reyes/RendererCL.cpp:115-135): starts a frame. */
int mp_prepare(mp_context* ctx);

/* Renderer::draw_patches (src/Reyes/RendererCL.cpp:170-200).  mv is the
 * column-major modelview (Scene.cpp:190), color the diffuse colour.
 * Every draw of one frame renders into the window with one projection: the
 * viewport must equal the window size and a later draw's fovy / near plane
 * must equal the first draw's (MP_E_INVALID otherwise; the reference could
 * change the projection between draws, RendererCL.cpp:172-178, which no
 * caller of it does: Scene::draw passes one projection per frame). */
int mp_draw_patches(mp_context* ctx, uint64_t handle, const float mv[16],
                    const mp_projection* projection, const float color[4]);

/* Renderer::finish (src/Reyes/RendererCL.cpp:138-154): completes the frame
 * (split, dice, shade, sample, resolve) and waits for it. */
int mp_finish(mp_context* ctx);

/* Framebuffer readback (src/Reyes/Framebuffer.cpp:107-124, display
 * transform shaders/tex_draw.frag:25-37).  Row-major W*H, row 0 = bottom
 * (OpenGL window convention).  Any pointer may be NULL.
 *   rgba8  : display colour, gamma 1/2.2, alpha 255
 *   depth  : FP32 depth of the visible sample (1.0 where empty)
 *   ids    : visible micropolygon id (0xFFFFFFFF where empty);
 *            id = range_index * N*N + v*N + u, range_index in canonical
 *            (object, patch id, umin, vmin) order
 *   rgba32 : linear RGBA32F framebuffer value (the reference's buffer) */
int mp_read_framebuffer(mp_context* ctx, uint8_t* rgba8, float* depth,
                        uint32_t* ids, float* rgba32);

/* Display conversion of the rgba8 readback (gamma 1/2.2 of
 * shaders/tex_draw.frag:25-37, computed with the host's powf and rounded to
 * 8 bits; no reference counterpart as a call).  thresholds[k-1] = the
 * smallest linear value whose byte is at least k, k = 1..255; a value's byte
 * is the number of thresholds <= it (0 for NaN and values <= 0).  The
 * resolve kernel converts with this table, so rgba8 equals the host
 * conversion bit for bit.  Host only. */
int mp_display_thresholds(float thresholds[255]);

/* Debug tap: DRAW ranges of the last frame in canonical order, as
 * (pid, object, depth, umin, vmin, umax, vmax) records; *count receives the
 * number of ranges.  Pass records == NULL to query the count. */
typedef struct mp_range_record {
    int32_t pid;
    int32_t object;
    int32_t depth;
    float   umin, vmin, umax, vmax;
} mp_range_record;
int mp_read_ranges(mp_context* ctx, mp_range_record* records, int64_t capacity,
                   int64_t* count);

/* Debug tap: the micropolygon grids of canonical ranges [first_range,
 * first_range + n_ranges) of the last frame, as the dice stage computes them
 * (kernels/dice.cl:75-79): per range (N+1)^2 points, index nu + nv*(N+1),
 *   pos    : eye-space position x, y, z, w (4 floats per point)
 *   pxlpos : sub-pixel coordinate x, y (2 int32 per point)
 *   depth  : p.z / p.w (1 float per point)
 * Range indices are those of mp_read_ranges.  Host pointers; allocates a
 * temporary device buffer. */
int mp_read_grids(mp_context* ctx, int64_t first_range, int64_t n_ranges, float* pos,
                  int32_t* pxlpos, float* depth);

/* Debug tap: per-pass DRAW counts of the last frame (BREADTH: per split
 * level; BOUNDED and CPU: per batch of reyes_patches_per_pass; LOCAL/BALANCED:
 * the DRAW total in slices of reyes_patches_per_pass). */
int mp_read_pass_counts(mp_context* ctx, int32_t* counts, int32_t capacity,
                        int32_t* n_passes);

/* Reyes::BoundNSplitCLCPU (src/Reyes/BoundNSplitCLCPU.cpp:31-159): the CPU
 * method's host split of one patch set run to completion, with no context
 * and no device (the split a CPU-method frame performs per object).
 * records receive every DRAW range in canonical (pid, umin, vmin) order
 * (object 0); pass NULL to query *count.  pass_counts receive the DRAW
 * count of each batch of cfg->reyes_patches_per_pass that the reference's
 * single stack would emit; *n_bounds the ranges bounded.  Any output
 * pointer may be NULL.  On failure mp_create_error() holds the reason. */
int mp_cpu_bound_n_split(const mp_config* cfg, const float* xyz, size_t n_points,
                         int32_t patch_type, const float mv[16],
                         const mp_projection* projection, mp_range_record* records,
                         int64_t capacity, int64_t* count, int32_t* pass_counts,
                         int32_t max_passes, int32_t* n_passes, int64_t* n_bounds);

/* Renderer::dump_trace (src/Reyes/Renderer.h:44): write the stages of the
 * last frame (run once more with events between its stages for LOCAL and
 * BALANCED frames; for CPU-method frames each batch's host split and device
 * raster as recorded; the whole frame otherwise) in the reference's trace
 * format, one line per stage,
 *   name@queue:queued:submit:start:end:id:dependencies
 * with host-clock nanoseconds (src/CL/Device.cpp:275-289). */
int mp_dump_trace(mp_context* ctx, const char* path);

/* Statistics::dump_stats (src/base/Statistics.cpp:196-218): `key = value;`
 * lines for the last frame; opencl_mem is the device memory this context
 * holds (by use), opengl_mem 0, bound_n_split_balance empty. */
int mp_dump_stats(mp_context* ctx, const char* path);

/* Statistics (src/base/Statistics.cpp:160-218) of the last frame. */
int mp_get_stats(mp_context* ctx, mp_stats* stats);

/* Displacement amplitude used by subsequent draws (overrides
 * mp_config.displacement_amplitude).  No reference counterpart. */
int mp_set_displacement_amplitude(mp_context* ctx, float amplitude);

/* ---- batching: many independent views/scenes rendered by one set of
 * persistent launches (no reference counterpart; the plain path is the
 * 1-view case of the same machinery). ---- */

/* Start a batch of n_views views, each with its own framebuffer of
 * projections[v].width x projections[v].height. */
int mp_batch_begin(mp_context* ctx, int32_t n_views, const mp_projection* projections);
/* Add an object to view `view` (same arguments as mp_draw_patches). */
int mp_batch_draw(mp_context* ctx, int32_t view, uint64_t handle, const float mv[16],
                  const float color[4]);
/* Render every view of the batch and wait. */
int mp_batch_finish(mp_context* ctx);
/* Read back view `view` of the last batch (same layout as mp_read_framebuffer). */
int mp_batch_read(mp_context* ctx, int32_t view, uint8_t* rgba8, float* depth,
                  uint32_t* ids, float* rgba32);

/* ---- device-resident benchmarking: replay the last frame or batch n times
 * on the device (inputs resident), returning the mean device ms per replay
 * measured with HIP events around the replays. ---- */
int mp_replay(mp_context* ctx, int32_t n, int32_t use_graph, double* ms_per_frame);

/* Scheduling of the device-driven frame (LOCAL / BALANCED).  None of these
 * choices changes any output; they trade launch count, load balance and
 * occupancy.  No reference counterpart (the reference fixes its schedule by
 * bound_n_split_method and work-group sizes). */
typedef struct mp_tuning {
    int32_t split_engine;    /* 0: persistent LDS-stack split, one launch;
                                1: one launch per split level + persistent tail;
                                2: forest split (default): a work-group splits
                                the whole trees of its roots with all its lanes
                                and, for frames of at most 4096 patches, ranks
                                their DRAW ranges itself (one order launch) */
    int32_t work_sharing;    /* persistent split: deep stacks spill to idle groups */
    int32_t raster_per_cu;   /* raster work-groups per CU (0 = occupancy limit) */
    int32_t raster_chunk;    /* raster range schedule: 0 static stride (no
                                atomics), -1 dynamic chunks sized from the
                                range count, > 0 dynamic chunks of that size,
                                <= -2 static chunks of -raster_chunk
                                consecutive ranges (no atomics) */
    int32_t split_levels;    /* split_engine 1: level launches before the
                                persistent tail (0 = as many as the frame needs) */
    int32_t split_budget;    /* LOCAL, split_engine 0, no work sharing: steps
                                a persistent group runs before its queued and
                                stacked ranges are dealt to all groups for a
                                second launch (0 = one launch) */
    int32_t split_waves;     /* persistent split kernel build: 2 = all registers,
                                fewest groups (latency-bound frames); 3 = capped
                                registers, one more group per CU (frames with
                                many ranges); 0 = 2 */
    double  frame_ms;        /* mp_autotune: replay time of the chosen setting */
} mp_tuning;
int mp_get_tuning(mp_context* ctx, mp_tuning* tuning);
/* Apply a setting; with a finished frame it is re-run once under it. */
int mp_set_tuning(mp_context* ctx, const mp_tuning* tuning);
/* Autotuner: replay the last finished LOCAL/BALANCED frame or batch under
 * candidate settings, `replays` graph replays each (best of two rounds):
 * first the raster schedule (work-groups per CU x static / dynamic ranges),
 * then the split (the forest split, persistent with / without work sharing
 * or a step budget, or a number of level launches before the persistent
 * tail); keep the fastest. */
int mp_autotune(mp_context* ctx, int32_t replays, mp_tuning* chosen);

/* Replay the last LOCAL/BALANCED frame n times with HIP events between its
 * stages on the library's stream; stage_ms[0..3] receive the mean device ms
 * of split, canonical ordering, raster (dice+shade+sample) and resolve. */
int mp_replay_stages(mp_context* ctx, int32_t n, double stage_ms[4]);

/* Measurement (no reference counterpart): the device's rate of 64-bit
 * atomic minimum updates to distinct addresses, each wave covering 64
 * consecutive keys (the raster's key update, relaxed, device scope, no
 * return), on `n` keys (best of three passes).  Feeds the frame's
 * HBM + atomic roofline; allocates and frees its own buffer. */
int mp_measure_atomic_rate(mp_context* ctx, uint32_t n, double* atomics_per_second);

/* ---- differentiable extension (no reference counterpart) ----
 * Reverse pass of dice + displacement + flat shade + z-resolve for the last
 * finished frame or batch, with visibility held fixed to its visible-id
 * buffers (silhouette terms ignored) and every grid recomputed per range.
 * All pointers are device pointers (HIP/torch allocations on this device).
 *   dl_dcolor      : dL/dC for every view, concatenated in view order, each
 *                    height*width*4 floats in the rgba32 layout (alpha unused)
 *   grad_cps       : per object in draw order, n_patches*16*3 floats
 *   grad_color     : per object 4 floats (albedo gradient, alpha 0)
 *   grad_amplitude : per object 1 float
 * Outputs are overwritten.  Bezier objects only.  Blocks until done. */
int mp_backward(mp_context* ctx, uint64_t frame_serial, const float* dl_dcolor, size_t dl_floats,
                float* grad_cps, size_t grad_cps_floats, float* grad_color, float* grad_amplitude,
                int32_t n_objects);
/*   frame_serial    : mp_frame_serial() of the frame whose visibility and
 *                     control points the gradients are taken for; MP_E_STATE
 *                     if the context has rendered another frame (or its
 *                     patches changed) since
 *   dl_floats       : floats in dl_dcolor, must equal 4 * the pixels of all
 *                     views; grad_cps_floats: floats in grad_cps, must equal
 *                     the objects' n_patches*16*3 summed; n_objects: entries
 *                     of grad_color (x4) and grad_amplitude, must equal the
 *                     draws of the frame (MP_E_INVALID otherwise) */

/* Serial number of the last finished frame or batch (0: none).  It changes
 * with every mp_finish / mp_batch_finish and with every change of a loaded
 * patch set (mp_load_patches, mp_unload_patches, mp_update_patches_device);
 * replays of a frame keep it. */
int mp_frame_serial(mp_context* ctx, uint64_t* serial);

/* Device-to-device readback of a view's linear RGBA32F colour and ids
 * (either may be NULL; needs write_rgba32 for the colour).  n_pixels is the
 * caller's buffer size in pixels (rgba32 holds 4 floats per pixel) and must
 * equal the view's width * height (MP_E_INVALID otherwise). */
int mp_batch_read_device(mp_context* ctx, int32_t view, float* rgba32, uint32_t* ids, size_t n_pixels);

/* Replace a loaded patch set's control points from device memory (xyz
 * floats, same count as loaded). */
int mp_update_patches_device(mp_context* ctx, uint64_t handle, const float* xyz, size_t n_points);

/* ---- allocation-free frame (no reference counterpart) ----
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The
 * allocation-free variant of the device-driven frame (mp_create +
 * mp_load_patches + mp_prepare + mp_draw_patches + mp_finish +
 * mp_read_framebuffer for one object, LOCAL or BALANCED): no context, no
 * hipMalloc / hipFree.  Split queues, range lists, per-patch counts, the
 * Lambert store and the visibility keys live in one 64 MiB device array of
 * the library (per device, held until the process ends); a frame that does
 * not fit it (window keys + patches + DRAW ranges) is refused with
 * MP_E_INVALID.  Every call on one device shares that array: calls that
 * overlap on one device (other host threads or streams) race.
 *   xyz                : device pointer, n_points * 3 floats
 *   rgba8, depth, ids  : device pointers, width * height each (layout of
 *                        mp_read_framebuffer)
 *   stream             : hipStream_t to run on (NULL: the null stream);
 *                        the call waits for it before returning
 * The image equals mp_finish's for the same configuration and draw.  On
 * failure mp_create_error() holds the reason. */
int mp_render_noalloc(const mp_config* cfg, const float* xyz, size_t n_points, int32_t patch_type,
                      const float mv[16], const mp_projection* projection, const float color[4],
                      uint8_t* rgba8, float* depth, uint32_t* ids, void* stream);

#ifdef __cplusplus
}
#endif
#endif
