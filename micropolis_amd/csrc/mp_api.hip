// Host side of the C ABI: patch index, frame orchestration, batching, replay.
//
// Mirrors Reyes::RendererCL (src/Reyes/RendererCL.cpp), the BoundNSplitCL*
// schedulers (src/Reyes/BoundNSplitCL{Breadth,Bounded,Local}.cpp),
// Reyes::PatchIndex (src/Reyes/PatchIndex.cpp), Reyes::Projection
// (src/base/Projection.cpp) and Reyes::Framebuffer (src/Reyes/Framebuffer.cpp).
//
// A frame is recorded by mp_prepare / mp_draw_patches and executed by
// mp_finish: all objects of all views are split together (one persistent
// launch for LOCAL/BALANCED, a host-driven level loop for BREADTH/BOUNDED),
// the DRAW ranges are put in canonical order, rasterised by one persistent
// fused dice/shade/sample launch and resolved per view.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/micropolis.h"
#include "mp_cpu_split.h"
#include "mp_host.h"
#include "mp_kernels.h"
#include "mp_types.h"

using namespace mp;

namespace {

std::string g_create_error;

struct PatchSet {
    float4* d_cps = nullptr;
    int type = 0;
    int n_patches = 0;
    std::vector<float> host;   // x,y,z,1 per point; kept for the CPU split method
};

struct ObjectReq {
    uint64_t handle;
    int view;
    float mv[16];
    float color[4];
    float amplitude;
};

struct ViewBuffers {
    int width = 0, height = 0;
    int grid_w = 0, grid_h = 0;
    unsigned long long* keys = nullptr;
    uint8_t* rgba8 = nullptr;
    float* depth = nullptr;
    uint32_t* ids = nullptr;
    float* rgba32 = nullptr;
};

// Display byte of a linear colour value: gamma 1/2.2 (shaders/tex_draw.frag:25-37)
// with the host's powf, rounded to 8 bits, 0 for NaN and values <= 0.
static uint32_t display_byte_host(float c)
{
    if (!(c > 0.0f)) return 0;
    const float g = powf(c, 0.454545f);
    if (!(g < 1.0f)) return 255;
    return (uint32_t)(g * 255.0f + 0.5f);
}

}  // namespace

struct mp_context {
    mp_config cfg;
    DevParams prm;
    std::string err;
    int device = 0;
    int num_cus = 1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float amplitude = 0.0f;

    std::unordered_map<uint64_t, PatchSet> patches;

    // recorded frame
    int state = 0;  // 0 idle, 1 frame, 2 batch
    std::vector<mp_projection> view_projs;
    std::vector<ObjectReq> objects;

    // per-view device buffers
    std::vector<ViewBuffers> views;

    // device tables (+ pinned host staging)
    DevView* d_views = nullptr; DevView* h_views = nullptr; int views_cap = 0;
    DevObject* d_objs = nullptr; DevObject* h_objs = nullptr; int objs_cap = 0;
    FrameCounters* d_fc = nullptr; FrameCounters* h_fc = nullptr;
    // the last frame's counters, copied by its resolve (k_resolve or
    // k_frame_end) before it resets d_fc for the next frame
    FrameCounters* d_fc_out = nullptr;
    // d_fc, the split's shared state and the root queues are ready for a
    // device frame to start with its split: armed by every frame end, and
    // the root queues written for init_sig's layout
    bool armed = false;
    struct QueueSig {
        int groups; uint32_t roots; int64_t stride; int balanced; const void* queue; const void* qcnt; const void* shq;
        const void* lcnt;   // the level counts (their last entry counts the ranges a step budget leaves)
        bool operator==(const QueueSig& o) const
        {
            return groups == o.groups && roots == o.roots && stride == o.stride && balanced == o.balanced &&
                   queue == o.queue && qcnt == o.qcnt && shq == o.shq && lcnt == o.lcnt;
        }
    } init_sig = { -1, 0, 0, 0, nullptr, nullptr, nullptr, nullptr };
    bool graph_inits = false;   // the captured frame writes the root queues itself
    bool frame_inits = false;   // the last device frame launched wrote them
    uint32_t* d_totals = nullptr; uint32_t* h_totals = nullptr;

    // persistent split queues
    RangeRec* d_queue = nullptr; size_t queue_entries = 0;
    int32_t* d_queue_cnt = nullptr; int queue_groups_cap = 0;
    RangeRec* d_shared_q = nullptr; size_t shared_entries = 0;
    int32_t* d_shared_cnt = nullptr;
    RangeRec* d_chunks = nullptr; uint32_t* d_chunk_count = nullptr; uint32_t* d_chunk_ready = nullptr;
    uint32_t chunk_cap = 0;
    ShareState* d_share = nullptr;
    unsigned long long* d_keys = nullptr;
    uint2* d_span = nullptr;

    // segment engine (BREADTH / BOUNDED)
    RangeRec* d_seg_a = nullptr; size_t seg_a_cap = 0;
    RangeRec* d_seg_b = nullptr; size_t seg_b_cap = 0;
    uint8_t* d_flags = nullptr; size_t flags_cap = 0;
    uint32_t* d_blk = nullptr; size_t blk_cap = 0;

    // canonical lists
    RangeRec* d_draw = nullptr; RangeRec* d_bucket = nullptr; RangeRec* d_canon = nullptr;
    unsigned long long* d_sort_keys = nullptr; uint32_t* d_sort_vals = nullptr;   // 2 x list_cap each
    // ranked forest frames: per-group output regions (kForestKeys entries per
    // group), in-patch ranks, group totals
    RangeRec* d_forest_out = nullptr; uint32_t* d_rank = nullptr; uint32_t* d_forest_total = nullptr;
    int forest_out_groups = 0;
    uint32_t list_cap = 0;
    uint32_t* d_counts = nullptr; uint32_t* d_offsets = nullptr; uint32_t* d_cursor = nullptr;
    uint32_t* d_part = nullptr;
    uint32_t gpid_cap = 0;
    float* d_quad = nullptr; size_t quad_cap = 0;
    float* d_display = nullptr;  // display byte thresholds (mp_display_thresholds)

    // last frame
    uint32_t n_gpids = 0;
    uint32_t total_roots = 0;
    int n_views = 0;
    int max_pixels = 0;
    int raster_grid_size = 0;
    int split_groups = 0;
    int split_waves = 2;         // persistent split kernel build: 2 or 3 waves per SIMD
    int stack_cap = 0;
    int split_spread = 2;        // persistent split: 0 three samples per lane; 1 one sample
                                 // per lane on narrow steps; 2 also two levels on the narrowest
    int split_budget = 8;        // LOCAL persistent split: steps before rebalancing (0: none)
    int fused_order = 1;         // canonical order: count scan inside the scatter launch
    mp_stats stats;
    std::vector<int32_t> pass_counts;
    bool frame_valid = false;
    uint64_t frame_serial = 0;   // the last finished frame (0: none); see mp_frame_serial
    uint64_t serial_counter = 0;
    // Per-patch DRAW counts and bucket cursors are zero between frames (the
    // rank kernel clears them).  A launch sequence that stops with an error
    // after the split has counted but before the rank ran leaves them stale;
    // it sets this flag and the next frame clears them first.
    bool counts_dirty = false;
    bool device_driven = false;  // the last frame can be replayed without host syncs
    hipGraphExec_t graph = nullptr;
    bool work_sharing = false;   // spill queue between split groups
    uint32_t grad_floats = 0;
    uint32_t* d_visible = nullptr; size_t visible_cap = 0;
    uint32_t* d_next = nullptr;
    int adjoint_grid_size = 0;
    // level-synchronous split engine
    int split_engine = 0;          // 0 persistent LDS-stack engine, 1 level-synchronous,
                                   // 2 forest (a group splits whole trees with all its lanes)
    int forest_cap = 1536;         // forest split: LDS stack entries per group
    int forest_per_cu = 0;         // forest split: groups per CU (0: occupancy)
    int patch_types = 0;           // bit 0: the frame has Bezier objects, bit 1: Gregory ones
    int forest_rank = 1;           // forest split: rank the DRAW ranges in the groups (0 after a
                                   // frame whose groups could not hold their ranges)
    bool frame_ranked = false;     // the frame being launched takes the ranked order
    RangeRec* d_level_a = nullptr; RangeRec* d_level_b = nullptr; uint32_t level_cap = 0;
    uint32_t* d_level_cnt = nullptr; uint32_t* h_level_cnt = nullptr;
    int level_launches = 0;        // levels launched before the persistent tail
    int level_grid_size = 0;
    int level_lanes = 1;           // lanes per range in the level kernels (1 or 9; 1 measured faster)
    int level_per_cu = 1;          // level-kernel work-groups per CU (0: occupancy)
    int level_limit = 0;           // level launches before the persistent tail (0: all needed)
    int spill_at = 192;
    int linger = -1;             // groups that wait for spilled work (-1: one per CU)
    // CPU split method
    CpuSplitFrame cpu_frame;
    int cpu_threads = 0;         // 0: the host's cores (at most 16)
    // its transfer ring (BoundNSplitCLCPU's batch records): bns_pipeline_length
    // pinned slots of reyes_patches_per_pass ranges; a slot is refilled once
    // the copy of its previous batch has completed
    RangeRec* h_ring = nullptr;
    int ring_len = 0;
    size_t ring_slot = 0;
    std::vector<hipEvent_t> ring_ev;
    // trace of the last CPU-method frame: per batch the host split interval
    // (ms from the frame's start) and events around its device raster
    struct BatchTrace { double split0, split1; int n; };
    std::vector<BatchTrace> cpu_trace;
    std::vector<hipEvent_t> trace_ev;   // 2 per batch
    hipEvent_t trace_ev0 = nullptr;
};

#define HIPCHK(ctx, expr)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                \
            return MP_E_HIP;                                                               \
        }                                                                                  \
    } while (0)

static int fail(mp_context* ctx, int code, const std::string& msg)
{
    ctx->err = msg;
    return code;
}

template <typename T>
static int dev_alloc(mp_context* ctx, T** p, size_t count)
{
    if (*p) { hipFree(*p); *p = nullptr; }
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        ctx->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return MP_E_NOMEM;
    }
    return MP_OK;
}

// grow a device array keeping its first `keep` elements
template <typename T>
static int dev_grow(mp_context* ctx, T** p, size_t* cap, size_t need, size_t keep)
{
    if (need <= *cap && *p) return MP_OK;
    size_t ncap = std::max(need, *cap * 2);
    T* np = nullptr;
    hipError_t e = hipMalloc((void**)&np, ncap * sizeof(T));
    if (e != hipSuccess) return fail(ctx, MP_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    if (*p && keep) {
        HIPCHK(ctx, hipMemcpyAsync(np, *p, keep * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (*p) hipFree(*p);
    *p = np;
    *cap = ncap;
    return MP_OK;
}

static void invalidate_graph(mp_context* ctx)
{
    if (ctx->graph) {
        hipGraphExecDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
}

// ---------------------------------------------------------------------------
// projection (src/base/Projection.cpp:23-50; glm::perspective, far 10000)

void mp::make_view(const mp_projection& pr, const float clear[4], int subpixel_bits, DevView& v)
{
    memset(&v, 0, sizeof(v));
    const int W = pr.width, H = pr.height;
    const float nearp = pr.near_plane;
    float aspect = (float)W / (float)H;
    float fovy = (float)((double)pr.fovy_deg * 3.14159265358979323846 / 180.0);
    float zfar = 10000.0f;
    float th = tanf(fovy / 2.0f);
    v.proj[0] = 1.0f / (aspect * th);
    v.proj[5] = 1.0f / th;
    v.proj[10] = -(zfar + nearp) / (zfar - nearp);
    v.proj[11] = -1.0f;
    v.proj[14] = -(2.0f * zfar * nearp) / (zfar - nearp);
    float fy = 1.0f / (float)tan((double)pr.fovy_deg * 3.14159265358979323846 / 360.0);
    float fx = fy / aspect;
    float vpx = (float)W / 2.0f, vpy = (float)H / 2.0f;
    fx *= vpx;
    fy *= vpy;
    v.f[0] = fx;
    v.f[1] = fy;
    v.screen[0] = vpx; v.screen[1] = 0.0f; v.screen[2] = 0.0f; v.screen[3] = vpy;
    v.near_ = nearp;
    v.far_ = 1000.0f;
    v.eps = (float)((double)nearp * 0.1);
    v.width = W;
    v.height = H;
    v.grid_w = (W + 7) / 8;
    v.grid_h = (H + 7) / 8;
    v.vs_x = W << subpixel_bits;
    v.vs_y = H << subpixel_bits;
    for (int k = 0; k < 3; ++k) v.clear_linear[k] = powf(clear[k], 2.2f);
    v.clear_linear[3] = 1000.0f;
    v.clear_rgba8 = display_byte_host(v.clear_linear[0]) | (display_byte_host(v.clear_linear[1]) << 8) |
                    (display_byte_host(v.clear_linear[2]) << 16) | (255u << 24);
    v.solo_obj = -1;
}

// ---------------------------------------------------------------------------

int mp::validate_config(const mp_config* c, std::string& why)
{
    if (c->reyes_patch_size != 8 && c->reyes_patch_size != 16 && c->reyes_patch_size != 32) {
        why = "reyes_patch_size must be 8, 16 or 32";
        return 0;
    }
    if (c->bound_sample_rate != 3) { why = "bound_sample_rate must be 3"; return 0; }
    if (c->max_split_depth < 1 || c->max_split_depth > 23) { why = "max_split_depth must be in [1,23]"; return 0; }
    if (c->framebuffer_tile_size != 8) { why = "framebuffer_tile_size must be 8"; return 0; }
    if (c->subpixel_bits < 0 || c->subpixel_bits > 8) { why = "subpixel_bits must be in [0,8]"; return 0; }
    if (c->reyes_patches_per_pass < 1) { why = "reyes_patches_per_pass must be positive"; return 0; }
    if (c->bound_n_split_method < MP_BNS_CPU || c->bound_n_split_method > MP_BNS_BALANCED) {
        why = "unknown bound_n_split_method";
        return 0;
    }
    if (c->window_width < 1 || c->window_height < 1) { why = "window_size must be positive"; return 0; }
    if (c->displacement && c->displacement_octaves < 1) { why = "displacement_octaves must be >= 1"; return 0; }
    return 1;
}

DevParams mp::device_params(const mp_config* cfg)
{
    DevParams p;
    memset(&p, 0, sizeof(p));
    p.patch_size = cfg->reyes_patch_size;
    p.max_split_depth = cfg->max_split_depth;
    p.cull_ribbon = cfg->cull_ribbon;
    p.backface_culling = cfg->backface_culling;
    p.subpixel_bits = cfg->subpixel_bits;
    p.patches_per_pass = cfg->reyes_patches_per_pass;
    p.displacement = cfg->displacement;
    p.disp_frequency = cfg->displacement_frequency;
    p.disp_seed = cfg->displacement_seed;
    p.disp_octaves = cfg->displacement_octaves;
    p.noise_wave_cells = 2;   // lane window (0 per lane, 1 wave-uniform cell)
    p.raster_chunk = 0;       // static range stride (see mp_tuning)
    p.raster_one_wave = 1;    // N = 16: one wave per work-group (no waits between waves)
    return p;
}

extern "C" int mp_config_defaults(mp_config* c)
{
    if (!c) return MP_E_INVALID;
    memset(c, 0, sizeof(*c));
    // reyes.options
    c->window_width = 1600;
    c->window_height = 1200;
    c->clear_color[0] = 0.1f; c->clear_color[1] = 0.1f; c->clear_color[2] = 0.7f; c->clear_color[3] = 0.0f;
    c->reyes_patch_size = 8;
    c->bound_n_split_limit = 8.0f;
    c->cull_ribbon = 32.0f;
    c->backface_culling = 1;
    c->bound_n_split_method = MP_BNS_BOUNDED;
    c->bound_sample_rate = 3;
    c->max_split_depth = 23;
    c->reyes_patches_per_pass = 100000;
    c->framebuffer_tile_size = 8;
    c->subpixel_bits = 6;
    c->dice_group_width = 8;
    c->bns_pipeline_length = 8;
    c->local_bns_work_groups = 128;
    c->displacement = 0;
    c->displacement_amplitude = 0.02f;
    c->displacement_frequency = 4.0f;
    c->displacement_seed = 1;
    c->displacement_octaves = 4;
    c->dummy_render = 0;
    c->write_rgba32 = 0;
    return MP_OK;
}

extern "C" const char* mp_create_error(void) { return g_create_error.c_str(); }

void mp::set_create_error(const std::string& msg) { g_create_error = msg; }

// tab[k - 1] = smallest positive float whose display byte is at least k
// (binary search over the bit patterns of positive floats, on which the
// byte is non-decreasing); tab[255] = NaN, never passed by a comparison.
void mp::display_thresholds(float tab[256])
{
    for (uint32_t k = 1; k <= 255; ++k) {
        uint32_t lo = 1u, hi = 0x7f800000u;   // smallest denormal .. +inf (byte 255)
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            float f;
            memcpy(&f, &mid, 4);
            if (display_byte_host(f) >= k) hi = mid;
            else lo = mid + 1;
        }
        memcpy(&tab[k - 1], &lo, 4);
    }
    const uint32_t qnan = 0x7fc00000u;
    memcpy(&tab[255], &qnan, 4);
}

extern "C" int mp_display_thresholds(float thresholds[255])
{
    if (!thresholds) return MP_E_INVALID;
    float tab[256];
    display_thresholds(tab);
    memcpy(thresholds, tab, 255 * sizeof(float));
    return MP_OK;
}

extern "C" mp_context* mp_create(const mp_config* cfg)
{
    if (!cfg) { g_create_error = "null config"; return nullptr; }
    std::string why;
    if (!validate_config(cfg, why)) { g_create_error = why; return nullptr; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_error = "no HIP device";
        return nullptr;
    }
    mp_context* ctx = new mp_context();
    ctx->cfg = *cfg;
    ctx->amplitude = cfg->displacement_amplitude;
    ctx->prm = device_params(cfg);

    hipDeviceProp_t prop;
    if (hipGetDevice(&ctx->device) != hipSuccess ||
        hipGetDeviceProperties(&prop, ctx->device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipMalloc((void**)&ctx->d_fc, kFrameCountersBytes) != hipSuccess ||   // counters + their stripes
        hipMalloc((void**)&ctx->d_fc_out, sizeof(FrameCounters)) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_fc, sizeof(FrameCounters), 0) != hipSuccess ||
        hipMalloc((void**)&ctx->d_totals, 4 * sizeof(uint32_t)) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_totals, 4 * sizeof(uint32_t), 0) != hipSuccess ||
        hipMalloc((void**)&ctx->d_display, 256 * sizeof(float)) != hipSuccess) {
        g_create_error = "HIP initialisation failed";
        mp_destroy(ctx);
        return nullptr;
    }
    {
        float tab[256];
        display_thresholds(tab);
        if (hipMemcpy(ctx->d_display, tab, sizeof(tab), hipMemcpyHostToDevice) != hipSuccess) {
            g_create_error = "HIP initialisation failed";
            mp_destroy(ctx);
            return nullptr;
        }
    }
    ctx->num_cus = prop.multiProcessorCount;
    if (const char* e = getenv("MP_RASTER_ONE_WAVE")) ctx->prm.raster_one_wave = atoi(e) != 0;
    if (const char* e = getenv("MP_RASTER_CHUNK")) ctx->prm.raster_chunk = atoi(e);
    if (const char* e = getenv("MP_NOISE_WAVE_CELLS")) ctx->prm.noise_wave_cells = atoi(e);
    ctx->raster_grid_size = raster_grid(cfg->reyes_patch_size, ctx->num_cus, ctx->prm);
    // LDS ring stack of the persistent split, depth-bounded: 64 * (max_split_depth + 2)
    // entries.  Fewer resident groups that own their SIMD finish the deepest
    // split trees sooner than more groups sharing it (measured on the torus frame).
    ctx->stack_cap = 64 * (cfg->max_split_depth + 2);
    // tuning knobs (defaults chosen by measurement; mp_autotune picks per
    // frame): LOCAL and BALANCED frames split fastest in the forest engine
    // (frames whose tables do not fit it take the persistent kernel)
    ctx->split_engine = 2;
    if (const char* e = getenv("MP_WORK_SHARING")) ctx->work_sharing = atoi(e) != 0;
    if (const char* e = getenv("MP_SPILL_AT")) ctx->spill_at = atoi(e);
    if (const char* e = getenv("MP_LINGER")) ctx->linger = atoi(e);
    if (const char* e = getenv("MP_STACK_CAP")) ctx->stack_cap = std::max(256, atoi(e));
    if (const char* e = getenv("MP_SPLIT_SPREAD")) ctx->split_spread = atoi(e);
    if (const char* e = getenv("MP_SPLIT_BUDGET")) ctx->split_budget = std::max(0, atoi(e));
    if (const char* e = getenv("MP_FUSED_ORDER")) ctx->fused_order = atoi(e) != 0;
    if (const char* e = getenv("MP_SPLIT_WAVES")) ctx->split_waves = atoi(e) == 3 ? 3 : 2;
    if (const char* e = getenv("MP_SPLIT_ENGINE")) ctx->split_engine = atoi(e);
    if (const char* e = getenv("MP_FOREST_CAP")) ctx->forest_cap = std::max(kForestMinStack, atoi(e));
    if (const char* e = getenv("MP_FOREST_PER_CU")) ctx->forest_per_cu = std::max(0, atoi(e));
    if (const char* e = getenv("MP_CPU_THREADS")) ctx->cpu_threads = atoi(e);
    if (const char* e = getenv("MP_LEVEL_LANES")) ctx->level_lanes = atoi(e) == 1 ? 1 : 9;
    if (const char* e = getenv("MP_LEVEL_PER_CU")) ctx->level_per_cu = atoi(e);
    if (const char* e = getenv("MP_SPLIT_LEVELS")) ctx->level_limit = std::max(0, atoi(e));
    ctx->level_launches = cfg->max_split_depth;
    if (const char* e = getenv("MP_RASTER_PER_CU")) ctx->raster_grid_size = ctx->num_cus * std::max(1, atoi(e));
    memset(&ctx->stats, 0, sizeof(ctx->stats));
    return ctx;
}

static void free_view(ViewBuffers& vb)
{
    hipFree(vb.keys); hipFree(vb.rgba8); hipFree(vb.depth); hipFree(vb.ids); hipFree(vb.rgba32);
    vb = ViewBuffers();
}

extern "C" void mp_destroy(mp_context* ctx)
{
    if (!ctx) return;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    invalidate_graph(ctx);
    for (auto& kv : ctx->patches) hipFree(kv.second.d_cps);
    for (auto& vb : ctx->views) free_view(vb);
    hipFree(ctx->d_views); hipHostFree(ctx->h_views);
    hipFree(ctx->d_objs); hipHostFree(ctx->h_objs);
    hipFree(ctx->d_fc); hipFree(ctx->d_fc_out); hipHostFree(ctx->h_fc);
    hipFree(ctx->d_totals); hipHostFree(ctx->h_totals);
    hipFree(ctx->d_display);
    hipFree(ctx->d_queue); hipFree(ctx->d_queue_cnt); hipFree(ctx->d_shared_q); hipFree(ctx->d_shared_cnt);
    hipFree(ctx->d_chunks); hipFree(ctx->d_chunk_count); hipFree(ctx->d_chunk_ready); hipFree(ctx->d_share);
    hipFree(ctx->d_keys); hipFree(ctx->d_span);
    hipFree(ctx->d_visible); hipFree(ctx->d_next);
    hipFree(ctx->d_level_a); hipFree(ctx->d_level_b); hipFree(ctx->d_level_cnt); hipHostFree(ctx->h_level_cnt);
    hipFree(ctx->d_seg_a); hipFree(ctx->d_seg_b); hipFree(ctx->d_flags); hipFree(ctx->d_blk);
    hipFree(ctx->d_draw); hipFree(ctx->d_bucket); hipFree(ctx->d_canon);
    hipFree(ctx->d_sort_keys); hipFree(ctx->d_sort_vals);
    hipFree(ctx->d_forest_out); hipFree(ctx->d_rank); hipFree(ctx->d_forest_total);
    hipFree(ctx->d_counts); hipFree(ctx->d_offsets); hipFree(ctx->d_cursor); hipFree(ctx->d_part);
    hipFree(ctx->d_quad);
    if (ctx->h_ring) hipHostFree(ctx->h_ring);
    for (hipEvent_t e : ctx->ring_ev) hipEventDestroy(e);
    for (hipEvent_t e : ctx->trace_ev) hipEventDestroy(e);
    if (ctx->trace_ev0) hipEventDestroy(ctx->trace_ev0);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char* mp_last_error(const mp_context* ctx)
{
    return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int mp_are_patches_loaded(mp_context* ctx, uint64_t handle)
{
    if (!ctx) return 0;
    return ctx->patches.count(handle) ? 1 : 0;
}

extern "C" int mp_load_patches(mp_context* ctx, uint64_t handle, const float* xyz, size_t n_points,
                               int32_t patch_type)
{
    if (!ctx) return MP_E_INVALID;
    if (patch_type != MP_BEZIER && patch_type != MP_GREGORY)
        return fail(ctx, MP_E_INVALID, "unknown patch type");
    const size_t per = patch_type == MP_GREGORY ? 20 : 16;
    if (n_points % per != 0) return fail(ctx, MP_E_INVALID, "point count is not a multiple of the patch size");
    if (n_points && !xyz) return fail(ctx, MP_E_INVALID, "null patch data");
    if (n_points / per >= (1u << 20))
        return fail(ctx, MP_E_INVALID, "at most 2^20 - 1 patches per handle");
    if (patch_type == MP_GREGORY && ctx->cfg.displacement)
        return fail(ctx, MP_E_INVALID, "displacement is defined for Bezier patches only");
    // PatchIndex::load_patches: vec3 -> vec4(p, 1)
    std::vector<float4> host(n_points);
    for (size_t i = 0; i < n_points; ++i)
        host[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 1.0f);
    PatchSet ps;
    ps.type = patch_type;
    ps.n_patches = (int)(n_points / per);
    if (ctx->cfg.bound_n_split_method == MP_BNS_CPU) {
        // BoundNSplitCLCPU keeps the host vector (PatchIndex::enable_retain_vector)
        ps.host.resize(n_points * 4);
        memcpy(ps.host.data(), host.data(), n_points * sizeof(float4));
    }
    if (n_points) {
        HIPCHK(ctx, hipMalloc((void**)&ps.d_cps, n_points * sizeof(float4)));
        HIPCHK(ctx, hipMemcpy(ps.d_cps, host.data(), n_points * sizeof(float4), hipMemcpyHostToDevice));
    }
    auto it = ctx->patches.find(handle);
    if (it != ctx->patches.end()) hipFree(it->second.d_cps);
    ctx->patches[handle] = ps;
    invalidate_graph(ctx);
    ctx->frame_valid = false;
    ctx->frame_serial = ++ctx->serial_counter;
    return MP_OK;
}

extern "C" int mp_unload_patches(mp_context* ctx, uint64_t handle)
{
    if (!ctx) return MP_E_INVALID;
    auto it = ctx->patches.find(handle);
    if (it == ctx->patches.end()) return fail(ctx, MP_E_NOTFOUND, "unknown patch handle");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(it->second.d_cps);
    ctx->patches.erase(it);
    invalidate_graph(ctx);
    ctx->frame_valid = false;
    ctx->frame_serial = ++ctx->serial_counter;
    return MP_OK;
}

extern "C" int mp_set_displacement_amplitude(mp_context* ctx, float amplitude)
{
    if (!ctx) return MP_E_INVALID;
    ctx->amplitude = amplitude;
    return MP_OK;
}

extern "C" int mp_prepare(mp_context* ctx)
{
    if (!ctx) return MP_E_INVALID;
    ctx->state = 1;
    ctx->objects.clear();
    ctx->view_projs.assign(1, mp_projection{ 60.0f, 0.01f, ctx->cfg.window_width, ctx->cfg.window_height });
    return MP_OK;
}

static int add_object(mp_context* ctx, int view, uint64_t handle, const float mv[16], const float color[4])
{
    if (!mv || !color) return fail(ctx, MP_E_INVALID, "null matrix or colour");
    if (!ctx->patches.count(handle)) return fail(ctx, MP_E_NOTFOUND, "unknown patch handle");
    ObjectReq o;
    o.handle = handle;
    o.view = view;
    memcpy(o.mv, mv, sizeof(o.mv));
    memcpy(o.color, color, sizeof(o.color));
    o.amplitude = ctx->amplitude;
    ctx->objects.push_back(o);
    return MP_OK;
}

extern "C" int mp_draw_patches(mp_context* ctx, uint64_t handle, const float mv[16],
                               const mp_projection* projection, const float color[4])
{
    if (!ctx) return MP_E_INVALID;
    if (ctx->state != 1) return fail(ctx, MP_E_STATE, "mp_draw_patches outside mp_prepare/mp_finish");
    if (!projection) return fail(ctx, MP_E_INVALID, "null projection");
    if (projection->width != ctx->cfg.window_width || projection->height != ctx->cfg.window_height)
        return fail(ctx, MP_E_INVALID, "projection viewport must equal the window size");
    if (ctx->objects.empty()) {
        ctx->view_projs[0] = *projection;
    } else if (projection->fovy_deg != ctx->view_projs[0].fovy_deg ||
               projection->near_plane != ctx->view_projs[0].near_plane) {
        // one projection per frame (Scene::draw passes the active camera's)
        return fail(ctx, MP_E_INVALID, "every draw of a frame must use the first draw's fovy and near plane");
    }
    return add_object(ctx, 0, handle, mv, color);
}

// ---------------------------------------------------------------------------
// frame execution

static int ensure_views(mp_context* ctx)
{
    const int nv = (int)ctx->view_projs.size();
    if ((int)ctx->views.size() < nv) ctx->views.resize(nv);
    ctx->max_pixels = 0;
    for (int v = 0; v < nv; ++v) {
        const mp_projection& pr = ctx->view_projs[v];
        ViewBuffers& vb = ctx->views[v];
        const bool want32 = ctx->cfg.write_rgba32 != 0;
        if (vb.width != pr.width || vb.height != pr.height || (want32 && !vb.rgba32)) {
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
            free_view(vb);
            vb.width = pr.width;
            vb.height = pr.height;
            vb.grid_w = (pr.width + 7) / 8;
            vb.grid_h = (pr.height + 7) / 8;
            const size_t nk = (size_t)vb.grid_w * vb.grid_h * 64;
            const size_t np = (size_t)pr.width * pr.height;
            int rc;
            if ((rc = dev_alloc(ctx, &vb.keys, nk)) || (rc = dev_alloc(ctx, &vb.rgba8, np * 4)) ||
                (rc = dev_alloc(ctx, &vb.depth, np)) || (rc = dev_alloc(ctx, &vb.ids, np)))
                return rc;
            if (want32 && (rc = dev_alloc(ctx, &vb.rgba32, np * 4))) return rc;
            HIPCHK(ctx, launch_fill_u64(vb.keys, nk, MP_EMPTY_KEY, ctx->stream));
            invalidate_graph(ctx);
        }
        ctx->max_pixels = std::max(ctx->max_pixels, vb.grid_w * vb.grid_h * 64);
    }
    return MP_OK;
}

static int ensure_lists(mp_context* ctx, uint32_t cap)
{
    if (cap <= ctx->list_cap && ctx->d_draw) return MP_OK;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    int rc;
    if ((rc = dev_alloc(ctx, &ctx->d_draw, cap)) || (rc = dev_alloc(ctx, &ctx->d_bucket, cap)) ||
        (rc = dev_alloc(ctx, &ctx->d_canon, cap)) || (rc = dev_alloc(ctx, &ctx->d_keys, cap)) ||
        (rc = dev_alloc(ctx, &ctx->d_span, cap)) || (rc = dev_alloc(ctx, &ctx->d_sort_keys, 2 * (size_t)cap)) ||
        (rc = dev_alloc(ctx, &ctx->d_sort_vals, 2 * (size_t)cap)))
        return rc;
    const size_t q = (size_t)cap * ctx->cfg.reyes_patch_size * ctx->cfg.reyes_patch_size;
    if ((rc = dev_alloc(ctx, &ctx->d_quad, q))) return rc;
    ctx->quad_cap = q;
    ctx->list_cap = cap;
    invalidate_graph(ctx);
    return MP_OK;
}

// grow the range lists to at least `need`, keeping the first `keep` DRAW
// records (the segment engine appends level by level)
static int grow_lists(mp_context* ctx, uint32_t need, uint32_t keep)
{
    if (need <= ctx->list_cap && ctx->d_draw) return MP_OK;
    size_t cap = ctx->list_cap;
    int rc;
    if ((rc = dev_grow(ctx, &ctx->d_draw, &cap, need, keep))) return rc;
    const uint32_t ncap = (uint32_t)cap;
    if ((rc = dev_alloc(ctx, &ctx->d_bucket, ncap)) || (rc = dev_alloc(ctx, &ctx->d_canon, ncap)) ||
        (rc = dev_alloc(ctx, &ctx->d_keys, ncap)) || (rc = dev_alloc(ctx, &ctx->d_span, ncap)) ||
        (rc = dev_alloc(ctx, &ctx->d_sort_keys, 2 * (size_t)ncap)) ||
        (rc = dev_alloc(ctx, &ctx->d_sort_vals, 2 * (size_t)ncap)))
        return rc;
    const size_t q = (size_t)ncap * ctx->cfg.reyes_patch_size * ctx->cfg.reyes_patch_size;
    if ((rc = dev_alloc(ctx, &ctx->d_quad, q))) return rc;
    ctx->quad_cap = q;
    ctx->list_cap = ncap;
    invalidate_graph(ctx);
    return MP_OK;
}

static int ensure_gpids(mp_context* ctx, uint32_t n)
{
    if (n + 1 <= ctx->gpid_cap && ctx->d_counts) return MP_OK;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t cap = std::max<uint32_t>(n + 1, 1024);
    int rc;
    if ((rc = dev_alloc(ctx, &ctx->d_counts, cap)) || (rc = dev_alloc(ctx, &ctx->d_offsets, cap)) ||
        (rc = dev_alloc(ctx, &ctx->d_cursor, cap)) || (rc = dev_alloc(ctx, &ctx->d_part, cap / 4096 + 2)))
        return rc;
    HIPCHK(ctx, launch_fill_u32(ctx->d_counts, cap, 0, ctx->stream));
    HIPCHK(ctx, launch_fill_u32(ctx->d_cursor, cap, 0, ctx->stream));
    ctx->gpid_cap = cap;
    invalidate_graph(ctx);
    return MP_OK;
}

static CanonBuffers canon_buffers(mp_context* ctx)
{
    CanonBuffers cb;
    cb.draw = ctx->d_draw;
    cb.bucket = ctx->d_bucket;
    cb.canon = ctx->d_canon;
    cb.keys = ctx->d_keys;
    cb.span = ctx->d_span;
    cb.cap = ctx->list_cap;
    cb.counts = ctx->d_counts;
    cb.offsets = ctx->d_offsets;
    cb.cursor = ctx->d_cursor;
    cb.part = ctx->d_part;
    cb.n_gpids = ctx->n_gpids;
    cb.sort_keys = ctx->d_sort_keys;
    cb.sort_vals = ctx->d_sort_vals;
    if (ctx->split_engine == 2 && ctx->forest_out_groups >= ctx->split_groups) {
        cb.forest_out = ctx->d_forest_out;
        cb.rank = ctx->d_rank;
        cb.forest_total = ctx->d_forest_total;
        cb.forest_groups = (uint32_t)ctx->split_groups;
    }
    return cb;
}

// Clears per-patch counts and bucket cursors that an interrupted launch
// sequence may have left non-zero (see mp_context::counts_dirty).
static int clear_stale_counts(mp_context* ctx)
{
    if (!ctx->counts_dirty || !ctx->d_counts) return MP_OK;
    HIPCHK(ctx, launch_fill_u32(ctx->d_counts, ctx->gpid_cap, 0u, ctx->stream));
    HIPCHK(ctx, launch_fill_u32(ctx->d_cursor, ctx->gpid_cap, 0u, ctx->stream));
    ctx->counts_dirty = false;
    return MP_OK;
}

static int upload_tables(mp_context* ctx)
{
    const int nv = (int)ctx->view_projs.size();
    const int no = (int)ctx->objects.size();
    if (nv > ctx->views_cap) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        hipFree(ctx->d_views); hipHostFree(ctx->h_views);
        ctx->d_views = nullptr; ctx->h_views = nullptr;
        HIPCHK(ctx, hipMalloc((void**)&ctx->d_views, nv * sizeof(DevView)));
        HIPCHK(ctx, hipHostMalloc((void**)&ctx->h_views, nv * sizeof(DevView), 0));
        ctx->views_cap = nv;
        invalidate_graph(ctx);
    }
    if (std::max(no, 1) > ctx->objs_cap) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        hipFree(ctx->d_objs); hipHostFree(ctx->h_objs);
        ctx->d_objs = nullptr; ctx->h_objs = nullptr;
        int cap = std::max(no, 1);
        HIPCHK(ctx, hipMalloc((void**)&ctx->d_objs, cap * sizeof(DevObject)));
        HIPCHK(ctx, hipHostMalloc((void**)&ctx->h_objs, cap * sizeof(DevObject), 0));
        ctx->objs_cap = cap;
        invalidate_graph(ctx);
    }
    // the previous frame may still read the staging buffers
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    // objects are ordered by view so that each view owns a contiguous
    // range of global patch indices (stable within a view = draw order)
    std::vector<int> order(no);
    for (int i = 0; i < no; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return ctx->objects[a].view < ctx->objects[b].view; });
    std::vector<uint32_t> first(nv, 0), end(nv, 0);
    std::vector<int> solo(nv, -1);
    uint32_t g = 0;
    ctx->patch_types = 0;
    int oi = 0;
    for (int v = 0; v < nv; ++v) {
        first[v] = g;
        const int oi0 = oi;
        while (oi < no && ctx->objects[order[oi]].view == v) {
            const ObjectReq& o = ctx->objects[order[oi]];
            const PatchSet& ps = ctx->patches[o.handle];
            DevObject& d = ctx->h_objs[oi];
            d.cps = ps.d_cps;
            memcpy(d.mv, o.mv, sizeof(d.mv));
            memcpy(d.color, o.color, sizeof(d.color));
            d.split_limit = ctx->cfg.bound_n_split_limit;
            d.disp_amplitude = o.amplitude;
            d.type = ps.type;
            ctx->patch_types |= ps.type == MP_GREGORY ? 2 : 1;
            d.n_patches = ps.n_patches;
            d.view = v;
            d.gpid_base = g;
            d.draw_index = order[oi];
            g += (uint32_t)ps.n_patches;
            ++oi;
        }
        if (oi == oi0 + 1) solo[v] = oi0;
        end[v] = g;
    }
    ctx->n_gpids = g;
    ctx->total_roots = g;
    // control-point gradient offsets in draw-call order
    {
        std::vector<uint32_t> off(no + 1, 0);
        for (int i = 0; i < no; ++i) {
            const PatchSet& ps = ctx->patches[ctx->objects[i].handle];
            off[i + 1] = off[i] + (uint32_t)ps.n_patches * (ps.type == MP_GREGORY ? 20u : 16u) * 3u;
        }
        for (int i = 0; i < no; ++i) ctx->h_objs[i].grad_offset = off[ctx->h_objs[i].draw_index];
        ctx->grad_floats = off[no];
    }
    uint32_t pix = 0;
    for (int v = 0; v < nv; ++v) {
        DevView& dv = ctx->h_views[v];
        make_view(ctx->view_projs[v], ctx->cfg.clear_color, ctx->cfg.subpixel_bits, dv);
        const ViewBuffers& vb = ctx->views[v];
        dv.keys = vb.keys;
        dv.rgba8 = vb.rgba8;
        dv.depth = vb.depth;
        dv.ids = vb.ids;
        dv.rgba32 = vb.rgba32;
        dv.first_gpid = first[v];
        dv.gpid_end = end[v];
        dv.pixel_offset = pix;
        dv.solo_obj = solo[v];
        pix += (uint32_t)dv.width * (uint32_t)dv.height;
    }
    if (nv) HIPCHK(ctx, hipMemcpyAsync(ctx->d_views, ctx->h_views, nv * sizeof(DevView), hipMemcpyHostToDevice, ctx->stream));
    if (no) HIPCHK(ctx, hipMemcpyAsync(ctx->d_objs, ctx->h_objs, no * sizeof(DevObject), hipMemcpyHostToDevice, ctx->stream));
    ctx->n_views = nv;
    return MP_OK;
}

// LOCAL persistent split in two launches: a step budget, then the left
// ranges dealt to all groups
static bool rebalancing(const mp_context* ctx)
{
    return ctx->split_engine == 0 && ctx->split_budget > 0 &&
           ctx->cfg.bound_n_split_method != MP_BNS_BALANCED && !ctx->work_sharing;
}

// BALANCED frames of the persistent engine share one root queue; every other
// device-driven split deals the roots to private queues
static bool shared_roots(const mp_context* ctx)
{
    return ctx->cfg.bound_n_split_method == MP_BNS_BALANCED && ctx->split_engine == 0;
}

static int ensure_queues(mp_context* ctx)
{
    int rc;
    const int balanced = shared_roots(ctx);
    int groups = ctx->cfg.local_bns_work_groups;
    if (groups <= 0 && ctx->split_engine == 2) {
        const int per_cu = ctx->forest_per_cu > 0 ? ctx->forest_per_cu
                                                  : split_forest_occupancy(split_forest_lds(ctx->forest_cap));
        groups = ctx->num_cus * per_cu;
    } else if (groups <= 0) {
        groups = ctx->num_cus * split_persistent_occupancy(split_persistent_lds(ctx->stack_cap), ctx->split_waves);
    }
    ctx->split_groups = groups;
    if (ctx->split_engine == 2 && groups > ctx->forest_out_groups) {
        // ranked forest frames: one output region of kForestKeys ranges per group
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        const size_t n = (size_t)groups * kForestKeys;
        if ((rc = dev_alloc(ctx, &ctx->d_forest_out, n)) || (rc = dev_alloc(ctx, &ctx->d_rank, n)) ||
            (rc = dev_alloc(ctx, &ctx->d_forest_total, (size_t)groups)))
            return rc;
        ctx->forest_out_groups = groups;
        invalidate_graph(ctx);
    }
    if (ctx->split_engine == 1 || rebalancing(ctx)) {
        // level lists: sized for the widest level; overflow re-renders bigger
        const uint32_t want = std::max<uint32_t>(ctx->level_cap, std::max<uint32_t>(65536, 8 * ctx->total_roots));
        if (want > ctx->level_cap || !ctx->d_level_a) {
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
            if ((rc = dev_alloc(ctx, &ctx->d_level_a, want)) || (rc = dev_alloc(ctx, &ctx->d_level_b, want))) return rc;
            ctx->level_cap = want;
            invalidate_graph(ctx);
        }
        if (!ctx->d_level_cnt) {
            if ((rc = dev_alloc(ctx, &ctx->d_level_cnt, 64))) return rc;
            HIPCHK(ctx, hipHostMalloc((void**)&ctx->h_level_cnt, 64 * sizeof(uint32_t), 0));
        }
    }
    // private root queues (ranges left by level launches or a step budget
    // are read from the level list in place)
    const size_t stride = balanced ? 1 : (ctx->total_roots + groups - 1) / groups;
    const size_t entries = std::max<size_t>(1, stride * groups);
    if (entries > ctx->queue_entries) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if ((rc = dev_alloc(ctx, &ctx->d_queue, entries))) return rc;
        ctx->queue_entries = entries;
        invalidate_graph(ctx);
    }
    if (groups > ctx->queue_groups_cap) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if ((rc = dev_alloc(ctx, &ctx->d_queue_cnt, groups))) return rc;
        ctx->queue_groups_cap = groups;
        invalidate_graph(ctx);
    }
    if (!ctx->d_shared_cnt && (rc = dev_alloc(ctx, &ctx->d_shared_cnt, 1))) return rc;
    // spilled-chunk pool for work sharing: enough for every range of a
    // frame four times over; running out only disables further spills
    const uint32_t want_chunks = std::max<uint32_t>(4096, (ctx->total_roots * 64u) / 64u + 4096);
    if (want_chunks > ctx->chunk_cap) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if ((rc = dev_alloc(ctx, &ctx->d_chunks, (size_t)want_chunks * 64)) ||
            (rc = dev_alloc(ctx, &ctx->d_chunk_count, want_chunks)) ||
            (rc = dev_alloc(ctx, &ctx->d_chunk_ready, want_chunks)))
            return rc;
        HIPCHK(ctx, launch_fill_u32(ctx->d_chunk_ready, want_chunks, 0u, ctx->stream));
        ctx->chunk_cap = want_chunks;
        invalidate_graph(ctx);
    }
    if (!ctx->d_share) {
        if ((rc = dev_alloc(ctx, &ctx->d_share, 1))) return rc;
        HIPCHK(ctx, launch_fill_u32(reinterpret_cast<uint32_t*>(ctx->d_share), 4, 0u, ctx->stream));
    }
    if (balanced && ctx->total_roots > ctx->shared_entries) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if ((rc = dev_alloc(ctx, &ctx->d_shared_q, ctx->total_roots))) return rc;
        ctx->shared_entries = ctx->total_roots;
        invalidate_graph(ctx);
    }
    return MP_OK;
}

static SplitQueues split_queues(mp_context* ctx)
{
    SplitQueues q;
    q.queue = ctx->d_queue;
    q.queue_cnt = ctx->d_queue_cnt;
    q.stride = shared_roots(ctx) ? 0 : (int64_t)((ctx->total_roots + ctx->split_groups - 1) / ctx->split_groups);
    q.shared_q = ctx->d_shared_q;
    q.shared_cnt = ctx->d_shared_cnt;
    q.n_groups = ctx->split_groups;
    q.stack_cap = ctx->split_engine == 2 ? ctx->forest_cap : ctx->stack_cap;
    q.chunks = ctx->d_chunks;
    q.chunk_count = ctx->d_chunk_count;
    q.chunk_ready = ctx->d_chunk_ready;
    q.chunk_cap = ctx->work_sharing ? ctx->chunk_cap : 0u;
    q.spill_at = ctx->spill_at;
    q.linger = ctx->linger >= 0 ? ctx->linger : ctx->num_cus;
    q.share = ctx->d_share;
    q.pending = nullptr;
    q.spread = ctx->split_spread;
    q.step_budget = 0;
    q.left = nullptr;
    q.left_cnt = nullptr;
    q.left_cap = 0;
    q.strided = nullptr;
    q.waves = ctx->split_waves;
    return q;
}

// The frame-end step of this context (FrameEnd): counters to d_fc_out, and
// the next device frame's counters and split state reset.
static FrameEnd frame_end(mp_context* ctx)
{
    FrameEnd fe;
    if (ctx->frame_ranked) {
        fe.counts = ctx->d_counts;
        fe.n_counts = ctx->n_gpids;
    }
    fe.fc = ctx->d_fc;
    fe.out = ctx->d_fc_out;
    fe.share = ctx->d_share;
    fe.left_cnt = ctx->d_level_cnt ? ctx->d_level_cnt + 63 : nullptr;
    if (ctx->cfg.bound_n_split_method == MP_BNS_BALANCED && ctx->d_shared_cnt) {
        fe.shared_cnt = ctx->d_shared_cnt;
        fe.shared_total = (int32_t)ctx->total_roots;
    }
    return fe;
}

// kernels that follow the split: canonical order, raster, resolve (which
// also ends the frame); an event is recorded after each stage when `ev` is
// given
static int launch_back_end(mp_context* ctx, hipEvent_t* ev = nullptr)
{
    CanonBuffers cb = canon_buffers(ctx);
    if (ctx->frame_ranked)
        HIPCHK(ctx, launch_place_ranked(ctx->d_views, ctx->n_views, cb, ctx->d_fc, ctx->num_cus,
                                        ctx->stream));
    else
        HIPCHK(ctx, launch_canonicalize(ctx->d_objs, ctx->d_views, ctx->n_views, cb, ctx->d_fc,
                                        ctx->num_cus, ctx->fused_order, ctx->stream));
    if (ev) HIPCHK(ctx, hipEventRecord(ev[0], ctx->stream));
    if (!ctx->cfg.dummy_render)
        HIPCHK(ctx, launch_raster(ctx->cfg.reyes_patch_size, ctx->d_canon, ctx->d_objs, ctx->d_views,
                                  ctx->prm, ctx->d_fc, ctx->d_quad, ctx->raster_grid_size, ctx->stream));
    if (ev) HIPCHK(ctx, hipEventRecord(ev[1], ctx->stream));
    if (!ctx->cfg.dummy_render)
        HIPCHK(ctx, launch_resolve(ctx->d_views, ctx->n_views, ctx->max_pixels, ctx->d_canon, ctx->d_objs,
                                   ctx->d_quad, ctx->cfg.reyes_patch_size, ctx->cfg.write_rgba32,
                                   ctx->d_display, ctx->stream, frame_end(ctx)));
    else
        HIPCHK(ctx, launch_frame_end(frame_end(ctx), ctx->stream));
    if (ev) HIPCHK(ctx, hipEventRecord(ev[2], ctx->stream));
    ctx->armed = true;
    return MP_OK;
}

// the device-driven frame (LOCAL / BALANCED): no host synchronisation.  The
// root queues are written (and the counters reset) only when the previous
// frame did not leave them ready: a frame ends by resetting what its split
// consumes, and the roots do not change between frames.
static mp_context::QueueSig queue_sig(const mp_context* ctx)
{
    const int balanced = shared_roots(ctx);
    const int64_t stride = balanced ? 0 : (int64_t)((ctx->total_roots + ctx->split_groups - 1) / ctx->split_groups);
    return { ctx->split_groups, ctx->total_roots, stride, balanced, ctx->d_queue, ctx->d_queue_cnt, ctx->d_shared_q,
             ctx->d_level_cnt };
}

// whether a device frame may start with its split
static bool frame_ready(const mp_context* ctx)
{
    return ctx->armed && queue_sig(ctx) == ctx->init_sig;
}

static int launch_device_frame(mp_context* ctx, hipEvent_t* ev = nullptr)
{
    const int balanced = shared_roots(ctx);
    const mp_context::QueueSig sig = queue_sig(ctx);
    const bool need_init = !frame_ready(ctx);
    ctx->frame_inits = need_init;   // read by a capture of this frame
    ctx->armed = false;   // until this frame's end
    ctx->frame_ranked = ctx->total_roots && ctx->split_engine == 2 && ctx->forest_rank &&
                        forest_rank_fits(ctx->n_gpids) && ctx->forest_out_groups >= ctx->split_groups;
    if (ev) HIPCHK(ctx, hipEventRecord(ev[0], ctx->stream));
    // the first split launch resets the frame counters
    if (!ctx->total_roots) HIPCHK(ctx, launch_reset(ctx->d_fc, ctx->d_share, ctx->stream));
    SplitQueues q = split_queues(ctx);
    if (ctx->total_roots && ctx->split_engine == 1) {
        if (!ctx->level_grid_size)
            ctx->level_grid_size = level_grid(ctx->num_cus, ctx->level_lanes, ctx->level_per_cu);
        HIPCHK(ctx, launch_split_levels(ctx->d_objs, (int)ctx->objects.size(), ctx->d_views, ctx->prm,
                                        ctx->total_roots, ctx->d_level_a, ctx->d_level_b, ctx->level_cap,
                                        ctx->d_level_cnt, ctx->level_launches, q, canon_buffers(ctx), ctx->d_fc,
                                        ctx->level_grid_size, ctx->level_lanes, ctx->d_share, ctx->stream));
    } else if (ctx->total_roots && ctx->split_engine == 2) {
        if (need_init) {
            HIPCHK(ctx, launch_init_queues(ctx->d_objs, (int)ctx->objects.size(), ctx->total_roots, q, 0,
                                           ctx->d_fc, ctx->d_share, ctx->stream));
            ctx->init_sig = sig;
        }
        HIPCHK(ctx, launch_split_forest(ctx->d_objs, ctx->d_views, ctx->prm, q, canon_buffers(ctx), ctx->d_fc,
                                        ctx->patch_types, (int)ctx->objects.size(), ctx->n_views,
                                        ctx->frame_ranked ? 1 : 0, ctx->stream));
    } else if (ctx->total_roots && rebalancing(ctx)) {
        q.step_budget = ctx->split_budget;
        q.left = ctx->d_level_a;
        q.left_cnt = ctx->d_level_cnt + 63;
        q.left_cap = ctx->level_cap;
        if (need_init) {
            HIPCHK(ctx, launch_init_queues(ctx->d_objs, (int)ctx->objects.size(), ctx->total_roots, q,
                                           balanced, ctx->d_fc, ctx->d_share, ctx->stream));
            ctx->init_sig = sig;
        }
        HIPCHK(ctx, launch_split_persistent(ctx->d_objs, ctx->d_views, ctx->prm, q,
                                            canon_buffers(ctx), ctx->d_fc, balanced, ctx->stream));
        HIPCHK(ctx, launch_split_rebalanced(ctx->d_objs, ctx->d_views, ctx->prm, q, canon_buffers(ctx),
                                            ctx->d_fc, ctx->stream));
    } else if (ctx->total_roots) {
        if (need_init) {
            HIPCHK(ctx, launch_init_queues(ctx->d_objs, (int)ctx->objects.size(), ctx->total_roots, q,
                                           balanced, ctx->d_fc, ctx->d_share, ctx->stream));
            ctx->init_sig = sig;
        }
        HIPCHK(ctx, launch_split_persistent(ctx->d_objs, ctx->d_views, ctx->prm, q,
                                            canon_buffers(ctx), ctx->d_fc, balanced, ctx->stream));
    }
    if (ev) HIPCHK(ctx, hipEventRecord(ev[1], ctx->stream));
    return launch_back_end(ctx, ev ? ev + 2 : nullptr);
}

// BREADTH / BOUNDED: host-driven loop per object (one sync per level/batch)
static int run_segment_split(mp_context* ctx, uint32_t& draw_total, int64_t& dropped, int64_t& max_patches)
{
    const bool breadth = ctx->cfg.bound_n_split_method == MP_BNS_BREADTH;
    const uint32_t ppp = (uint32_t)ctx->cfg.reyes_patches_per_pass;
    ctx->pass_counts.clear();
    int rc;
    for (int oi = 0; oi < (int)ctx->objects.size(); ++oi) {
        const DevObject& ob = ctx->h_objs[oi];
        const uint32_t P = (uint32_t)ob.n_patches;
        if (P == 0) continue;
        if (breadth) {
            if ((rc = dev_grow(ctx, &ctx->d_seg_a, &ctx->seg_a_cap, P, 0))) return rc;
            HIPCHK(ctx, launch_init_roots(oi, P, ctx->d_seg_a, ctx->stream));
            uint32_t n = P;
            while (n > 0) {
                max_patches = std::max<int64_t>(max_patches, n);
                const uint32_t nb = segment_blocks(n);
                if ((rc = dev_grow(ctx, &ctx->d_seg_b, &ctx->seg_b_cap, (size_t)2 * n, 0)) ||
                    (rc = dev_grow(ctx, &ctx->d_flags, &ctx->flags_cap, n, 0)) ||
                    (rc = dev_grow(ctx, &ctx->d_blk, &ctx->blk_cap, (size_t)2 * nb, 0)))
                    return rc;
                if ((rc = grow_lists(ctx, draw_total + std::min(n, ppp), draw_total))) return rc;
                HIPCHK(ctx, launch_bound_segment(ctx->d_objs, ctx->d_views, ctx->prm, ctx->d_seg_a,
                                                 ctx->d_seg_a, n, ctx->d_flags, ctx->d_blk, ctx->d_fc,
                                                 ctx->stream));
                HIPCHK(ctx, launch_scan_blocks(ctx->d_blk, nb, ctx->d_totals, ctx->stream));
                HIPCHK(ctx, launch_move_segment(ctx->d_objs, ctx->d_seg_a, n, ctx->d_flags, ctx->d_blk,
                                                ctx->d_draw, draw_total, ppp, ctx->d_seg_b,
                                                ctx->d_counts, ctx->stream));
                HIPCHK(ctx, hipMemcpyAsync(ctx->h_totals, ctx->d_totals, 2 * sizeof(uint32_t),
                                           hipMemcpyDeviceToHost, ctx->stream));
                HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
                const uint32_t draws = ctx->h_totals[0], splits = ctx->h_totals[1];
                ctx->pass_counts.push_back((int32_t)draws);
                const uint32_t emitted = std::min(draws, ppp);
                dropped += draws - emitted;
                draw_total += emitted;
                n = 2 * splits;
                std::swap(ctx->d_seg_a, ctx->d_seg_b);
                std::swap(ctx->seg_a_cap, ctx->seg_b_cap);
            }
        } else {
            // BOUNDED: LIFO stack, batches of reyes_patches_per_pass
            const size_t scap = (size_t)P + (size_t)2 * ppp;
            if ((rc = dev_grow(ctx, &ctx->d_seg_a, &ctx->seg_a_cap, scap, 0)) ||
                (rc = dev_grow(ctx, &ctx->d_seg_b, &ctx->seg_b_cap, ppp, 0)))
                return rc;
            HIPCHK(ctx, launch_init_roots(oi, P, ctx->d_seg_a, ctx->stream));
            uint32_t height = P;
            while (height > 0) {
                max_patches = std::max<int64_t>(max_patches, height);
                const uint32_t b = std::min(height, ppp);
                height -= b;
                const uint32_t nb = segment_blocks(b);
                if ((rc = dev_grow(ctx, &ctx->d_seg_a, &ctx->seg_a_cap, (size_t)height + 2 * b, height + b)) ||
                    (rc = dev_grow(ctx, &ctx->d_flags, &ctx->flags_cap, b, 0)) ||
                    (rc = dev_grow(ctx, &ctx->d_blk, &ctx->blk_cap, (size_t)2 * nb, 0)))
                    return rc;
                if ((rc = grow_lists(ctx, draw_total + b, draw_total))) return rc;
                HIPCHK(ctx, launch_bound_segment(ctx->d_objs, ctx->d_views, ctx->prm, ctx->d_seg_a + height,
                                                 ctx->d_seg_b, b, ctx->d_flags, ctx->d_blk, ctx->d_fc,
                                                 ctx->stream));
                HIPCHK(ctx, launch_scan_blocks(ctx->d_blk, nb, ctx->d_totals, ctx->stream));
                HIPCHK(ctx, launch_move_segment(ctx->d_objs, ctx->d_seg_b, b, ctx->d_flags, ctx->d_blk,
                                                ctx->d_draw, draw_total, 0xFFFFFFFFu,
                                                ctx->d_seg_a + height, ctx->d_counts, ctx->stream));
                HIPCHK(ctx, hipMemcpyAsync(ctx->h_totals, ctx->d_totals, 2 * sizeof(uint32_t),
                                           hipMemcpyDeviceToHost, ctx->stream));
                HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
                const uint32_t draws = ctx->h_totals[0], splits = ctx->h_totals[1];
                ctx->pass_counts.push_back((int32_t)draws);
                draw_total += draws;
                height += 2 * splits;
            }
        }
    }
    return MP_OK;
}

// Grows the canonical list (and the Lambert store) to hold `need` ranges,
// keeping the first `keep` of both; the other lists are reallocated empty.
static int grow_canon(mp_context* ctx, uint32_t need, uint32_t keep)
{
    if (need <= ctx->list_cap) return MP_OK;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    size_t cap = ctx->list_cap;
    int rc;
    if ((rc = dev_grow(ctx, &ctx->d_canon, &cap, need, keep))) return rc;
    const size_t N2 = (size_t)ctx->cfg.reyes_patch_size * ctx->cfg.reyes_patch_size;
    float* quad = nullptr;
    if (hipMalloc((void**)&quad, cap * N2 * sizeof(float)) != hipSuccess)
        return fail(ctx, MP_E_NOMEM, "hipMalloc: Lambert store");
    if (keep) HIPCHK(ctx, hipMemcpy(quad, ctx->d_quad, (size_t)keep * N2 * sizeof(float), hipMemcpyDeviceToDevice));
    hipFree(ctx->d_quad);
    ctx->d_quad = quad;
    ctx->quad_cap = cap * N2;
    const uint32_t ncap = (uint32_t)cap;
    if ((rc = dev_alloc(ctx, &ctx->d_draw, ncap)) || (rc = dev_alloc(ctx, &ctx->d_bucket, ncap)) ||
        (rc = dev_alloc(ctx, &ctx->d_keys, ncap)) || (rc = dev_alloc(ctx, &ctx->d_span, ncap)) ||
        (rc = dev_alloc(ctx, &ctx->d_sort_keys, 2 * (size_t)ncap)) ||
        (rc = dev_alloc(ctx, &ctx->d_sort_vals, 2 * (size_t)ncap)))
        return rc;
    ctx->list_cap = ncap;
    invalidate_graph(ctx);
    return MP_OK;
}

static int hip_event(mp_context* ctx, std::vector<hipEvent_t>& pool, size_t k)
{
    while (pool.size() <= k) {
        hipEvent_t e = nullptr;
        HIPCHK(ctx, hipEventCreate(&e));
        pool.push_back(e);
    }
    return MP_OK;
}

// CPU method (BoundNSplitCLCPU.cpp:75-228): the host splits the patches on
// its cores and hands the DRAW ranges over in canonical order (patches in
// order, each patch's ranges sorted on the host), so every batch of
// reyes_patches_per_pass ranges goes straight into the canonical list at its
// final place: it is copied from a slot of the pinned ring of
// bns_pipeline_length buffers and rasterised by a launch of its own while the
// host splits the next patches.  The ring slot is reused once its copy has
// completed.  The resolve follows the last batch.
static int run_cpu_frame(mp_context* ctx, uint32_t& draw_total, int64_t& max_patches)
{
    const int no = (int)ctx->objects.size();
    std::vector<CpuSplitObject> objs(no);
    for (int oi = 0; oi < no; ++oi) {
        const DevObject& d = ctx->h_objs[oi];
        const PatchSet& ps = ctx->patches[ctx->objects[d.draw_index].handle];
        const DevView& v = ctx->h_views[d.view];
        CpuSplitObject& o = objs[oi];
        o.cps = ps.host.data();
        o.type = d.type;
        o.n_patches = d.n_patches;
        memcpy(o.mv, d.mv, sizeof(o.mv));
        o.width = v.width;
        o.height = v.height;
        o.fx = v.f[0];
        o.fy = v.f[1];
        o.near_ = v.near_;
        o.split_limit = d.split_limit;
        o.gpid_base = d.gpid_base;
        o.slot = (uint32_t)oi;
    }
    const size_t slot_cap = (size_t)std::max(1, ctx->cfg.reyes_patches_per_pass);
    const int ring_len = std::max(1, ctx->cfg.bns_pipeline_length);
    if (!ctx->h_ring || ctx->ring_len != ring_len || ctx->ring_slot != slot_cap) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->h_ring) hipHostFree(ctx->h_ring);
        ctx->h_ring = nullptr;
        HIPCHK(ctx, hipHostMalloc((void**)&ctx->h_ring, (size_t)ring_len * slot_cap * sizeof(RangeRec), 0));
        ctx->ring_len = ring_len;
        ctx->ring_slot = slot_cap;
    }
    int rc;
    if ((rc = hip_event(ctx, ctx->ring_ev, (size_t)ring_len - 1))) return rc;
    if (!ctx->trace_ev0) HIPCHK(ctx, hipEventCreate(&ctx->trace_ev0));
    std::vector<bool> busy((size_t)ring_len, false);
    DevParams prm = ctx->prm;
    prm.raster_chunk = 0;   // windows take the static range schedule
    const bool render = !ctx->cfg.dummy_render;

    ctx->cpu_trace.clear();
    const auto t0 = std::chrono::steady_clock::now();
    auto now_ms = [&] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    HIPCHK(ctx, hipEventRecord(ctx->trace_ev0, ctx->stream));
    uint32_t shipped = 0;
    size_t fill = 0;
    int slot = 0, cur_view = -1;
    double batch_t0 = 0.0;
    int err = MP_OK;
    auto ship = [&]() -> int {
        if (!fill) return MP_OK;
        int r2;
        if ((r2 = grow_canon(ctx, shipped + (uint32_t)fill, shipped))) return r2;
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_canon + shipped, ctx->h_ring + (size_t)slot * slot_cap,
                                   fill * sizeof(RangeRec), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipEventRecord(ctx->ring_ev[slot], ctx->stream));
        busy[slot] = true;
        const size_t b = ctx->cpu_trace.size();
        if ((r2 = hip_event(ctx, ctx->trace_ev, 2 * b + 1))) return r2;
        ctx->cpu_trace.push_back(mp_context::BatchTrace{ batch_t0, now_ms(), (int)fill });
        HIPCHK(ctx, hipEventRecord(ctx->trace_ev[2 * b], ctx->stream));
        if (render)
            HIPCHK(ctx, launch_raster(ctx->cfg.reyes_patch_size, ctx->d_canon, ctx->d_objs, ctx->d_views, prm,
                                      ctx->d_fc, ctx->d_quad, ctx->raster_grid_size, ctx->stream, shipped,
                                      (uint32_t)fill));
        HIPCHK(ctx, hipEventRecord(ctx->trace_ev[2 * b + 1], ctx->stream));
        shipped += (uint32_t)fill;
        fill = 0;
        slot = (slot + 1) % ring_len;
        if (busy[slot]) HIPCHK(ctx, hipEventSynchronize(ctx->ring_ev[slot]));   // its copy has read the slot
        busy[slot] = false;
        batch_t0 = now_ms();
        return MP_OK;
    };
    auto emit = [&](const RangeRec* r, size_t n) {
        for (size_t i = 0; i < n && err == MP_OK; ++i) {
            const int v = ctx->h_objs[r[i].obj_depth & 0xffffffu].view;
            while (cur_view < v && err == MP_OK) {
                ++cur_view;   // the view's first range: its canonical base
                if (launch_set_view_base(ctx->d_views, cur_view, shipped + (uint32_t)fill, nullptr, 0u,
                                         ctx->stream) != hipSuccess)
                    err = fail(ctx, MP_E_HIP, "launch_set_view_base failed");
            }
            ctx->h_ring[(size_t)slot * slot_cap + fill++] = r[i];
            if (fill == slot_cap && err == MP_OK) err = ship();
        }
    };
    CpuSplitFrame& fr = ctx->cpu_frame;
    cpu_split_stream(objs, ctx->n_gpids, ctx->cfg.cull_ribbon, ctx->cfg.max_split_depth,
                     ctx->cfg.reyes_patches_per_pass, ctx->cpu_threads, fr, emit);
    if (err) return err;
    if ((rc = ship())) return rc;
    // views with no range, and the frame's range total
    while (cur_view + 1 < ctx->n_views) {
        ++cur_view;
        HIPCHK(ctx, launch_set_view_base(ctx->d_views, cur_view, shipped, nullptr, 0u, ctx->stream));
    }
    HIPCHK(ctx, launch_set_view_base(ctx->d_views, -1, 0u, ctx->d_fc, shipped, ctx->stream));
    if (render)
        HIPCHK(ctx, launch_resolve(ctx->d_views, ctx->n_views, ctx->max_pixels, ctx->d_canon, ctx->d_objs,
                                   ctx->d_quad, ctx->cfg.reyes_patch_size, ctx->cfg.write_rgba32,
                                   ctx->d_display, ctx->stream, frame_end(ctx)));
    else
        HIPCHK(ctx, launch_frame_end(frame_end(ctx), ctx->stream));
    ctx->armed = true;
    draw_total = shipped;
    ctx->pass_counts = fr.pass_counts;
    max_patches = fr.max_stack;
    return MP_OK;
}

// Level engine: launch only as many levels as the last frame needed (+1);
// the persistent tail still finishes anything deeper.
static int update_level_launches(mp_context* ctx, int64_t* max_patches)
{
    HIPCHK(ctx, hipMemcpy(ctx->h_level_cnt, ctx->d_level_cnt, 64 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    int deepest = 0;
    for (int L = 0; L <= ctx->level_launches && L < 63; ++L) {
        if (ctx->h_level_cnt[L]) deepest = L;
        if (max_patches) *max_patches = std::max<int64_t>(*max_patches, ctx->h_level_cnt[L]);
    }
    ctx->level_launches = std::min(ctx->cfg.max_split_depth, deepest + 2);
    if (ctx->level_limit > 0) ctx->level_launches = std::min(ctx->level_launches, ctx->level_limit);
    return MP_OK;
}

static int execute_frame(mp_context* ctx)
{
    int rc;
    ctx->frame_valid = false;
    ctx->frame_ranked = false;   // set by a device frame that takes the ranked order
    invalidate_graph(ctx);  // a captured graph bakes the previous frame's launch sizes
    if ((rc = ensure_views(ctx))) return rc;
    if ((rc = upload_tables(ctx))) return rc;
    if ((rc = ensure_gpids(ctx, ctx->n_gpids))) return rc;
    if (ctx->list_cap == 0 && (rc = ensure_lists(ctx, std::max<uint32_t>(65536, 4 * ctx->total_roots)))) return rc;
    if ((rc = clear_stale_counts(ctx))) return rc;
    ctx->counts_dirty = true;   // until the frame has completed
    const int method = ctx->cfg.bound_n_split_method;
    const bool device_driven = (method == MP_BNS_LOCAL || method == MP_BNS_BALANCED);
    int64_t relaunches = 0, dropped = 0, max_patches = 0;
    uint32_t draw_total = 0;

    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (device_driven) {
        // the forest split stages the frame's tables in LDS: bigger tables
        // take the persistent engine
        if (ctx->split_engine == 2 && !split_forest_fits((int)ctx->objects.size(), ctx->n_views))
            ctx->split_engine = 0;
        if ((rc = ensure_queues(ctx))) return rc;
        for (;;) {
            if ((rc = launch_device_frame(ctx))) return rc;
            HIPCHK(ctx, hipMemcpyAsync(ctx->h_fc, ctx->d_fc_out, sizeof(FrameCounters), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
            const uint32_t ovf = ctx->h_fc->overflow;
            if (ovf & 4u) return fail(ctx, MP_E_HIP, "split work sharing timed out");
            if (!(ovf & 43u)) break;
            if (ovf & 32u) {
                // a forest group held more DRAW ranges than it can rank:
                // the frame takes the unranked order
                ctx->forest_rank = 0;
                ctx->counts_dirty = true;
                if ((rc = clear_stale_counts(ctx))) return rc;
            }
            if (ovf & 8u) {
                ctx->level_cap *= 4;   // a split level outgrew its list
                HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
                if ((rc = dev_alloc(ctx, &ctx->d_level_a, ctx->level_cap)) ||
                    (rc = dev_alloc(ctx, &ctx->d_level_b, ctx->level_cap)))
                    return rc;
                if ((rc = ensure_queues(ctx))) return rc;
            }
            ++relaunches;
            if ((ovf & 2u) && ctx->split_engine == 2) {
                // a forest group's trees outgrew its LDS stack: the
                // persistent engine (which spreads ranges over groups)
                ctx->split_engine = 0;
                if ((rc = ensure_queues(ctx))) return rc;
            } else if (ovf & 2u) {
                const int full = 64 * (ctx->cfg.max_split_depth + 2);
                if (ctx->stack_cap >= full) return fail(ctx, MP_E_INVALID, "split stack overflow");
                ctx->stack_cap = full;   // deeper LDS stack; fewer resident groups
                if ((rc = ensure_queues(ctx))) return rc;
            }
            // draw list too small: grow and render the frame again
            if ((ovf & 1u) && (rc = ensure_lists(ctx, ctx->list_cap * 2))) return rc;
            HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        }
        draw_total = ctx->h_fc->total_ranges;
        const int64_t ppp = ctx->cfg.reyes_patches_per_pass;
        ctx->pass_counts.clear();
        int64_t left = draw_total;
        do {
            ctx->pass_counts.push_back((int32_t)std::min<int64_t>(left, ppp));
            left -= ppp;
        } while (left > 0);
        max_patches = ctx->h_fc->max_stack;
        if (ctx->split_engine == 1 && (rc = update_level_launches(ctx, &max_patches))) return rc;
    } else if (method == MP_BNS_CPU) {
        HIPCHK(ctx, launch_reset(ctx->d_fc, ctx->d_share, ctx->stream));
        if ((rc = run_cpu_frame(ctx, draw_total, max_patches))) return rc;
        HIPCHK(ctx, hipMemcpyAsync(ctx->h_fc, ctx->d_fc_out, sizeof(FrameCounters), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        HIPCHK(ctx, launch_reset(ctx->d_fc, ctx->d_share, ctx->stream));
        if ((rc = run_segment_split(ctx, draw_total, dropped, max_patches))) return rc;
        if ((rc = launch_back_end(ctx))) return rc;
        HIPCHK(ctx, hipMemcpyAsync(ctx->h_fc, ctx->d_fc_out, sizeof(FrameCounters), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    // the canonical order's scan must see exactly this frame's DRAW ranges
    const uint32_t emitted = device_driven ? ctx->h_fc->draw_count : draw_total;
    if ((ctx->h_fc->overflow & 16u) || ctx->h_fc->total_ranges != emitted)
        return fail(ctx, MP_E_STATE, "per-patch range counts disagree with the draw list; frame discarded");
    ctx->counts_dirty = false;

    mp_stats& st = ctx->stats;
    memset(&st, 0, sizeof(st));
    const int64_t N2 = (int64_t)ctx->cfg.reyes_patch_size * ctx->cfg.reyes_patch_size;
    st.diced_ranges = ctx->h_fc->total_ranges;
    st.dropped_ranges = dropped;
    st.patches_per_frame = st.diced_ranges + dropped;
    st.bounded_patches = method == MP_BNS_CPU ? ctx->cpu_frame.bounds : (int64_t)ctx->h_fc->bounds;
    st.micropolygons = ctx->cfg.dummy_render ? 0 : N2 * st.diced_ranges;
    st.pass_count = (int64_t)ctx->pass_counts.size();
    st.max_patches = max_patches;
    st.slow_blocks = ctx->h_fc->slow_blocks;
    st.overflow_relaunches = relaunches;
    st.coverage_samples = (int64_t)ctx->h_fc->atomics;
    st.sorted_ranges = (int64_t)ctx->h_fc->big_ranges;
    st.frame_ms = ms;
    ctx->device_driven = device_driven;
    ctx->frame_valid = true;
    ctx->frame_serial = ++ctx->serial_counter;
    return MP_OK;
}

extern "C" int mp_cpu_bound_n_split(const mp_config* cfg, const float* xyz, size_t n_points,
                                    int32_t patch_type, const float mv[16],
                                    const mp_projection* projection, mp_range_record* records,
                                    int64_t capacity, int64_t* count, int32_t* pass_counts,
                                    int32_t max_passes, int32_t* n_passes, int64_t* n_bounds)
{
    if (!cfg || !mv || !projection) { g_create_error = "null argument"; return MP_E_INVALID; }
    std::string why;
    if (!validate_config(cfg, why)) { g_create_error = why; return MP_E_INVALID; }
    if (patch_type != MP_BEZIER && patch_type != MP_GREGORY) { g_create_error = "unknown patch type"; return MP_E_INVALID; }
    const size_t per = patch_type == MP_GREGORY ? 20 : 16;
    if (n_points % per != 0 || (n_points && !xyz) || n_points / per >= (1u << 20)) {
        g_create_error = "bad patch data";
        return MP_E_INVALID;
    }
    if (projection->width < 1 || projection->height < 1) { g_create_error = "view size must be positive"; return MP_E_INVALID; }
    std::vector<float> host(n_points * 4);
    for (size_t i = 0; i < n_points; ++i) {
        host[4 * i] = xyz[3 * i]; host[4 * i + 1] = xyz[3 * i + 1]; host[4 * i + 2] = xyz[3 * i + 2];
        host[4 * i + 3] = 1.0f;
    }
    DevView v;
    make_view(*projection, cfg->clear_color, cfg->subpixel_bits, v);
    CpuSplitObject o;
    o.cps = host.data();
    o.type = patch_type;
    o.n_patches = (int)(n_points / per);
    memcpy(o.mv, mv, sizeof(o.mv));
    o.width = v.width; o.height = v.height;
    o.fx = v.f[0]; o.fy = v.f[1]; o.near_ = v.near_;
    o.split_limit = cfg->bound_n_split_limit;
    o.gpid_base = 0;
    o.slot = 0;
    CpuSplitFrame fr;
    const char* e = getenv("MP_CPU_THREADS");
    cpu_split_frame(std::vector<CpuSplitObject>(1, o), (uint32_t)o.n_patches, cfg->cull_ribbon,
                    cfg->max_split_depth, cfg->reyes_patches_per_pass, e ? atoi(e) : 0, fr);
    // fr.draws is in canonical order: (pid, umin, vmin)
    const int64_t n = (int64_t)fr.draws.size();
    if (count) *count = n;
    if (records) {
        if (capacity < n) { g_create_error = "record capacity too small"; return MP_E_INVALID; }
        for (int64_t i = 0; i < n; ++i) {
            const RangeRec& r = fr.draws[(size_t)i];
            records[i] = mp_range_record{ (int32_t)r.pid, 0, (int32_t)(r.obj_depth >> 24),
                                          r.umin, r.vmin, r.umax, r.vmax };
        }
    }
    if (n_passes) *n_passes = (int32_t)fr.pass_counts.size();
    if (pass_counts)
        for (int32_t k = 0; k < max_passes && k < (int32_t)fr.pass_counts.size(); ++k) pass_counts[k] = fr.pass_counts[k];
    if (n_bounds) *n_bounds = fr.bounds;
    return MP_OK;
}

extern "C" int mp_finish(mp_context* ctx)
{
    if (!ctx) return MP_E_INVALID;
    if (ctx->state != 1) return fail(ctx, MP_E_STATE, "mp_finish without mp_prepare");
    ctx->state = 0;
    return execute_frame(ctx);
}

extern "C" int mp_batch_begin(mp_context* ctx, int32_t n_views, const mp_projection* projections)
{
    if (!ctx) return MP_E_INVALID;
    if (n_views < 1 || !projections) return fail(ctx, MP_E_INVALID, "need at least one view");
    for (int v = 0; v < n_views; ++v)
        if (projections[v].width < 1 || projections[v].height < 1)
            return fail(ctx, MP_E_INVALID, "view size must be positive");
    ctx->state = 2;
    ctx->objects.clear();
    ctx->view_projs.assign(projections, projections + n_views);
    return MP_OK;
}

extern "C" int mp_batch_draw(mp_context* ctx, int32_t view, uint64_t handle, const float mv[16],
                             const float color[4])
{
    if (!ctx) return MP_E_INVALID;
    if (ctx->state != 2) return fail(ctx, MP_E_STATE, "mp_batch_draw outside a batch");
    if (view < 0 || view >= (int)ctx->view_projs.size()) return fail(ctx, MP_E_INVALID, "bad view index");
    return add_object(ctx, view, handle, mv, color);
}

extern "C" int mp_batch_finish(mp_context* ctx)
{
    if (!ctx) return MP_E_INVALID;
    if (ctx->state != 2) return fail(ctx, MP_E_STATE, "mp_batch_finish without mp_batch_begin");
    ctx->state = 0;
    return execute_frame(ctx);
}

static int read_view(mp_context* ctx, int view, uint8_t* rgba8, float* depth, uint32_t* ids, float* rgba32)
{
    if (!ctx->frame_valid) return fail(ctx, MP_E_STATE, "no finished frame");
    if (view < 0 || view >= ctx->n_views) return fail(ctx, MP_E_INVALID, "bad view index");
    if (ctx->cfg.dummy_render) return fail(ctx, MP_E_STATE, "dummy_render frames have no framebuffer");
    const ViewBuffers& vb = ctx->views[view];
    const size_t np = (size_t)vb.width * vb.height;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (rgba8) HIPCHK(ctx, hipMemcpy(rgba8, vb.rgba8, np * 4, hipMemcpyDeviceToHost));
    if (depth) HIPCHK(ctx, hipMemcpy(depth, vb.depth, np * sizeof(float), hipMemcpyDeviceToHost));
    if (ids) HIPCHK(ctx, hipMemcpy(ids, vb.ids, np * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (rgba32) {
        if (!vb.rgba32) return fail(ctx, MP_E_STATE, "write_rgba32 is off");
        HIPCHK(ctx, hipMemcpy(rgba32, vb.rgba32, np * 4 * sizeof(float), hipMemcpyDeviceToHost));
    }
    return MP_OK;
}

extern "C" int mp_read_framebuffer(mp_context* ctx, uint8_t* rgba8, float* depth, uint32_t* ids,
                                   float* rgba32)
{
    if (!ctx) return MP_E_INVALID;
    return read_view(ctx, 0, rgba8, depth, ids, rgba32);
}

extern "C" int mp_batch_read(mp_context* ctx, int32_t view, uint8_t* rgba8, float* depth,
                             uint32_t* ids, float* rgba32)
{
    if (!ctx) return MP_E_INVALID;
    return read_view(ctx, view, rgba8, depth, ids, rgba32);
}

extern "C" int mp_read_ranges(mp_context* ctx, mp_range_record* records, int64_t capacity,
                              int64_t* count)
{
    if (!ctx) return MP_E_INVALID;
    if (!ctx->frame_valid) return fail(ctx, MP_E_STATE, "no finished frame");
    const uint32_t n = ctx->h_fc->total_ranges;
    if (count) *count = n;
    if (!records) return MP_OK;
    if (capacity < (int64_t)n) return fail(ctx, MP_E_INVALID, "capacity too small");
    std::vector<RangeRec> host(n);
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (n) HIPCHK(ctx, hipMemcpy(host.data(), ctx->d_canon, n * sizeof(RangeRec), hipMemcpyDeviceToHost));
    // object slots are view-ordered on the device; report draw-call order
    std::vector<int> order(ctx->objects.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return ctx->objects[a].view < ctx->objects[b].view; });
    for (uint32_t i = 0; i < n; ++i) {
        records[i].pid = (int32_t)host[i].pid;
        records[i].object = order[host[i].obj_depth & 0xffffffu];
        records[i].depth = (int32_t)(host[i].obj_depth >> 24);
        records[i].umin = host[i].umin;
        records[i].vmin = host[i].vmin;
        records[i].umax = host[i].umax;
        records[i].vmax = host[i].vmax;
    }
    return MP_OK;
}

extern "C" int mp_read_grids(mp_context* ctx, int64_t first_range, int64_t n_ranges, float* pos,
                             int32_t* pxlpos, float* depth)
{
    if (!ctx) return MP_E_INVALID;
    if (!ctx->frame_valid) return fail(ctx, MP_E_STATE, "no finished frame");
    const int64_t total = ctx->h_fc->total_ranges;
    if (first_range < 0 || n_ranges < 0 || first_range + n_ranges > total)
        return fail(ctx, MP_E_INVALID, "range interval outside the frame's ranges");
    if (!pos || !pxlpos || !depth) return fail(ctx, MP_E_INVALID, "null output");
    if (n_ranges == 0) return MP_OK;
    const size_t G = (size_t)(ctx->cfg.reyes_patch_size + 1) * (ctx->cfg.reyes_patch_size + 1);
    const size_t pts = (size_t)n_ranges * G;
    char* d = nullptr;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (hipMalloc((void**)&d, pts * (16 + 8 + 4)) != hipSuccess) return fail(ctx, MP_E_NOMEM, "grid tap buffer");
    float4* d_pos = reinterpret_cast<float4*>(d);
    int2* d_pxl = reinterpret_cast<int2*>(d + pts * 16);
    float* d_dep = reinterpret_cast<float*>(d + pts * 24);
    hipError_t e = launch_read_grids(ctx->cfg.reyes_patch_size, ctx->d_canon, ctx->d_objs, ctx->d_views, ctx->prm,
                                     (uint32_t)first_range, (uint32_t)n_ranges, d_pos, d_pxl, d_dep, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(pos, d_pos, pts * 16, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(pxlpos, d_pxl, pts * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(depth, d_dep, pts * 4, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return fail(ctx, MP_E_HIP, std::string("grid tap: ") + hipGetErrorString(e));
    return MP_OK;
}

extern "C" int mp_read_pass_counts(mp_context* ctx, int32_t* counts, int32_t capacity, int32_t* n_passes)
{
    if (!ctx) return MP_E_INVALID;
    if (!ctx->frame_valid) return fail(ctx, MP_E_STATE, "no finished frame");
    const int32_t n = (int32_t)ctx->pass_counts.size();
    if (n_passes) *n_passes = n;
    if (counts) {
        if (capacity < n) return fail(ctx, MP_E_INVALID, "capacity too small");
        memcpy(counts, ctx->pass_counts.data(), n * sizeof(int32_t));
    }
    return MP_OK;
}

extern "C" int mp_get_stats(mp_context* ctx, mp_stats* stats)
{
    if (!ctx || !stats) return MP_E_INVALID;
    *stats = ctx->stats;
    return MP_OK;
}

// ---------------------------------------------------------------------------
// trace and statistics files in the reference's formats

// Device memory held by the context, by use (Statistics::opencl_memory_by_use).
static std::vector<std::pair<std::string, size_t>> memory_by_use(const mp_context* ctx)
{
    size_t patches = 0, fb = 0;
    for (const auto& kv : ctx->patches)
        patches += (size_t)kv.second.n_patches * (kv.second.type == MP_GREGORY ? 20 : 16) * sizeof(float4);
    for (const ViewBuffers& vb : ctx->views) {
        const size_t px = (size_t)vb.width * vb.height;
        fb += (size_t)vb.grid_w * vb.grid_h * 64 * 8 + px * (4 + 4 + 4) + (vb.rgba32 ? px * 16 : 0);
    }
    const size_t split = ctx->queue_entries * sizeof(RangeRec) + (size_t)ctx->queue_groups_cap * 4 +
                         ctx->shared_entries * sizeof(RangeRec) + (size_t)ctx->chunk_cap * (64 * sizeof(RangeRec) + 8) +
                         (size_t)ctx->level_cap * 2 * sizeof(RangeRec) + (ctx->seg_a_cap + ctx->seg_b_cap) * sizeof(RangeRec) +
                         ctx->flags_cap + ctx->blk_cap * 4;
    const size_t lists = (size_t)ctx->list_cap * (3 * sizeof(RangeRec) + 8 + 8 + 24) + ctx->quad_cap * sizeof(float) +
                         (size_t)ctx->gpid_cap * 12 + ctx->visible_cap * 4;
    return { { "patch data", patches }, { "framebuffer", fb }, { "bound&split", split }, { "range lists", lists } };
}

extern "C" int mp_dump_stats(mp_context* ctx, const char* path)
{
    if (!ctx || !path) return MP_E_INVALID;
    if (!ctx->frame_valid) return fail(ctx, MP_E_STATE, "no finished frame");
    std::ofstream fs(path);
    if (!fs) return fail(ctx, MP_E_INVALID, std::string("cannot write ") + path);
    const auto mem = memory_by_use(ctx);
    size_t total = 0;
    for (const auto& m : mem) total += m.second;
    int64_t inputs = 0;
    for (const ObjectReq& o : ctx->objects) inputs += ctx->patches[o.handle].n_patches;
    const mp_stats& st = ctx->stats;
    fs << "opencl_mem = " << total << ";\n";
    for (const auto& m : mem) fs << "opencl_mem@" << m.first << " = " << m.second << ";\n";
    fs << "opengl_mem = 0;\n";
    fs << "max_patches = " << st.max_patches << ";\n";
    fs << "patches_per_frame = " << st.patches_per_frame << ";\n";
    fs << "processed_patches_per_frame = " << st.bounded_patches << ";\n";
    fs << "pass_count = " << st.pass_count << ";\n";
    fs << "total_input_patches = " << inputs << ";\n";
    fs << "bound_n_split_balance = ;";
    return fs.good() ? MP_OK : fail(ctx, MP_E_INVALID, "write failed");
}

extern "C" int mp_dump_trace(mp_context* ctx, const char* path)
{
    if (!ctx || !path) return MP_E_INVALID;
    if (!ctx->frame_valid) return fail(ctx, MP_E_STATE, "no finished frame");
    std::ofstream fs(path);
    if (!fs) return fail(ctx, MP_E_INVALID, std::string("cannot write ") + path);
    const uint64_t base = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                              std::chrono::steady_clock::now().time_since_epoch()).count();
    struct Rec { const char* name; double t0, t1; };
    std::vector<Rec> recs;
    if (ctx->device_driven) {
        // one more run of the frame with an event between its stages
        hipEvent_t ev[5] = { nullptr, nullptr, nullptr, nullptr, nullptr };
        int rc = clear_stale_counts(ctx);
        ctx->counts_dirty = true;
        for (auto& e : ev)
            if (rc == MP_OK && hipEventCreate(&e) != hipSuccess) rc = fail(ctx, MP_E_HIP, "hipEventCreate failed");
        if (rc == MP_OK) rc = launch_device_frame(ctx, ev);
        if (rc == MP_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ctx, MP_E_HIP, "sync failed");
        if (rc == MP_OK) ctx->counts_dirty = false;
        // names from the reference's trace categories (tools/trace_summary.py):
        // split -> subdivision, canonical order (scan + scatter) -> subdivision,
        // fused dice/shade/sample -> sample, resolve (also clears) -> clear
        static const char* names[4] = { "bound & split", "accumulate", "sample", "clear framebuffer" };
        for (int k = 0; k < 4 && rc == MP_OK; ++k) {
            float a = 0.0f, b = 0.0f;
            hipEventElapsedTime(&a, ev[0], ev[k]);
            hipEventElapsedTime(&b, ev[0], ev[k + 1]);
            recs.push_back(Rec{ names[k], a, b });
        }
        for (auto& e : ev)
            if (e) hipEventDestroy(e);
        if (rc) return rc;
    } else if (ctx->cfg.bound_n_split_method == MP_BNS_CPU && !ctx->cpu_trace.empty()) {
        // per batch: the host split that filled it (host clock) and its device
        // raster (HIP events, from the frame's first event, which the stream
        // reaches right after the host clock starts); a batch's raster depends
        // on its split, the next split runs meanwhile
        std::vector<std::string> lines;
        for (size_t k = 0; k < ctx->cpu_trace.size(); ++k) {
            float a = 0.0f, b = 0.0f;
            if (hipEventElapsedTime(&a, ctx->trace_ev0, ctx->trace_ev[2 * k]) != hipSuccess ||
                hipEventElapsedTime(&b, ctx->trace_ev0, ctx->trace_ev[2 * k + 1]) != hipSuccess)
                return fail(ctx, MP_E_HIP, "trace events");
            const uint64_t h0 = base + (uint64_t)(ctx->cpu_trace[k].split0 * 1e6);
            const uint64_t h1 = base + (uint64_t)(ctx->cpu_trace[k].split1 * 1e6);
            const uint64_t d0 = base + (uint64_t)((double)a * 1e6), d1 = base + (uint64_t)((double)b * 1e6);
            fs << "bound & split@host:" << h0 << ":" << h0 << ":" << h0 << ":" << h1 << ":" << 2 * k << ":\n";
            fs << "sample@device:" << h1 << ":" << h1 << ":" << d0 << ":" << d1 << ":" << 2 * k + 1 << ":" << 2 * k
               << "\n";
        }
        return fs.good() ? MP_OK : fail(ctx, MP_E_INVALID, "write failed");
    } else {
        recs.push_back(Rec{ "frame", 0.0, ctx->stats.frame_ms });
    }
    // name@queue:queued:submit:start:end:id:dependencies (src/CL/Device.cpp:275-289), ns
    for (size_t k = 0; k < recs.size(); ++k) {
        const uint64_t t0 = base + (uint64_t)(recs[k].t0 * 1e6), t1 = base + (uint64_t)(recs[k].t1 * 1e6);
        fs << recs[k].name << "@device:" << t0 << ":" << t0 << ":" << t0 << ":" << t1 << ":" << k << ":";
        if (k > 0) fs << (k - 1);
        fs << "\n";
    }
    return fs.good() ? MP_OK : fail(ctx, MP_E_INVALID, "write failed");
}

extern "C" int mp_replay(mp_context* ctx, int32_t n, int32_t use_graph, double* ms_per_frame)
{
    if (!ctx || n < 1) return MP_E_INVALID;
    if (!ctx->frame_valid) return fail(ctx, MP_E_STATE, "no finished frame to replay");
    int rc;
    if ((rc = clear_stale_counts(ctx))) return rc;
    ctx->counts_dirty = true;   // until the replays have completed
    if (!ctx->device_driven) {
        if (use_graph) return fail(ctx, MP_E_INVALID, "graph replay needs a device-driven split method");
        HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        for (int i = 0; i < n; ++i) {
            int64_t dropped = 0, maxp = 0;
            uint32_t dt = 0;
            HIPCHK(ctx, launch_reset(ctx->d_fc, ctx->d_share, ctx->stream));
            if (ctx->cfg.bound_n_split_method == MP_BNS_CPU) {
                rc = run_cpu_frame(ctx, dt, maxp);   // split, ring, rasters and resolve
            } else {
                rc = run_segment_split(ctx, dt, dropped, maxp);
                if (!rc) rc = launch_back_end(ctx);
            }
            if (rc) return rc;
        }
    } else if (use_graph) {
        // a frame captured without the root-queue writes relies on the state
        // the previous frame left
        if (ctx->graph && !ctx->graph_inits && !frame_ready(ctx)) invalidate_graph(ctx);
        if (!ctx->graph) {
            hipGraph_t g = nullptr;
            const bool was_armed = ctx->armed;   // a capture runs nothing
            HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            rc = launch_device_frame(ctx);
            hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            ctx->armed = was_armed;
            if (rc) { if (g) hipGraphDestroy(g); return rc; }
            HIPCHK(ctx, e);
            e = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
            ctx->graph_inits = ctx->frame_inits;
            hipGraphDestroy(g);
            HIPCHK(ctx, e);
        }
        ctx->armed = false;   // until the launches below have been enqueued
        HIPCHK(ctx, hipGraphLaunch(ctx->graph, ctx->stream));  // warm-up
        HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        for (int i = 0; i < n; ++i) HIPCHK(ctx, hipGraphLaunch(ctx->graph, ctx->stream));
        ctx->armed = true;    // every replay ends with its frame-end step
    } else {
        HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        for (int i = 0; i < n; ++i)
            if ((rc = launch_device_frame(ctx))) return rc;
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    HIPCHK(ctx, hipMemcpy(ctx->h_fc, ctx->d_fc_out, sizeof(FrameCounters), hipMemcpyDeviceToHost));
    if (ctx->h_fc->overflow) return fail(ctx, MP_E_STATE, "replay overflowed the draw list");
    ctx->counts_dirty = false;
    if (ms_per_frame) *ms_per_frame = (double)ms / n;
    return MP_OK;
}

// ---------------------------------------------------------------------------
// scheduling choices and the autotuner

extern "C" int mp_get_tuning(mp_context* ctx, mp_tuning* t)
{
    if (!ctx || !t) return MP_E_INVALID;
    t->split_engine = ctx->split_engine;
    t->work_sharing = ctx->work_sharing ? 1 : 0;
    t->raster_per_cu = ctx->raster_grid_size / std::max(1, ctx->num_cus);
    t->raster_chunk = ctx->prm.raster_chunk;
    t->split_levels = ctx->level_limit;
    t->split_budget = ctx->split_budget;
    t->split_waves = ctx->split_waves;
    t->frame_ms = 0.0;
    return MP_OK;
}

// Applies a scheduling choice and, for a finished device-driven frame, runs
// it once so that the level engine knows the frame's depth.
static int apply_tuning(mp_context* ctx, const mp_tuning& t)
{
    ctx->split_engine = (t.split_engine == 1 || t.split_engine == 2) ? t.split_engine : 0;
    if (ctx->split_engine == 2 && !split_forest_fits((int)ctx->objects.size(), ctx->n_views))
        return fail(ctx, MP_E_INVALID, "the forest split takes frames of at most 64 objects and views");
    ctx->work_sharing = t.work_sharing != 0;
    ctx->raster_grid_size = ctx->num_cus * std::max(1, t.raster_per_cu);
    ctx->prm.raster_chunk = t.raster_chunk;
    ctx->level_limit = std::max(0, t.split_levels);
    ctx->split_budget = std::max(0, t.split_budget);
    ctx->split_waves = t.split_waves == 3 ? 3 : 2;
    ctx->level_grid_size = 0;
    if (ctx->split_engine == 1)
        ctx->level_launches = ctx->level_limit > 0 ? std::min(ctx->level_limit, ctx->cfg.max_split_depth)
                                                   : ctx->cfg.max_split_depth;
    invalidate_graph(ctx);
    if (!ctx->frame_valid || !ctx->device_driven) return MP_OK;
    int rc;
    if ((rc = ensure_queues(ctx))) return rc;
    for (int attempt = 0;; ++attempt) {
        if ((rc = clear_stale_counts(ctx))) return rc;
        ctx->counts_dirty = true;
        if ((rc = launch_device_frame(ctx))) return rc;
        HIPCHK(ctx, hipMemcpyAsync(ctx->h_fc, ctx->d_fc_out, sizeof(FrameCounters), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->h_fc->overflow == 32u && attempt == 0) {
            ctx->forest_rank = 0;   // the forest groups cannot rank this frame
            continue;
        }
        break;
    }
    if (ctx->h_fc->overflow) return fail(ctx, MP_E_STATE, "frame overflowed under this tuning");
    ctx->counts_dirty = false;
    if (ctx->split_engine == 1 && (rc = update_level_launches(ctx, nullptr))) return rc;
    return MP_OK;
}

extern "C" int mp_set_tuning(mp_context* ctx, const mp_tuning* t)
{
    if (!ctx || !t) return MP_E_INVALID;
    if (t->split_engine < 0 || t->split_engine > 2 || t->raster_per_cu < 0 || t->raster_per_cu > 64 ||
        t->raster_chunk < -4096 || t->raster_chunk > 4096 || t->split_levels < 0 || t->split_levels > 23 ||
        t->split_budget < 0 || t->split_budget > 4096 ||
        (t->split_waves != 0 && t->split_waves != 2 && t->split_waves != 3))
        return fail(ctx, MP_E_INVALID, "tuning out of range");
    mp_tuning u = *t;
    if (u.raster_per_cu == 0) u.raster_per_cu = raster_grid(ctx->cfg.reyes_patch_size, ctx->num_cus, ctx->prm) / ctx->num_cus;
    return apply_tuning(ctx, u);
}

// Best of two rounds of `replays` graph replays per candidate; returns the
// index of the fastest candidate that ran, or -1.
static int pick_fastest(mp_context* ctx, std::vector<mp_tuning>& cand, int replays)
{
    for (mp_tuning& c : cand) c.frame_ms = 1e30;
    for (int round = 0; round < 2; ++round) {
        for (mp_tuning& c : cand) {
            if (apply_tuning(ctx, c) != MP_OK) continue;
            double ms = 0.0;
            if (mp_replay(ctx, replays, 1, &ms) != MP_OK) continue;
            c.frame_ms = std::min(c.frame_ms, ms);
        }
    }
    int best = -1;
    for (int i = 0; i < (int)cand.size(); ++i)
        if (cand[i].frame_ms < 1e29 && (best < 0 || cand[i].frame_ms < cand[best].frame_ms)) best = i;
    return best;
}

extern "C" int mp_autotune(mp_context* ctx, int32_t replays, mp_tuning* chosen)
{
    if (!ctx || replays < 1) return MP_E_INVALID;
    if (!ctx->frame_valid || !ctx->device_driven)
        return fail(ctx, MP_E_STATE, "autotune needs a finished LOCAL or BALANCED frame or batch");
    mp_tuning cur;
    mp_get_tuning(ctx, &cur);
    // stage 1: raster schedule (graph replays include the launch costs)
    const int occ = raster_grid(ctx->cfg.reyes_patch_size, ctx->num_cus, ctx->prm) / ctx->num_cus;
    std::vector<mp_tuning> cand;
    for (int per_cu : { occ, std::max(1, occ / 2) }) {
        for (int chunk : { 0, -1, -16 }) {   // stride, dynamic, static chunks
            mp_tuning c = cur;
            c.raster_per_cu = per_cu;
            c.raster_chunk = chunk;
            cand.push_back(c);
        }
        if (occ < 2) break;
    }
    int b = pick_fastest(ctx, cand, replays);
    if (b < 0) return fail(ctx, MP_E_STATE, "no tuning candidate ran");
    cur = cand[b];
    // stage 2: split schedule
    cand.clear();
    for (int ws : { 0, 1 }) {           // persistent split
        for (int budget : { 0, 5, 8 }) {    // (a budget applies to LOCAL frames without sharing)
            if (budget && (ws || ctx->cfg.bound_n_split_method == MP_BNS_BALANCED)) continue;
            for (int waves : { 2, 3 }) {    // fat groups / more groups
                if (waves == 3 && (ws || budget == 5)) continue;
                mp_tuning c = cur;
                c.split_engine = 0;
                c.work_sharing = ws;
                c.split_levels = 0;
                c.split_budget = budget;
                c.split_waves = waves;
                cand.push_back(c);
            }
        }
    }
    {                                   // forest split
        mp_tuning c = cur;
        c.split_engine = 2;
        c.work_sharing = 0;
        c.split_levels = 0;
        c.split_budget = 0;
        c.split_waves = 2;
        cand.push_back(c);
    }
    for (int lv : { 0, 3, 5 }) {         // level launches, then the persistent tail
        mp_tuning c = cur;
        c.split_engine = 1;
        c.work_sharing = 0;
        c.split_levels = lv;
        c.split_budget = 0;
        c.split_waves = 2;
        cand.push_back(c);
    }
    b = pick_fastest(ctx, cand, replays);
    if (b < 0) return fail(ctx, MP_E_STATE, "no tuning candidate ran");
    const mp_tuning best = cand[b];
    int rc;
    if ((rc = apply_tuning(ctx, best))) return rc;
    if (chosen) *chosen = best;
    return MP_OK;
}

extern "C" int mp_measure_atomic_rate(mp_context* ctx, uint32_t n, double* atomics_per_second)
{
    if (!ctx || !atomics_per_second || n == 0 || n > (1u << 30)) return MP_E_INVALID;
    unsigned long long* keys = nullptr;
    if (hipMalloc((void**)&keys, (size_t)n * sizeof(unsigned long long)) != hipSuccess)
        return fail(ctx, MP_E_NOMEM, "atomic-rate buffer");
    double best = 0.0;
    int rc = MP_OK;
    for (int pass = 0; pass < 4 && rc == MP_OK; ++pass) {
        float ms = 0.0f;
        if (launch_fill_u64(keys, n, MP_EMPTY_KEY, ctx->stream) != hipSuccess ||
            hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess ||
            launch_atomic_rate(keys, n, ctx->stream) != hipSuccess ||
            hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess ||
            hipEventSynchronize(ctx->ev1) != hipSuccess ||
            hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) {
            rc = fail(ctx, MP_E_HIP, "atomic-rate kernel");
            break;
        }
        if (pass > 0 && ms > 0.0f) best = std::max(best, (double)n / ((double)ms * 1e-3));   // pass 0 warms up
    }
    hipFree(keys);
    if (rc == MP_OK) *atomics_per_second = best;
    return rc;
}

extern "C" int mp_replay_stages(mp_context* ctx, int32_t n, double stage_ms[4])
{
    if (!ctx || n < 1 || !stage_ms) return MP_E_INVALID;
    if (!ctx->frame_valid) return fail(ctx, MP_E_STATE, "no finished frame to replay");
    if (!ctx->device_driven) return fail(ctx, MP_E_INVALID, "stage replay needs a device-driven split method");
    std::vector<hipEvent_t> ev((size_t)n * 5, nullptr);
    int rc = clear_stale_counts(ctx);
    ctx->counts_dirty = true;
    for (auto& e : ev)
        if (rc == MP_OK && hipEventCreate(&e) != hipSuccess) { rc = fail(ctx, MP_E_HIP, "hipEventCreate failed"); break; }
    for (int i = 0; i < n && rc == MP_OK; ++i) rc = launch_device_frame(ctx, &ev[(size_t)i * 5]);
    double acc[4] = { 0, 0, 0, 0 };
    if (rc == MP_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ctx, MP_E_HIP, "sync failed");
    if (rc == MP_OK) ctx->counts_dirty = false;
    for (int i = 0; i < n && rc == MP_OK; ++i)
        for (int k = 0; k < 4; ++k) {
            float ms = 0.0f;
            hipEventElapsedTime(&ms, ev[(size_t)i * 5 + k], ev[(size_t)i * 5 + k + 1]);
            acc[k] += ms;
        }
    for (auto& e : ev)
        if (e) hipEventDestroy(e);
    if (rc) return rc;
    for (int k = 0; k < 4; ++k) stage_ms[k] = acc[k] / n;
    return MP_OK;
}

extern "C" int mp_frame_serial(mp_context* ctx, uint64_t* serial)
{
    if (!ctx || !serial) return MP_E_INVALID;
    *serial = ctx->frame_valid ? ctx->frame_serial : 0u;
    return MP_OK;
}

extern "C" int mp_backward(mp_context* ctx, uint64_t frame_serial, const float* dl_dcolor, size_t dl_floats,
                           float* grad_cps, size_t grad_cps_floats, float* grad_color, float* grad_amplitude,
                           int32_t n_objects)
{
    if (!ctx) return MP_E_INVALID;
    if (!ctx->frame_valid) return fail(ctx, MP_E_STATE, "no finished frame");
    if (frame_serial != ctx->frame_serial)
        return fail(ctx, MP_E_STATE, "the context rendered another frame (or its patches changed) since that frame");
    if (!dl_dcolor || !grad_cps || !grad_color || !grad_amplitude) return fail(ctx, MP_E_INVALID, "null pointer");
    size_t pixels = 0;
    for (int v = 0; v < ctx->n_views; ++v) pixels += (size_t)ctx->views[v].width * ctx->views[v].height;
    if (dl_floats != 4 * pixels) return fail(ctx, MP_E_INVALID, "dl_dcolor must hold 4 floats per pixel of every view");
    if (grad_cps_floats != ctx->grad_floats)
        return fail(ctx, MP_E_INVALID, "grad_cps must hold n_patches*16*3 floats per object of the frame");
    if (n_objects != (int32_t)ctx->objects.size())
        return fail(ctx, MP_E_INVALID, "n_objects must equal the draws of the frame");
    if (ctx->cfg.reyes_patch_size != 8 && ctx->cfg.reyes_patch_size != 16)
        return fail(ctx, MP_E_INVALID, "mp_backward supports reyes_patch_size 8 and 16");
    for (const ObjectReq& o : ctx->objects)
        if (ctx->patches[o.handle].type != MP_BEZIER) return fail(ctx, MP_E_INVALID, "mp_backward needs Bezier patches");
    int rc;
    if (ctx->list_cap > ctx->visible_cap) {
        // per-range visibility flags, then the list of visible ranges
        if ((rc = dev_alloc(ctx, &ctx->d_visible, 2 * (size_t)ctx->list_cap))) return rc;
        ctx->visible_cap = ctx->list_cap;
    }
    if (!ctx->d_next && (rc = dev_alloc(ctx, &ctx->d_next, 1))) return rc;
    if (!ctx->adjoint_grid_size) ctx->adjoint_grid_size = adjoint_grid(ctx->cfg.reyes_patch_size, ctx->num_cus);
    const int no = (int)ctx->objects.size();
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(ctx, launch_fill_u32(ctx->d_visible, ctx->visible_cap, 0u, ctx->stream));
    HIPCHK(ctx, launch_fill_u32(ctx->d_next, 1, 0u, ctx->stream));
    HIPCHK(ctx, launch_fill_u32(reinterpret_cast<uint32_t*>(grad_cps), ctx->grad_floats, 0u, ctx->stream));
    HIPCHK(ctx, launch_fill_u32(reinterpret_cast<uint32_t*>(grad_color), (size_t)no * 4, 0u, ctx->stream));
    HIPCHK(ctx, launch_fill_u32(reinterpret_cast<uint32_t*>(grad_amplitude), (size_t)no, 0u, ctx->stream));
    HIPCHK(ctx, launch_adjoint(ctx->cfg.reyes_patch_size, ctx->d_canon, ctx->d_objs, ctx->d_views, ctx->n_views,
                               ctx->max_pixels, ctx->prm, ctx->d_fc_out, ctx->d_quad, dl_dcolor, ctx->d_visible,
                               ctx->d_visible + ctx->visible_cap, ctx->d_next, grad_cps, grad_color,
                               grad_amplitude, ctx->adjoint_grid_size, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->stats.backward_ms = ms;
    return MP_OK;
}

extern "C" int mp_batch_read_device(mp_context* ctx, int32_t view, float* rgba32, uint32_t* ids, size_t n_pixels)
{
    if (!ctx) return MP_E_INVALID;
    if (!ctx->frame_valid) return fail(ctx, MP_E_STATE, "no finished frame");
    if (view < 0 || view >= ctx->n_views) return fail(ctx, MP_E_INVALID, "bad view index");
    const ViewBuffers& vb = ctx->views[view];
    const size_t np = (size_t)vb.width * vb.height;
    if (n_pixels != np) return fail(ctx, MP_E_INVALID, "buffer size differs from the view's width * height");
    if (rgba32) {
        if (!vb.rgba32) return fail(ctx, MP_E_STATE, "write_rgba32 is off");
        HIPCHK(ctx, hipMemcpyAsync(rgba32, vb.rgba32, np * 16, hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (ids) HIPCHK(ctx, hipMemcpyAsync(ids, vb.ids, np * 4, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MP_OK;
}

extern "C" int mp_update_patches_device(mp_context* ctx, uint64_t handle, const float* xyz, size_t n_points)
{
    if (!ctx) return MP_E_INVALID;
    auto it = ctx->patches.find(handle);
    if (it == ctx->patches.end()) return fail(ctx, MP_E_NOTFOUND, "unknown patch handle");
    const size_t per = it->second.type == MP_GREGORY ? 20 : 16;
    if (n_points != per * (size_t)it->second.n_patches) return fail(ctx, MP_E_INVALID, "point count differs from the loaded set");
    if (n_points && !xyz) return fail(ctx, MP_E_INVALID, "null patch data");
    HIPCHK(ctx, launch_xyz_to_cps(xyz, it->second.d_cps, n_points, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->frame_serial = ++ctx->serial_counter;   // the last frame no longer matches its patches
    return MP_OK;
}
