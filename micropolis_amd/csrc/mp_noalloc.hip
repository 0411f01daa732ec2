// Allocation-free frame: mp_render_noalloc (include/micropolis.h).
//
// The same device-driven frame as mp_finish for LOCAL / BALANCED (persistent
// split, canonical order, fused dice/shade/sample, resolve) for one object,
// with every intermediate buffer -- the object / view tables, the control
// points, split queues, per-patch counts, range lists, Lambert store and
// visibility keys -- carved from one file-scope device array instead of
// hipMalloc'd context buffers.  For callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid allocation); they pass device-resident inputs and outputs.
//
// The array is kScratchBytes per device, held until the process ends.  Every
// call on one device shares it and leaves its contents behind; each call
// clears what must start at zero (per-patch counts, bucket cursors, frame
// counters) and fills the visibility keys before use.  Calls that overlap on
// one device (two host threads, or two streams) race on it.
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/micropolis.h"
#include "mp_host.h"
#include "mp_kernels.h"
#include "mp_types.h"

namespace {

constexpr size_t kScratchBytes = 64ull << 20;   // 64 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up once per
// device, with that device current).  The array's name is only a host-side
// stand-in: kernels are given this address.
hipError_t scratch_address(int device, unsigned char** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
    return hipSuccess;
}

// Byte offsets of the carved regions, each 256-byte aligned.
struct Carve {
    size_t off = 0;
    size_t take(size_t bytes)
    {
        const size_t o = off;
        off = (off + bytes + 255u) & ~(size_t)255u;
        return o;
    }
};

struct Tables {
    mp::DevObject obj;
    mp::DevView view;
    float tab[256];
};

// The frame's object / view records and display table, passed by value and
// written into the scratch array (no host-to-device copy).
__global__ void k_write_tables(Tables t, mp::DevObject* obj, mp::DevView* view, float* tab)
{
    if (threadIdx.x == 0) {
        *obj = t.obj;
        *view = t.view;
    }
    tab[threadIdx.x] = t.tab[threadIdx.x];
}

int fail_noalloc(int code, const std::string& msg)
{
    mp::set_create_error(msg);
    return code;
}

}  // namespace

using namespace mp;

#define NA_HIP(expr)                                                                     \
    do {                                                                                 \
        const hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                            \
            return fail_noalloc(MP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" int mp_render_noalloc(const mp_config* cfg, const float* xyz, size_t n_points, int32_t patch_type,
                                 const float mv[16], const mp_projection* projection, const float color[4],
                                 uint8_t* rgba8, float* depth, uint32_t* ids, void* stream)
{
    if (!cfg || !mv || !projection || !color) return fail_noalloc(MP_E_INVALID, "null argument");
    std::string why;
    if (!validate_config(cfg, why)) return fail_noalloc(MP_E_INVALID, why);
    const int method = cfg->bound_n_split_method;
    if (method != MP_BNS_LOCAL && method != MP_BNS_BALANCED)
        return fail_noalloc(MP_E_INVALID, "mp_render_noalloc renders LOCAL and BALANCED frames only");
    if (patch_type != MP_BEZIER && patch_type != MP_GREGORY) return fail_noalloc(MP_E_INVALID, "unknown patch type");
    const size_t per = patch_type == MP_GREGORY ? 20 : 16;
    if (n_points % per != 0) return fail_noalloc(MP_E_INVALID, "point count is not a multiple of the patch size");
    if (n_points && !xyz) return fail_noalloc(MP_E_INVALID, "null patch data");
    if (n_points / per >= (1u << 20)) return fail_noalloc(MP_E_INVALID, "at most 2^20 - 1 patches");
    if (patch_type == MP_GREGORY && cfg->displacement)
        return fail_noalloc(MP_E_INVALID, "displacement is defined for Bezier patches only");
    if (projection->width != cfg->window_width || projection->height != cfg->window_height)
        return fail_noalloc(MP_E_INVALID, "projection viewport must equal the window size");
    if (!rgba8 || !depth || !ids) return fail_noalloc(MP_E_INVALID, "null output buffer");
    hipStream_t s = static_cast<hipStream_t>(stream);

    int device = 0, num_cus = 1;
    NA_HIP(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) return fail_noalloc(MP_E_INVALID, "device index out of range");
    NA_HIP(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, device));
    unsigned char* base = nullptr;
    NA_HIP(scratch_address(device, &base));

    // ---- sizes and the carve of the scratch array ----
    const uint32_t P = (uint32_t)(n_points / per);
    const int N = cfg->reyes_patch_size;
    const int balanced = method == MP_BNS_BALANCED;
    const int stack_cap = 64 * (cfg->max_split_depth + 2);
    int groups = cfg->local_bns_work_groups;
    if (groups <= 0) groups = num_cus * split_persistent_occupancy(split_persistent_lds(stack_cap), 2);
    const size_t stride = balanced ? 1 : (P + (uint32_t)groups - 1u) / (uint32_t)groups;
    const int grid_w = (projection->width + 7) / 8, grid_h = (projection->height + 7) / 8;
    const size_t n_keys = (size_t)grid_w * grid_h * 64;

    Carve c;
    const size_t o_obj = c.take(sizeof(DevObject)), o_view = c.take(sizeof(DevView));
    const size_t o_fc = c.take(kFrameCountersBytes), o_share = c.take(sizeof(ShareState));
    const size_t o_tab = c.take(256 * sizeof(float));
    const size_t o_cps = c.take(n_points * sizeof(float4));
    const size_t o_queue = c.take(std::max<size_t>(1, stride * (size_t)groups) * sizeof(RangeRec));
    const size_t o_qcnt = c.take((size_t)groups * sizeof(int32_t));
    const size_t o_shq = c.take((balanced ? std::max<uint32_t>(P, 1u) : 1u) * sizeof(RangeRec));
    const size_t o_shc = c.take(sizeof(int32_t));
    const size_t o_counts = c.take(((size_t)P + 1) * 4), o_offsets = c.take(((size_t)P + 1) * 4);
    const size_t o_cursor = c.take(((size_t)P + 1) * 4), o_part = c.take(((size_t)P / 4096 + 2) * 4);
    const size_t o_keys = c.take(n_keys * 8);
    // range lists: draw, bucket and canonical records, in-patch keys, spans,
    // Lambert terms; the rest of the array (less the eight regions' alignment)
    // (+ the radix-sort lists of patches with more than 4096 ranges: 2 keys and
    // 2 slot indices per range)
    const size_t per_range = 3 * sizeof(RangeRec) + 8 + 8 + 4 * (size_t)N * N + 2 * 8 + 2 * 4;
    const size_t fixed = c.off + 8 * 256;
    if (fixed >= kScratchBytes || (kScratchBytes - fixed) / per_range < std::max<size_t>(P, 256))
        return fail_noalloc(MP_E_INVALID,
                            "frame too large for the 64 MiB scratch array (" + std::to_string(P) + " patches, " +
                            std::to_string(projection->width) + "x" + std::to_string(projection->height) +
                            " window)");
    const uint32_t cap = (uint32_t)std::min<size_t>((kScratchBytes - fixed) / per_range, 0x40000000u);
    const size_t o_draw = c.take((size_t)cap * sizeof(RangeRec));
    const size_t o_bucket = c.take((size_t)cap * sizeof(RangeRec));
    const size_t o_canon = c.take((size_t)cap * sizeof(RangeRec));
    const size_t o_rkeys = c.take((size_t)cap * 8);
    const size_t o_span = c.take((size_t)cap * 8);
    const size_t o_quad = c.take((size_t)cap * N * N * sizeof(float));
    const size_t o_skeys = c.take((size_t)cap * 2 * 8);
    const size_t o_svals = c.take((size_t)cap * 2 * 4);
    if (c.off > kScratchBytes) return fail_noalloc(MP_E_INVALID, "scratch carve exceeds the array");

    DevObject* d_obj = reinterpret_cast<DevObject*>(base + o_obj);
    DevView* d_view = reinterpret_cast<DevView*>(base + o_view);
    FrameCounters* d_fc = reinterpret_cast<FrameCounters*>(base + o_fc);
    ShareState* d_share = reinterpret_cast<ShareState*>(base + o_share);
    float* d_tab = reinterpret_cast<float*>(base + o_tab);
    float4* d_cps = reinterpret_cast<float4*>(base + o_cps);

    // ---- tables (host-side values, written by a kernel) ----
    static float s_tab[256];
    static std::once_flag s_tab_once;
    std::call_once(s_tab_once, [] { display_thresholds(s_tab); });
    Tables t;
    memset(&t, 0, sizeof(t));
    DevObject& ob = t.obj;
    ob.cps = d_cps;
    memcpy(ob.mv, mv, sizeof(ob.mv));
    memcpy(ob.color, color, sizeof(ob.color));
    ob.split_limit = cfg->bound_n_split_limit;
    ob.disp_amplitude = cfg->displacement_amplitude;
    ob.type = patch_type;
    ob.n_patches = (int32_t)P;
    ob.view = 0;
    ob.gpid_base = 0;
    ob.grad_offset = 0;
    ob.draw_index = 0;
    DevView& vw = t.view;
    make_view(*projection, cfg->clear_color, cfg->subpixel_bits, vw);
    vw.keys = reinterpret_cast<unsigned long long*>(base + o_keys);
    vw.rgba8 = rgba8;
    vw.depth = depth;
    vw.ids = ids;
    vw.rgba32 = nullptr;
    vw.first_gpid = 0;
    vw.gpid_end = P;
    vw.pixel_offset = 0;
    vw.solo_obj = 0;
    memcpy(t.tab, s_tab, sizeof(t.tab));
    const DevParams prm = device_params(cfg);

    SplitQueues q;
    memset(&q, 0, sizeof(q));
    q.queue = reinterpret_cast<RangeRec*>(base + o_queue);
    q.queue_cnt = reinterpret_cast<int32_t*>(base + o_qcnt);
    q.stride = balanced ? 0 : (int64_t)stride;
    q.shared_q = reinterpret_cast<RangeRec*>(base + o_shq);
    q.shared_cnt = reinterpret_cast<int32_t*>(base + o_shc);
    q.n_groups = groups;
    q.stack_cap = stack_cap;
    q.chunk_cap = 0;            // no work sharing: no chunk pool
    q.share = d_share;
    q.spread = 2;
    q.waves = 2;

    CanonBuffers cb;
    cb.draw = reinterpret_cast<RangeRec*>(base + o_draw);
    cb.bucket = reinterpret_cast<RangeRec*>(base + o_bucket);
    cb.canon = reinterpret_cast<RangeRec*>(base + o_canon);
    cb.keys = reinterpret_cast<unsigned long long*>(base + o_rkeys);
    cb.span = reinterpret_cast<uint2*>(base + o_span);
    cb.cap = cap;
    cb.counts = reinterpret_cast<uint32_t*>(base + o_counts);
    cb.offsets = reinterpret_cast<uint32_t*>(base + o_offsets);
    cb.cursor = reinterpret_cast<uint32_t*>(base + o_cursor);
    cb.part = reinterpret_cast<uint32_t*>(base + o_part);
    cb.n_gpids = P;
    cb.sort_keys = reinterpret_cast<unsigned long long*>(base + o_skeys);
    cb.sort_vals = reinterpret_cast<uint32_t*>(base + o_svals);
    float* d_quad = reinterpret_cast<float*>(base + o_quad);

    // ---- the frame ----
    hipLaunchKernelGGL(k_write_tables, dim3(1), dim3(256), 0, s, t, d_obj, d_view, d_tab);
    NA_HIP(hipGetLastError());
    if (n_points) NA_HIP(launch_xyz_to_cps(xyz, d_cps, n_points, s));
    // what must start at zero / empty: a previous call leaves its contents
    NA_HIP(launch_fill_u32(cb.counts, (size_t)P + 1, 0u, s));
    NA_HIP(launch_fill_u32(cb.cursor, (size_t)P + 1, 0u, s));
    NA_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(d_share), sizeof(ShareState) / 4, 0u, s));
    NA_HIP(launch_fill_u64(vw.keys, n_keys, MP_EMPTY_KEY, s));
    if (P) {
        NA_HIP(launch_init_queues(d_obj, 1, P, q, balanced, d_fc, d_share, s));
        NA_HIP(launch_split_persistent(d_obj, d_view, prm, q, cb, d_fc, balanced, s));
    } else {
        NA_HIP(launch_reset(d_fc, d_share, s));
    }
    NA_HIP(launch_canonicalize(d_obj, d_view, 1, cb, d_fc, num_cus, 1, s));
    NA_HIP(launch_raster(N, cb.canon, d_obj, d_view, prm, d_fc, d_quad, raster_grid(N, num_cus, prm), s));
    NA_HIP(launch_resolve(d_view, 1, (int)n_keys, cb.canon, d_obj, d_quad, N, 0, d_tab, s));
    FrameCounters h_fc;
    NA_HIP(hipMemcpyAsync(&h_fc, d_fc, sizeof(h_fc), hipMemcpyDeviceToHost, s));
    NA_HIP(hipStreamSynchronize(s));
    if (h_fc.overflow & 1u)
        return fail_noalloc(MP_E_INVALID, "frame has more DRAW ranges than the scratch array holds (" +
                                              std::to_string(cap) + "); output incomplete");
    if (h_fc.overflow)
        return fail_noalloc(MP_E_STATE, "split or ordering overflow (flags " + std::to_string(h_fc.overflow) + ")");
    if (h_fc.total_ranges != h_fc.draw_count)
        return fail_noalloc(MP_E_STATE, "range counts disagree with the draw list");
    return MP_OK;
}
