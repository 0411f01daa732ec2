// HIP kernels of the REYES hot path for gfx950 (CDNA4, wave64).
//
//  split      k_split_persistent   LOCAL / BALANCED: persistent one-wave
//                                  work-groups, LDS range stack, wave-ballot
//                                  compaction (kernels/bound_n_split_local.cl,
//                                  kernels/bound_n_split_balanced.cl)
//             k_bound_segment +    BREADTH / BOUNDED: one level or batch,
//             k_scan_blocks +      stable compaction by ballot + LDS prefix
//             k_move_segment       (kernels/bound_n_split_breadthfirst.cl,
//                                  kernels/bound_n_split_multipass.cl,
//                                  kernels/prefix_sum.cl)
//  order      k_scan_counts +      canonical (object, patch, umin, vmin) order
//             k_scatter + k_rank   of the frame's DRAW ranges
//  raster     k_raster<N>          fused dice + shade + sample of one range
//                                  per work-group (kernels/dice.cl,
//                                  kernels/reyes.cl); 64-bit depth|id atomicMin
//  resolve    k_resolve            key buffer -> RGBA8 / depth / ids, and clear
#include <hip/hip_runtime.h>

#include "mp_device.h"
#include "mp_kernels.h"

namespace mp {

// ------------------------------------------------------------------------
// small utilities

__global__ void k_reset(FrameCounters* fc)
{
    if (threadIdx.x == 0) {
        fc->draw_count = 0;
        fc->overflow = 0;
        fc->bounds = 0;
        fc->slow_blocks = 0;
        fc->max_stack = 0;
        fc->total_ranges = 0;
        fc->atomics = 0;
    }
}

__global__ void k_fill_u64(unsigned long long* p, size_t n, unsigned long long v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = v;
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = v;
}

hipError_t launch_reset(FrameCounters* fc, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, fc);
    return hipGetLastError();
}

static unsigned fill_grid(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (unsigned)b;
}

hipError_t launch_fill_u64(unsigned long long* p, size_t n, unsigned long long v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u64, dim3(fill_grid(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

MP_DEV uint32_t gpid_of(const DevObject* objs, const RangeRec& r)
{
    return objs[r.obj_depth & 0xffffffu].gpid_base + r.pid;
}

MP_DEV int bound_rec(const DevObject* __restrict__ objs, const DevView* __restrict__ views,
                     const DevParams& prm, const RangeRec& rec)
{
    const DevObject& ob = objs[rec.obj_depth & 0xffffffu];
    const int ncp = ob.type == MP_TYPE_GREGORY ? 20 : 16;
    return bound3(ob.cps + (size_t)rec.pid * ncp, ob.type, (int)(rec.obj_depth >> 24), rec.umin,
                  rec.vmin, rec.umax, rec.vmax, ob.mv, views[ob.view], prm, ob.split_limit);
}

MP_DEV uint32_t lanes_below(uint64_t mask)
{
    return (uint32_t)__popcll(mask & __lanemask_lt());
}

// ------------------------------------------------------------------------
// LOCAL / BALANCED: persistent split

__global__ void k_init_queues(const DevObject* __restrict__ objs, int n_objs, uint32_t total,
                              SplitQueues q, int balanced)
{
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < (uint32_t)q.n_groups)
        q.queue_cnt[g] = balanced ? 0 : (int32_t)((total + (uint32_t)(q.n_groups - 1) - g) / (uint32_t)q.n_groups);
    if (balanced && g == 0) *q.shared_cnt = (int32_t)total;
    if (g >= total) return;
    int lo = 0, hi = n_objs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (objs[mid].gpid_base <= g) lo = mid; else hi = mid - 1;
    }
    RangeRec r;
    r.pid = g - objs[lo].gpid_base;
    r.obj_depth = (uint32_t)lo;
    r.umin = 0.0f; r.vmin = 0.0f; r.umax = 1.0f; r.vmax = 1.0f;
    if (balanced) q.shared_q[g] = r;
    else q.queue[(int64_t)(g % (uint32_t)q.n_groups) * q.stride + g / (uint32_t)q.n_groups] = r;
}

hipError_t launch_init_queues(const DevObject* objs, int n_objs, uint32_t total_roots,
                              const SplitQueues& q, int balanced, hipStream_t s)
{
    uint32_t n = total_roots > (uint32_t)q.n_groups ? total_roots : (uint32_t)q.n_groups;
    hipLaunchKernelGGL(k_init_queues, dim3((n + 255) / 256), dim3(256), 0, s, objs, n_objs,
                       total_roots, q, balanced);
    return hipGetLastError();
}

// One wave per work-group.  Pops up to 64 ranges (private queue, then the
// shared queue, then the LDS stack), bounds them, pushes split children on
// the LDS stack (ballot rank) and appends DRAW ranges to the frame's draw
// list (one atomic per wave).  Exits when every source is empty.
__global__ __launch_bounds__(64) void k_split_persistent(const DevObject* __restrict__ objs,
                                                         const DevView* __restrict__ views,
                                                         DevParams prm, SplitQueues q,
                                                         CanonBuffers cb, FrameCounters* fc,
                                                         int balanced)
{
    extern __shared__ RangeRec stack[];
    const int lane = threadIdx.x;
    const int wid = blockIdx.x;
    RangeRec* myq = q.queue + (int64_t)wid * q.stride;
    int sh = 0;
    int qcnt = q.queue_cnt[wid];
    uint32_t nbounds = 0, maxsh = 0;

    for (;;) {
        int cnt = 0, start = 0, scnt, src = 0;
        if (sh < 64) {
            const int want = 64 - sh;
            if (qcnt > 0) {
                cnt = min(want, qcnt);
                qcnt -= cnt;
                start = qcnt;
            } else if (balanced) {
                int top = 0;
                if (lane == 0) top = atomicSub(q.shared_cnt, want);
                top = __shfl(top, 0);
                cnt = min(want, max(0, top));
                start = top - cnt;
                src = 1;
            }
            scnt = min(sh, 64 - cnt);
        } else {
            scnt = 64;
        }
        sh -= scnt;
        if (cnt == 0 && scnt == 0) break;

        RangeRec rec;
        bool occ = false;
        if (lane < cnt) {
            rec = src ? q.shared_q[start + lane] : myq[start + lane];
            occ = true;
        } else if (lane < cnt + scnt) {
            rec = stack[sh + lane - cnt];
            occ = true;
        }
        __syncthreads();
        int flags = 0;
        if (occ) flags = bound_rec(objs, views, prm, rec);
        nbounds += (uint32_t)(cnt + scnt);

        const uint64_t sb = __ballot(flags & 2);
        const int ns = __popcll(sb);
        if (sh + 2 * ns > q.stack_cap) {
            if (lane == 0) atomicOr(&fc->overflow, 2u);
            break;
        }
        if (flags & 2) {
            const uint32_t rk = lanes_below(sb);
            RangeRec c0, c1;
            split_children(rec, flags, c0, c1);
            stack[sh + 2 * rk] = c0;
            stack[sh + 2 * rk + 1] = c1;
        }
        sh += 2 * ns;
        maxsh = max(maxsh, (uint32_t)sh);

        const uint64_t db = __ballot(flags & 1);
        const int nd = __popcll(db);
        if (nd) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&fc->draw_count, (uint32_t)nd);
            base = __shfl(base, 0);
            if (flags & 1) {
                const uint32_t idx = base + lanes_below(db);
                if (idx < cb.cap) {
                    cb.draw[idx] = rec;
                    atomicAdd(&cb.counts[gpid_of(objs, rec)], 1u);
                }
            }
            if (base + (uint32_t)nd > cb.cap) {
                if (lane == 0) atomicOr(&fc->overflow, 1u);
                break;
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        atomicAdd(&fc->bounds, (unsigned long long)nbounds);
        atomicMax(&fc->max_stack, maxsh);
    }
}

hipError_t launch_split_persistent(const DevObject* objs, const DevView* views,
                                   const DevParams& prm, const SplitQueues& q,
                                   const CanonBuffers& cb, FrameCounters* fc, int balanced,
                                   hipStream_t s)
{
    size_t lds = (size_t)q.stack_cap * sizeof(RangeRec);
    hipLaunchKernelGGL(k_split_persistent, dim3(q.n_groups), dim3(64), lds, s, objs, views, prm,
                       q, cb, fc, balanced);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// BREADTH / BOUNDED: one level (batch) = bound + block scan + stable move

constexpr int SEG = 256;

uint32_t segment_blocks(uint32_t n) { return (n + SEG - 1) / SEG; }

__global__ __launch_bounds__(SEG) void k_bound_segment(const DevObject* __restrict__ objs,
                                                       const DevView* __restrict__ views,
                                                       DevParams prm, const RangeRec* in,
                                                       RangeRec* pad, uint32_t n,
                                                       uint8_t* __restrict__ flags_out,
                                                       uint32_t* __restrict__ blk_counts,
                                                       FrameCounters* fc)
{
    __shared__ uint32_t sd[SEG / 64], ss[SEG / 64];
    const uint32_t i = blockIdx.x * SEG + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int f = 0;
    if (i < n) {
        RangeRec rec = in[i];
        if (pad != in) pad[i] = rec;
        f = bound_rec(objs, views, prm, rec);
        flags_out[i] = (uint8_t)f;
    }
    const uint64_t db = __ballot(f & 1), sb = __ballot(f & 2);
    if (lane == 0) { sd[wave] = __popcll(db); ss[wave] = __popcll(sb); }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = 0, b = 0;
        for (int w = 0; w < SEG / 64; ++w) { a += sd[w]; b += ss[w]; }
        blk_counts[2 * blockIdx.x] = a;
        blk_counts[2 * blockIdx.x + 1] = b;
        uint32_t m = n - blockIdx.x * SEG;
        atomicAdd(&fc->bounds, (unsigned long long)(m < SEG ? m : SEG));
    }
}

hipError_t launch_bound_segment(const DevObject* objs, const DevView* views,
                                const DevParams& prm, const RangeRec* in, RangeRec* pad,
                                uint32_t n, uint8_t* flags, uint32_t* blk_counts,
                                FrameCounters* fc, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bound_segment, dim3(segment_blocks(n)), dim3(SEG), 0, s, objs, views,
                       prm, in, pad, n, flags, blk_counts, fc);
    return hipGetLastError();
}

// exclusive block scan of 1024 values held one per thread; returns the total
MP_DEV uint32_t block_exclusive_scan_1024(uint32_t x, uint32_t* lds_waves, uint32_t& total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(v, o);
        if (lane >= o) v += y;
    }
    if (lane == 63) lds_waves[wave] = v;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < 16 ? lds_waves[lane] : 0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            uint32_t y = __shfl_up(w, o);
            if (lane >= o) w += y;
        }
        if (lane < 16) lds_waves[16 + lane] = w;
    }
    __syncthreads();
    total = lds_waves[16 + 15];
    uint32_t wave_excl = wave ? lds_waves[16 + wave - 1] : 0;
    uint32_t r = wave_excl + v - x;
    __syncthreads();
    return r;
}

// Exclusive scan of n pairs (draw, split) in place; totals[0..1] = sums.
__global__ __launch_bounds__(1024) void k_scan_blocks(uint32_t* blk, uint32_t n, uint32_t* totals)
{
    __shared__ uint32_t lw[32];
    const uint32_t t = threadIdx.x;
    const uint32_t chunk = (n + 1023) / 1024;
    const uint32_t lo = min(n, t * chunk), hi = min(n, lo + chunk);
    for (int c = 0; c < 2; ++c) {
        uint32_t s = 0;
        for (uint32_t i = lo; i < hi; ++i) s += blk[2 * i + c];
        uint32_t tot;
        uint32_t run = block_exclusive_scan_1024(s, lw, tot);
        for (uint32_t i = lo; i < hi; ++i) {
            uint32_t v = blk[2 * i + c];
            blk[2 * i + c] = run;
            run += v;
        }
        if (t == 0) totals[c] = tot;
        __syncthreads();
    }
}

hipError_t launch_scan_blocks(uint32_t* blk_counts, uint32_t n_blocks, uint32_t* totals,
                              hipStream_t s)
{
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, blk_counts, n_blocks, totals);
    return hipGetLastError();
}

__global__ __launch_bounds__(SEG) void k_move_segment(const DevObject* __restrict__ objs,
                                                      const RangeRec* __restrict__ pad, uint32_t n,
                                                      const uint8_t* __restrict__ flags,
                                                      const uint32_t* __restrict__ blk_off,
                                                      RangeRec* __restrict__ draw_out,
                                                      uint32_t draw_base, uint32_t draw_limit,
                                                      RangeRec* __restrict__ split_out,
                                                      uint32_t* __restrict__ counts)
{
    __shared__ uint32_t sd[SEG / 64], ss[SEG / 64];
    const uint32_t i = blockIdx.x * SEG + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int f = (i < n) ? flags[i] : 0;
    const uint64_t db = __ballot(f & 1), sb = __ballot(f & 2);
    if (lane == 0) { sd[wave] = __popcll(db); ss[wave] = __popcll(sb); }
    __syncthreads();
    uint32_t pd = blk_off[2 * blockIdx.x], ps = blk_off[2 * blockIdx.x + 1];
    for (int w = 0; w < wave; ++w) { pd += sd[w]; ps += ss[w]; }
    if (i >= n) return;
    if (f & 1) {
        const uint32_t rk = pd + lanes_below(db);
        if (rk < draw_limit) {
            const RangeRec rec = pad[i];
            draw_out[draw_base + rk] = rec;
            atomicAdd(&counts[gpid_of(objs, rec)], 1u);
        }
    } else if (f & 2) {
        const uint32_t rk = ps + lanes_below(sb);
        RangeRec c0, c1;
        split_children(pad[i], f, c0, c1);
        split_out[2 * rk] = c0;
        split_out[2 * rk + 1] = c1;
    }
}

hipError_t launch_move_segment(const DevObject* objs, const RangeRec* pad, uint32_t n,
                               const uint8_t* flags, const uint32_t* blk_offsets,
                               RangeRec* draw_out, uint32_t draw_base, uint32_t draw_limit,
                               RangeRec* split_out, uint32_t* counts, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_move_segment, dim3(segment_blocks(n)), dim3(SEG), 0, s, objs, pad, n,
                       flags, blk_offsets, draw_out, draw_base, draw_limit, split_out, counts);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// canonical order: bucket by global patch index, rank inside the bucket

__global__ __launch_bounds__(1024) void k_scan_counts(CanonBuffers cb, DevView* views, int n_views,
                                                      FrameCounters* fc)
{
    __shared__ uint32_t lw[32];
    const uint32_t t = threadIdx.x;
    const uint32_t n = cb.n_gpids;
    const uint32_t chunk = (n + 1023) / 1024;
    const uint32_t lo = min(n, t * chunk), hi = min(n, lo + chunk);
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += cb.counts[i];
    uint32_t tot;
    uint32_t run = block_exclusive_scan_1024(s, lw, tot);
    for (uint32_t i = lo; i < hi; ++i) {
        uint32_t c = cb.counts[i];
        cb.offsets[i] = run;
        run += c;
        cb.counts[i] = 0;
        cb.cursor[i] = 0;
    }
    if (t == 0) {
        cb.offsets[n] = tot;
        fc->total_ranges = tot;
    }
    __threadfence_block();
    __syncthreads();
    for (int v = (int)t; v < n_views; v += 1024) {
        uint32_t b = cb.offsets[views[v].first_gpid];
        views[v].range_base = b;
        views[v].range_count = cb.offsets[views[v].gpid_end] - b;
    }
}

__global__ __launch_bounds__(256) void k_scatter(const DevObject* __restrict__ objs, CanonBuffers cb,
                                                 const FrameCounters* __restrict__ fc)
{
    const uint32_t total = fc->total_ranges;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256) {
        const RangeRec r = cb.draw[i];
        const uint32_t g = gpid_of(objs, r);
        const uint32_t slot = cb.offsets[g] + atomicAdd(&cb.cursor[g], 1u);
        cb.bucket[slot] = r;
    }
}

__global__ __launch_bounds__(256) void k_rank(const DevObject* __restrict__ objs, CanonBuffers cb,
                                              const FrameCounters* __restrict__ fc)
{
    const uint32_t total = fc->total_ranges;
    for (uint32_t j = blockIdx.x * 256 + threadIdx.x; j < total; j += gridDim.x * 256) {
        const RangeRec r = cb.bucket[j];
        const uint32_t g = gpid_of(objs, r);
        const uint32_t lo = cb.offsets[g], hi = cb.offsets[g + 1];
        const uint64_t key = canonical_key(r);
        uint32_t rank = 0;
        for (uint32_t k = lo; k < hi; ++k) rank += (canonical_key(cb.bucket[k]) < key) ? 1u : 0u;
        cb.canon[lo + rank] = r;
    }
}

hipError_t launch_canonicalize(const DevObject* objs, DevView* views, int n_views,
                               const CanonBuffers& cb, FrameCounters* fc, int num_cus,
                               hipStream_t s)
{
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, cb, views, n_views, fc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int grid = num_cus * 4;
    hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(256), 0, s, objs, cb, fc);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rank, dim3(grid), dim3(256), 0, s, objs, cb, fc);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// fused dice + shade + sample

MP_DEV int wave_min_i(int x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o));
    return x;
}
MP_DEV int wave_max_i(int x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o));
    return x;
}

MP_DEV int wave_sum_i(int x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

MP_DEV void normalize3(float& x, float& y, float& z)
{
    const float inv = 1.0f / sqrtf(dot3(x, y, z, x, y, z));
    x = x * inv; y = y * inv; z = z * inv;
}

// Flat Lambert term of one micropolygon (reyes.cl:172-191); NaN when the
// reference colour is NaN through the zero-weighted specular term.
MP_DEV float quad_lambert(const float4& p0, const float4& p1, const float4& p2, const float4& p3)
{
    float dux = ((p1.x - p0.x) + (p3.x - p2.x)) * 0.5f;
    float duy = ((p1.y - p0.y) + (p3.y - p2.y)) * 0.5f;
    float duz = ((p1.z - p0.z) + (p3.z - p2.z)) * 0.5f;
    float dvx = ((p2.x - p0.x) + (p3.x - p1.x)) * 0.5f;
    float dvy = ((p2.y - p0.y) + (p3.y - p1.y)) * 0.5f;
    float dvz = ((p2.z - p0.z) + (p3.z - p1.z)) * 0.5f;
    // n = cross(dv, du)
    float nx = dvy * duz - dvz * duy;
    float ny = dvz * dux - dvx * duz;
    float nzz = dvx * duy - dvy * dux;
    normalize3(nx, ny, nzz);
    float lx = 4.0f, ly = 3.0f, lz = 8.0f;
    normalize3(lx, ly, lz);
    float sx = ((p0.x + p1.x) + p2.x) + p3.x;
    float sy = ((p0.y + p1.y) + p2.y) + p3.y;
    float sz = ((p0.z + p1.z) + p2.z) + p3.z;
    normalize3(sx, sy, sz);
    float hx = lx + (-sx), hy = ly + (-sy), hz = lz + (-sz);
    normalize3(hx, hy, hz);
    const float spec = cl_max(dot3(nx, ny, nzz, hx, hy, hz), 0.0f);
    const float s = cl_max(dot3(nx, ny, nzz, lx, ly, lz), 0.0f);
    return (spec != spec) ? spec : s;
}

template <int N>
__global__ __launch_bounds__(N * N) void k_raster(const RangeRec* __restrict__ canon,
                                                  const DevObject* __restrict__ objs,
                                                  const DevView* __restrict__ views,
                                                  DevParams prm, FrameCounters* fc,
                                                  float* __restrict__ quad_store)
{
    constexpr int G1 = N + 1;
    constexpr int G = G1 * G1;
    constexpr int Q = N * N;
    constexpr int B = N / 8;
    __shared__ float4 s_cp[20];
    __shared__ float4 s_pos[G];
    __shared__ int2 s_pxl[G];
    __shared__ float s_dep[G];

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int u = (wave % B) * 8 + (lane & 7);
    const int v = (wave / B) * 8 + (lane >> 3);
    const int g0 = u + v * G1;
    const int sh = prm.subpixel_bits;
    const uint32_t total = fc->total_ranges;
    uint32_t slow = 0, natom = 0;

    for (uint32_t r = blockIdx.x; r < total; r += gridDim.x) {
        const RangeRec rec = canon[r];
        const DevObject& ob = objs[rec.obj_depth & 0xffffffu];
        const DevView& vw = views[ob.view];
        const int ncp = ob.type == MP_TYPE_GREGORY ? 20 : 16;
        if (tid < ncp) s_cp[tid] = ob.cps[(size_t)rec.pid * ncp + tid];
        __syncthreads();

        // ---- dice (dice.cl:38-79) ----
        for (int g = tid; g < G; g += Q) {
            const int nu = g % G1, nv = g / G1;
            const float tu = (float)nu / (float)N, tv = (float)nv / (float)N;
            const float uu = rec.umin + (rec.umax - rec.umin) * tu;
            const float vv = rec.vmin + (rec.vmax - rec.vmin) * tv;
            float P[4], e[4], q[4];
            displaced_point(s_cp, ob.type, uu, vv, prm, ob.disp_amplitude, P);
            mat4_mul(ob.mv, P, e);
            mat4_mul(vw.proj, e, q);
            s_pos[g] = make_float4(e[0], e[1], e[2], e[3]);
            const int32_t px = f2i_sat((q[0] / q[3]) * (float)vw.vs_x / 2.0f + (float)(vw.vs_x / 2));
            const int32_t py = f2i_sat((q[1] / q[3]) * (float)vw.vs_y / 2.0f + (float)(vw.vs_y / 2));
            s_pxl[g] = make_int2(px, py);
            s_dep[g] = q[2] / q[3];
        }
        __syncthreads();

        // ---- shade (reyes.cl:84-199); one wave = one 8x8 block ----
        const float4 p0 = s_pos[g0], p1 = s_pos[g0 + 1], p2 = s_pos[g0 + G1], p3 = s_pos[g0 + G1 + 1];
        const int2 x0 = s_pxl[g0], x1 = s_pxl[g0 + 1], x2 = s_pxl[g0 + G1], x3 = s_pxl[g0 + G1 + 1];
        const float d0 = s_dep[g0], d1 = s_dep[g0 + 1], d2 = s_dep[g0 + G1], d3 = s_dep[g0 + G1 + 1];
        const bool zero = (p0.z == 0.0f) || (p1.z == 0.0f) || (p2.z == 0.0f) || (p3.z == 0.0f);
        const bool allnormal = !__any(zero);
        const int32_t vmaxx = (vw.width << sh) - 1, vmaxy = (vw.height << sh) - 1;
        bool ff = true;
        if (prm.backface_culling) {
            const int32_t d1x = wsub(x1.x, x0.x), d1y = wsub(x1.y, x0.y);
            const int32_t d2x = wsub(x2.x, x0.x), d2y = wsub(x2.y, x0.y);
            const int32_t d3x = wsub(x3.x, x0.x), d3y = wsub(x3.y, x0.y);
            ff = (wsub(wmul(d1x, d3y), wmul(d3x, d1y)) < 0) || (wsub(wmul(d3x, d2y), wmul(d2x, d3y)) < 0);
        }
        const int32_t qminx = min(min(x0.x, x1.x), min(x2.x, x3.x));
        const int32_t qminy = min(min(x0.y, x1.y), min(x2.y, x3.y));
        const int32_t qmaxx = max(max(x0.x, x1.x), max(x2.x, x3.x));
        const int32_t qmaxy = max(max(x0.y, x1.y), max(x2.y, x3.y));
        int32_t xmin = wave_min_i(ff ? min(vmaxx, qminx) : vmaxx);
        int32_t ymin = wave_min_i(ff ? min(vmaxy, qminy) : vmaxy);
        int32_t xmax = wave_max_i(ff ? max(0, qmaxx) : 0);
        int32_t ymax = wave_max_i(ff ? max(0, qmaxy) : 0);
        xmin = max(0, xmin); ymin = max(0, ymin);
        xmax = min(vmaxx, xmax); ymax = min(vmaxy, ymax);
        if (!allnormal) { xmin = 1; ymin = 1; xmax = -1; ymax = -1; }

        quad_store[(size_t)r * Q + (u + v * N)] = quad_lambert(p0, p1, p2, p3);

        // ---- sample (reyes.cl:270-398) ----
        const int32_t mintx = xmin >> (sh + 3), minty = ymin >> (sh + 3);
        const int32_t maxtx = xmax >> (sh + 3), maxty = ymax >> (sh + 3);
        if (mintx <= maxtx && minty <= maxty) {
            const uint32_t id = (r - vw.range_base) * (uint32_t)Q + (uint32_t)(u + v * N);
            unsigned long long* keys = vw.keys;
            const int gw = vw.grid_w;
            // pixel-candidate bounds with the reference's initialisers
            const int32_t mingx = min(vmaxx + 1, qminx), mingy = min(vmaxy + 1, qminy);
            const int32_t maxgx = max(-1, qmaxx), maxgy = max(-1, qmaxy);
            // Exact-integer regime: all coordinates within +-2^18 and edges
            // shorter than 2^11 sub-pixels keep every edge product inside
            // int32, so quads that are not front-facing and pixels outside a
            // quad's bounding box can never pass the edge test.
            const int32_t lim = 1 << 18;
            const bool small = (qminx > -lim) && (qmaxx < lim) && (qminy > -lim) && (qmaxy < lim) &&
                               (qmaxx - qminx < (1 << 11)) && (qmaxy - qminy < (1 << 11));
            if (__all(small)) {
                if (ff) {
                    const int32_t X0 = max(mingx >> sh, mintx * 8), X1 = min(maxgx >> sh, maxtx * 8 + 7);
                    const int32_t Y0 = max(mingy >> sh, minty * 8), Y1 = min(maxgy >> sh, maxty * 8 + 7);
                    for (int32_t Y = Y0; Y <= Y1; ++Y) {
                        for (int32_t X = X0; X <= X1; ++X) {
                            const int32_t tx = X << sh, ty = Y << sh;
                            float dd = 1.0f;
                            const bool in1 = inside_tri(x0.x, x0.y, x1.x, x1.y, x3.x, x3.y, d0, d1, d3, tx, ty, dd);
                            const bool in2 = inside_tri(x0.x, x0.y, x3.x, x3.y, x2.x, x2.y, d0, d3, d2, tx, ty, dd);
                            if (in1 || in2) {
                                const unsigned long long key = ((unsigned long long)mono_bits(dd) << 32) | id;
                                atomicMin(&keys[fb_index(X, Y, gw)], key);
                                ++natom;
                            }
                        }
                    }
                }
            } else {
                // literal per-tile loops of the reference (wide coordinates)
                if (lane == 0) ++slow;
                const int32_t maxl = (8 << sh) - 1;
                for (int32_t ty = minty; ty <= maxty; ++ty) {
                    for (int32_t tx = mintx; tx <= maxtx; ++tx) {
                        const int32_t osx = (tx * 8) << sh, osy = (ty * 8) << sh;
                        const int32_t minpx = min(max(wsub(mingx, osx), 0), maxl) >> sh;
                        const int32_t minpy = min(max(wsub(mingy, osy), 0), maxl) >> sh;
                        const int32_t maxpx = min(max(wsub(maxgx, osx), 0), maxl) >> sh;
                        const int32_t maxpy = min(max(wsub(maxgy, osy), 0), maxl) >> sh;
                        for (int32_t y = minpy; y <= maxpy; ++y) {
                            for (int32_t x = minpx; x <= maxpx; ++x) {
                                const int32_t tpx = (x << sh) + osx, tpy = (y << sh) + osy;
                                float dd = 1.0f;
                                const bool in1 = inside_tri(x0.x, x0.y, x1.x, x1.y, x3.x, x3.y, d0, d1, d3, tpx, tpy, dd);
                                const bool in2 = inside_tri(x0.x, x0.y, x3.x, x3.y, x2.x, x2.y, d0, d3, d2, tpx, tpy, dd);
                                if (in1 || in2) {
                                    const unsigned long long key = ((unsigned long long)mono_bits(dd) << 32) | id;
                                    atomicMin(&keys[fb_index(tx * 8 + x, ty * 8 + y, gw)], key);
                                    ++natom;
                                }
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    if (slow) atomicAdd(&fc->slow_blocks, slow);
    natom = (uint32_t)wave_sum_i((int)natom);
    if (lane == 0 && natom) atomicAdd(&fc->atomics, (unsigned long long)natom);
}

template <int N>
static int raster_occupancy()
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_raster<N>, N * N, 0) != hipSuccess)
        blocks = 1;
    return blocks < 1 ? 1 : blocks;
}

int raster_grid(int patch_size, int num_cus)
{
    int per_cu = 1;
    switch (patch_size) {
    case 8: per_cu = raster_occupancy<8>(); break;
    case 16: per_cu = raster_occupancy<16>(); break;
    case 32: per_cu = raster_occupancy<32>(); break;
    default: break;
    }
    return per_cu * num_cus;
}

hipError_t launch_raster(int patch_size, const RangeRec* canon, const DevObject* objs,
                         const DevView* views, const DevParams& prm, FrameCounters* fc,
                         float* quad_store, int grid, hipStream_t s)
{
    switch (patch_size) {
    case 8:
        hipLaunchKernelGGL(k_raster<8>, dim3(grid), dim3(64), 0, s, canon, objs, views, prm, fc, quad_store);
        break;
    case 16:
        hipLaunchKernelGGL(k_raster<16>, dim3(grid), dim3(256), 0, s, canon, objs, views, prm, fc, quad_store);
        break;
    case 32:
        hipLaunchKernelGGL(k_raster<32>, dim3(grid), dim3(1024), 0, s, canon, objs, views, prm, fc, quad_store);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// resolve: key -> display colour, FP32 depth, id; clears the key

MP_DEV uint8_t to_srgb8(float c)
{
    if (!(c > 0.0f)) return 0;
    const float g = powf(c, 0.454545f);
    if (!(g < 1.0f)) return 255;
    return (uint8_t)(g * 255.0f + 0.5f);
}

__global__ __launch_bounds__(256) void k_resolve(const DevView* __restrict__ views,
                                                 const RangeRec* __restrict__ canon,
                                                 const DevObject* __restrict__ objs,
                                                 const float* __restrict__ quad_store, int Q,
                                                 int write_rgba32)
{
    const DevView& vw = views[blockIdx.y];
    const uint32_t n = (uint32_t)vw.grid_w * (uint32_t)vw.grid_h * 64u;
    const float ac[4] = { 0.015f, 0.015f, 0.015f, 1.0f };
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const unsigned long long key = vw.keys[i];
        vw.keys[i] = MP_EMPTY_KEY;
        const uint32_t t = i >> 6, l = i & 63;
        const int X = (int)(t % (uint32_t)vw.grid_w) * 8 + (int)(l & 7);
        const int Y = (int)(t / (uint32_t)vw.grid_w) * 8 + (int)(l >> 3);
        if (X >= vw.width || Y >= vw.height) continue;
        const size_t o = (size_t)Y * vw.width + X;
        float c[4];
        float d = 1.0f;
        uint32_t id = 0xFFFFFFFFu;
        if (key == MP_EMPTY_KEY) {
            c[0] = vw.clear_linear[0]; c[1] = vw.clear_linear[1];
            c[2] = vw.clear_linear[2]; c[3] = vw.clear_linear[3];
        } else {
            id = (uint32_t)key;
            d = unmono_bits((uint32_t)(key >> 32));
            const uint32_t rr = vw.range_base + id / (uint32_t)Q;
            const float s = quad_store[(size_t)vw.range_base * Q + id];
            const DevObject& ob = objs[canon[rr].obj_depth & 0xffffffu];
#pragma unroll
            for (int k = 0; k < 4; ++k) c[k] = (ac[k] * ob.color[k] + s * ob.color[k]) + 0.0f;
        }
        uchar4 px;
        px.x = to_srgb8(c[0]); px.y = to_srgb8(c[1]); px.z = to_srgb8(c[2]); px.w = 255;
        reinterpret_cast<uchar4*>(vw.rgba8)[o] = px;
        vw.depth[o] = d;
        vw.ids[o] = id;
        if (write_rgba32) reinterpret_cast<float4*>(vw.rgba32)[o] = make_float4(c[0], c[1], c[2], c[3]);
    }
}

hipError_t launch_resolve(const DevView* views, int n_views, int max_pixels,
                          const RangeRec* canon, const DevObject* objs, const float* quad_store,
                          int patch_size, int write_rgba32, hipStream_t s)
{
    int bx = (max_pixels + 255) / 256;
    if (bx > 2048) bx = 2048;
    if (bx < 1) bx = 1;
    hipLaunchKernelGGL(k_resolve, dim3(bx, n_views), dim3(256), 0, s, views, canon, objs,
                       quad_store, patch_size * patch_size, write_rgba32);
    return hipGetLastError();
}

}  // namespace mp

namespace mp {

__global__ void k_init_roots(int obj, uint32_t n, RangeRec* out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    RangeRec r;
    r.pid = i;
    r.obj_depth = (uint32_t)obj;
    r.umin = 0.0f; r.vmin = 0.0f; r.umax = 1.0f; r.vmax = 1.0f;
    out[i] = r;
}

hipError_t launch_init_roots(int obj, uint32_t n, RangeRec* out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_init_roots, dim3((n + 255) / 256), dim3(256), 0, s, obj, n, out);
    return hipGetLastError();
}

int split_persistent_occupancy(size_t lds_bytes)
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_split_persistent, 64, lds_bytes) != hipSuccess)
        blocks = 1;
    return blocks < 1 ? 1 : blocks;
}

}  // namespace mp
