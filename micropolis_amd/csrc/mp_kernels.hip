// HIP kernels of the REYES hot path for gfx950 (CDNA4, wave64).
//
//  split      k_split_persistent   LOCAL / BALANCED: persistent work-groups
//                                  (controller wave + sample waves), LDS
//                                  range stack, wave-ballot
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
#include <algorithm>
#include <type_traits>

#include "mp_device.h"
#include "mp_kernels.h"

namespace mp {

// ------------------------------------------------------------------------
// small utilities

// Per-patch count scans done by every work-group of a launch into LDS
// (k_scan_scatter, k_place_ranked): frames of at most this many patches.
constexpr uint32_t kFusedScan = 4096;

// Word w of the stripe of work-group key (kStripes stripes after the counters)
MP_DEV unsigned long long* stripe(FrameCounters* fc, uint32_t key, int w)
{
    return reinterpret_cast<unsigned long long*>(fc + 1) + (size_t)(key % (uint32_t)kStripes) * kStripeWords + w;
}

// One whole wave: the stripes added into the counters and cleared.
MP_DEV void fold_stripes(FrameCounters* fc)
{
    const int lane = threadIdx.x & 63;
    static_assert(kStripes == 64, "one stripe per lane");
    unsigned long long* st = stripe(fc, (uint32_t)lane, 0);
    unsigned long long a = st[kStripeAtomics], b = st[kStripeBounds], m = st[kStripeMaxStack];
    st[kStripeAtomics] = 0ull;
    st[kStripeBounds] = 0ull;
    st[kStripeMaxStack] = 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
        const unsigned long long mo = __shfl_xor(m, o);
        m = mo > m ? mo : m;
    }
    if (lane == 0) {
        fc->atomics += a;
        fc->bounds += b;
        fc->max_stack = max(fc->max_stack, (uint32_t)m);
    }
}

// Start-of-frame counters (one thread).
MP_DEV void reset_counters(FrameCounters* fc, ShareState* share)
{
    for (int k = 0; k < kStripes; ++k)
        for (int w = 0; w < kStripeUsed; ++w) *stripe(fc, (uint32_t)k, w) = 0ull;
    if (share) {
        share->head = 0;
        share->tail = 0;
        share->active = 0;
        share->epoch = share->epoch + 1u;
    }
    fc->draw_count = 0;
    fc->overflow = 0;
    fc->bounds = 0;
    fc->slow_blocks = 0;
    fc->max_stack = 0;
    fc->total_ranges = 0;
    fc->big_ranges = 0;
    fc->atomics = 0;
    fc->raster_next = 0;
}

__global__ void k_reset(FrameCounters* fc, ShareState* share)
{
    if (threadIdx.x == 0) reset_counters(fc, share);
}

// One thread, after fold_stripes: FrameEnd's copy and resets (the frame's
// kernels before it have finished with the counters).
MP_DEV void end_frame(const FrameEnd& fe)
{
    if (!fe.fc) return;
    for (uint32_t i = 0; i < fe.n_counts && !fe.resolve_clears; ++i) fe.counts[i] = 0u;
    *fe.out = *fe.fc;
    reset_counters(fe.fc, fe.share);
    if (fe.left_cnt) *fe.left_cnt = 0u;
    if (fe.shared_cnt) *fe.shared_cnt = fe.shared_total;
}

__global__ void k_frame_end(FrameEnd fe)
{
    if (!fe.fc) return;
    fold_stripes(fe.fc);
    if (threadIdx.x == 0) end_frame(fe);
}

hipError_t launch_frame_end(const FrameEnd& fe, hipStream_t s)
{
    hipLaunchKernelGGL(k_frame_end, dim3(1), dim3(64), 0, s, fe);
    return hipGetLastError();
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

hipError_t launch_reset(FrameCounters* fc, ShareState* share, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, fc, share);
    return hipGetLastError();
}

static unsigned fill_grid(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (unsigned)b;
}

// 64-bit key_min over n distinct keys, one per thread (a measurement kernel)
__global__ __launch_bounds__(256) void k_atomic_rate(unsigned long long* keys, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) key_min(keys + i, ((unsigned long long)(i * 2654435761u) << 32) | i);
}

hipError_t launch_atomic_rate(unsigned long long* keys, uint32_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_atomic_rate, dim3((n + 255u) / 256u), dim3(256), 0, s, keys, n);
    return hipGetLastError();
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
                              SplitQueues q, int balanced, FrameCounters* reset_fc, ShareState* reset_share)
{
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0 && reset_fc) reset_counters(reset_fc, reset_share);
    if (g < (uint32_t)q.n_groups)
        q.queue_cnt[g] = balanced ? 0 : (int32_t)((total + (uint32_t)(q.n_groups - 1) - g) / (uint32_t)q.n_groups);
    if (balanced && g == 0) *q.shared_cnt = (int32_t)total;
    if (g == 0 && q.left_cnt) *q.left_cnt = 0u;
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
                              const SplitQueues& q, int balanced, FrameCounters* reset_fc,
                              ShareState* reset_share, hipStream_t s)
{
    uint32_t n = total_roots > (uint32_t)q.n_groups ? total_roots : (uint32_t)q.n_groups;
    hipLaunchKernelGGL(k_init_queues, dim3((n + 255) / 256), dim3(256), 0, s, objs, n_objs,
                       total_roots, q, balanced, reset_fc, reset_share);
    return hipGetLastError();
}

// Per-wave aggregated atomicAdd(&counts[key], 1) over the lanes with pred.
MP_DEV void wave_count_add(uint32_t* counts, uint32_t key, bool pred)
{
    const int lane = threadIdx.x & 63;
    uint64_t m = __ballot(pred);
    while (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        const uint32_t lk = __shfl(key, leader);
        const uint64_t same = __ballot(pred && key == lk) & m;
        if (lane == leader) atomicAdd(&counts[lk], (uint32_t)__popcll(same));
        m &= ~same;
    }
}

// Orders one wave's LDS accesses across its lanes (a wave's LDS operations
// complete in issue order; this keeps the compiler from moving them).
MP_DEV void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Persistent split work-group: one controller wave (queues, stack, draws)
// and kSplitWaves - 1 sample waves; wave w evaluates column iu = w of
// bound()'s 3x3 samples.  A step of at most kNarrowStep ranges gives each
// range three lanes per wave (one sample per lane: one patch evaluation of
// latency); a wider step gives each range one lane per wave (three samples
// per lane) instead of nine samples on one lane.
constexpr int kSplitWaves = 3;
constexpr int kNarrowStep = 21;

// Two levels per step on the narrowest steps (at most kSpecStep ranges): each
// range gets 21 lanes over the three waves for its 9 samples plus the 12
// samples its children add in either split direction (a child's other 6
// samples are the parent's: the ranges are dyadic, k / 2^d with d <= 21 for
// any range that is evaluated, so u = umin + (umax - umin) * t is exact for
// t = 0, 1/2, 1 and a child's edge samples repeat the parent's bit for bit).
// Five lanes per range then take the parent's and the four possible
// children's decisions at once; the children of the direction the parent
// chose continue the step as if they had been popped.
constexpr int kSpecStep = 9;
constexpr int kSpecSamples = 21;

// Range and grid position of speculative sample s (0..20) of parent p:
// 0..8 the parent's (iu, iv); 9..11 / 12..14 the middle column of the u-split
// children; 15..17 / 18..20 the middle row of the v-split children.
MP_DEV void spec_sample(const RangeRec& p, int s, RangeRec& c, int& iu, int& iv)
{
    c = p;
    if (s < 9) { iu = s / 3; iv = s - 3 * iu; return; }
    const float cu = (p.umin + p.umax) * 0.5f, cv = (p.vmin + p.vmax) * 0.5f;   // as split_children
    c.obj_depth = p.obj_depth + (1u << 24);
    if (s < 12)      { c.umax = cu; iu = 1; iv = s - 9; }
    else if (s < 15) { c.umin = cu; iu = 1; iv = s - 12; }
    else if (s < 18) { c.vmax = cv; iu = s - 15; iv = 1; }
    else             { c.vmin = cv; iu = s - 18; iv = 1; }
}

// Speculative sample index of sample (iu, iv) of decision d: 0 the parent,
// 1 / 2 the u-split children (umax = cu / umin = cu), 3 / 4 the v-split ones.
MP_DEV int spec_index(int d, int iu, int iv)
{
    switch (d) {
    case 0: return iu * 3 + iv;
    case 1: return iu == 0 ? iv : (iu == 1 ? 9 + iv : 3 + iv);
    case 2: return iu == 0 ? 3 + iv : (iu == 1 ? 12 + iv : 6 + iv);
    case 3: return iv == 0 ? iu * 3 : (iv == 1 ? 15 + iu : iu * 3 + 1);
    default: return iv == 0 ? iu * 3 + 1 : (iv == 1 ? 18 + iu : iu * 3 + 2);
    }
}

MP_DEV void spec_range(const RangeRec& p, int d, RangeRec& c)
{
    c = p;
    if (d == 0) return;
    const float cu = (p.umin + p.umax) * 0.5f, cv = (p.vmin + p.vmax) * 0.5f;
    c.obj_depth = p.obj_depth + (1u << 24);
    if (d == 1) c.umax = cu;
    else if (d == 2) c.umin = cu;
    else if (d == 3) c.vmax = cv;
    else c.vmin = cv;
}

// Calls f(object, view) for the ranges a wave holds (valid lanes): with a
// wave-uniform object index when every valid lane has the same object, so
// the tables are read by scalar loads, else per lane.
template <typename F>
MP_DEV void with_object(const DevObject* __restrict__ objs, const DevView* __restrict__ views,
                        uint32_t obj, bool valid, F&& f)
{
    const uint32_t o0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)obj);
    if (wave_all(!valid || obj == o0)) {
        const DevObject& ob = objs[o0];
        f(ob, views[ob.view]);
    } else {
        const DevObject& ob = objs[obj];
        f(ob, views[ob.view]);
    }
}

// Sample (iu, iv) of bound()'s samples of rec: eye-space x, y, z, screen x, y.
MP_DEV void bound_one(const DevObject& ob, const DevView& P, const RangeRec& rec, int iu, int iv,
                      float (&v)[5])
{
    float e[3];
    if (ob.type == MP_TYPE_GREGORY) {
        const float4* cp = ob.cps + (size_t)rec.pid * 20;
        float4 C[20];
#pragma unroll
        for (int k = 0; k < 20; ++k) C[k] = ldg4(cp + k);
        bound_sample<20>(C, iu, iv, rec.umin, rec.vmin, rec.umax, rec.vmax, ob.mv, P, e, v[3], v[4]);
    } else {
        const float4* cp = ob.cps + (size_t)rec.pid * 16;
        float4 C[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) C[k] = ldg4(cp + k);
        bound_sample<16>(C, iu, iv, rec.umin, rec.vmin, rec.umax, rec.vmax, ob.mv, P, e, v[3], v[4]);
    }
    v[0] = e[0]; v[1] = e[1]; v[2] = e[2];
}

// Column iu (samples iu*3 + 0..2) of bound()'s samples of rec.
MP_DEV void bound_column(const DevObject& ob, const DevView& P, const RangeRec& rec, int iu,
                         float (&v)[15])
{
    float e[3];
    if (ob.type == MP_TYPE_GREGORY) {
        const float4* cp = ob.cps + (size_t)rec.pid * 20;
        float4 C[20];
#pragma unroll
        for (int k = 0; k < 20; ++k) C[k] = ldg4(cp + k);
#pragma unroll
        for (int iv = 0; iv < 3; ++iv) {
            bound_sample<20>(C, iu, iv, rec.umin, rec.vmin, rec.umax, rec.vmax, ob.mv, P, e, v[9 + iv], v[12 + iv]);
            v[iv] = e[0]; v[3 + iv] = e[1]; v[6 + iv] = e[2];
        }
    } else {
        const float4* cp = ob.cps + (size_t)rec.pid * 16;
        float4 C[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) C[k] = ldg4(cp + k);
#pragma unroll
        for (int iv = 0; iv < 3; ++iv) {
            bound_sample<16>(C, iu, iv, rec.umin, rec.vmin, rec.umax, rec.vmax, ob.mv, P, e, v[9 + iv], v[12 + iv]);
            v[iv] = e[0]; v[3 + iv] = e[1]; v[6 + iv] = e[2];
        }
    }
}

// DRAW ranges a split work-group buffers in LDS before appending them to the
// frame's draw list: the list's counter is one address that every
// work-group's returning atomic serialises on, so a group appends once per
// kDrawBuf draws (usually once, at its end) instead of once per step.
constexpr int kDrawBuf = 32;

// Writes the buffered draws (lanes < dfill) and this step's draws (lanes
// with is_draw, at slot among nd) to the draw list under one counter atomic
// and counts them per patch.  Returns false when the list overflows.
MP_DEV bool flush_draws(const DevObject* __restrict__ objs, const CanonBuffers& cb,
                        FrameCounters* fc, const RangeRec* dbuf, int dfill,
                        const RangeRec& rec, bool is_draw, uint32_t slot, int nd)
{
    const int lane = threadIdx.x & 63;
    wave_sync();   // the buffer's LDS writes by other lanes
    uint32_t dbase = 0;
    if (lane == 0) dbase = atomicAdd(&fc->draw_count, (uint32_t)(dfill + nd));
    dbase = __shfl(dbase, 0);
    {
        const bool have = lane < dfill;
        RangeRec b;
        if (have) b = dbuf[lane];
        const uint32_t idx = dbase + (uint32_t)lane;
        const bool put = have && idx < cb.cap;
        if (put) cb.draw[idx] = b;
        wave_count_add(cb.counts, put ? gpid_of(objs, b) : 0u, put);
    }
    if (nd) {
        const uint32_t idx = dbase + (uint32_t)dfill + slot;
        const bool put = is_draw && idx < cb.cap;
        if (put) cb.draw[idx] = rec;
        wave_count_add(cb.counts, put ? gpid_of(objs, rec) : 0u, put);
    }
    if (dbase + (uint32_t)(dfill + nd) > cb.cap) {
        if (lane == 0) atomicOr(&fc->overflow, 1u);
        return false;
    }
    return true;
}

constexpr uint32_t kFlagSpin = 1u << 20;     // bound on waiting for a reserved chunk
constexpr uint32_t kLingerPolls = 1u << 14;  // bound on waiting for spilled work

// Lane 0 only: reserve the next published chunk; -1 if none.  The wait for
// the publishing store is bounded: the publisher is a running wave that
// writes the chunk right after reserving it.
MP_DEV int try_pop_chunk(const SplitQueues& q, uint32_t epoch, FrameCounters* fc)
{
    ShareState* st = q.share;
    for (;;) {
        const uint32_t h = __hip_atomic_load(&st->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t t = __hip_atomic_load(&st->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        t = min(t, q.chunk_cap);
        if (h >= t) return -1;
        if (atomicCAS(&st->head, h, h + 1) == h) {
            for (uint32_t spin = 0;; ++spin) {
                if (__hip_atomic_load(&q.chunk_ready[h], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == epoch)
                    return (int)h;
                if (spin > kFlagSpin) {
                    atomicOr(&fc->overflow, 4u);
                    return -1;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
}

// Following kernels/bound_n_split_local.cl, the controller wave pops up to
// 64 ranges (private queue, then the shared BALANCED queue, then spilled
// chunks, then the LDS stack), bounds them (with the sample waves, or with
// 3 or 9 of its own lanes per range on narrow steps), pushes split children
// on the LDS ring stack (ballot rank) and appends DRAW ranges to the frame's
// draw list (buffered in LDS, one atomic per kDrawBuf draws).  A stack that
// grows past spill_at publishes its 64 oldest (shallowest) ranges as a chunk
// that idle work-groups pick up.  A work-group with no work keeps polling
// for chunks while others are active (bounded); every chunk is popped
// because its publisher drains the chunk queue before it exits.  The sample
// waves wait at the work-group barrier between the controller's requests;
// a negative request count releases them.
// WPE: waves per SIMD the kernel is built for.  2 keeps every register (the
// fewest, fattest groups: latency-bound frames); 3 caps the registers (some
// spills) for one more group per CU (frames with plenty of ranges).
template <int WPE>
__global__ __launch_bounds__(64 * kSplitWaves) __attribute__((amdgpu_waves_per_eu(WPE))) void k_split_persistent(const DevObject* __restrict__ objs,
                                                                       const DevView* __restrict__ views,
                                                                       DevParams prm, SplitQueues q,
                                                                       CanonBuffers cb, FrameCounters* fc,
                                                                       int balanced)
{
    extern __shared__ RangeRec ring[];   // stack_cap stack entries, then kDrawBuf draws
    __shared__ RangeRec s_stage[64];                  // ranges of a sampled step
    __shared__ float s_xch[kSplitWaves - 1][15][64];  // sample waves' columns
    __shared__ int s_cmd;                             // ranges in the step; < 0: exit
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wid = blockIdx.x;
    if (q.pending && *q.pending == 0u) return;   // nothing left by the level launches (whole group)
    if (wave > 0) {
        for (;;) {
            __syncthreads();   // request posted
            const int req = s_cmd;
            if (req < 0) return;
            const int cmd = req & 0xffff, mode = req >> 16;
            if (mode == 2) {
                float* s_spec = &s_xch[0][0][0];
                const int L = wave * 64 + lane, g = L / kSpecSamples, sidx = L - kSpecSamples * g;
                const bool valid = g < cmd;
                const RangeRec pr = s_stage[min(g, 63)];
                RangeRec c;
                int iu, iv;
                spec_sample(pr, sidx, c, iu, iv);
                const bool live = valid && (int)(c.obj_depth >> 24) < prm.max_split_depth - 1;
                float v[5] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
                with_object(objs, views, pr.obj_depth & 0xffffffu, valid,
                            [&](const DevObject& ob, const DevView& P) {
                                if (live) bound_one(ob, P, c, iu, iv, v);
                            });
                if (valid)
#pragma unroll
                    for (int t = 0; t < 5; ++t) s_spec[(g * kSpecSamples + sidx) * 5 + t] = v[t];
            } else if (mode == 1) {
                const int g = lane / 3;
                const RangeRec rec = s_stage[min(g, 63)];
                float v[5] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
                const bool live = g < cmd && (int)(rec.obj_depth >> 24) < prm.max_split_depth - 1;
                with_object(objs, views, rec.obj_depth & 0xffffffu, g < cmd,
                            [&](const DevObject& ob, const DevView& P) {
                                if (live) bound_one(ob, P, rec, wave, lane - 3 * g, v);
                            });
#pragma unroll
                for (int t = 0; t < 5; ++t) s_xch[wave - 1][t][lane] = v[t];
            } else {
                const RangeRec rec = s_stage[lane];
                float v[15];
#pragma unroll
                for (int t = 0; t < 15; ++t) v[t] = 0.0f;
                const bool live = lane < cmd && (int)(rec.obj_depth >> 24) < prm.max_split_depth - 1;
                with_object(objs, views, rec.obj_depth & 0xffffffu, lane < cmd,
                            [&](const DevObject& ob, const DevView& P) {
                                if (live) bound_column(ob, P, rec, wave, v);
                            });
#pragma unroll
                for (int t = 0; t < 15; ++t) s_xch[wave - 1][t][lane] = v[t];
            }
            __syncthreads();   // samples written
        }
    }
    const int cap = q.stack_cap;
    RangeRec* dbuf = ring + cap;
    int dfill = 0;
    bool ok = true;
    RangeRec* myq = q.queue + (int64_t)wid * q.stride;
    int base = 0, sh = 0;
    int qcnt;
    if (q.strided) {
        const uint32_t nl = min(*q.pending, q.left_cap), G = (uint32_t)q.n_groups;
        qcnt = nl > (uint32_t)wid ? (int)((nl - (uint32_t)wid + G - 1u) / G) : 0;
    } else {
        qcnt = q.queue_cnt[wid];
    }
    uint32_t nbounds = 0, maxsh = 0;
    const uint32_t epoch = q.share ? q.share->epoch : 0u;
    const bool sharing = q.chunk_cap > 0;
    if (sharing && lane == 0) atomicAdd(&q.share->active, 1u);
    bool active = true;
    int steps = 0;

    for (;;) {
        if (q.step_budget > 0 && steps == q.step_budget) {
            // budget spent: hand the queued and stacked ranges to the next launch
            const int total = qcnt + sh;
            uint32_t lb = 0;
            if (lane == 0 && total) lb = atomicAdd(q.left_cnt, (uint32_t)total);
            lb = __shfl(lb, 0);
            for (int i = lane; i < total; i += 64) {
                RangeRec r;
                if (i < qcnt) {
                    r = myq[i];
                } else {
                    int k = base + i - qcnt;
                    if (k >= cap) k -= cap;
                    r = ring[k];
                }
                if (lb + (uint32_t)i < q.left_cap) q.left[lb + (uint32_t)i] = r;
            }
            if (lane == 0 && lb + (uint32_t)total > q.left_cap) atomicOr(&fc->overflow, 8u);
            break;
        }
        ++steps;
        int cnt = 0, start = 0, scnt = 0, src = 0;
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
            if (cnt == 0 && sh == 0 && sharing) {
                int c = -1;
                if (lane == 0) c = try_pop_chunk(q, epoch, fc);
                c = __shfl(c, 0);
                if (c >= 0) {
                    cnt = (int)q.chunk_count[c];
                    start = c * 64;
                    src = 2;
                }
            }
            scnt = min(sh, 64 - cnt);
        } else {
            scnt = 64;
        }
        sh -= scnt;
        if (cnt == 0 && scnt == 0) {
            // only the first `linger` groups wait for spilled work: a bounded
            // set of slow pollers cannot flood the shared state lines
            if (!sharing || wid >= q.linger) {
                if (sharing && active && lane == 0) atomicSub(&q.share->active, 1u);
                active = false;
                break;
            }
            // idle: wait (bounded) for chunks while other work-groups hold work
            if (active && lane == 0) atomicSub(&q.share->active, 1u);
            active = false;
            int c = -1;
            if (lane == 0) {
                for (uint32_t it = 0; it < kLingerPolls; ++it) {
                    const uint32_t h = __hip_atomic_load(&q.share->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const uint32_t t = min(__hip_atomic_load(&q.share->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), q.chunk_cap);
                    if (h < t) {
                        atomicAdd(&q.share->active, 1u);
                        c = try_pop_chunk(q, epoch, fc);
                        if (c >= 0) break;
                        atomicSub(&q.share->active, 1u);
                    } else if (__hip_atomic_load(&q.share->active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
                        break;
                    }
                    __builtin_amdgcn_s_sleep(127);
                }
            }
            c = __shfl(c, 0);
            if (c < 0) break;
            active = true;
            cnt = (int)q.chunk_count[c];
            start = c * 64;
            src = 2;
            scnt = 0;
        }

        RangeRec rec;
        rec.pid = 0; rec.obj_depth = 0; rec.umin = rec.vmin = rec.umax = rec.vmax = 0.0f;
        bool occ = false;
        if (lane < cnt) {
            rec = (src == 2)  ? q.chunks[start + lane]
                : (src == 1)  ? q.shared_q[start + lane]
                : q.strided   ? q.strided[(size_t)(start + lane) * (uint32_t)q.n_groups + (uint32_t)wid]
                              : myq[start + lane];
            occ = true;
        } else if (lane < cnt + scnt) {
            int k = base + sh + lane - cnt;
            rec = ring[k >= cap ? k - cap : k];
            occ = true;
        }
        wave_sync();   // stack reads before this step's pushes
        int flags = 0;
        const int n = cnt + scnt;
        if (q.spread >= 2 && n <= kSpecStep) {
            // narrowest step: the ranges and, speculatively, their children
            float* s_spec = &s_xch[0][0][0];
            s_stage[lane] = rec;
            if (lane == 0) s_cmd = n | (2 << 16);
            __syncthreads();   // request posted
            {
                const int g = lane / kSpecSamples, sidx = lane - kSpecSamples * g;
                const bool valid = g < n;
                const RangeRec pr = s_stage[g];
                RangeRec c;
                int iu, iv;
                spec_sample(pr, sidx, c, iu, iv);
                const bool live = valid && (int)(c.obj_depth >> 24) < prm.max_split_depth - 1;
                float v[5] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
                with_object(objs, views, pr.obj_depth & 0xffffffu, valid,
                            [&](const DevObject& ob, const DevView& P) {
                                if (live) bound_one(ob, P, c, iu, iv, v);
                            });
                if (valid)
#pragma unroll
                    for (int t = 0; t < 5; ++t) s_spec[(g * kSpecSamples + sidx) * 5 + t] = v[t];
            }
            __syncthreads();   // samples written
            // decision d of range g on lane 5g + d
            const int g = lane / 5, d = lane - 5 * g;
            const bool valid = g < n;
            const RangeRec pr = s_stage[min(g, 63)];
            RangeRec cr;
            spec_range(pr, d, cr);
            const bool live = valid && (int)(cr.obj_depth >> 24) < prm.max_split_depth - 1;
            float ex[9], ey[9], ez[9], sx[9], sy[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const int o = valid ? (g * kSpecSamples + spec_index(d, k / 3, k % 3)) * 5 : 0;
                ex[k] = s_spec[o]; ey[k] = s_spec[o + 1]; ez[k] = s_spec[o + 2];
                sx[k] = s_spec[o + 3]; sy[k] = s_spec[o + 4];
            }
            int f = 0;
            with_object(objs, views, pr.obj_depth & 0xffffffu, valid,
                        [&](const DevObject& ob, const DevView& P) {
                            if (live)
                                f = bound_decide(ex, ey, ez, sx, sy, cr.umin, cr.vmin, cr.umax, cr.vmax, P,
                                                 prm, ob.split_limit);
                        });
            // the parent's decision picks which lanes continue the step
            const int pf = __shfl(f, min(5 * g, 63));
            const bool psplit = (pf & 2) != 0, udir = (pf & 4) != 0;
            int out = 0;
            if (valid) {
                if (d == 0) out = pf == MP_DRAW ? MP_DRAW : 0;
                else if (d <= 2) out = (psplit && udir) ? f : 0;
                else out = (psplit && !udir) ? f : 0;
            }
            rec = cr;
            flags = out;
            occ = out != 0;
            // the children of a split parent are bounded in this step too
            nbounds += 2u * (uint32_t)__popcll(__ballot(valid && d == 0 && psplit));
        } else if (q.spread && n <= kNarrowStep) {
            // narrow step (the deep tail of a split tree): nine lanes per
            // range over the three waves, one sample each
            s_stage[lane] = rec;
            if (lane == 0) s_cmd = n | (1 << 16);
            __syncthreads();   // request posted
            const int g = lane / 3, iv = lane - 3 * g;
            const RangeRec rg = s_stage[min(g, 63)];
            const bool valid = g < n;
            const bool live = valid && (int)(rg.obj_depth >> 24) < prm.max_split_depth - 1;
            float v[5] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
            with_object(objs, views, rg.obj_depth & 0xffffffu, valid,
                        [&](const DevObject& ob, const DevView& P) {
                            if (live) bound_one(ob, P, rg, 0, iv, v);
                        });
            __syncthreads();   // samples written
            // sample (iu, iv) of range g: wave iu, lane 3g + iv
            float ex[9], ey[9], ez[9], sx[9], sy[9];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int src = min(3 * g + j, 63);
                ex[j] = __shfl(v[0], src); ey[j] = __shfl(v[1], src); ez[j] = __shfl(v[2], src);
                sx[j] = __shfl(v[3], src); sy[j] = __shfl(v[4], src);
#pragma unroll
                for (int w = 1; w < kSplitWaves; ++w) {
                    const int k = w * 3 + j;
                    ex[k] = s_xch[w - 1][0][src]; ey[k] = s_xch[w - 1][1][src];
                    ez[k] = s_xch[w - 1][2][src]; sx[k] = s_xch[w - 1][3][src];
                    sy[k] = s_xch[w - 1][4][src];
                }
            }
            rec = rg;
            occ = valid && iv == 0;
            with_object(objs, views, rec.obj_depth & 0xffffffu, occ,
                        [&](const DevObject& ob, const DevView& P) {
                            if (occ && live)
                                flags = bound_decide(ex, ey, ez, sx, sy, rec.umin, rec.vmin, rec.umax,
                                                     rec.vmax, P, prm, ob.split_limit);
                        });
        } else {
            // one sample column per wave
            s_stage[lane] = rec;
            if (lane == 0) s_cmd = n;
            __syncthreads();   // request posted
            const bool live = occ && (int)(rec.obj_depth >> 24) < prm.max_split_depth - 1;
            float c0[15];
#pragma unroll
            for (int t = 0; t < 15; ++t) c0[t] = 0.0f;
            with_object(objs, views, rec.obj_depth & 0xffffffu, occ,
                        [&](const DevObject& ob, const DevView& P) {
                            if (live) bound_column(ob, P, rec, 0, c0);
                        });
            __syncthreads();   // samples written
            float ex[9], ey[9], ez[9], sx[9], sy[9];
#pragma unroll
            for (int iv = 0; iv < 3; ++iv) {
                ex[iv] = c0[iv]; ey[iv] = c0[3 + iv]; ez[iv] = c0[6 + iv];
                sx[iv] = c0[9 + iv]; sy[iv] = c0[12 + iv];
            }
#pragma unroll
            for (int w = 1; w < kSplitWaves; ++w)
#pragma unroll
                for (int iv = 0; iv < 3; ++iv) {
                    const int k = w * 3 + iv;
                    ex[k] = s_xch[w - 1][iv][lane];
                    ey[k] = s_xch[w - 1][3 + iv][lane];
                    ez[k] = s_xch[w - 1][6 + iv][lane];
                    sx[k] = s_xch[w - 1][9 + iv][lane];
                    sy[k] = s_xch[w - 1][12 + iv][lane];
                }
            with_object(objs, views, rec.obj_depth & 0xffffffu, occ,
                        [&](const DevObject& ob, const DevView& P) {
                            if (live)
                                flags = bound_decide(ex, ey, ez, sx, sy, rec.umin, rec.vmin, rec.umax,
                                                     rec.vmax, P, prm, ob.split_limit);
                        });
        }
        nbounds += (uint32_t)n;

        const uint64_t sb = __ballot(flags & 2);
        const int ns = __popcll(sb);
        if (sh + 2 * ns > cap) {
            if (lane == 0) atomicOr(&fc->overflow, 2u);
            break;
        }
        if (flags & 2) {
            const uint32_t rk = lanes_below(sb);
            RangeRec c0, c1;
            split_children(rec, flags, c0, c1);
            int k0 = base + sh + 2 * (int)rk;
            if (k0 >= cap) k0 -= cap;
            int k1 = k0 + 1;
            if (k1 >= cap) k1 -= cap;
            ring[k0] = c0;
            ring[k1] = c1;
        }
        sh += 2 * ns;
        maxsh = max(maxsh, (uint32_t)sh);

        const uint64_t db = __ballot(flags & 1);
        const int nd = __popcll(db);
        if (nd) {
            const uint32_t slot = lanes_below(db);
            if (dfill + nd <= kDrawBuf) {
                if (flags & 1) dbuf[dfill + (int)slot] = rec;
                dfill += nd;
            } else {
                ok = flush_draws(objs, cb, fc, dbuf, dfill, rec, (flags & 1) != 0, slot, nd);
                dfill = 0;
                if (!ok) break;
            }
        }
        wave_sync();

        // publish the 64 oldest stack entries while the stack is deep
        while (sharing && sh > q.spill_at) {
            uint32_t idx = 0;
            if (lane == 0) idx = atomicAdd(&q.share->tail, 1u);
            idx = __shfl(idx, 0);
            if (idx >= q.chunk_cap) break;
            int k = base + lane;
            if (k >= cap) k -= cap;
            q.chunks[(size_t)idx * 64 + lane] = ring[k];
            if (lane == 0) q.chunk_count[idx] = 64u;
            __threadfence();
            wave_sync();
            if (lane == 0) __hip_atomic_store(&q.chunk_ready[idx], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            base += 64;
            if (base >= cap) base -= cap;
            sh -= 64;
        }
    }
    if (ok && dfill > 0) {
        RangeRec none;
        flush_draws(objs, cb, fc, dbuf, dfill, none, false, 0u, 0);
    }
    if (lane == 0) s_cmd = -1;
    __syncthreads();   // release the sample waves
    if (lane == 0) {
        if (sharing && active) atomicSub(&q.share->active, 1u);
        atomicAdd(stripe(fc, (uint32_t)wid, kStripeBounds), (unsigned long long)nbounds);
        atomicMax(stripe(fc, (uint32_t)wid, kStripeMaxStack), (unsigned long long)maxsh);
    }
}

hipError_t launch_split_persistent(const DevObject* objs, const DevView* views,
                                   const DevParams& prm, const SplitQueues& q,
                                   const CanonBuffers& cb, FrameCounters* fc, int balanced,
                                   hipStream_t s)
{
    size_t lds = split_persistent_lds(q.stack_cap);
    if (q.waves == 3)
        hipLaunchKernelGGL(k_split_persistent<3>, dim3(q.n_groups), dim3(64 * kSplitWaves), lds, s, objs, views,
                           prm, q, cb, fc, balanced);
    else
        hipLaunchKernelGGL(k_split_persistent<2>, dim3(q.n_groups), dim3(64 * kSplitWaves), lds, s, objs, views, prm,
                       q, cb, fc, balanced);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// LOCAL / BALANCED: forest split
//
// A work-group owns the whole split trees of its roots (root g of the frame
// belongs to group g mod G, the private queue layout of k_init_queues) and
// works on them with all NT lanes: a step pops up to kForestStep ranges from
// the group's LDS stack (roots are taken from the private queue while the
// stack holds fewer), spreads bound()'s nine samples of every popped range
// over the lanes (on the narrowest steps the 21 samples of the range and of
// its possible children, deciding two levels at once as the persistent
// kernel's spec_sample steps do), decides each range on one lane from the
// samples in LDS and pushes the children of split ranges back with one
// work-group prefix count.  No range leaves its group, so a step touches
// global memory only for control points and the buffered DRAW appends, and a
// group runs about one step per level of its trees.  Every lane runs the same
// sample code and the same decision code in every step (one copy of each in
// the kernel: the code of the whole kernel stays inside the instruction
// cache), with the object and view tables read per lane; PT selects the
// patch types compiled in (0 Bezier, 1 Gregory, 2 both).
constexpr int kForestThreads = 768;

// spec_sample and spec_range with every field chosen by selects (a record
// whose fields are stored by a runtime index goes through scratch memory)
MP_DEV void spec_sample_sel(const RangeRec& p, int j, RangeRec& c, int& iu, int& iv)
{
    const float cu = (p.umin + p.umax) * 0.5f, cv = (p.vmin + p.vmax) * 0.5f;   // as split_children
    c.pid = p.pid;
    c.obj_depth = p.obj_depth + (j >= 9 ? (1u << 24) : 0u);
    c.umin = (j >= 12 && j < 15) ? cu : p.umin;
    c.umax = (j >= 9 && j < 12) ? cu : p.umax;
    c.vmin = j >= 18 ? cv : p.vmin;
    c.vmax = (j >= 15 && j < 18) ? cv : p.vmax;
    const int r = j < 9 ? j / 3 : (j < 15 ? 1 : (j - 15) % 3);
    iv = j < 9 ? j - 3 * r : (j < 15 ? (j - 9) % 3 : 1);
    iu = r;
}

MP_DEV void spec_range_sel(const RangeRec& p, int d, RangeRec& c)
{
    const float cu = (p.umin + p.umax) * 0.5f, cv = (p.vmin + p.vmax) * 0.5f;
    c.pid = p.pid;
    c.obj_depth = p.obj_depth + (d != 0 ? (1u << 24) : 0u);
    c.umin = d == 2 ? cu : p.umin;
    c.umax = d == 1 ? cu : p.umax;
    c.vmin = d == 4 ? cv : p.vmin;
    c.vmax = d == 3 ? cv : p.vmax;
}

// The object and view fields the split reads, staged in LDS once per
// work-group (frames of at most kForestTables objects and views).
constexpr int kForestTables = 64;
struct ForestObj {
    float mv[16];
    const float4* cps;
    float split_limit;
    int32_t type;
    int32_t view;
    uint32_t gpid_base;
};
struct ForestView {
    float proj[16];
    float screen[4];
    float f[2];
    float near_, far_, eps;
    int32_t width, height;
};
// Control points of a group's roots staged in LDS: root slot k (the group's
// k-th root, global patch index wid + k * G) at k * stride float4.
constexpr int kForestCps = 1024;   // float4
// Ranked frames: roots per group (more: the frame runs unranked); the DRAW
// ranges a group can rank are kForestKeys (its LDS key store and output region).
constexpr int kForestRankRoots = 64;
constexpr int kRankBins = 4096;   // columns of the column-sorted ranking

// Exclusive scan of in[0 .. n) into out[0 .. n) (LDS) by the NT threads of
// a work-group, each taking a contiguous chunk; s_wsum holds NT / 64 words.
// Ends with a barrier.
template <int NT>
MP_DEV void lds_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* s_wsum)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t per = (n + NT - 1u) / NT;
    const uint32_t b0 = min(tid * per, n), b1 = min(b0 + per, n);
    uint32_t sum = 0;
    for (uint32_t i = b0; i < b1; ++i) sum += in[i];
    uint32_t v = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(v, o);
        if ((int)lane >= o) v += y;
    }
    if (lane == 63) s_wsum[wave] = v;
    __syncthreads();
    uint32_t run = v - sum;
    for (uint32_t w = 0; w < wave; ++w) run += s_wsum[w];
    for (uint32_t i = b0; i < b1; ++i) {
        const uint32_t x = in[i];
        out[i] = run;
        run += x;
    }
    __syncthreads();
}

// bound() sample (iu, iv) of rec: eye-space x, y, z and screen x, y; the
// range's control points from the group's LDS copy when its root has a slot
template <int PT, typename O, typename V>
MP_DEV void forest_sample(const O& ob, const V& P, const float4* s_cps, int slots, int stride, const RangeRec& rec,
                          int iu, int iv, float (&v)[5])
{
    const uint32_t k = (ob.gpid_base + (uint32_t)rec.pid - (uint32_t)blockIdx.x) / gridDim.x;
    const bool cached = k < (uint32_t)slots;
    float e[3];
    if (PT == 1 || (PT == 2 && ob.type == MP_TYPE_GREGORY)) {
        const float4* cp = ob.cps + (size_t)rec.pid * 20;
        float4 C[20];
        if (cached) {
#pragma unroll
            for (int i = 0; i < 20; ++i) C[i] = s_cps[k * stride + i];
        } else {
#pragma unroll
            for (int i = 0; i < 20; ++i) C[i] = ldg4(cp + i);
        }
        bound_sample<20>(C, iu, iv, rec.umin, rec.vmin, rec.umax, rec.vmax, ob.mv, P, e, v[3], v[4]);
    } else {
        const float4* cp = ob.cps + (size_t)rec.pid * 16;
        float4 C[16];
        if (cached) {
#pragma unroll
            for (int i = 0; i < 16; ++i) C[i] = s_cps[k * stride + i];
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i) C[i] = ldg4(cp + i);
        }
        bound_sample<16>(C, iu, iv, rec.umin, rec.vmin, rec.umax, rec.vmax, ob.mv, P, e, v[3], v[4]);
    }
    v[0] = e[0]; v[1] = e[1]; v[2] = e[2];
}

// Flushes the group's buffered DRAW ranges (buf[0 .. cnt), cnt <= NT) to the
// frame's draw list under one counter atomic and counts them per patch; all
// NT threads call it.  Returns false (uniformly) when the list overflows.
template <int NT>
MP_DEV bool forest_flush(const DevObject* __restrict__ objs, const CanonBuffers& cb, FrameCounters* fc,
                         const RangeRec* buf, int cnt, uint32_t* s_base)
{
    const int tid = threadIdx.x;
    if (tid == 0) *s_base = atomicAdd(&fc->draw_count, (uint32_t)cnt);
    __syncthreads();
    const uint32_t base = *s_base;
    const bool have = tid < cnt;
    const RangeRec r = buf[min(tid, max(cnt - 1, 0))];
    const uint32_t idx = base + (uint32_t)tid;
    const bool put = have && idx < cb.cap;
    if (put) cb.draw[idx] = r;
    wave_count_add(cb.counts, put ? gpid_of(objs, r) : 0u, put);
    __syncthreads();   // buffer and base read before they are reused
    if (base + (uint32_t)cnt > cb.cap) {
        if (tid == 0) atomicOr(&fc->overflow, 1u);
        return false;
    }
    return true;
}

// SO: a frame of one object (and view): its tables are read from the kernel
// arguments' arrays at index 0 (wave-uniform: scalar registers, no per-step
// loads) instead of the LDS copies indexed per range.
template <int NT, int PT, bool SO>
__global__ __launch_bounds__(NT) void k_split_forest(const DevObject* __restrict__ objs,
                                                     const DevView* __restrict__ views, DevParams prm,
                                                     SplitQueues q, CanonBuffers cb, FrameCounters* fc,
                                                     int n_objs, int n_views, int ranked)
{
    static_assert(NT % 64 == 0 && NT >= 256, "whole waves");
    constexpr int NW = NT / 64;
    constexpr int STEP = NT / 2 < 256 ? NT / 2 : 256;   // ranges per step
    // ranges of a two-level step: five candidate decisions each on one lane
    // pass (a wider two-level step costs more sample passes than the level
    // it saves)
    constexpr int SPEC = NT / 5;
    constexpr int XCH = (kSpecSamples * SPEC > 9 * STEP ? kSpecSamples * SPEC : 9 * STEP) * 5;
    static_assert(5 * SPEC <= NT, "two-level candidates");
    constexpr int DBUF = NT;                       // buffered DRAW ranges
    extern __shared__ RangeRec s_stack[];          // q.stack_cap entries
    __shared__ __align__(16) float s_xch[XCH];    // the step's samples, 5 floats each
    __shared__ RangeRec s_dbuf[DBUF];              // DRAW ranges not yet appended
    __shared__ int s_wcnt[NW][2];                  // per wave: split / draw lanes
    __shared__ int s_pf[SPEC + 1];                 // two-level steps: the ranges' decisions
    __shared__ uint32_t s_base;
    __shared__ ForestObj s_obj[kForestTables];
    __shared__ ForestView s_view[kForestTables];
    __shared__ float4 s_cps[kForestCps];
    // ranked frames: every DRAW range of the group as (root slot << 48 |
    // canonical key), in emission order (its record is at the same index of
    // the group's output region)
    __shared__ unsigned long long s_dk[kForestKeys];
    __shared__ uint32_t s_hist[kForestRankRoots];
    __shared__ uint32_t s_wcnt_scan[2 * kForestRankRoots];
    __shared__ uint32_t s_wsum[NW];
    static_assert(kForestRankRoots == 64, "one wave scans the bucket sizes");
    static_assert(kForestKeys * 12 + kRankBins * 8 <= (int)sizeof(s_xch), "rank lists fit the sample buffer");
    static_assert(5 * SPEC <= NT, "one lane per two-level candidate");
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wid = blockIdx.x;
    const int cap = q.stack_cap;
    const int dlim = prm.max_split_depth - 1;      // bound() culls ranges this deep
    const RangeRec* myq = q.queue + (int64_t)wid * q.stride;
    int qcnt = q.queue_cnt[wid];
    const int nroots = qcnt;
    int sh = 0, dfill = 0, maxsh = 0, kfill = 0;
    uint32_t nbounds = 0;
    bool ok = true;
    bool keys_ok = ranked != 0 && nroots <= kForestRankRoots;

    // tables and the roots' control points to LDS
    for (int i = tid; i < n_objs; i += NT) {
        const DevObject& o = objs[i];
        ForestObj f;
#pragma unroll
        for (int k = 0; k < 16; ++k) f.mv[k] = o.mv[k];
        f.cps = o.cps;
        f.split_limit = o.split_limit;
        f.type = o.type;
        f.view = o.view;
        f.gpid_base = o.gpid_base;
        s_obj[i] = f;
    }
    for (int i = tid; i < n_views; i += NT) {
        const DevView& v = views[i];
        ForestView f;
#pragma unroll
        for (int k = 0; k < 16; ++k) f.proj[k] = v.proj[k];
#pragma unroll
        for (int k = 0; k < 4; ++k) f.screen[k] = v.screen[k];
        f.f[0] = v.f[0]; f.f[1] = v.f[1];
        f.near_ = v.near_; f.far_ = v.far_; f.eps = v.eps;
        f.width = v.width; f.height = v.height;
        s_view[i] = f;
    }
    constexpr int CS = PT == 0 ? 16 : 20;   // float4 per root slot
    const int slots = min(qcnt, kForestCps / CS);
    for (int i = tid; i < slots * CS; i += NT) {
        const int k = i / CS, c = i - CS * (i / CS);
        // slot k: the root of global patch index wid + k * G (k_init_queues)
        const RangeRec rt = myq[k];
        const DevObject& o = objs[rt.obj_depth & 0xffffffu];
        const int ncp = o.type == MP_TYPE_GREGORY ? 20 : 16;
        if (c < ncp) s_cps[i] = ldg4(o.cps + (size_t)rt.pid * ncp + c);
    }

    for (;;) {
        // roots from the private queue while the step would not be full
        if (sh < STEP && qcnt > 0) {
            const int take = min(STEP - sh, qcnt);
            qcnt -= take;
            if (tid < take) s_stack[sh + tid] = myq[qcnt + tid];
            sh += take;
        }
        __syncthreads();   // roots and the last step's pushes visible
        if (sh == 0) break;
        const int n = min(sh, STEP);
        const int b = sh - n;   // the step takes s_stack[b .. b + n)
        const bool two = n <= SPEC;
        const int per = two ? kSpecSamples : 9;

        // ---- samples: sample j of range g at s_xch[(g * per + j) * 5] ----
        for (int t = tid; t < per * n; t += NT) {
            const int g = t / per, j = t - per * g;
            const RangeRec pr = s_stack[b + g];
            RangeRec c = pr;
            int iu = j / 3, iv = j - 3 * (j / 3);
            if (two) spec_sample_sel(pr, j, c, iu, iv);
            float v[5] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
            if ((int)(c.obj_depth >> 24) < dlim) {
                if (SO) {
                    forest_sample<PT>(objs[0], views[0], s_cps, slots, CS, c, iu, iv, v);
                } else {
                    const ForestObj& ob = s_obj[c.obj_depth & 0xffffffu];
                    forest_sample<PT>(ob, s_view[ob.view], s_cps, slots, CS, c, iu, iv, v);
                }
            }
#pragma unroll
            for (int k = 0; k < 5; ++k) s_xch[t * 5 + k] = v[k];
        }
        __syncthreads();   // samples written

        // ---- decisions: lane c decides candidate c: range c, or on a
        // two-level step candidate d = c % 5 (0 the range, 1..4 its possible
        // children) of range g = c / 5 ----
        const int g = two ? tid / 5 : tid;
        const int d = two ? tid - 5 * g : 0;
        const bool valid = g < n;
        const RangeRec pr = s_stack[b + min(g, n - 1)];
        RangeRec rec;
        spec_range_sel(pr, d, rec);
        const bool live = valid && (int)(rec.obj_depth >> 24) < dlim;
        int f = 0;
        {
            float ex[9], ey[9], ez[9], sx[9], sy[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const int j = two ? spec_index(d, k / 3, k % 3) : k;
                // (issued before the candidate's record is known)
                const int o = (min(g, n - 1) * per + j) * 5;
                ex[k] = s_xch[o]; ey[k] = s_xch[o + 1]; ez[k] = s_xch[o + 2];
                sx[k] = s_xch[o + 3]; sy[k] = s_xch[o + 4];
            }
            if (live) {
                if (SO) {
                    f = bound_decide(ex, ey, ez, sx, sy, rec.umin, rec.vmin, rec.umax, rec.vmax, views[0], prm,
                                     objs[0].split_limit);
                } else {
                    const ForestObj& ob = s_obj[rec.obj_depth & 0xffffffu];
                    f = bound_decide(ex, ey, ez, sx, sy, rec.umin, rec.vmin, rec.umax, rec.vmax, s_view[ob.view],
                                     prm, ob.split_limit);
                }
            }
        }
        int flags = f;
        if (two) {
            if (valid && d == 0) s_pf[g] = f;
            __syncthreads();   // the ranges' decisions
            const int pf = valid ? s_pf[g] : 0;
            const bool psplit = (pf & 2) != 0, udir = (pf & 4) != 0;
            flags = 0;
            if (valid) {
                if (d == 0) flags = pf == MP_DRAW ? MP_DRAW : 0;
                else if (d <= 2) flags = (psplit && udir) ? f : 0;
                else flags = (psplit && !udir) ? f : 0;
                if (d == 0) nbounds += psplit ? 3u : 1u;
            }
        } else if (valid) {
            nbounds += 1u;
        }
        if (!valid) flags = 0;

        // ---- compaction: children onto the stack, DRAW ranges to the buffer ----
        const uint64_t sb = __ballot((flags & 2) != 0);
        const uint64_t db = __ballot((flags & 1) != 0);
        if (lane == 0) {
            s_wcnt[wave][0] = __popcll(sb);
            s_wcnt[wave][1] = __popcll(db);
        }
        __syncthreads();   // counts posted; every read of the popped ranges done
        int so = 0, dof = 0, ns = 0, nd = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const int a = s_wcnt[w][0], c = s_wcnt[w][1];
            so += w < wave ? a : 0;
            dof += w < wave ? c : 0;
            ns += a;
            nd += c;
        }
        if (b + 2 * ns > cap) {
            if (tid == 0) atomicOr(&fc->overflow, 2u);
            ok = false;
            break;
        }
        if (flags & 2) {
            const int rk = so + (int)lanes_below(sb);
            RangeRec c0, c1;
            split_children(rec, flags, c0, c1);
            s_stack[b + 2 * rk] = c0;
            s_stack[b + 2 * rk + 1] = c1;
        }
        sh = b + 2 * ns;
        maxsh = max(maxsh, sh);
        if (ranked) {
            // the group's own output region: no shared counter
            if (kfill + nd > kForestKeys) keys_ok = false;
            if ((flags & 1) && keys_ok) {
                const int at = kfill + dof + (int)lanes_below(db);
                const uint32_t gb = SO ? objs[0].gpid_base : s_obj[rec.obj_depth & 0xffffffu].gpid_base;
                const uint32_t slot = (gb + (uint32_t)rec.pid - (uint32_t)wid) / (uint32_t)gridDim.x;
                s_dk[at] = ((unsigned long long)slot << 48) | canonical_key(rec);
                cb.forest_out[(size_t)wid * kForestKeys + at] = rec;
            }
            if (keys_ok) kfill += nd;
            continue;
        }
        if (dfill + nd > DBUF) {
            ok = forest_flush<NT>(objs, cb, fc, s_dbuf, dfill, &s_base);
            dfill = 0;
            if (!ok) break;
        }
        if (flags & 1) s_dbuf[dfill + dof + (int)lanes_below(db)] = rec;
        dfill += nd;
    }
    __syncthreads();   // the last DRAW buffer writes
    if (ok && dfill > 0) ok = forest_flush<NT>(objs, cb, fc, s_dbuf, dfill, &s_base);
    if (ranked && ok) {
        if (!keys_ok) {
            if (tid == 0) atomicOr(&fc->overflow, 32u);
        } else {
            // rank of each DRAW range among its patch's (the group owns the
            // patch's whole tree), and every root's DRAW count.  Keys are
            // (slot, u, v): with g = the fewest trailing zero bits of any u
            // (the finest split along u), (slot, u >> g) numbers the group's
            // columns in key order; when there are at most kRankBins of them
            // a counting sort by column leaves each key's rank as its
            // column's offset plus the smaller keys of its column (a few).
            // Otherwise the keys are bucketed by slot and each counts the
            // smaller keys of its bucket.
            unsigned long long* s_bk = reinterpret_cast<unsigned long long*>(s_xch);
            uint32_t* s_bg = reinterpret_cast<uint32_t*>(s_bk + kForestKeys);
            uint32_t* s_col = s_bg + kForestKeys;          // kRankBins column offsets
            uint32_t* s_cur = s_col + kRankBins;           // kRankBins cursors
            if (tid == 0) s_base = 24u;
            __syncthreads();
            {
                uint32_t tz = 24u;
                for (int i = tid; i < kfill; i += NT) {
                    const uint32_t u = (uint32_t)(s_dk[i] >> 24) & 0xffffffu;
                    tz = min(tz, u ? (uint32_t)__builtin_ctz(u) : 24u);
                }
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) tz = min(tz, (uint32_t)__shfl_xor((int)tz, o));
                if (lane == 0) atomicMin(&s_base, tz);
            }
            __syncthreads();
            const uint32_t gsh = s_base;                       // column = u >> gsh
            const uint32_t ncol = 1u << (24u - gsh);           // columns per root
            const bool columns = (uint64_t)nroots * ncol <= (uint64_t)kRankBins;
            if (columns) {
                const uint32_t nb = (uint32_t)nroots * ncol;
                for (uint32_t i = tid; i < nb; i += NT) s_cur[i] = 0u;
                __syncthreads();
                for (int i = tid; i < kfill; i += NT) {
                    const unsigned long long K = s_dk[i];
                    const uint32_t bin = (uint32_t)(K >> 48) * ncol + (((uint32_t)(K >> 24) & 0xffffffu) >> gsh);
                    atomicAdd(&s_cur[bin], 1u);
                }
                __syncthreads();
                lds_exclusive_scan<NT>(s_cur, s_col, nb, s_wsum);   // s_col = offsets; s_cur keeps the sizes
                for (uint32_t i = tid; i < nb; i += NT) s_cur[i] = s_col[i];
                __syncthreads();
                for (int i = tid; i < kfill; i += NT) {
                    const unsigned long long K = s_dk[i];
                    const uint32_t bin = (uint32_t)(K >> 48) * ncol + (((uint32_t)(K >> 24) & 0xffffffu) >> gsh);
                    const uint32_t at = atomicAdd(&s_cur[bin], 1u);
                    s_bk[at] = K;
                    s_bg[at] = (uint32_t)i;
                }
                __syncthreads();
                for (int p = tid; p < kfill; p += NT) {
                    const unsigned long long K = s_bk[p];
                    const uint32_t sl = (uint32_t)(K >> 48);
                    const uint32_t bin = sl * ncol + (((uint32_t)(K >> 24) & 0xffffffu) >> gsh);
                    // s_cur[bin] is now the column's end
                    const uint32_t j0 = s_col[bin], j1 = s_cur[bin];
                    uint32_t r = j0 - s_col[sl * ncol];
                    for (uint32_t j = j0; j < j1; ++j) r += s_bk[j] < K ? 1u : 0u;
                    cb.rank[(size_t)wid * kForestKeys + s_bg[p]] = (((uint32_t)wid + sl * gridDim.x) << 16) | r;
                }
                for (int k = tid; k < nroots; k += NT) {
                    const uint32_t a = s_col[(uint32_t)k * ncol];
                    const uint32_t e = (uint32_t)(k + 1) < (uint32_t)nroots ? s_col[(uint32_t)(k + 1) * ncol]
                                                                          : (uint32_t)kfill;
                    cb.counts[(uint32_t)wid + (uint32_t)k * gridDim.x] = e - a;
                }
            } else {
                uint32_t* s_bs = s_wcnt_scan;
                for (int i = tid; i < kForestRankRoots; i += NT) s_hist[i] = 0u;
                __syncthreads();
                for (int i = tid; i < kfill; i += NT) atomicAdd(&s_hist[(uint32_t)(s_dk[i] >> 48)], 1u);
                __syncthreads();
                if (tid < 64) {
                    // exclusive scan of the kForestRankRoots (64) bucket sizes
                    const uint32_t x = s_hist[tid];
                    uint32_t v = x;
#pragma unroll
                    for (int o = 1; o < 64; o <<= 1) {
                        const uint32_t y = __shfl_up(v, o);
                        if (lane >= o) v += y;
                    }
                    s_bs[tid] = v - x;          // bucket start
                    s_bs[64 + tid] = v - x;     // fill cursor
                }
                __syncthreads();
                for (int i = tid; i < kfill; i += NT) {
                    const unsigned long long K = s_dk[i];
                    const uint32_t at = atomicAdd(&s_bs[64 + (uint32_t)(K >> 48)], 1u);
                    s_bk[at] = K;
                    s_bg[at] = (uint32_t)i;
                }
                __syncthreads();
                for (int p = tid; p < kfill; p += NT) {
                    const unsigned long long K = s_bk[p];
                    const uint32_t sl = (uint32_t)(K >> 48);
                    const int j0 = (int)s_bs[sl], j1 = j0 + (int)s_hist[sl];
                    uint32_t r = 0;
                    int j = j0;
                    for (; j + 8 <= j1; j += 8) {
                        uint32_t c = 0;
#pragma unroll
                        for (int q8 = 0; q8 < 8; ++q8) c += s_bk[j + q8] < K ? 1u : 0u;
                        r += c;
                    }
                    for (; j < j1; ++j) r += s_bk[j] < K ? 1u : 0u;
                    cb.rank[(size_t)wid * kForestKeys + s_bg[p]] = (((uint32_t)wid + sl * gridDim.x) << 16) | r;
                }
                for (int k = tid; k < nroots; k += NT)
                    cb.counts[(uint32_t)wid + (uint32_t)k * gridDim.x] = s_hist[k];
            }
            if (tid == 0) cb.forest_total[wid] = (uint32_t)kfill;
        }
    }
    uint32_t nb = nbounds;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nb += (uint32_t)__shfl_xor((int)nb, o);
    if (lane == 0) atomicAdd(stripe(fc, (uint32_t)(wid * NW + wave), kStripeBounds), (unsigned long long)nb);
    if (tid == 0) atomicMax(stripe(fc, (uint32_t)wid, kStripeMaxStack), (unsigned long long)maxsh);
}

size_t split_forest_lds(int stack_cap)
{
    return (size_t)stack_cap * sizeof(RangeRec);
}

int split_forest_occupancy(size_t lds_bytes)
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_split_forest<kForestThreads, 2, false>, kForestThreads,
                                                     lds_bytes) != hipSuccess)
        blocks = 1;
    return blocks < 1 ? 1 : blocks;
}

bool split_forest_fits(int n_objs, int n_views)
{
    return n_objs <= kForestTables && n_views <= kForestTables;
}

bool forest_rank_fits(uint32_t n_gpids)
{
    return n_gpids <= kFusedScan;
}

hipError_t launch_split_forest(const DevObject* objs, const DevView* views, const DevParams& prm,
                               const SplitQueues& q, const CanonBuffers& cb, FrameCounters* fc, int patch_types,
                               int n_objs, int n_views, int ranked, hipStream_t s)
{
    if (!split_forest_fits(n_objs, n_views)) return hipErrorInvalidValue;
    const size_t lds = split_forest_lds(q.stack_cap);
#define MP_FOREST(PT, SO)                                                                                     \
    hipLaunchKernelGGL((k_split_forest<kForestThreads, PT, SO>), dim3(q.n_groups), dim3(kForestThreads), lds, s, \
                       objs, views, prm, q, cb, fc, n_objs, n_views, ranked)
    const bool one = n_objs == 1;
    if (patch_types == 1) {
        if (one) MP_FOREST(0, true); else MP_FOREST(0, false);
    } else if (patch_types == 2) {
        if (one) MP_FOREST(1, true); else MP_FOREST(1, false);
    } else {
        MP_FOREST(2, false);
    }
#undef MP_FOREST
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
        atomicAdd(stripe(fc, blockIdx.x, kStripeBounds), (unsigned long long)(m < SEG ? m : SEG));
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

MP_DEV uint32_t block_exclusive_scan_1024(uint32_t x, uint32_t* lds_waves, uint32_t& total);
MP_DEV uint32_t block_sum_1024(uint32_t x, uint32_t* lds_waves)
{
    uint32_t total;
    block_exclusive_scan_1024(x, lds_waves, total);
    return total;
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
        // more counted ranges than the list holds: stale counts (flagged, clamped)
        if (tot > cb.cap) atomicOr(&fc->overflow, 16u);
        fc->total_ranges = min(tot, cb.cap);
    }
}

// Many patches (batches): the same scan over blocks of kScanBlock counts.
// k_count_sums: one sum per block; k_scan_parts: exclusive scan of the sums
// (one work-group, running carry); k_scan_apply: each block scans its counts
// from its block offset, writes the offsets and clears counts and cursors.
constexpr uint32_t kScanBlock = 4096;   // 1024 threads x 4 counts

__global__ __launch_bounds__(1024) void k_count_sums(CanonBuffers cb)
{
    __shared__ uint32_t lw[32];
    const uint32_t n = cb.n_gpids;
    const uint32_t i0 = blockIdx.x * kScanBlock + threadIdx.x * 4u;
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
        if (i0 + k < n) s += cb.counts[i0 + k];
    uint32_t tot;
    block_exclusive_scan_1024(s, lw, tot);
    if (threadIdx.x == 0) cb.part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_parts(CanonBuffers cb, uint32_t nb, FrameCounters* fc)
{
    __shared__ uint32_t lw[32];
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += 1024u) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t x = i < nb ? cb.part[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan_1024(x, lw, tot);
        if (i < nb) cb.part[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        cb.offsets[cb.n_gpids] = carry;
        if (carry > cb.cap) atomicOr(&fc->overflow, 16u);
        fc->total_ranges = min(carry, cb.cap);
    }
}

__global__ __launch_bounds__(1024) void k_scan_apply(CanonBuffers cb)
{
    __shared__ uint32_t lw[32];
    const uint32_t n = cb.n_gpids;
    const uint32_t i0 = blockIdx.x * kScanBlock + threadIdx.x * 4u;
    uint32_t c[4], s = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        c[k] = i0 + k < n ? cb.counts[i0 + k] : 0u;
        s += c[k];
    }
    uint32_t tot;
    uint32_t run = cb.part[blockIdx.x] + block_exclusive_scan_1024(s, lw, tot);
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (i0 + k < n) {
            cb.offsets[i0 + k] = run;
            run += c[k];
            cb.counts[i0 + k] = 0;
            cb.cursor[i0 + k] = 0;
        }
    }
}

__global__ __launch_bounds__(256) void k_scatter(const DevObject* __restrict__ objs, CanonBuffers cb,
                                                 const FrameCounters* __restrict__ fc, DevView* views,
                                                 int n_views)
{
    const uint32_t total = fc->total_ranges;
    if (blockIdx.x == 0) {
        // per-view range bases (the offsets are final)
        for (int v = (int)threadIdx.x; v < n_views; v += 256) {
            const uint32_t b = cb.offsets[views[v].first_gpid];
            views[v].range_base = b;
            views[v].range_count = cb.offsets[views[v].gpid_end] - b;
        }
    }
    const int lane = threadIdx.x & 63;
    const uint32_t stride = gridDim.x * 256;
    const uint32_t lim = (total + 255u) & ~255u;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < lim; i += stride) {
        const bool valid = i < total;
        RangeRec r;
        uint32_t g = 0;
        if (valid) {
            r = cb.draw[i];
            g = gpid_of(objs, r);
        }
        // wave-aggregated bucket cursors: the lanes of one patch elect a
        // leader, then every leader adds its count in one vector atomic (the
        // returns arrive together, not one round trip per patch)
        uint64_t m = __ballot(valid);
        int my_leader = 0;
        uint32_t my_rank = 0, lead_cnt = 0;
        while (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            const uint32_t lg = __shfl(g, leader);
            const uint64_t same = __ballot(valid && g == lg) & m;
            if ((same >> lane) & 1ull) { my_leader = leader; my_rank = lanes_below(same); }
            if (lane == leader) lead_cnt = (uint32_t)__popcll(same);
            m &= ~same;
        }
        uint32_t b = 0;
        if (lead_cnt) b = cb.offsets[g] + atomicAdd(&cb.cursor[g], lead_cnt);
        const uint32_t slot = (uint32_t)__shfl((int)b, my_leader) + my_rank;
        if (valid && slot < cb.cap) {
            cb.bucket[slot] = r;
            cb.keys[slot] = canonical_key(r);
            cb.span[slot] = make_uint2(cb.offsets[g], min(cb.offsets[g + 1], cb.cap));
        }
    }
}

// Frames of up to kFusedScan patches: every work-group scans the per-patch
// counts itself (four per thread, offsets kept in LDS) and scatters its share
// of the draw list from there, as k_scatter does: one launch instead of a scan
// launch and a scatter launch.  It relies on the counts being those of this
// frame only and on zero cursors: k_rank clears both after every scatter.

__global__ __launch_bounds__(1024) void k_scan_scatter(const DevObject* __restrict__ objs, CanonBuffers cb,
                                                       FrameCounters* fc, DevView* views, int n_views)
{
    __shared__ uint32_t s_off[kFusedScan + 1];
    __shared__ uint32_t lw[32];
    const uint32_t n = cb.n_gpids;
    const uint32_t t = threadIdx.x;
    uint32_t c[4], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t i = t * 4u + k;
        c[k] = i < n ? cb.counts[i] : 0u;
        sum += c[k];
    }
    uint32_t counted;
    uint32_t run = block_exclusive_scan_1024(sum, lw, counted);
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t i = t * 4u + k;
        if (i < n) s_off[i] = run;
        run += c[k];
    }
    if (t == 0) s_off[n] = counted;
    __syncthreads();
    // more counted ranges than the list holds: stale counts (flagged, clamped)
    const uint32_t total = min(counted, cb.cap);
    if (blockIdx.x == 0) {
        if (t == 0) {
            fc->total_ranges = total;
            if (counted > cb.cap) atomicOr(&fc->overflow, 16u);
        }
        for (int v = (int)t; v < n_views; v += 1024) {
            const uint32_t b = s_off[views[v].first_gpid];
            views[v].range_base = b;
            views[v].range_count = s_off[views[v].gpid_end] - b;
        }
    }
    const int lane = t & 63;
    const uint32_t stride = gridDim.x * 1024u;
    const uint32_t lim = (total + 1023u) & ~1023u;
    for (uint32_t i = blockIdx.x * 1024u + t; i < lim; i += stride) {
        const bool valid = i < total;
        RangeRec r;
        uint32_t g = 0;
        if (valid) {
            r = cb.draw[i];
            g = gpid_of(objs, r);
        }
        uint64_t m = __ballot(valid);
        int my_leader = 0;
        uint32_t my_rank = 0, lead_cnt = 0;
        while (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            const uint32_t lg = __shfl(g, leader);
            const uint64_t same = __ballot(valid && g == lg) & m;
            if ((same >> lane) & 1ull) { my_leader = leader; my_rank = lanes_below(same); }
            if (lane == leader) lead_cnt = (uint32_t)__popcll(same);
            m &= ~same;
        }
        uint32_t b = 0;
        if (lead_cnt) b = s_off[g] + atomicAdd(&cb.cursor[g], lead_cnt);
        const uint32_t slot = (uint32_t)__shfl((int)b, my_leader) + my_rank;
        if (valid && slot < cb.cap) {
            cb.bucket[slot] = r;
            cb.keys[slot] = canonical_key(r);
            cb.span[slot] = make_uint2(s_off[g], min(s_off[g + 1], cb.cap));
        }
    }
}

// Canonical order of a ranked forest frame (k_split_forest with ranked):
// work-group g places forest group g's DRAW ranges at their patch's offset +
// their rank inside the patch, the offsets from its own scan of the
// per-patch DRAW counts (kFusedScan at most, 16 per thread in LDS).
constexpr int kPlaceThreads = 256;

// the forest's rank words hold a global patch index and a rank, 16 bits each
static_assert(kFusedScan <= 65536 && kForestKeys <= 65536, "patch index and rank share a word");

__global__ __launch_bounds__(kPlaceThreads) void k_place_ranked(CanonBuffers cb, FrameCounters* fc, DevView* views,
                                                                int n_views)
{
    constexpr uint32_t PER = kFusedScan / kPlaceThreads;   // counts per thread
    __shared__ uint32_t s_off[kFusedScan + 1];
    __shared__ uint32_t s_w[2][kPlaceThreads / 64];
    const uint32_t n = cb.n_gpids;
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint32_t g = blockIdx.x;
    const uint32_t cnt = min(cb.forest_total[g], (uint32_t)kForestKeys);
    const uint32_t ovf = fc->overflow;   // final: the split has ended
    // the group's records and ranks load now, in flight during the scan
    // (read only once the frame is known whole)
    constexpr uint32_t PR = (kForestKeys + kPlaceThreads - 1) / kPlaceThreads;
    // the records as six words each (a RangeRec array would live in scratch)
    uint32_t w[PR][6], rk[PR];
#pragma unroll
    for (uint32_t j = 0; j < PR; ++j) {
        const uint32_t i = t + j * kPlaceThreads;
        if (i < cnt) {
            const size_t e = (size_t)g * kForestKeys + i;
            const uint32_t* src = reinterpret_cast<const uint32_t*>(cb.forest_out + e);
#pragma unroll
            for (int f = 0; f < 6; ++f) w[j][f] = src[f];
            rk[j] = cb.rank[e];
        }
    }
    uint32_t c[PER], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t i = t * PER + k;
        c[k] = i < n ? cb.counts[i] : 0u;
        sum += c[k];
    }
    // block exclusive scan of the per-thread sums (and, for work-group 0,
    // the sum of the forest groups' totals)
    uint32_t emitted = 0;
    if (g == 0)
        for (uint32_t q = t; q < cb.forest_groups; q += kPlaceThreads) emitted += cb.forest_total[q];
    uint32_t v = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(v, o);
        if ((int)lane >= o) v += y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) emitted += (uint32_t)__shfl_xor((int)emitted, o);
    if (lane == 63) s_w[0][wave] = v;
    if (lane == 0) s_w[1][wave] = emitted;
    __syncthreads();
    uint32_t run = v - sum, counted = 0, emitted_all = 0;
#pragma unroll
    for (uint32_t w = 0; w < kPlaceThreads / 64; ++w) {
        run += w < wave ? s_w[0][w] : 0u;
        counted += s_w[0][w];
        emitted_all += s_w[1][w];
    }
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        const uint32_t i = t * PER + k;
        if (i < n) s_off[i] = run;
        run += c[k];
    }
    if (t == 0) s_off[n] = counted;
    __syncthreads();
    // a split that overflowed a stack or a group's key store left ranks and
    // counts that do not match its output, and a frame with more ranges than
    // the canonical list holds must grow it: nothing is placed (the host
    // renders the frame again)
    const bool whole = (ovf & 34u) == 0u && counted <= cb.cap;
    if (g == 0) {
        if (t == 0) {
            fc->total_ranges = whole ? counted : 0u;
            fc->draw_count = whole ? emitted_all : 0u;
            if (counted > cb.cap) atomicOr(&fc->overflow, 1u);
        }
        for (int vv = (int)t; vv < n_views; vv += kPlaceThreads) {
            const uint32_t b = whole ? s_off[views[vv].first_gpid] : 0u;
            views[vv].range_base = b;
            views[vv].range_count = whole ? s_off[views[vv].gpid_end] - b : 0u;
        }
    }
    if (!whole) return;
#pragma unroll
    for (uint32_t j = 0; j < PR; ++j) {
        if (t + j * kPlaceThreads < cnt) {
            // rk: the range's global patch index << 16 | its rank in the patch
            const uint32_t slot = s_off[rk[j] >> 16] + (rk[j] & 0xffffu);
            if (slot < cb.cap) {
                uint32_t* dst = reinterpret_cast<uint32_t*>(cb.canon + slot);
#pragma unroll
                for (int f = 0; f < 6; ++f) dst[f] = w[j][f];
            }
        }
    }
}

hipError_t launch_place_ranked(DevView* views, int n_views, const CanonBuffers& cb,
                               FrameCounters* fc, int num_cus, hipStream_t s)
{
    if (cb.n_gpids > kFusedScan || !cb.rank || !cb.forest_out || !cb.forest_total || !cb.forest_groups)
        return hipErrorInvalidValue;
    (void)num_cus;
    hipLaunchKernelGGL(k_place_ranked, dim3(cb.forest_groups), dim3(kPlaceThreads), 0, s, cb, fc, views, n_views);
    return hipGetLastError();
}

// Rank of each bucketed range among its patch's ranges (its canonical
// index = the patch's first slot + the number of its ranges with a smaller
// key).  A work-group takes 256 consecutive bucketed ranges; the keys of
// every patch they belong to form one contiguous span, staged in LDS once
// and counted from there (spans above kRankStage keys count from global).
// A bucket of more than kRankStage ranges is not counted (that would cost
// its size squared) but sorted by the work-group whose 256 ranges include
// the bucket's first slot: a stable LSD radix sort of (key, slot) pairs,
// linear in the bucket size.
constexpr uint32_t kRankStage = 4096;

// Stable LSD radix sort of the n keys at keys[0..n) (48-bit, unique within
// a bucket) by one 256-thread work-group, 8-bit digits, ping-ponging through
// ka/va and kb/vb; returns the slot indices in key order in vb.  Within a
// tile of 256 keys a key's place among the equal digits before it is its rank
// among its wave's peers (eight ballots) plus the counts of the earlier waves.
MP_DEV void bucket_radix_sort(const unsigned long long* __restrict__ keys, uint32_t n, unsigned long long* ka,
                              uint32_t* va, unsigned long long* kb, uint32_t* vb, uint32_t* s_base,
                              uint32_t (*s_wcnt)[256])
{
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    constexpr int kPasses = 6;   // 48 key bits
    for (int pass = 0; pass < kPasses; ++pass) {
        const int shift = 8 * pass;
        // pass 0 reads the bucket's keys; the passes alternate a -> b -> a ...
        // and the sixth writes b
        const unsigned long long* ksrc = pass == 0 ? keys : ((pass & 1) ? ka : kb);
        const uint32_t* vsrc = (pass & 1) ? va : vb;
        unsigned long long* kdst = (pass & 1) ? kb : ka;
        uint32_t* vdst = (pass & 1) ? vb : va;
        s_base[tid] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < n; i += 256u)
            atomicAdd(&s_base[(uint32_t)(ksrc[i] >> shift) & 255u], 1u);
        __syncthreads();
        {
            // exclusive scan of the 256 bins (one per thread)
            const uint32_t x = s_base[tid];
            uint32_t v = x;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(v, o);
                if ((int)lane >= o) v += y;
            }
            if (lane == 63) s_wcnt[0][wave] = v;
            __syncthreads();
            uint32_t before = 0;
            for (uint32_t w = 0; w < wave; ++w) before += s_wcnt[0][w];
            __syncthreads();
            s_base[tid] = before + v - x;
        }
        __syncthreads();
        for (uint32_t t0 = 0; t0 < n; t0 += 256u) {
            const uint32_t i = t0 + tid;
            const bool valid = i < n;
            unsigned long long k = 0;
            uint32_t v = 0, d = 0;
            if (valid) {
                k = ksrc[i];
                v = pass == 0 ? i : vsrc[i];
                d = (uint32_t)(k >> shift) & 255u;
            }
#pragma unroll
            for (int w = 0; w < 4; ++w) s_wcnt[w][tid] = 0;
            __syncthreads();
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const uint64_t m = __ballot(valid && ((d >> b) & 1u));
                peers &= ((d >> b) & 1u) ? m : ~m;
            }
            if (valid && (uint32_t)(__ffsll((unsigned long long)peers) - 1) == lane)
                s_wcnt[wave][d] = (uint32_t)__popcll(peers);
            __syncthreads();
            if (valid) {
                uint32_t off = s_base[d] + lanes_below(peers);
                for (uint32_t w = 0; w < wave; ++w) off += s_wcnt[w][d];
                kdst[off] = k;
                vdst[off] = v;
            }
            __syncthreads();
            s_base[tid] += s_wcnt[0][tid] + s_wcnt[1][tid] + s_wcnt[2][tid] + s_wcnt[3][tid];
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(256) void k_rank(CanonBuffers cb, FrameCounters* __restrict__ fc)
{
    __shared__ unsigned long long s_key[kRankStage];
    __shared__ uint32_t s_lo[4], s_hi[4];
    __shared__ uint32_t s_big_lo, s_big_hi;
    const uint32_t total = fc->total_ranges;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    // every path leaves the per-patch counts and bucket cursors zero for the
    // next frame (the scatter, fused or not, has used them)
    for (uint32_t i = blockIdx.x * 256u + tid; i < cb.n_gpids; i += gridDim.x * 256u) {
        cb.counts[i] = 0;
        cb.cursor[i] = 0;
    }
    for (uint32_t j0 = blockIdx.x * 256u; j0 < total; j0 += gridDim.x * 256u) {
        const uint32_t j = j0 + tid;
        const bool valid = j < total;
        uint2 sp = make_uint2(0xffffffffu, 0u);
        unsigned long long key = 0ull;
        RangeRec r;
        if (valid) {
            sp = cb.span[j];
            // inside the lists even for spans a stale count produced
            sp.x = min(sp.x, cb.cap);
            sp.y = min(max(sp.y, sp.x), cb.cap);
            key = cb.keys[j];
            r = cb.bucket[j];
        }
        // a range of a bucket above kRankStage is placed by the bucket's sort
        const bool big = valid && sp.y - sp.x > kRankStage;
        if (tid == 0) s_big_hi = 0u;
        uint32_t lo = big ? 0xffffffffu : sp.x, hi = big ? 0u : sp.y;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lo = min(lo, (uint32_t)__shfl_xor((int)lo, o));
            hi = max(hi, (uint32_t)__shfl_xor((int)hi, o));
        }
        if (lane == 0) { s_lo[wave] = lo; s_hi[wave] = hi; }
        __syncthreads();
        if (big && j == sp.x) { s_big_lo = sp.x; s_big_hi = sp.y; }   // this window holds its first slot
        const uint32_t L = min(min(s_lo[0], s_lo[1]), min(s_lo[2], s_lo[3]));
        const uint32_t H = max(max(s_hi[0], s_hi[1]), max(s_hi[2], s_hi[3]));
        const bool staged = H - L <= kRankStage;
        if (staged)
            for (uint32_t k = L + tid; k < H; k += 256u) s_key[k - L] = cb.keys[k];
        __syncthreads();
        if (valid && !big) {
            uint32_t rank = 0;
            if (staged) {
                // eight LDS reads in flight per step
                const unsigned long long* b = s_key + (sp.x - L);
                const uint32_t n = sp.y - sp.x;
                uint32_t k = 0;
                for (; k + 8u <= n; k += 8u) {
                    uint32_t c = 0;
#pragma unroll
                    for (int t = 0; t < 8; ++t) c += b[k + t] < key ? 1u : 0u;
                    rank += c;
                }
                for (; k < n; ++k) rank += b[k] < key ? 1u : 0u;
            } else {
                for (uint32_t k = sp.x; k < sp.y; ++k) rank += cb.keys[k] < key ? 1u : 0u;
            }
            if (sp.x + rank < cb.cap) cb.canon[sp.x + rank] = r;
        }
        __syncthreads();   // s_key / s_lo reused by the next 256 ranges
        if (s_big_hi) {
            // sort the big bucket whose first slot lies in this window
            const uint32_t b0 = s_big_lo, n = s_big_hi - s_big_lo;
            uint32_t* s_base = reinterpret_cast<uint32_t*>(s_key);
            uint32_t (*s_wcnt)[256] = reinterpret_cast<uint32_t (*)[256]>(s_base + 256);
            bucket_radix_sort(cb.keys + b0, n, cb.sort_keys + b0, cb.sort_vals + b0, cb.sort_keys + cb.cap + b0,
                              cb.sort_vals + cb.cap + b0, s_base, s_wcnt);
            const uint32_t* order = cb.sort_vals + cb.cap + b0;
            for (uint32_t i = tid; i < n; i += 256u) cb.canon[b0 + i] = cb.bucket[b0 + order[i]];
            if (tid == 0) atomicAdd(&fc->big_ranges, n);
            __syncthreads();
        }
    }
}

// rank work-groups: one per 256 list entries the lists can hold, at most four
// per CU (each stages up to 32 KB of keys)
static int rank_grid(const CanonBuffers& cb, int num_cus)
{
    const uint32_t want = std::max(1u, (cb.cap + 255u) / 256u);
    return (int)std::min<uint32_t>(want, (uint32_t)num_cus * 4u);
}

hipError_t launch_canonicalize(const DevObject* objs, DevView* views, int n_views,
                               const CanonBuffers& cb, FrameCounters* fc, int num_cus, int fused,
                               hipStream_t s)
{
    hipError_t e;
    if (fused && cb.n_gpids <= kFusedScan) {
        const uint32_t want = (cb.cap + 1023u) / 1024u;
        const int grid = (int)std::max(1u, std::min<uint32_t>((uint32_t)num_cus, want));
        hipLaunchKernelGGL(k_scan_scatter, dim3(grid), dim3(1024), 0, s, objs, cb, fc, views, n_views);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(k_rank, dim3(rank_grid(cb, num_cus)), dim3(256), 0, s, cb, fc);
        return hipGetLastError();
    }
    if (cb.n_gpids <= 2u * kScanBlock) {
        hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, cb, views, n_views, fc);
    } else {
        const uint32_t nb = (cb.n_gpids + kScanBlock - 1u) / kScanBlock;
        hipLaunchKernelGGL(k_count_sums, dim3(nb), dim3(1024), 0, s, cb);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(1024), 0, s, cb, nb, fc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(1024), 0, s, cb);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int grid = num_cus * 4;
    hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(256), 0, s, objs, cb, fc, views, n_views);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rank, dim3(rank_grid(cb, num_cus)), dim3(256), 0, s, cb, fc);
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
// normalize3 of per-lane vectors (the same value, inv_sqrt_ieee's steps)
MP_DEV void normalize3_wave(float& x, float& y, float& z)
{
    const float inv = inv_sqrt_ieee(dot3(x, y, z, x, y, z));
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
    float nx = mad_diff(dvy, duz, dvz, duy);
    float ny = mad_diff(dvz, dux, dvx, duz);
    float nzz = mad_diff(dvx, duy, dvy, dux);
    normalize3_wave(nx, ny, nzz);
    float lx = 4.0f, ly = 3.0f, lz = 8.0f;
    normalize3(lx, ly, lz);
    const float s = cl_max(dot3(nx, ny, nzz, lx, ly, lz), 0.0f);
    float sx = ((p0.x + p1.x) + p2.x) + p3.x;
    float sy = ((p0.y + p1.y) + p2.y) + p3.y;
    float sz = ((p0.z + p1.z) + p2.z) + p3.z;
    // The specular term only matters when it is NaN.  With n finite, a
    // position sum in front of the eye (z < 0) and away from under- and
    // overflow, the unit sum has z < 0, l - unit sum has z > l.z > 0, so the
    // half vector and its dot product with n are finite: the result is s.
    const float lo = 1.0e-18f, hi = 1.0e18f;
    if (fabsf(nx) < INFINITY && fabsf(ny) < INFINITY && fabsf(nzz) < INFINITY && sz < -lo && sz > -hi &&
        fabsf(sx) < hi && fabsf(sy) < hi)
        return s;
    normalize3(sx, sy, sz);
    float hx = lx + (-sx), hy = ly + (-sy), hz = lz + (-sz);
    normalize3(hx, hy, hz);
    const float spec = cl_max(dot3(nx, ny, nzz, hx, hy, hz), 0.0f);
    return (spec != spec) ? spec : s;
}

// MODE 0: no displacement; 1: displacement (value noise by lane windows).
// Shade (reyes.cl:84-199) and sample (reyes.cl:270-398) one 8x8 block of a
// diced range: the calling wave's lane (u, v) owns quad (u, v); s_pos /
// s_pxl / s_dep hold the range's (N+1)^2 grid; r is its canonical index.
// Eye-space grid position as the shading reads it: x, y, z (w unused).
struct EyePos {
    float x, y, z;
};
MP_DEV float4 eye4(const EyePos& p) { return make_float4(p.x, p.y, p.z, 0.0f); }
MP_DEV float4 eye4(const float4& p) { return p; }
MP_DEV void set_pos(EyePos& d, float x, float y, float z, float) { d.x = x; d.y = y; d.z = z; }
MP_DEV void set_pos(float4& d, float x, float y, float z, float w) { d = make_float4(x, y, z, w); }

template <int N, typename PT>
MP_DEV void shade_sample(const PT* s_pos, const int2* s_pxl, const float* s_dep, int u, int v, int lane,
                         const DevView& vw, const DevParams& prm, uint32_t r, float* __restrict__ quad_store,
                         uint32_t& slow, uint32_t& natom)
{
    constexpr int G1 = N + 1;
    constexpr int Q = N * N;
    const int g0 = u + v * G1;
    const int sh = prm.subpixel_bits;
    // ---- shade (reyes.cl:84-199); one wave = one 8x8 block ----
    const float z0 = s_pos[g0].z, z1 = s_pos[g0 + 1].z, z2 = s_pos[g0 + G1].z, z3 = s_pos[g0 + G1 + 1].z;
    const int2 x0 = s_pxl[g0], x1 = s_pxl[g0 + 1], x2 = s_pxl[g0 + G1], x3 = s_pxl[g0 + G1 + 1];
    const float d0 = s_dep[g0], d1 = s_dep[g0 + 1], d2 = s_dep[g0 + G1], d3 = s_dep[g0 + G1 + 1];
    const bool allnormal = (lane_mask(z0 == 0.0f) | lane_mask(z1 == 0.0f) | lane_mask(z2 == 0.0f) |
                            lane_mask(z3 == 0.0f)) == 0;
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

    // Exact-integer regime: all coordinates within +-2^18 and edges
    // shorter than 2^11 sub-pixels keep every edge product inside
    // int32, so quads that are not front-facing and pixels outside a
    // quad's bounding box can never pass the edge test.
    const int32_t lim = 1 << 18;
    const uint64_t small = lane_mask(qminx > -lim) & lane_mask(qmaxx < lim) & lane_mask(qminy > -lim) &
                           lane_mask(qmaxy < lim) & lane_mask(qmaxx - qminx < (1 << 11)) &
                           lane_mask(qmaxy - qminy < (1 << 11));
    const bool exact = small == __builtin_amdgcn_read_exec();

    // ---- sample (reyes.cl:270-398) ----; true for a quad that covers a sample
    auto sample = [&]() -> bool {
        bool hit = false;
        const uint32_t id = (r - vw.range_base) * (uint32_t)Q + (uint32_t)(u + v * N);
        unsigned long long* keys = vw.keys;
        const int gw = vw.grid_w;
        // pixel-candidate bounds with the reference's initialisers
        const int32_t mingx = min(vmaxx + 1, qminx), mingy = min(vmaxy + 1, qminy);
        const int32_t maxgx = max(-1, qmaxx), maxgy = max(-1, qmaxy);
        if (exact) {
            if (ff) {
                // The block's tile box (the wave's clamped quad boxes) only
                // narrows a front-facing quad's pixels to the viewport and the
                // tile after its last column / row: pixel X samples at its corner
                // X << sh, inside the box only for X >= ceil(mingx / 2^sh) (the
                // reference's loop starts at the floor, whose corner lies outside
                // the box unless mingx is a whole pixel), from column 0, up to
                // maxgx >> sh and the last column of the viewport's last tile.
                // The box contains every front-facing quad's clamped box, so no
                // wave reduction is needed here.
                const int32_t rnd = (1 << sh) - 1;
                const int32_t X0 = max((mingx + rnd) >> sh, 0), X1 = min(maxgx >> sh, (vw.width - 1) | 7);
                const int32_t Y0 = max((mingy + rnd) >> sh, 0), Y1 = min(maxgy >> sh, (vw.height - 1) | 7);
                SmallTri t1, t2;
                small_tri_setup(t1, x0.x, x0.y, x1.x, x1.y, x3.x, x3.y);
                small_tri_setup(t2, x0.x, x0.y, x3.x, x3.y, x2.x, x2.y);
                for (int32_t Y = Y0; Y <= Y1; ++Y) {
                    for (int32_t X = X0; X <= X1; ++X) {
                        const int32_t tx = X << sh, ty = Y << sh;
                        float dd = 1.0f;
                        const bool in1 = small_tri_test(t1, tx, ty, d0, d1, d3, dd);
                        const bool in2 = small_tri_test(t2, tx, ty, d0, d3, d2, dd);
                        if (in1 || in2) {
                            hit = true;
                            const unsigned long long key = ((unsigned long long)mono_bits(dd) << 32) | id;
                            key_min(&keys[fb_index(X, Y, gw)], key);
                            ++natom;
                        }
                    }
                }
            }
            return hit;
        }
        // the block's tile box: front-facing quads' boxes, clamped to the viewport
        int32_t xmin = wave_min_i(ff ? min(vmaxx, qminx) : vmaxx);
        int32_t ymin = wave_min_i(ff ? min(vmaxy, qminy) : vmaxy);
        int32_t xmax = wave_max_i(ff ? max(0, qmaxx) : 0);
        int32_t ymax = wave_max_i(ff ? max(0, qmaxy) : 0);
        xmin = max(0, xmin); ymin = max(0, ymin);
        xmax = min(vmaxx, xmax); ymax = min(vmaxy, ymax);
        const int32_t mintx = xmin >> (sh + 3), minty = ymin >> (sh + 3);
        const int32_t maxtx = xmax >> (sh + 3), maxty = ymax >> (sh + 3);
        if (mintx <= maxtx && minty <= maxty) {
            {
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
                                    hit = true;
                                    const unsigned long long key = ((unsigned long long)mono_bits(dd) << 32) | id;
                                    key_min(&keys[fb_index(tx * 8 + x, ty * 8 + y, gw)], key);
                                    ++natom;
                                }
                            }
                        }
                    }
                }
            }
        }
        return hit;
    };
    // The reference empties the box of a block with a zero eye depth.  The
    // Lambert term is read back (resolve, adjoint) only for micropolygons
    // that own a pixel.  8x8 grids (small ranges, many quads without a
    // sample) shade after sampling, only the quads that cover a sample;
    // larger grids, whose quads nearly all cover one, shade first every quad
    // that may (front-facing in the exact-integer regime).  The other entries
    // of the store are left as they were.
    const size_t qs = (size_t)r * Q + (u + v * N);
    auto shade = [&]() {
        return quad_lambert(eye4(s_pos[g0]), eye4(s_pos[g0 + 1]), eye4(s_pos[g0 + G1]), eye4(s_pos[g0 + G1 + 1]));
    };
    if constexpr (N == 8) {
        if (allnormal && sample()) quad_store[qs] = shade();
    } else {
        if (allnormal && (ff || !exact)) quad_store[qs] = shade();
        if (allnormal) sample();
    }
}

// Dice one range (dice.cl:38-79) into LDS: eye-space position, sub-pixel
// coordinate and depth of each of the (N+1)^2 grid points, displaced by the
// value noise in MODE 1.  For a displaced Bezier patch the per-range tables
// come first: the displacement normal's column sums (s_col, 3 * 8 * (N+1)
// floats, component-major) and the Bernstein weights of every u column and
// v row (s_wt, 3 * (N+1) float4: bern_u of column c at s_wt[c], bern_v and
// dbern of row r at s_wt[G1 + 2r] and s_wt[G1 + 2r + 1]), the values the
// point evaluation would compute from the same u and v.  s_cp holds the
// range's control points.  Every lane of each of the NT / 64 waves runs every
// pass (the lane-window noise exchanges values across the wave); lanes past
// the last point repeat it and store nothing.  nw: lattice windows the
// calling wave keeps across calls.
template <int N, int MODE, int NT, typename PT>
MP_DEV void dice_grid(const float4* __restrict__ s_cp, float* __restrict__ s_col, float4* __restrict__ s_wt,
                      const RangeRec& rec, const DevObject& ob, const DevView& vw, const DevParams& prm,
                      PT* __restrict__ s_pos, int2* __restrict__ s_pxl, float* __restrict__ s_dep,
                      NoiseWindows& nw)
{
    constexpr int G1 = N + 1;
    constexpr int G = G1 * G1;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const bool disp_cols = MODE == 1 && ob.type == MP_TYPE_BEZIER;
    // umin + (umax - umin) * t (dice.cl:47): the width is a power of two
    // and t = i / N has at most 6 significant bits, so the product is exact
    // and the fused form equals the written one
    const float du = rec.umax - rec.umin, dv = rec.vmax - rec.vmin;
    if (disp_cols) {
        for (int t = tid; t < G1 * 9; t += NT) {
            if (t < G1 * 8) {
                // displacement normal: column sums once per u column
                const int c = t >> 3, k = t & 7;
                const float uc = mad(du, (float)c / (float)N, rec.umin);
                float w[4];
                const float4 v = normal_column(s_cp, uc, k, w);
                s_col[(0 * 8 + k) * G1 + c] = v.x;
                s_col[(1 * 8 + k) * G1 + c] = v.y;
                s_col[(2 * 8 + k) * G1 + c] = v.z;
                if (k == 0) s_wt[c] = make_float4(w[0], w[1], w[2], w[3]);
            } else {
                const int r = t - G1 * 8;
                const float vr = mad(dv, (float)r / (float)N, rec.vmin);
                const float sv = 1.0f - vr;
                float b[4], db[4];
                bern_v(vr, sv, b);
                dbern(vr, sv, db);
                s_wt[G1 + 2 * r] = make_float4(b[0], b[1], b[2], b[3]);
                s_wt[G1 + 2 * r + 1] = make_float4(db[0], db[1], db[2], db[3]);
            }
        }
        __syncthreads();
    }
    // eye position, sub-pixel coordinate and depth of grid point g from its
    // patch point P; stored when the point exists (store)
    auto project = [&](int g, const float P[4], bool store) {
        float e[4], q[4];
        mat4_mul(ob.mv, P, e);
        perspective_mul(vw.proj, e, q);
        // (x * W) / 2 + W / 2 (dice.cl:71-72): halving is exact, and a product
        // too small for it to be is absorbed by the addend (>= 32), so the
        // fused form equals the written one
        const int32_t px = f2i_sat(mad((q[0] / q[3]) * (float)vw.vs_x, 0.5f, (float)(vw.vs_x / 2)));
        const int32_t py = f2i_sat(mad((q[1] / q[3]) * (float)vw.vs_y, 0.5f, (float)(vw.vs_y / 2)));
        if (store) {
            set_pos(s_pos[g], e[0], e[1], e[2], e[3]);
            s_pxl[g] = make_int2(px, py);
            s_dep[g] = q[2] / q[3];
        }
    };
    if (MODE == 0 && ob.type == MP_TYPE_BEZIER) {
        // undisplaced Bezier: each lane takes points gb + lane and, while the
        // grid has them, gb + NT + lane in one pass, every control point read
        // from LDS once for both
        for (int gb = wave * 64; gb < G; gb += 2 * NT) {
            const int g0 = min(gb + lane, G - 1);
            const float u0 = mad(du, (float)(g0 % G1) / (float)N, rec.umin);
            const float v0 = mad(dv, (float)(g0 / G1) / (float)N, rec.vmin);
            float P0[4];
            if (gb + NT < G) {
                const int g1 = min(gb + NT + lane, G - 1);
                const float u1 = mad(du, (float)(g1 % G1) / (float)N, rec.umin);
                const float v1 = mad(dv, (float)(g1 / G1) / (float)N, rec.vmin);
                float P1[4];
                eval_bezier2(s_cp, s_cp, u0, v0, u1, v1, P0, P1);
                project(g0, P0, gb + lane < G);
                project(g1, P1, gb + NT + lane < G);
            } else {
                eval_bezier(s_cp, u0, v0, P0);
                project(g0, P0, gb + lane < G);
            }
        }
        return;
    }
    for (int gb = wave * 64; gb < G; gb += NT) {
        // control points re-read from LDS each pass (registers for the kept
        // lattice windows)
        asm volatile("" ::: "memory");
        const int g = min(gb + lane, G - 1);
        const int nu = g % G1, nv = g / G1;
        float P[4];
        if (disp_cols) {
            const float4 A = s_wt[nu], Bv = s_wt[G1 + 2 * nv], Dv = s_wt[G1 + 2 * nv + 1];
            const float a[4] = { A.x, A.y, A.z, A.w }, b[4] = { Bv.x, Bv.y, Bv.z, Bv.w };
            const float db[4] = { Dv.x, Dv.y, Dv.z, Dv.w };
            displaced_point_lds_cols<G1>(s_cp, s_col, nu, a, b, db, prm, ob.disp_amplitude, P, nw);
        } else {
            const float uu = mad(du, (float)nu / (float)N, rec.umin);
            const float vv = mad(dv, (float)nv / (float)N, rec.vmin);
            if (MODE == 1) displaced_point(s_cp, ob.type, uu, vv, prm, ob.disp_amplitude, P);
            else eval_patch(s_cp, ob.type, uu, vv, P);
        }
        project(g, P, gb + lane < G);
    }
}

// Undisplaced dice (dice.cl:38-79) of the k (1 to 3) consecutive ranges a,
// b, c of one object by one wave into the flat grid arrays: range j's points
// at [jG, (j + 1)G).  A grid of 64 < G <= 85 points takes, for one range,
// one lane pass of two points per lane (a Bezier patch's control points read
// once for both); for two ranges, a's and b's first 64 points side by side,
// then the G - 64 left of each (a's on lanes 0..31, b's on 32..63); for
// three, then c's first 64 points beside the three ranges' last G - 64 (a's
// on the first lanes, b's after them, c's after those): two passes of two
// points per lane for three ranges.
template <int N, typename PT>
MP_DEV void dice_group(const float4* cpa, const float4* cpb, const float4* cpc, const RangeRec& ra,
                       const RangeRec& rb, const RangeRec& rc, int k, const DevObject& ob, const DevView& vw,
                       PT* __restrict__ s_pos, int2* __restrict__ s_pxl, float* __restrict__ s_dep)
{
    constexpr int G1 = N + 1;
    constexpr int G = G1 * G1;
    constexpr int T = G - 64;   // points past the first 64 of a grid
    static_assert(G > 64 && 3 * T <= 64, "grid of 64 < G <= 85 points");
    const int lane = threadIdx.x & 63;
    auto project = [&](int f, const float P[4], bool store) {
        float e[4], q[4];
        mat4_mul(ob.mv, P, e);
        perspective_mul(vw.proj, e, q);
        // the fused form of (x * W) / 2 + W / 2 (dice.cl:71-72), as in dice_grid
        const int32_t px = f2i_sat(mad((q[0] / q[3]) * (float)vw.vs_x, 0.5f, (float)(vw.vs_x / 2)));
        const int32_t py = f2i_sat(mad((q[1] / q[3]) * (float)vw.vs_y, 0.5f, (float)(vw.vs_y / 2)));
        if (store) {
            set_pos(s_pos[f], e[0], e[1], e[2], e[3]);
            s_pxl[f] = make_int2(px, py);
            s_dep[f] = q[2] / q[3];
        }
    };
    // umin + (umax - umin) * i / N, fused as in dice_grid
    const float dua = ra.umax - ra.umin, dva = ra.vmax - ra.vmin;
    const float dub = rb.umax - rb.umin, dvb = rb.vmax - rb.vmin;
    const float duc = rc.umax - rc.umin, dvc = rc.vmax - rc.vmin;
    const bool greg = ob.type == MP_TYPE_GREGORY;
    float P0[4], P1[4];
    auto eval2 = [&](const float4* c0, const float4* c1, float u0, float v0, float u1, float v1) {
        if (greg) {
            eval_gregory(c0, u0, v0, P0);
            eval_gregory(c1, u1, v1, P1);
        } else {
            eval_bezier2(c0, c1, u0, v0, u1, v1, P0, P1);
        }
    };
    if (k == 1) {
        const int g1 = min(64 + lane, G - 1);
        eval2(cpa, cpa, mad(dua, (float)(lane % G1) / (float)N, ra.umin), mad(dva, (float)(lane / G1) / (float)N, ra.vmin),
              mad(dua, (float)(g1 % G1) / (float)N, ra.umin), mad(dva, (float)(g1 / G1) / (float)N, ra.vmin));
        project(lane, P0, true);
        project(g1, P1, 64 + lane < G);
        return;
    }
    const float tu = (float)(lane % G1) / (float)N, tv = (float)(lane / G1) / (float)N;
    eval2(cpa, cpb, mad(dua, tu, ra.umin), mad(dva, tv, ra.vmin), mad(dub, tu, rb.umin), mad(dvb, tv, rb.vmin));
    project(lane, P0, true);
    project(G + lane, P1, true);
    if (k == 2) {
        const bool hb = lane >= 32;
        const int j = lane & 31;
        const int g = 64 + min(j, T - 1);
        const float du = hb ? dub : dua, dv = hb ? dvb : dva;
        const float umin = hb ? rb.umin : ra.umin, vmin = hb ? rb.vmin : ra.vmin;
        const float u = mad(du, (float)(g % G1) / (float)N, umin), v = mad(dv, (float)(g / G1) / (float)N, vmin);
        if (greg) eval_gregory(hb ? cpb : cpa, u, v, P0);
        else eval_bezier(hb ? cpb : cpa, u, v, P0);
        project((hb ? G : 0) + g, P0, j < T);
        return;
    }
    // range t's last T points on lanes tT .. tT + T - 1
    const int t = lane < T ? 0 : lane < 2 * T ? 1 : 2;
    const int j = lane - T * t;
    const int g = 64 + min(j, T - 1);
    const float du = t == 0 ? dua : t == 1 ? dub : duc, dv = t == 0 ? dva : t == 1 ? dvb : dvc;
    const float umin = t == 0 ? ra.umin : t == 1 ? rb.umin : rc.umin;
    const float vmin = t == 0 ? ra.vmin : t == 1 ? rb.vmin : rc.vmin;
    eval2(cpc, t == 0 ? cpa : t == 1 ? cpb : cpc, mad(duc, tu, rc.umin), mad(dvc, tv, rc.vmin),
          mad(du, (float)(g % G1) / (float)N, umin), mad(dv, (float)(g / G1) / (float)N, vmin));
    project(2 * G + lane, P0, true);
    project(t * G + g, P1, j < T);
}

// NT threads per work-group: N*N (one wave per 8x8 block of quads) or 64
// (one wave dices the whole grid and shades the blocks in turn: no waiting
// between waves at the range's barriers).
template <int N, int MODE, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_raster(const RangeRec* __restrict__ canon,
                                                  const DevObject* __restrict__ objs,
                                                  const DevView* __restrict__ views,
                                                  DevParams prm, FrameCounters* fc,
                                                  float* __restrict__ quad_store, uint32_t first,
                                                  uint32_t count)
{
    constexpr int G1 = N + 1;
    constexpr int G = G1 * G1;
    constexpr int Q = N * N;
    constexpr int B = N / 8;
    // One-wave work-groups of 9 x 9 undisplaced grids dice up to three
    // consecutive ranges of one object together (three ranges in two lane
    // passes of two points per lane, two in three lane passes).
    constexpr bool kPair = MODE == 0 && NT == 64 && G > 64 && 3 * (G - 64) <= 64;
    constexpr int GS = kPair ? 3 * G : G;
    __shared__ float4 s_cpb[kPair ? 1 : 2][20];   // control points: current range / next range
    __shared__ float4 s_cq[kPair ? 2 : 1][3][20];  // groups: [buffer][range of the group]
    // eye-space x, y, z per grid point
    __shared__ EyePos s_pos[GS];
    __shared__ int2 s_pxl[GS];
    __shared__ float s_dep[GS];
    __shared__ float s_col[MODE == 1 ? 3 * 8 * G1 : 1];   // normal column sums per u column
    __shared__ float4 s_wt[MODE == 1 ? 3 * G1 : 1];       // Bernstein weights per u column / v row
    __shared__ uint4 s_win[NT / 64][kKeptOctaves];         // kept lattice windows per wave

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    // the canonical ranges [first, first + count), or all of the frame's (count 0)
    const uint32_t total = count ? first + count : fc->total_ranges;
    uint32_t slow = 0, natom = 0;
    __shared__ uint32_t s_next;

    // Range schedule.  Static (prm.raster_chunk == 0): work-group g takes
    // ranges g, g + grid, ... with no atomics, the next range known ahead.
    // Static chunks (raster_chunk <= -2): work-group g takes the chunks of
    // -raster_chunk consecutive ranges g, g + grid, ... (neighbouring ranges
    // share patches and objects in the caches), no atomics.  Dynamic: chunks
    // claimed from one counter, one atomic per chunk (a returning atomic on
    // one address serialises across the device, so the chunk must stay large
    // against the range count); raster_chunk == -1 sizes it from the range
    // count, > 0 fixes it.
    const bool blk = prm.raster_chunk <= -2 && count == 0;
    const bool dyn = prm.raster_chunk != 0 && !blk && count == 0;
    const uint32_t chunk = blk ? (uint32_t)(-prm.raster_chunk)
                         : prm.raster_chunk > 0 ? (uint32_t)prm.raster_chunk
                                                : max(8u, min(64u, total / (gridDim.x * 4u)));
    uint32_t r, r_end, cstart = blockIdx.x * chunk;
    if (dyn) {
        if (tid == 0) s_next = atomicAdd(&fc->raster_next, chunk);
        __syncthreads();
        r = s_next;
        r_end = min(r + chunk, total);
    } else if (blk) {
        r = cstart;
        r_end = min(cstart + chunk, total);
    } else {
        r = first + blockIdx.x;
        r_end = total;
    }
    NoiseWindows nw;   // lattice windows of the low octaves, kept across passes and ranges
    if (MODE == 1) nw.reset(s_win[wave]);
    // the range after x (chunk end x_end) in this work-group's schedule
    auto next = [&](uint32_t& x, uint32_t& x_end) {
        uint32_t xn = (dyn || blk) ? x + 1 : x + gridDim.x;
        if (dyn && xn >= x_end) {
            if (tid == 0) s_next = atomicAdd(&fc->raster_next, chunk);
            __syncthreads();
            xn = s_next;
            x_end = min(xn + chunk, total);
        } else if (blk && xn >= x_end) {
            cstart += gridDim.x * chunk;
            xn = cstart;
            x_end = min(cstart + chunk, total);
        }
        x = xn;
    };
    // control points of ranges x (lanes 0..19), y (20..39) and z (40..59)
    // into pf, for slot lane / 20, entry lane % 20
    const int fslot = lane / 20, fent = lane - 20 * (lane / 20);
    auto fetch3 = [&](uint32_t x, uint32_t y, uint32_t z, float4& pf) {
        const uint32_t w = fslot == 0 ? x : fslot == 1 ? y : fslot == 2 ? z : total;
        if (w < total) {
            const RangeRec rw = canon[w];
            const DevObject& ow = objs[rw.obj_depth & 0xffffffu];
            const int ncp = ow.type == MP_TYPE_GREGORY ? 20 : 16;
            if (fent < ncp) {
                pf = ldg4(ow.cps + (size_t)rw.pid * ncp + fent);
                return true;
            }
        }
        return false;
    };
    if constexpr (kPair) {
        // the window: w0 the current range, w1 and w2 the next two in the
        // schedule (total: none), e2 the end of w2's chunk
        uint32_t w0 = r, w1 = total, w2 = total, e2 = r_end;
        if (w0 < total) {
            w1 = w0;
            next(w1, e2);
        }
        if (w1 < total) {
            w2 = w1;
            next(w2, e2);
        }
        float4 pf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (fetch3(w0, w1, w2, pf)) s_cq[0][fslot][fent] = pf;
        __syncthreads();
        int qb = 0;
        while (w0 < total) {
            const RangeRec ra = canon[w0];
            const uint32_t oa = ra.obj_depth & 0xffffffu;
            const DevObject& ob = objs[oa];
            const DevView& vw = views[ob.view];
            RangeRec rb = ra, rc = ra;
            int k = 1;   // consecutive ranges of this object diced together
            if (w1 < total) {
                rb = canon[w1];
                if ((rb.obj_depth & 0xffffffu) == oa) {
                    k = 2;
                    if (w2 < total) {
                        rc = canon[w2];
                        if ((rc.obj_depth & 0xffffffu) == oa) k = 3;
                    }
                }
            }
            // the next window (the ranges after the k diced now); its control
            // points load now and go to the other buffer after the dice
            uint32_t x = w2;
            auto adv = [&]() {
                if (x < total) next(x, e2);
                else x = total;
                return x;
            };
            const uint32_t m3 = adv();
            const uint32_t m4 = k >= 2 ? adv() : total;
            const uint32_t m5 = k == 3 ? adv() : total;
            const uint32_t n0 = k == 1 ? w1 : k == 2 ? w2 : m3;
            const uint32_t n1 = k == 1 ? w2 : k == 2 ? m3 : m4;
            const uint32_t n2 = k == 1 ? m3 : k == 2 ? m4 : m5;
            const bool have_pf = fetch3(n0, n1, n2, pf);
            dice_group<N>(s_cq[qb][0], s_cq[qb][1], s_cq[qb][2], ra, rb, rc, k, ob, vw, s_pos, s_pxl, s_dep);
            __syncthreads();
            if (have_pf) s_cq[qb ^ 1][fslot][fent] = pf;
            // one shading code path for the ranges (instruction cache)
#pragma nounroll
            for (int j = 0; j < k; ++j)
                shade_sample<N>(s_pos + j * G, s_pxl + j * G, s_dep + j * G, lane & 7, lane >> 3, lane, vw, prm,
                                j == 0 ? w0 : j == 1 ? w1 : w2, quad_store, slow, natom);
            w0 = n0;
            w1 = n1;
            w2 = n2;
            qb ^= 1;
            __syncthreads();
        }
        r = total;   // the loop below has nothing left
    }
    if (!kPair && r < total) {
        const RangeRec rec0 = canon[r];
        const DevObject& ob0 = objs[rec0.obj_depth & 0xffffffu];
        const int ncp0 = ob0.type == MP_TYPE_GREGORY ? 20 : 16;
        if (tid < ncp0) s_cpb[0][tid] = ldg4(ob0.cps + (size_t)rec0.pid * ncp0 + tid);
    }
    __syncthreads();
    int buf = 0;
    float4 pf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    while (r < total) {
        // next range (the following chunk is claimed one range ahead) and its
        // control points: loaded now, stored to the other LDS buffer after
        // this range, so the load latency hides behind this range's work
        uint32_t rn = r, rn_end = r_end;
        next(rn, rn_end);
        bool have_pf = false;
        if (rn < total) {
            const RangeRec nrec = canon[rn];
            const DevObject& nob = objs[nrec.obj_depth & 0xffffffu];
            const int nncp = nob.type == MP_TYPE_GREGORY ? 20 : 16;
            if (tid < nncp) {
                pf = ldg4(nob.cps + (size_t)nrec.pid * nncp + tid);
                have_pf = true;
            }
        }
        const float4* s_cp = s_cpb[buf];
        const RangeRec rec = canon[r];
        const DevObject& ob = objs[rec.obj_depth & 0xffffffu];
        const DevView& vw = views[ob.view];

        // ---- dice (dice.cl:38-79) ----
        dice_grid<N, MODE, NT>(s_cp, s_col, s_wt, rec, ob, vw, prm, s_pos, s_pxl, s_dep, nw);
        __syncthreads();

        // The next range's control points go to the other buffer now: the
        // load was issued before the dice and has landed, and no atomic of
        // this range is in flight yet (the wait counter for vector memory
        // is in order, so a later wait would also wait for the atomics).
        if (have_pf) s_cpb[buf ^ 1][tid] = pf;

        // one wave per 8x8 block; a work-group of fewer waves takes several
        for (int blk = wave; blk < Q / 64; blk += NT / 64) {
            const int u = (blk % B) * 8 + (lane & 7);
            const int v = (blk / B) * 8 + (lane >> 3);
            shade_sample<N>(s_pos, s_pxl, s_dep, u, v, lane, vw, prm, r, quad_store, slow, natom);
        }

        buf ^= 1;
        r = rn;
        r_end = rn_end;
        __syncthreads();
    }
    if (slow) atomicAdd(&fc->slow_blocks, slow);
    natom = (uint32_t)wave_sum_i((int)natom);
    if (lane == 0 && natom) atomicAdd(stripe(fc, blockIdx.x, kStripeAtomics), (unsigned long long)natom);
}

template <int N, int MODE, int NT>
static int raster_occupancy()
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_raster<N, MODE, NT>, NT, 0) != hipSuccess)
        blocks = 1;
    return blocks < 1 ? 1 : blocks;
}

static int raster_mode(const DevParams& prm)
{
    return prm.displacement ? 1 : 0;
}

int raster_grid(int patch_size, int num_cus, const DevParams& prm)
{
    const int mode = raster_mode(prm);
    int per_cu = 1;
#define MP_OCC(NN, NT)                                                                     \
    per_cu = mode == 0 ? raster_occupancy<NN, 0, NT>() : raster_occupancy<NN, 1, NT>();
    switch (patch_size) {
    case 8: MP_OCC(8, 64) break;
    case 16:
        if (prm.raster_one_wave) { MP_OCC(16, 64) } else { MP_OCC(16, 256) }
        break;
    case 32: MP_OCC(32, 1024) break;
    default: break;
    }
#undef MP_OCC
    return per_cu * num_cus;
}

template <int N, int NT>
static void launch_raster_n(int mode, int grid, hipStream_t s, const RangeRec* canon,
                            const DevObject* objs, const DevView* views, const DevParams& prm,
                            FrameCounters* fc, float* quad_store, uint32_t first, uint32_t count)
{
    if (mode == 0)
        hipLaunchKernelGGL((k_raster<N, 0, NT>), dim3(grid), dim3(NT), 0, s, canon, objs, views, prm, fc, quad_store,
                           first, count);
    else
        hipLaunchKernelGGL((k_raster<N, 1, NT>), dim3(grid), dim3(NT), 0, s, canon, objs, views, prm, fc, quad_store,
                           first, count);
}

hipError_t launch_raster(int patch_size, const RangeRec* canon, const DevObject* objs,
                         const DevView* views, const DevParams& prm, FrameCounters* fc,
                         float* quad_store, int grid, hipStream_t s, uint32_t first, uint32_t count)
{
    const int mode = raster_mode(prm);
    if (count) grid = (int)std::min<uint32_t>((uint32_t)grid, count);   // a window needs no more groups than ranges
    switch (patch_size) {
    case 8: launch_raster_n<8, 64>(mode, grid, s, canon, objs, views, prm, fc, quad_store, first, count); break;
    case 16:
        if (prm.raster_one_wave)
            launch_raster_n<16, 64>(mode, grid, s, canon, objs, views, prm, fc, quad_store, first, count);
        else
            launch_raster_n<16, 256>(mode, grid, s, canon, objs, views, prm, fc, quad_store, first, count);
        break;
    case 32: launch_raster_n<32, 1024>(mode, grid, s, canon, objs, views, prm, fc, quad_store, first, count); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The canonical range base of one view, and the frame's range total (the
// CPU split method's host order replaces the device ordering stage).
__global__ void k_set_view_base(DevView* views, int view, uint32_t base, FrameCounters* fc, uint32_t total)
{
    if (threadIdx.x == 0) {
        if (view >= 0) views[view].range_base = base;
        if (fc) fc->total_ranges = total;
    }
}

hipError_t launch_set_view_base(DevView* views, int view, uint32_t base, FrameCounters* fc, uint32_t total,
                                hipStream_t s)
{
    hipLaunchKernelGGL(k_set_view_base, dim3(1), dim3(64), 0, s, views, view, base, fc, total);
    return hipGetLastError();
}

// Debug tap: the diced grids of canonical ranges [first, first + n) of the
// last frame, computed by dice_grid as the raster does (one 64-lane wave per
// range), written as eye-space position (x, y, z, w), sub-pixel coordinate
// and depth per grid point, range-major, point index nu + nv * (N + 1)
// (dice.cl:21-24, 75-79).
template <int N, int MODE>
__global__ __launch_bounds__(64) void k_read_grids(const RangeRec* __restrict__ canon,
                                                   const DevObject* __restrict__ objs,
                                                   const DevView* __restrict__ views, DevParams prm,
                                                   uint32_t first, uint32_t n, float4* __restrict__ pos,
                                                   int2* __restrict__ pxl, float* __restrict__ dep)
{
    constexpr int G1 = N + 1;
    constexpr int G = G1 * G1;
    __shared__ float4 s_cp[20];
    __shared__ float4 s_pos[G];
    __shared__ int2 s_pxl[G];
    __shared__ float s_dep[G];
    __shared__ float s_col[MODE == 1 ? 3 * 8 * G1 : 1];
    __shared__ float4 s_wt[MODE == 1 ? 3 * G1 : 1];
    __shared__ uint4 s_win[kKeptOctaves];
    const int tid = threadIdx.x;
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
        const RangeRec rec = canon[first + k];
        const DevObject& ob = objs[rec.obj_depth & 0xffffffu];
        const int ncp = ob.type == MP_TYPE_GREGORY ? 20 : 16;
        if (tid < ncp) s_cp[tid] = ldg4(ob.cps + (size_t)rec.pid * ncp + tid);
        __syncthreads();
        NoiseWindows nw;
        if (MODE == 1) nw.reset(s_win);
        dice_grid<N, MODE, 64>(s_cp, s_col, s_wt, rec, ob, views[ob.view], prm, s_pos, s_pxl, s_dep, nw);
        __syncthreads();
        for (int g = tid; g < G; g += 64) {
            pos[(size_t)k * G + g] = s_pos[g];
            pxl[(size_t)k * G + g] = s_pxl[g];
            dep[(size_t)k * G + g] = s_dep[g];
        }
        __syncthreads();
    }
}

hipError_t launch_read_grids(int patch_size, const RangeRec* canon, const DevObject* objs, const DevView* views,
                             const DevParams& prm, uint32_t first, uint32_t n, float4* pos, int2* pxl, float* dep,
                             hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const dim3 grid(std::min<uint32_t>(n, 4096u)), block(64);
    const bool disp = prm.displacement != 0;
#define MP_GRIDS(NN)                                                                                        \
    if (disp) hipLaunchKernelGGL((k_read_grids<NN, 1>), grid, block, 0, s, canon, objs, views, prm, first, n, \
                                 pos, pxl, dep);                                                            \
    else hipLaunchKernelGGL((k_read_grids<NN, 0>), grid, block, 0, s, canon, objs, views, prm, first, n, pos,  \
                            pxl, dep);
    switch (patch_size) {
    case 8: MP_GRIDS(8) break;
    case 16: MP_GRIDS(16) break;
    case 32: MP_GRIDS(32) break;
    default: return hipErrorInvalidValue;
    }
#undef MP_GRIDS
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// resolve: key -> display colour, FP32 depth, id; clears the key

// Display byte of a linear value: tab[k - 1] is the smallest value whose
// byte is at least k (k = 1..255, increasing), so the byte is the k with
// tab[k - 1] <= c < tab[k]; NaN and values below tab[0] give 0.  The hardware
// log2 / exp2 estimate the byte, and the two loops walk to the k the table
// defines from any start (within one step in practice: usually no step).
MP_DEV uint32_t display_byte(float c, const float* tab)
{
    if (!(c >= tab[0])) return 0;
    if (c >= tab[254]) return 255;
    const float g = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(c) * 0.454545f);
    int k = min(max((int)(g * 255.0f + 0.5f), 1), 254);
    while (c < tab[k - 1]) --k;    // stops at k >= 1: c >= tab[0]
    while (c >= tab[k]) ++k;       // stops at k <= 254: c < tab[254]
    return (uint32_t)k;
}

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_resolve(const DevView* __restrict__ views,
                                                 const RangeRec* __restrict__ canon,
                                                 const DevObject* __restrict__ objs,
                                                 const float* __restrict__ quad_store, int Q,
                                                 int write_rgba32, const float* __restrict__ display_tab,
                                                 FrameEnd fe)
{
    __shared__ float s_tab[256];
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 64 && fe.fc) {   // the resolve reads no counter
        fold_stripes(fe.fc);
        if (threadIdx.x == 0) end_frame(fe);
    }
    if (blockIdx.y == 0) {
        // per-patch counts of a ranked frame (nothing reads them after the order)
        const uint32_t i = blockIdx.x * 256u + threadIdx.x;
        for (uint32_t k = i; k < fe.n_counts; k += gridDim.x * 256u) fe.counts[k] = 0u;
    }
    if (threadIdx.x < 255) s_tab[threadIdx.x] = display_tab[threadIdx.x];
    __syncthreads();
    const DevView& vw = views[blockIdx.y];
    // four consecutive keys per thread (half a tile row): the key loads and
    // then the gathers of the four pixels are in flight together
    const uint32_t n4 = (uint32_t)vw.grid_w * (uint32_t)vw.grid_h * 16u;
    const float ac[4] = { 0.015f, 0.015f, 0.015f, 1.0f };
    const uint32_t qshift = (uint32_t)__builtin_ctz((unsigned)Q);   // Q = N*N is a power of two
    const size_t qbase = (size_t)vw.range_base * Q;
    // a view of one object takes its colour without the range -> object gather
    const int solo = vw.solo_obj;
    // whole-row stores of the four pixels need 16-byte aligned rows
    const bool vec = ((reinterpret_cast<uintptr_t>(vw.rgba8) | reinterpret_cast<uintptr_t>(vw.depth) |
                       reinterpret_cast<uintptr_t>(vw.ids)) & 15u) == 0 && (vw.width & 3) == 0;
    for (uint32_t j = blockIdx.x * 256 + threadIdx.x; j < n4; j += gridDim.x * 256) {
        const uint32_t i = j * 4u;
        MP_GLOBAL u64x2* kp = gptr(reinterpret_cast<u64x2*>(vw.keys + i));
        const u64x2 ka = kp[0], kb = kp[1];
        const u64x2 empty = { MP_EMPTY_KEY, MP_EMPTY_KEY };
        // only key pairs a sample wrote are cleared (most of a frame's keys
        // stay empty: their writes would be most of the kernel's traffic)
        if ((ka.x & ka.y) != MP_EMPTY_KEY) kp[0] = empty;
        if ((kb.x & kb.y) != MP_EMPTY_KEY) kp[1] = empty;
        const unsigned long long key[4] = { ka.x, ka.y, kb.x, kb.y };
        const uint32_t t = i >> 6, l = i & 63;
        const int X = (int)(t % (uint32_t)vw.grid_w) * 8 + (int)(l & 7);
        const int Y = (int)(t / (uint32_t)vw.grid_w) * 8 + (int)(l >> 3);
        if (X >= vw.width || Y >= vw.height) continue;
        float s[4];
        int obj[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            s[k] = 0.0f;
            obj[k] = solo;
            if (key[k] != MP_EMPTY_KEY) {
                const uint32_t id = (uint32_t)key[k];
                s[k] = quad_store[qbase + id];
                if (solo < 0) obj[k] = (int)(canon[vw.range_base + (id >> qshift)].obj_depth & 0xffffffu);
            }
        }
        uint32_t bytes[4], ids[4];
        float dep[4], c[4][4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            dep[k] = 1.0f;
            ids[k] = 0xFFFFFFFFu;
            bytes[k] = vw.clear_rgba8;
            if (key[k] == MP_EMPTY_KEY) {
#pragma unroll
                for (int m = 0; m < 4; ++m) c[k][m] = vw.clear_linear[m];
            } else {
                ids[k] = (uint32_t)key[k];
                dep[k] = unmono_bits((uint32_t)(key[k] >> 32));
                const DevObject& ob = objs[obj[k]];
#pragma unroll
                for (int m = 0; m < 4; ++m) c[k][m] = mad(s[k], ob.color[m], ac[m] * ob.color[m]) + 0.0f;
                bytes[k] = display_byte(c[k][0], s_tab) | (display_byte(c[k][1], s_tab) << 8) |
                           (display_byte(c[k][2], s_tab) << 16) | (255u << 24);
            }
        }
        const size_t o = (size_t)Y * vw.width + X;
        if (vec) {
            // width % 4 == 0 and X % 4 == 0: all four pixels are in the row
            const u32x4 b4 = { bytes[0], bytes[1], bytes[2], bytes[3] };
            const f32x4 d4 = { dep[0], dep[1], dep[2], dep[3] };
            const u32x4 i4 = { ids[0], ids[1], ids[2], ids[3] };
            *gptr(reinterpret_cast<u32x4*>(vw.rgba8 + 4 * o)) = b4;
            *gptr(reinterpret_cast<f32x4*>(vw.depth + o)) = d4;
            *gptr(reinterpret_cast<u32x4*>(vw.ids + o)) = i4;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (X + k >= vw.width) break;
                *gptr(reinterpret_cast<uint32_t*>(vw.rgba8) + o + k) = bytes[k];
                gptr(vw.depth)[o + k] = dep[k];
                gptr(vw.ids)[o + k] = ids[k];
            }
        }
        if (write_rgba32) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (X + k < vw.width)
                    stg4(reinterpret_cast<float4*>(vw.rgba32) + o + k, make_float4(c[k][0], c[k][1], c[k][2], c[k][3]));
        }
    }
}

hipError_t launch_resolve(const DevView* views, int n_views, int max_pixels,
                          const RangeRec* canon, const DevObject* objs, const float* quad_store,
                          int patch_size, int write_rgba32, const float* display_tab, hipStream_t s,
                          const FrameEnd& fe)
{
    // four keys per thread: every pixel's gather chain (range, object,
    // Lambert term) is in flight at once
    int bx = ((max_pixels + 3) / 4 + 255) / 256;
    if (bx > 65535) bx = 65535;
    if (bx < 1) bx = 1;
    FrameEnd fr = fe;
    fr.resolve_clears = true;
    hipLaunchKernelGGL(k_resolve, dim3(bx, n_views), dim3(256), 0, s, views, canon, objs,
                       quad_store, patch_size * patch_size, write_rgba32, display_tab, fr);
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

size_t split_persistent_lds(int stack_cap)
{
    return ((size_t)stack_cap + kDrawBuf) * sizeof(RangeRec);
}

int split_persistent_occupancy(size_t lds_bytes, int waves)
{
    int blocks = 0;
    const hipError_t e = waves == 3
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_split_persistent<3>, 64 * kSplitWaves, lds_bytes)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_split_persistent<2>, 64 * kSplitWaves, lds_bytes);
    if (e != hipSuccess)
        blocks = 1;
    return blocks < 1 ? 1 : blocks;
}

}  // namespace mp

// ---------------------------------------------------------------------------
// Reverse pass of the differentiable extension.  Loss gradient dL/dC per
// pixel (linear RGB) -> control points, displacement amplitude and albedo,
// with visibility held fixed to the forward visible-id buffer and the grids
// recomputed per range (no stored forward state):
//   C_c = (0.015 + s) * albedo_c,  s = max(n.l, 0),  n = unit(cross(dv, du))
//   corner e = MV * (P + amp * noise(P) * unit(cross(dP/dv, dP/du)))

namespace mp {

__global__ void k_xyz_to_cps(const float* __restrict__ xyz, float4* __restrict__ cps, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cps[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 1.0f);
}

hipError_t launch_xyz_to_cps(const float* xyz, float4* cps, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_xyz_to_cps, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, xyz, cps, n);
    return hipGetLastError();
}

// flag every range that owns a visible micropolygon
__global__ __launch_bounds__(256) void k_adj_mark(const DevView* __restrict__ views, int Q,
                                                  uint32_t* __restrict__ visible)
{
    const DevView& vw = views[blockIdx.y];
    const uint32_t n = (uint32_t)vw.width * (uint32_t)vw.height;
    for (uint32_t o = blockIdx.x * 256 + threadIdx.x; o < n; o += gridDim.x * 256) {
        const uint32_t id = gptr(vw.ids)[o];
        if (id != 0xFFFFFFFFu) visible[vw.range_base + id / (uint32_t)Q] = 1u;
    }
}

// list the ranges that own a visible micropolygon (any order): one counter
// atomic per wave
__global__ __launch_bounds__(256) void k_adj_compact(const uint32_t* __restrict__ visible,
                                                     const FrameCounters* __restrict__ fc,
                                                     uint32_t* __restrict__ list, uint32_t* __restrict__ count)
{
    const uint32_t total = fc->total_ranges;
    const int lane = threadIdx.x & 63;
    for (uint32_t base = blockIdx.x * 256u; base < total; base += gridDim.x * 256u) {
        const uint32_t r = base + threadIdx.x;
        const bool v = r < total && visible[r] != 0u;
        const uint64_t m = __ballot(v);
        uint32_t slot = 0;
        if (lane == 0 && m) slot = atomicAdd(count, (uint32_t)__popcll(m));
        slot = (uint32_t)__shfl((int)slot, 0);
        if (v) list[slot + lanes_below(m)] = r;
    }
}

struct V3 { float x, y, z; };
MP_DEV V3 mk(float x, float y, float z) { V3 r = { x, y, z }; return r; }
MP_DEV V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
MP_DEV V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
MP_DEV V3 operator*(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
MP_DEV float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MP_DEV V3 vcross(V3 a, V3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// noise value and gradient (same lattice and octave sum as noise3).
// full_wave: every lane of the wave runs the call, so the lattice corners can
// come from the lane window (one hash per lane per octave, window_corners);
// otherwise, or where a cell falls outside the window, each lane hashes its
// eight corners.
template <bool full_wave>
MP_DEV float noise3_grad(const DevParams& prm, float x, float y, float z, V3& grad)
{
    float acc = 0.0f, amp = 1.0f, norm = 0.0f, freq = prm.disp_frequency;
    V3 g = mk(0.f, 0.f, 0.f);
    for (int k = 0; k < prm.disp_octaves; ++k) {
        const uint32_t seed = prm.disp_seed + (uint32_t)k;
        const float px = x * freq, py = y * freq, pz = z * freq;
        const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
        const int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
        const float tx = px - fx, ty = py - fy, tz = pz - fz;
        const float wx = fade(tx), wy = fade(ty), wz = fade(tz);
        float c[8];
        if (!full_wave || !window_corners(ix, iy, iz, seed, c)) {
            c[0] = lattice(ix, iy, iz, seed);         c[1] = lattice(ix + 1, iy, iz, seed);
            c[2] = lattice(ix, iy + 1, iz, seed);     c[3] = lattice(ix + 1, iy + 1, iz, seed);
            c[4] = lattice(ix, iy, iz + 1, seed);     c[5] = lattice(ix + 1, iy, iz + 1, seed);
            c[6] = lattice(ix, iy + 1, iz + 1, seed); c[7] = lattice(ix + 1, iy + 1, iz + 1, seed);
        }
        const float c000 = c[0], c100 = c[1], c010 = c[2], c110 = c[3];
        const float c001 = c[4], c101 = c[5], c011 = c[6], c111 = c[7];
        const float x00 = lerpf_(c000, c100, wx), x10 = lerpf_(c010, c110, wx);
        const float x01 = lerpf_(c001, c101, wx), x11 = lerpf_(c011, c111, wx);
        const float y0 = lerpf_(x00, x10, wy), y1 = lerpf_(x01, x11, wy);
        acc = acc + amp * lerpf_(y0, y1, wz);
        const float dwx = lerpf_(lerpf_(c100 - c000, c110 - c010, wy), lerpf_(c101 - c001, c111 - c011, wy), wz);
        const float dwy = lerpf_(x10 - x00, x11 - x01, wz);
        const float dwz = y1 - y0;
        const float s = amp * freq;
        g = g + mk(dwx * (6.0f * tx * (1.0f - tx)), dwy * (6.0f * ty * (1.0f - ty)),
                   dwz * (6.0f * tz * (1.0f - tz))) * s;
        norm = norm + amp;
        amp = amp * 0.5f;
        freq = freq * 2.0f;
    }
    grad = g * (1.0f / norm);
    return acc / norm;
}

template <int N>
__global__ __launch_bounds__(N * N) void k_adjoint(const RangeRec* __restrict__ canon,
                                                   const DevObject* __restrict__ objs,
                                                   const DevView* __restrict__ views, DevParams prm,
                                                   const FrameCounters* __restrict__ fc,
                                                   const float* __restrict__ quad_store,
                                                   const float* __restrict__ dl_dcolor,
                                                   const uint32_t* __restrict__ vis_list,
                                                   const uint32_t* __restrict__ vis_count,
                                                   float* __restrict__ grad_cps, float* __restrict__ grad_color,
                                                   float* __restrict__ grad_amp)
{
    constexpr int G1 = N + 1, G = G1 * G1, Q = N * N, W = Q / 64, PPT = (G + Q - 1) / Q;
    __shared__ float4 s_cp[16];
    __shared__ float4 s_e[G];
    __shared__ int2 s_pxl[G];
    __shared__ float s_ge[G][3];
    __shared__ float s_red[W][4];
    // per grid point: the forward state the reverse pass needs (dP/du,
    // dP/dv, unit normal, noise gradient, |normal|, noise, displacement),
    // then, in the same row, dL/d(P, dP/du, dP/dv)
    __shared__ float s_gp[G][16];
    __shared__ float s_row[G1][4][6];     // per grid row and i: the u sums
    __shared__ float s_wu[G1][8];         // a_i(u), a_i'(u) per grid column
    __shared__ float s_wv[G1][8];         // b_j(v), b_j'(v) per grid row
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sh = prm.subpixel_bits;
    const V3 L = [] { float lx = 4.0f, ly = 3.0f, lz = 8.0f; const float inv = 1.0f / sqrtf(89.0f);
                      return mk(lx * inv, ly * inv, lz * inv); }();

    // visible ranges by static stride over their list; the next one's index
    // is loaded a range ahead
    const uint32_t count = *vis_count;
    uint32_t i = blockIdx.x;
    uint32_t r_ahead = i < count ? vis_list[i] : 0u;
    for (; i < count; i += gridDim.x) {
        const uint32_t r = r_ahead;
        if (i + gridDim.x < count) r_ahead = vis_list[i + gridDim.x];
        const RangeRec rec = canon[r];
        const DevObject& ob = objs[rec.obj_depth & 0xffffffu];
        const DevView& vw = views[ob.view];
        if (tid < 16) s_cp[tid] = ldg4(ob.cps + (size_t)rec.pid * 16 + tid);
        if (tid < G1) {
            const float uu = rec.umin + (rec.umax - rec.umin) * ((float)tid / (float)N);
            const float su = 1.0f - uu;
            bern_u(uu, su, &s_wu[tid][0]);
            dbern(uu, su, &s_wu[tid][4]);
        } else if (tid >= 64 - G1 && tid < 64) {
            const int nv = tid - (64 - G1);
            const float vv = rec.vmin + (rec.vmax - rec.vmin) * ((float)nv / (float)N);
            const float sv = 1.0f - vv;
            bern_v(vv, sv, &s_wv[nv][0]);
            dbern(vv, sv, &s_wv[nv][4]);
        }
        __syncthreads();

        // forward state of the grid points owned by this thread (to LDS)
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            // every lane of a wave that owns any point runs (lanes past the
            // last point repeat it and store nothing): the noise lattice
            // comes from the lane window
            const int g = tid + k * Q;
            // the control points are read again from LDS for each point
            // instead of being held in 64 registers across the loop
            asm volatile("" ::: "memory");
            if ((tid & ~63) + k * Q < G) {
                const int gc = min(g, G - 1);
                const int nu = gc % G1, nv = gc / G1;
                const float tu = (float)nu / (float)N, tv = (float)nv / (float)N;
                const float uu = rec.umin + (rec.umax - rec.umin) * tu;
                const float vv = rec.vmin + (rec.vmax - rec.vmin) * tv;
                float P[4];
                eval_bezier(s_cp, uu, vv, P);
                float su = 1.0f - uu, sv = 1.0f - vv, a[4], b[4], da[4], db[4];
                bern_u(uu, su, a); bern_v(vv, sv, b); dbern(uu, su, da); dbern(vv, sv, db);
                V3 pu = mk(0.f, 0.f, 0.f), pv = mk(0.f, 0.f, 0.f);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const float4 C = s_cp[4 * i + j];
                        pu = pu + mk(C.x, C.y, C.z) * (da[i] * b[j]);
                        pv = pv + mk(C.x, C.y, C.z) * (a[i] * db[j]);
                    }
                const V3 Nn = vcross(pv, pu);
                const float len = sqrtf(vdot(Nn, Nn));
                const V3 h = len > 0.0f ? Nn * (1.0f / len) : mk(0.f, 0.f, 0.f);
                V3 gz = mk(0.f, 0.f, 0.f);
                float nzz = 0.0f, d = 0.0f;
                if (prm.displacement) {
                    nzz = noise3_grad<true>(prm, P[0], P[1], P[2], gz);
                    d = ob.disp_amplitude * nzz;
                }
                if (g < G) {
                    float* st = s_gp[g];
                    st[0] = pu.x; st[1] = pu.y; st[2] = pu.z; st[3] = pv.x; st[4] = pv.y; st[5] = pv.z;
                    st[6] = h.x; st[7] = h.y; st[8] = h.z; st[9] = gz.x; st[10] = gz.y; st[11] = gz.z;
                    st[12] = len; st[13] = nzz; st[14] = d;
                }
                // forward displaced eye position and pixel position: P + d * unit
                // normal with the same operations as the forward dice
                float Pd[4] = { P[0], P[1], P[2], P[3] };
                if (prm.displacement) { Pd[0] = mad(d, h.x, P[0]); Pd[1] = mad(d, h.y, P[1]); Pd[2] = mad(d, h.z, P[2]); }
                float e[4], q[4];
                mat4_mul(ob.mv, Pd, e);
                mat4_mul(vw.proj, e, q);
                if (g < G) {
                    s_e[g] = make_float4(e[0], e[1], e[2], e[3]);
                    s_pxl[g] = make_int2(f2i_sat((q[0] / q[3]) * (float)vw.vs_x / 2.0f + (float)(vw.vs_x / 2)),
                                         f2i_sat((q[1] / q[3]) * (float)vw.vs_y / 2.0f + (float)(vw.vs_y / 2)));
                    s_ge[g][0] = s_ge[g][1] = s_ge[g][2] = 0.0f;
                }
            }
        }
        __syncthreads();

        // micropolygon: gather dL/dC from the pixels it is visible in
        const int u = tid % N, v = tid / N, g0 = u + v * G1;
        const uint32_t myid = (r - vw.range_base) * (uint32_t)Q + (uint32_t)tid;
        // the forward Lambert term (the store holds it for every
        // micropolygon that owns a pixel; other entries are stale and unused)
        const float s_fwd = quad_store[(size_t)r * Q + tid];
        float gs = 0.0f, ga0 = 0.0f, ga1 = 0.0f, ga2 = 0.0f;
        bool owns = false;
        if (s_fwd == s_fwd) {
            const int2 x0 = s_pxl[g0], x1 = s_pxl[g0 + 1], x2 = s_pxl[g0 + G1], x3 = s_pxl[g0 + G1 + 1];
            const int32_t qminx = min(min(x0.x, x1.x), min(x2.x, x3.x)), qmaxx = max(max(x0.x, x1.x), max(x2.x, x3.x));
            const int32_t qminy = min(min(x0.y, x1.y), min(x2.y, x3.y)), qmaxy = max(max(x0.y, x1.y), max(x2.y, x3.y));
            // in the sampler's exact-integer regime only pixel corners inside
            // the quad's box can be covered: start at the first one
            const int32_t lim = 1 << 18;
            const bool small = (qminx > -lim) && (qmaxx < lim) && (qminy > -lim) && (qmaxy < lim) &&
                               (qmaxx - qminx < (1 << 11)) && (qmaxy - qminy < (1 << 11));
            const int32_t rnd = small ? (1 << sh) - 1 : 0;
            const int X0 = max(0, (qminx + rnd) >> sh);
            const int X1 = min(vw.width - 1, qmaxx >> sh);
            const int Y0 = max(0, (qminy + rnd) >> sh);
            const int Y1 = min(vw.height - 1, qmaxy >> sh);
            {
                for (int Y = Y0; Y <= Y1; ++Y)
                    for (int X = X0; X <= X1; ++X) {
                        const size_t o = (size_t)Y * vw.width + X;
                        if (gptr(vw.ids)[o] != myid) continue;
                        owns = true;
                        const float* dl = dl_dcolor + ((size_t)vw.pixel_offset + o) * 4;
                        gs += dl[0] * ob.color[0] + dl[1] * ob.color[1] + dl[2] * ob.color[2];
                        ga0 += dl[0]; ga1 += dl[1]; ga2 += dl[2];
                    }
            }
        }
        const float acs = owns ? 0.015f + s_fwd : 0.0f;
        ga0 *= acs; ga1 *= acs; ga2 *= acs;
        if (gs != 0.0f) {
            const float4 a0 = s_e[g0], a1 = s_e[g0 + 1], a2 = s_e[g0 + G1], a3 = s_e[g0 + G1 + 1];
            const V3 p0 = mk(a0.x, a0.y, a0.z), p1 = mk(a1.x, a1.y, a1.z);
            const V3 p2 = mk(a2.x, a2.y, a2.z), p3 = mk(a3.x, a3.y, a3.z);
            const V3 du = ((p1 - p0) + (p3 - p2)) * 0.5f;
            const V3 dv = ((p2 - p0) + (p3 - p1)) * 0.5f;
            const V3 m = vcross(dv, du);
            const float ml = sqrtf(vdot(m, m));
            if (ml > 0.0f) {
                const V3 n = m * (1.0f / ml);
                if (vdot(n, L) > 0.0f) {
                    const V3 gm = (L - n * vdot(n, L)) * (gs / ml);
                    const V3 gdv = vcross(du, gm), gdu = vcross(gm, dv);
                    const V3 c0 = (gdu + gdv) * -0.5f, c1 = (gdu - gdv) * 0.5f;
                    const V3 c2 = (gdv - gdu) * 0.5f, c3 = (gdu + gdv) * 0.5f;
                    const int gi[4] = { g0, g0 + 1, g0 + G1, g0 + G1 + 1 };
                    const V3 cc[4] = { c0, c1, c2, c3 };
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        atomicAdd(&s_ge[gi[k]][0], cc[k].x);
                        atomicAdd(&s_ge[gi[k]][1], cc[k].y);
                        atomicAdd(&s_ge[gi[k]][2], cc[k].z);
                    }
                }
            }
        }
        __syncthreads();

        // grid points: eye -> displaced -> surface; the gradients with
        // respect to the surface point and its two derivatives go to LDS
        float gamp = 0.0f;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int g = tid + k * Q;
            if (g >= G) continue;
            const V3 ge = mk(s_ge[g][0], s_ge[g][1], s_ge[g][2]);
            float* st = s_gp[g];
            const V3 Pu = mk(st[0], st[1], st[2]), Pv = mk(st[3], st[4], st[5]);
            const V3 nh = mk(st[6], st[7], st[8]), gn = mk(st[9], st[10], st[11]);
            const float nlen = st[12], nzv = st[13], dd = st[14];
            V3 gP = mk(0.f, 0.f, 0.f), gPu = mk(0.f, 0.f, 0.f), gPv = mk(0.f, 0.f, 0.f);
            if (ge.x != 0.0f || ge.y != 0.0f || ge.z != 0.0f) {
                const float* M = ob.mv;
                const V3 gPd = mk(M[0] * ge.x + M[1] * ge.y + M[2] * ge.z, M[4] * ge.x + M[5] * ge.y + M[6] * ge.z,
                                  M[8] * ge.x + M[9] * ge.y + M[10] * ge.z);
                gP = gPd;
                if (prm.displacement) {
                    const float nd = vdot(nh, gPd);
                    gamp += nzv * nd;
                    gP = gP + gn * (ob.disp_amplitude * nd);
                    if (nlen > 0.0f) {
                        const V3 gnh = gPd * dd;
                        const V3 gN = (gnh - nh * vdot(nh, gnh)) * (1.0f / nlen);
                        gPv = vcross(Pu, gN);
                        gPu = vcross(gN, Pv);
                    }
                }
            }
            st[0] = gP.x;  st[1] = gP.y;  st[2] = gP.z;
            st[3] = gPu.x; st[4] = gPu.y; st[5] = gPu.z;
            st[6] = gPv.x; st[7] = gPv.y; st[8] = gPv.z;
        }
        __syncthreads();

        // control points, separably over the grid's rows and columns:
        //   dL/dP_ij = sum_v b_j(v) [sum_u a_i(u) gP + a_i'(u) gPu] + b_j'(v) [sum_u a_i(u) gPv]
        // rows first: s_row[nv][i][q], q < 3 the bracket, q >= 3 the gPv sum
        for (int o = tid; o < G1 * 24; o += Q) {
            const int nv = o / 24, rem = o % 24, i = rem / 6, q = rem % 6;
            const float* gp = &s_gp[nv * G1][0];
            float t = 0.0f;
            if (q < 3) {
                for (int nu = 0; nu < G1; ++nu)
                    t += s_wu[nu][i] * gp[nu * 16 + q] + s_wu[nu][4 + i] * gp[nu * 16 + 3 + q];
            } else {
                for (int nu = 0; nu < G1; ++nu) t += s_wu[nu][i] * gp[nu * 16 + 3 + q];
            }
            s_row[nv][i][q] = t;
        }
        // amplitude and albedo sums of each wave, four chains side by side
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            gamp += __shfl_xor(gamp, o);
            ga0 += __shfl_xor(ga0, o);
            ga1 += __shfl_xor(ga1, o);
            ga2 += __shfl_xor(ga2, o);
        }
        if (lane == 0) { s_red[wave][0] = gamp; s_red[wave][1] = ga0; s_red[wave][2] = ga1; s_red[wave][3] = ga2; }
        __syncthreads();
        if (tid < 48) {
            const int c = tid / 3, comp = tid % 3, i = c >> 2, j = c & 3;
            float t = 0.0f;
            for (int nv = 0; nv < G1; ++nv)
                t += s_wv[nv][j] * s_row[nv][i][comp] + s_wv[nv][4 + j] * s_row[nv][i][3 + comp];
            if (t != 0.0f) atomicAdd(&grad_cps[(size_t)ob.grad_offset + (size_t)rec.pid * 48 + tid], t);
        } else if (tid < 52) {
            float t = 0.0f;
            for (int w = 0; w < W; ++w) t += s_red[w][tid - 48];
            if (t != 0.0f) {
                if (tid == 48) atomicAdd(&grad_amp[ob.draw_index], t);
                else atomicAdd(&grad_color[ob.draw_index * 4 + (tid - 49)], t);
            }
        }
        __syncthreads();
    }
}

int adjoint_grid(int patch_size, int num_cus)
{
    int blocks = 0;
    hipError_t e = hipErrorInvalidValue;
    if (patch_size == 8) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_adjoint<8>, 64, 0);
    if (patch_size == 16) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_adjoint<16>, 256, 0);
    if (e != hipSuccess || blocks < 1) blocks = 1;
    return blocks * num_cus;
}

hipError_t launch_adjoint(int patch_size, const RangeRec* canon, const DevObject* objs,
                          const DevView* views, int n_views, int max_pixels, const DevParams& prm,
                          const FrameCounters* fc, const float* quad_store, const float* dl_dcolor,
                          uint32_t* visible, uint32_t* vis_list, uint32_t* vis_count, float* grad_cps,
                          float* grad_color, float* grad_amp, int grid, hipStream_t s)
{
    const int Q = patch_size * patch_size;
    int bx = (max_pixels + 255) / 256;
    bx = bx > 1024 ? 1024 : (bx < 1 ? 1 : bx);
    hipLaunchKernelGGL(k_adj_mark, dim3(bx, n_views), dim3(256), 0, s, views, Q, visible);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_adj_compact, dim3(1024), dim3(256), 0, s, visible, fc, vis_list, vis_count);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (patch_size == 8)
        hipLaunchKernelGGL(k_adjoint<8>, dim3(grid), dim3(64), 0, s, canon, objs, views, prm, fc, quad_store,
                           dl_dcolor, vis_list, vis_count, grad_cps, grad_color, grad_amp);
    else if (patch_size == 16)
        hipLaunchKernelGGL(k_adjoint<16>, dim3(grid), dim3(256), 0, s, canon, objs, views, prm, fc, quad_store,
                           dl_dcolor, vis_list, vis_count, grad_cps, grad_color, grad_amp);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace mp

// ---------------------------------------------------------------------------
// Level-synchronous split engine: every range of split level L (all objects,
// all views) is bounded by one launch; split children are appended to level
// L+1 and DRAW ranges to the frame's draw list with one atomic per wave and
// target.  The DRAW set is the same as the persistent engine's; the frame's
// canonical ordering makes the emission order irrelevant.

namespace mp {

__global__ void k_level_roots(const DevObject* __restrict__ objs, int n_objs, uint32_t total,
                              RangeRec* out, uint32_t* level_cnt, int n_levels, FrameCounters* reset_fc,
                              ShareState* reset_share)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0 && reset_fc) reset_counters(reset_fc, reset_share);
    if (g < (uint32_t)n_levels + 1) level_cnt[g] = (g == 0) ? total : 0u;
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
    out[g] = r;
}

__global__ __launch_bounds__(256) void k_level(const DevObject* __restrict__ objs,
                                               const DevView* __restrict__ views, DevParams prm,
                                               const RangeRec* __restrict__ cur, RangeRec* __restrict__ nxt,
                                               uint32_t* level_cnt, int L, uint32_t nxt_cap,
                                               CanonBuffers cb, FrameCounters* fc)
{
    const uint32_t n = min(level_cnt[L], nxt_cap);
    const int lane = threadIdx.x & 63;
    const uint32_t stride = gridDim.x * 256;
    const uint32_t lim = (n + 63u) & ~63u;
    uint32_t nb = 0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < lim; i += stride) {
        const bool valid = i < n;
        RangeRec rec;
        int flags = 0;
        if (valid) {
            rec = cur[i];
            flags = bound_rec(objs, views, prm, rec);
        }
        const uint64_t sb = __ballot(flags & 2);
        const int ns = __popcll(sb);
        if (ns) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&level_cnt[L + 1], 2u * (uint32_t)ns);
            base = __shfl(base, 0);
            if (flags & 2) {
                const uint32_t k = base + 2u * lanes_below(sb);
                if (k + 1 < nxt_cap) {
                    RangeRec c0, c1;
                    split_children(rec, flags, c0, c1);
                    nxt[k] = c0;
                    nxt[k + 1] = c1;
                }
            }
            if (base + 2u * (uint32_t)ns > nxt_cap && lane == 0) atomicOr(&fc->overflow, 8u);
        }
        const uint64_t db = __ballot(flags & 1);
        const int nd = __popcll(db);
        if (nd) {
            uint32_t dbase = 0;
            if (lane == 0) dbase = atomicAdd(&fc->draw_count, (uint32_t)nd);
            dbase = __shfl(dbase, 0);
            const uint32_t idx = dbase + lanes_below(db);
            const bool put = (flags & 1) && idx < cb.cap;
            if (put) cb.draw[idx] = rec;
            wave_count_add(cb.counts, put ? gpid_of(objs, rec) : 0u, put);
            if (dbase + (uint32_t)nd > cb.cap && lane == 0) atomicOr(&fc->overflow, 1u);
        }
        nb += (uint32_t)__popcll(__ballot(valid));
    }
    if (lane == 0 && nb) atomicAdd(stripe(fc, blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6), kStripeBounds),
                                   (unsigned long long)nb);
}

// Same level step with nine lanes per range: lane 9g+s of a wave evaluates
// sample s = iu*3+iv of the wave's range g (7 ranges per wave), and the
// group's first lane gathers the nine samples and takes bound()'s decision
// in the sequential sample order.  One range costs one patch evaluation of
// latency instead of nine.
__global__ __launch_bounds__(256) void k_level9(const DevObject* __restrict__ objs,
                                                const DevView* __restrict__ views, DevParams prm,
                                                const RangeRec* __restrict__ cur, RangeRec* __restrict__ nxt,
                                                uint32_t* level_cnt, int L, uint32_t nxt_cap,
                                                CanonBuffers cb, FrameCounters* fc)
{
    const uint32_t n = min(level_cnt[L], nxt_cap);
    const int lane = threadIdx.x & 63;
    const int grp = lane / 9, smp = lane - 9 * grp;
    const int iu = smp / 3, iv = smp - 3 * iu;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t n_waves = gridDim.x * 4u;
    uint32_t nb = 0;
    for (uint32_t base = wave * 7u; base < n; base += n_waves * 7u) {
        const uint32_t i = base + (uint32_t)grp;
        const bool valid = grp < 7 && i < n;
        RangeRec rec;
        rec.pid = 0; rec.obj_depth = 0; rec.umin = rec.vmin = rec.umax = rec.vmax = 0.0f;
        float e[3] = { 0.0f, 0.0f, 0.0f };
        float sx = 0.0f, sy = 0.0f;
        bool live = false;
        if (valid) {
            rec = cur[i];
            const DevObject& ob = objs[rec.obj_depth & 0xffffffu];
            live = (int)(rec.obj_depth >> 24) < prm.max_split_depth - 1;
            if (live) {
                const DevView& P = views[ob.view];
                if (ob.type == MP_TYPE_GREGORY) {
                    const float4* cp = ob.cps + (size_t)rec.pid * 20;
                    float4 C[20];
#pragma unroll
                    for (int k = 0; k < 20; ++k) C[k] = ldg4(cp + k);
                    bound_sample<20>(C, iu, iv, rec.umin, rec.vmin, rec.umax, rec.vmax, ob.mv, P, e, sx, sy);
                } else {
                    const float4* cp = ob.cps + (size_t)rec.pid * 16;
                    float4 C[16];
#pragma unroll
                    for (int k = 0; k < 16; ++k) C[k] = ldg4(cp + k);
                    bound_sample<16>(C, iu, iv, rec.umin, rec.vmin, rec.umax, rec.vmax, ob.mv, P, e, sx, sy);
                }
            }
        }
        float gx[9], gy[9], gz[9], gsx[9], gsy[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int src = min(grp * 9 + k, 63);
            gx[k] = __shfl(e[0], src);
            gy[k] = __shfl(e[1], src);
            gz[k] = __shfl(e[2], src);
            gsx[k] = __shfl(sx, src);
            gsy[k] = __shfl(sy, src);
        }
        const bool leader = valid && smp == 0;
        int flags = 0;
        if (leader && live) {
            const DevObject& ob = objs[rec.obj_depth & 0xffffffu];
            flags = bound_decide(gx, gy, gz, gsx, gsy, rec.umin, rec.vmin, rec.umax, rec.vmax, views[ob.view],
                                 prm, ob.split_limit);
        }
        const uint64_t sb = __ballot(flags & 2);
        const int ns = __popcll(sb);
        if (ns) {
            uint32_t nbase = 0;
            if (lane == 0) nbase = atomicAdd(&level_cnt[L + 1], 2u * (uint32_t)ns);
            nbase = __shfl(nbase, 0);
            if (flags & 2) {
                const uint32_t k = nbase + 2u * lanes_below(sb);
                if (k + 1 < nxt_cap) {
                    RangeRec c0, c1;
                    split_children(rec, flags, c0, c1);
                    nxt[k] = c0;
                    nxt[k + 1] = c1;
                }
            }
            if (nbase + 2u * (uint32_t)ns > nxt_cap && lane == 0) atomicOr(&fc->overflow, 8u);
        }
        const uint64_t db = __ballot(flags & 1);
        const int nd = __popcll(db);
        if (nd) {
            uint32_t dbase = 0;
            if (lane == 0) dbase = atomicAdd(&fc->draw_count, (uint32_t)nd);
            dbase = __shfl(dbase, 0);
            const uint32_t idx = dbase + lanes_below(db);
            const bool put = (flags & 1) && idx < cb.cap;
            if (put) cb.draw[idx] = rec;
            wave_count_add(cb.counts, put ? gpid_of(objs, rec) : 0u, put);
            if (dbase + (uint32_t)nd > cb.cap && lane == 0) atomicOr(&fc->overflow, 1u);
        }
        nb += (uint32_t)__popcll(__ballot(leader));
    }
    if (lane == 0 && nb) atomicAdd(stripe(fc, blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6), kStripeBounds),
                                   (unsigned long long)nb);
}

hipError_t launch_split_rebalanced(const DevObject* objs, const DevView* views, const DevParams& prm,
                                   const SplitQueues& q, const CanonBuffers& cb, FrameCounters* fc,
                                   hipStream_t s)
{
    SplitQueues qt = q;
    qt.pending = q.left_cnt;
    qt.strided = q.left;
    qt.step_budget = 0;
    qt.left_cnt = nullptr;
    return launch_split_persistent(objs, views, prm, qt, cb, fc, 0, s);
}

int level_grid(int num_cus, int lanes_per_range, int per_cu)
{
    int blocks = 0;
    const hipError_t e = lanes_per_range == 9
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_level9, 256, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_level, 256, 0);
    if (e != hipSuccess || blocks < 1) blocks = 1;
    // few work-groups per CU: a level launch costs its dispatch, and levels
    // are small (thousands of ranges)
    if (per_cu > 0 && per_cu < blocks) blocks = per_cu;
    return blocks * num_cus;
}

hipError_t launch_split_levels(const DevObject* objs, int n_objs, const DevView* views, const DevParams& prm,
                               uint32_t total_roots, RangeRec* list_a, RangeRec* list_b, uint32_t list_cap,
                               uint32_t* level_cnt, int n_levels, const SplitQueues& q,
                               const CanonBuffers& cb, FrameCounters* fc, int grid, int lanes_per_range,
                               ShareState* reset_share, hipStream_t s)
{
    const uint32_t n0 = total_roots > (uint32_t)(n_levels + 1) ? total_roots : (uint32_t)(n_levels + 1);
    hipLaunchKernelGGL(k_level_roots, dim3((n0 + 255) / 256), dim3(256), 0, s, objs, n_objs, total_roots,
                       list_a, level_cnt, n_levels, fc, reset_share);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    for (int L = 0; L < n_levels; ++L) {
        const RangeRec* cur = (L & 1) ? list_b : list_a;
        RangeRec* nxt = (L & 1) ? list_a : list_b;
        if (lanes_per_range == 9)
            hipLaunchKernelGGL(k_level9, dim3(grid), dim3(256), 0, s, objs, views, prm, cur, nxt, level_cnt, L,
                               list_cap, cb, fc);
        else
            hipLaunchKernelGGL(k_level, dim3(grid), dim3(256), 0, s, objs, views, prm, cur, nxt, level_cnt, L,
                               list_cap, cb, fc);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // anything deeper finishes in the persistent engine, group g taking
    // left[g], left[g + G], ...
    SplitQueues qt = q;
    qt.pending = level_cnt + n_levels;
    qt.strided = (n_levels & 1) ? list_b : list_a;
    qt.left_cap = list_cap;
    return launch_split_persistent(objs, views, prm, qt, cb, fc, 0, s);
}

}  // namespace mp
