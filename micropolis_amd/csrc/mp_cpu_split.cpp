// CPU bound&split method on the host cores (see mp_cpu_split.h).
//
// Arithmetic follows the reference's host code as written: eval_spline and
// eval_patch / eval_gregory_patch (src/base/Patch.cpp:107-211),
// bound_patch_range (src/Reyes/BoundNSplitCLCPU.cpp:249-291), BBox
// (src/base/utility.h:96-135) and Projection::bound
// (src/base/Projection.cpp:57-106).  Those are glm expressions; glm 0.9.9
// evaluates mat4 * vec4 as (c0*x + c1*y) + (c2*z + c3*w), dot as
// (x*x + y*y) + z*z and vec / scalar per component.  Built with
// -ffp-contract=off, so no product is fused into an add.
#include "mp_cpu_split.h"

#include <math.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>

namespace mp {
namespace {

struct V3 { float x, y, z; };

inline V3 ld(const float* p) { return V3{ p[0], p[1], p[2] }; }

// s*s*s*p0 + 3*s*s*t*p1 + 3*s*t*t*p2 + t*t*t*p3, s = 1 - t
inline V3 spline(V3 p0, V3 p1, V3 p2, V3 p3, float t)
{
    const float s = 1.0f - t;
    const float a = (s * s) * s, b = ((3.0f * s) * s) * t, c = ((3.0f * s) * t) * t, d = (t * t) * t;
    return V3{ ((a * p0.x + b * p1.x) + c * p2.x) + d * p3.x,
               ((a * p0.y + b * p1.y) + c * p2.y) + d * p3.y,
               ((a * p0.z + b * p1.z) + c * p2.z) + d * p3.z };
}

// eval_patch(patch, t, s): rows P[i][0..3] along t, the four results along s
V3 eval_bezier_host(const float* cp, float t, float s)
{
    V3 r[4];
    for (int i = 0; i < 4; ++i)
        r[i] = spline(ld(cp + 16 * i), ld(cp + 16 * i + 4), ld(cp + 16 * i + 8), ld(cp + 16 * i + 12), t);
    return spline(r[0], r[1], r[2], r[3], s);
}

inline V3 rational(float a, V3 p, float b, V3 q, float den)
{
    return V3{ (a * p.x + b * q.x) / den, (a * p.y + b * q.y) / den, (a * p.z + b * q.z) / den };
}

// eval_gregory_patch(p, u, v): u and v swap on entry; the interior F terms
// are 0/0 (NaN) at the corners, as in the reference host code.
V3 eval_gregory_host(const float* P, float u, float v)
{
    std::swap(u, v);
    auto pt = [&](int k) { return ld(P + 4 * k); };
    // corners 0-3, ep 4-7, em 8-11, fp 12-15, fm 16-19
    const V3 F0 = rational(u, pt(12), v, pt(16), u + v);
    const V3 F1 = rational(1.0f - u, pt(17), v, pt(13), (1.0f - u) + v);
    const V3 F2 = rational(1.0f - u, pt(18), 1.0f - v, pt(14), (2.0f - u) - v);
    const V3 F3 = rational(u, pt(19), 1.0f - v, pt(15), (1.0f + u) - v);
    const V3 r0 = spline(pt(0), pt(8), pt(7), pt(3), u);
    const V3 r1 = spline(pt(4), F0, F3, pt(11), u);
    const V3 r2 = spline(pt(9), F1, F2, pt(6), u);
    const V3 r3 = spline(pt(1), pt(5), pt(10), pt(2), u);
    return spline(r0, r1, r2, r3, v);
}

inline float distance(V3 a, V3 b)
{
    const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    return sqrtf((dx * dx + dy * dy) + dz * dz);
}

}  // namespace

int cpu_bound_range(const CpuSplitObject& o, const float* cp, int depth, float umin, float vmin,
                    float umax, float vmax, float cull_ribbon, int max_split_depth)
{
    V3 ps[3][3];
    V3 bmin{ INFINITY, INFINITY, INFINITY }, bmax{ -INFINITY, -INFINITY, -INFINITY };
    const float* m = o.mv;
    for (int iu = 0; iu < 3; ++iu) {
        for (int iv = 0; iv < 3; ++iv) {
            // range.min.x + (range.max.x - range.min.x) * iu * (1.0f / (RES-1))
            const float v = umin + ((umax - umin) * (float)iu) * 0.5f;
            const float u = vmin + ((vmax - vmin) * (float)iv) * 0.5f;
            const V3 p = o.type == MP_TYPE_GREGORY ? eval_gregory_host(cp, u, v) : eval_bezier_host(cp, u, v);
            const V3 e{ (m[0] * p.x + m[4] * p.y) + (m[8] * p.z + m[12] * 1.0f),
                        (m[1] * p.x + m[5] * p.y) + (m[9] * p.z + m[13] * 1.0f),
                        (m[2] * p.x + m[6] * p.y) + (m[10] * p.z + m[14] * 1.0f) };
            ps[iu][iv] = e;
            bmin.x = e.x < bmin.x ? e.x : bmin.x;  bmax.x = e.x > bmax.x ? e.x : bmax.x;
            bmin.y = e.y < bmin.y ? e.y : bmin.y;  bmax.y = e.y > bmax.y ? e.y : bmax.y;
            bmin.z = e.z < bmin.z ? e.z : bmin.z;  bmax.z = e.z > bmax.z ? e.z : bmax.z;
        }
    }
    float vlen = 0.0f, hlen = 0.0f;
    for (int i = 0; i < 3; ++i) {
        float h = 0.0f, v = 0.0f;
        for (int j = 0; j < 2; ++j) {
            v = v + distance(ps[j][i], ps[j + 1][i]);
            h = h + distance(ps[i][j], ps[i][j + 1]);
        }
        vlen = v > vlen ? v : vlen;
        hlen = h > hlen ? h : hlen;
    }
    // Projection::bound
    float n = -bmax.z;
    const float f = -bmin.z;
    if (f < o.near_) return MP_CULL;
    n = o.near_ > n ? o.near_ : n;
    const float vpx = (float)o.width / 2.0f, vpy = (float)o.height / 2.0f;
    const float mnx = bmin.x < 0 ? bmin.x / n * o.fx + vpx : bmin.x / f * o.fx + vpx;
    const float mny = bmin.y < 0 ? bmin.y / n * o.fy + vpy : bmin.y / f * o.fy + vpy;
    const float mxx = bmax.x > 0 ? bmax.x / n * o.fx + vpx : bmax.x / f * o.fx + vpx;
    const float mxy = bmax.y > 0 ? bmax.y / n * o.fy + vpy : bmax.y / f * o.fy + vpy;
    if (mnx > (float)(o.width - 1) + cull_ribbon || mxx < -cull_ribbon ||
        mny > (float)(o.height - 1) + cull_ribbon || mxy < -cull_ribbon)
        return MP_CULL;
    const float sx = mxx - mnx, sy = mxy - mny;
    if (bmin.z < 0 && sx < o.split_limit && sy < o.split_limit) return MP_DRAW;
    if (depth > max_split_depth) return MP_CPU_DROP;
    return vlen < hlen ? MP_HSPLIT : MP_VSPLIT;
}

namespace {

struct RootStats {
    uint32_t draws;
    uint32_t max_local;   // largest stack height of this root's traversal
    uint8_t pending;      // entries left on the root's stack at its last DRAW
};

struct Chunk { int obj; int p0, p1; };

void split_roots(const CpuSplitObject& o, int p0, int p1, float ribbon, int msd,
                 std::vector<RangeRec>& out, RootStats* stats, int64_t& bounds)
{
    const int per = o.type == MP_TYPE_GREGORY ? 20 : 16;
    std::vector<RangeRec> st;
    st.reserve(64);
    for (int pid = p0; pid < p1; ++pid) {
        const float* cp = o.cps + (size_t)pid * per * 4;
        RootStats rs{ 0, 1, 0 };
        const size_t first = out.size();
        st.clear();
        st.push_back(RangeRec{ (uint32_t)pid, o.slot, 0.0f, 0.0f, 1.0f, 1.0f });
        while (!st.empty()) {
            const RangeRec r = st.back();
            st.pop_back();
            const int depth = (int)(r.obj_depth >> 24);
            const int fl = cpu_bound_range(o, cp, depth, r.umin, r.vmin, r.umax, r.vmax, ribbon, msd);
            ++bounds;
            if (fl == MP_DRAW) {
                out.push_back(r);
                ++rs.draws;
                rs.pending = st.empty() ? 0 : 1;
            } else if (fl == MP_HSPLIT || fl == MP_VSPLIT) {
                RangeRec c0 = r, c1 = r;
                c0.obj_depth = c1.obj_depth = o.slot | ((uint32_t)(depth + 1) << 24);
                if (fl == MP_VSPLIT) {
                    const float cx = (r.umin + r.umax) * 0.5f;
                    c0.umax = cx; c1.umin = cx;
                } else {
                    const float cy = (r.vmin + r.vmax) * 0.5f;
                    c0.vmax = cy; c1.vmin = cy;
                }
                st.push_back(c0);   // the upper half is popped first
                st.push_back(c1);
                rs.max_local = std::max<uint32_t>(rs.max_local, (uint32_t)st.size());
            }
        }
        // the patch's DRAW ranges in canonical order: (umin, vmin), dyadic
        // values compared as floats
        std::sort(out.begin() + (std::ptrdiff_t)first, out.end(), [](const RangeRec& a, const RangeRec& b) {
            return a.umin != b.umin ? a.umin < b.umin : a.vmin < b.vmin;
        });
        stats[pid] = rs;
    }
}

}  // namespace

void cpu_split_stream(const std::vector<CpuSplitObject>& objs, uint32_t n_gpids, float cull_ribbon,
                      int max_split_depth, int patches_per_pass, int n_threads, CpuSplitFrame& out,
                      const std::function<void(const RangeRec*, size_t)>& emit)
{
    out.draws.clear();
    out.counts.assign(n_gpids, 0u);
    out.pass_counts.clear();
    out.bounds = 0;
    out.max_stack = 0;
    std::vector<RootStats> stats(n_gpids);
    std::vector<Chunk> chunks;
    const int CH = 64;
    for (int oi = 0; oi < (int)objs.size(); ++oi)
        for (int p = 0; p < objs[oi].n_patches; p += CH)
            chunks.push_back(Chunk{ oi, p, std::min(objs[oi].n_patches, p + CH) });

    if (n_threads <= 0) n_threads = std::max(1, std::min((int)std::thread::hardware_concurrency(), 16));
    n_threads = std::max(1, std::min<int>(n_threads, (int)chunks.size()));

    // chunk c's DRAW ranges, ready once done[c] is set
    std::vector<std::vector<RangeRec>> part(chunks.size());
    std::vector<int64_t> chunk_bounds(chunks.size(), 0);
    std::unique_ptr<std::atomic<int>[]> done(new std::atomic<int>[chunks.size() + 1]);
    for (size_t c = 0; c < chunks.size(); ++c) done[c].store(0);
    std::mutex m;
    std::condition_variable cv;
    std::atomic<size_t> next{ 0 };
    auto split_chunk = [&](size_t c) {
        const Chunk& ck = chunks[c];
        const CpuSplitObject& o = objs[ck.obj];
        split_roots(o, ck.p0, ck.p1, cull_ribbon, max_split_depth, part[c], stats.data() + o.gpid_base,
                    chunk_bounds[c]);
        {
            std::lock_guard<std::mutex> lk(m);
            done[c].store(1, std::memory_order_release);
        }
        cv.notify_all();
    };
    std::vector<std::thread> pool;
    if (n_threads > 1) {
        for (int t = 0; t < n_threads; ++t)
            pool.emplace_back([&] {
                for (;;) {
                    const size_t c = next.fetch_add(1, std::memory_order_relaxed);
                    if (c >= chunks.size()) break;
                    split_chunk(c);
                }
            });
    }
    // chunks in order: object slots ascending, patches ascending
    for (size_t c = 0; c < chunks.size(); ++c) {
        if (n_threads > 1) {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return done[c].load(std::memory_order_acquire) != 0; });
        } else {
            split_chunk(c);
        }
        if (!part[c].empty()) emit(part[c].data(), part[c].size());
        out.bounds += chunk_bounds[c];
        std::vector<RangeRec>().swap(part[c]);
    }
    for (auto& th : pool) th.join();

    // Per-pass counts and stack height of the reference's one stack: roots are
    // popped from the back, so while root p is processed the roots 0..p-1 are
    // still below it.  A batch closes at patches_per_pass DRAW ranges; one more
    // (empty) batch follows when entries remain after the last DRAW.
    const int64_t ppp = std::max(1, patches_per_pass);
    for (const CpuSplitObject& o : objs) {
        int64_t draws = 0;
        int last = -1;
        for (int p = 0; p < o.n_patches; ++p) {
            const RootStats& rs = stats[o.gpid_base + p];
            out.counts[o.gpid_base + p] = rs.draws;
            draws += rs.draws;
            if (rs.draws && last < 0) last = p;   // lowest root with draws = last processed
            out.max_stack = std::max<int64_t>(out.max_stack, (int64_t)p + rs.max_local);
        }
        if (o.n_patches == 0) continue;
        for (int64_t left = draws; left >= ppp; left -= ppp) out.pass_counts.push_back((int32_t)ppp);
        const int64_t rem = draws % ppp;
        const bool pending_after_last =
            last < 0 || last > 0 || stats[o.gpid_base + last].pending;
        if (rem > 0 || draws == 0 || pending_after_last) out.pass_counts.push_back((int32_t)rem);
    }
}

void cpu_split_frame(const std::vector<CpuSplitObject>& objs, uint32_t n_gpids,
                     float cull_ribbon, int max_split_depth, int patches_per_pass,
                     int n_threads, CpuSplitFrame& out)
{
    std::vector<RangeRec> all;
    cpu_split_stream(objs, n_gpids, cull_ribbon, max_split_depth, patches_per_pass, n_threads, out,
                     [&](const RangeRec* r, size_t n) { all.insert(all.end(), r, r + n); });
    out.draws.swap(all);
}

}  // namespace mp
