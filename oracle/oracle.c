/*
 * CPU oracle for the micropolis REYES hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Scalar IEEE-strict restatement of the reference kernels and schedulers.
 * Build: gcc -O2 -std=c99 -ffp-contract=off -fno-fast-math (see oracle/Makefile).
 * Only the test-suite, the smoke check and the CPU-baseline leg of bench.py
 * load this library; the product HIP library never does.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   or_eval_patch      FP64 Bernstein closed form, corner interpolation,
 *                      Bezier->Gregory mapping (tools/bez2mscene_greg.py:111)
 *   or_bound           analytic screen lengths of planar patches, frustum/
 *                      eye-split/depth-cap cases built by hand
 *   or_split           brute-force recursive predicate (canonical DRAW set),
 *                      BREADTH level order, BOUNDED LIFO recurrence, literal
 *                      LOCAL kernel emulation (same multiset)
 *   or_dice            closed-form projection of known points
 *   or_shade/or_sample hand-enumerated micropolygons, exact rational edge
 *                      tests, top-left rule watertightness, brute-force
 *                      reference raster written independently in the test
 *   or_prefix_sum      the reference's own scan self-test (main.cpp:343-400,
 *                      srand(43), rand()%8+1) with glibc's rand sequence
 *   or_noise           the seeded lattice value at integer points, smoothstep
 *                      fade with flat ends, trilinear corner mean at cell
 *                      centres, octave halving / normalisation / reseeding
 *   quad_lambert       planar parallelograms of known orientation:
 *   (or_shade's term)  max(n . normalize(4,3,8), 0) from the FP64 plane
 *                      normal, back-lit clamp, NaN for a zero corner sum
 *   or_resolve         FP64 0.015*dc + s*dc colour, clear (rgb^2.2, 1000),
 *                      depth decode and ids of a hand-built key buffer
 *   or_make_view       glm::perspective's published entries, frustum planes
 *                      at NDC +-1, fx = fy = H / (2 tan(fovy/2))
 *   or_to_srgb8        FP64 pow(c, 0.454545) * 255 + 0.5 off byte boundaries
 *   or_framebuffer_*   tiled bijection of the padded window
 *   or_bound_cpu       FP64 de Casteljau rows vs the glm-order evaluation,
 *   or_eval_patch_cpu  analytic sizes of planar patches through
 *   or_projection_bound Projection::bound, hand-built cull/drop cases, the
 *   or_split (CPU)     split direction of elongated patches, brute-force
 *                      recursion of the predicate, batch recurrence
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* OpenCL builtin semantics, spelled out (kernels use -cl-fast-relaxed-math
 * -cl-mad-enable, src/CL/Program.cpp:118; this build fixes IEEE semantics
 * with one contraction rule for the kernel arithmetic: s + p*q is
 * fmaf(p, q, s) and p*q + r*s (or -) is fmaf(r, s, p*q), the left product
 * rounded -- the same places the HIP library fuses.  The CPU method's host
 * code below is not contracted).                                             */

/* s + p*q and p*q - r*s under the contraction rule */
static inline float mad(float p, float q, float s) { return fmaf(p, q, s); }
static inline float mad_diff(float p, float q, float r, float s) { return fmaf(-r, s, p * q); }

/* OpenCL min(x,y): "Returns y if y < x, otherwise it returns x." */
static inline float cl_minf(float x, float y) { return (y < x) ? y : x; }
/* OpenCL max(x,y): "Returns y if x < y, otherwise it returns x." */
static inline float cl_maxf(float x, float y) { return (x < y) ? y : x; }
static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int32_t iclamp(int32_t x, int32_t lo, int32_t hi) { return imin(imax(x, lo), hi); }

/* wrapping 32-bit integer arithmetic (OpenCL int overflow wraps on every
 * device the reference ran on; C makes it explicit) */
static inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }

/* float -> int32 conversion with defined out-of-range behaviour:
 * NaN -> 0, saturate to [INT32_MIN, INT32_MAX], else truncate toward zero */
static inline int32_t f2i_sat(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)x;
}

static inline float dot3(const float* a, const float* b)
{
    return mad(a[2], b[2], mad(a[1], b[1], a[0] * b[0]));
}

static inline void cross3(const float* a, const float* b, float* o)
{
    o[0] = mad_diff(a[1], b[2], a[2], b[1]);
    o[1] = mad_diff(a[2], b[0], a[0], b[2]);
    o[2] = mad_diff(a[0], b[1], a[1], b[0]);
}

/* normalize(v) = v * (1/sqrt(dot(v,v))) */
static inline void normalize3(const float* v, float* o)
{
    float inv = 1.0f / sqrtf(dot3(v, v));
    o[0] = v[0] * inv; o[1] = v[1] * inv; o[2] = v[2] * inv;
}

/* column-major 4x4 matrix times vector, each row summed left to right
 * (mul_pm4v4 / mul_m44v4, kernels/utility.h:164-180) */
static inline void mat4_mul(const float* m, const float* v, float* o)
{
    for (int i = 0; i < 4; ++i)
        o[i] = mad(m[12 + i], v[3], mad(m[8 + i], v[2], mad(m[4 + i], v[1], m[0 + i] * v[0])));
}

/* ------------------------------------------------------------------------ */
/* Patch evaluation (kernels/utility.h:41-126).                               */

/* Bernstein row/column weights in the reference's written order:
 * s^3, 3 s^2 t, 3 s t^2, t^3 with products left to right. */
static inline void bern_u(float t, float s, float* a)
{
    a[0] = (s * s) * s;
    a[1] = ((3.0f * s) * s) * t;
    a[2] = ((3.0f * s) * t) * t;
    a[3] = (t * t) * t;
}
static inline void bern_v(float t, float s, float* b)
{
    b[0] = (s * s) * s;            /* 1*s.y*s.y*s.y, 1*s.y == s.y exactly */
    b[1] = ((3.0f * t) * s) * s;
    b[2] = ((3.0f * t) * t) * s;
    b[3] = (t * t) * t;
}

/* p += (a*b) * P, componentwise (utility.h:48 etc.) */
static inline void acc4(float* p, float w, const float* P)
{
    p[0] = mad(w, P[0], p[0]);
    p[1] = mad(w, P[1], p[1]);
    p[2] = mad(w, P[2], p[2]);
    p[3] = mad(w, P[3], p[3]);
}

static void eval_bezier(const float* cp, float u, float v, float* p)
{
    float su = 1.0f - u, sv = 1.0f - v;
    float a[4], b[4];
    bern_u(u, su, a);
    bern_v(v, sv, b);
    p[0] = p[1] = p[2] = p[3] = 0.0f;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            acc4(p, a[i] * b[j], cp + 4 * (4 * i + j));
}

static inline void nan_to_zero4(float* f)
{
    for (int k = 0; k < 4; ++k)
        if (f[k] != f[k]) f[k] = 0.0f;
}

/* Gregory patch: corners P0-3, ep 4-7, em 8-11, fp 12-15, fm 16-19
 * (utility.h:75-126).  F terms are rational with NaN -> 0 (utility.h:95-98). */
static void eval_gregory(const float* P, float u, float v, float* p)
{
    float F0[4], F1[4], F2[4], F3[4];
    float d0 = u + v;
    float d1 = (1.0f - u) + v;
    float d2 = (2.0f - u) - v;
    float d3 = (1.0f + u) - v;
    for (int k = 0; k < 4; ++k) {
        F0[k] = mad(v, P[4 * 16 + k], u * P[4 * 12 + k]) / d0;
        F1[k] = mad(v, P[4 * 13 + k], (1.0f - u) * P[4 * 17 + k]) / d1;
        F2[k] = mad(1.0f - v, P[4 * 18 + k], (1.0f - u) * P[4 * 14 + k]) / d2;
        F3[k] = mad(1.0f - v, P[4 * 15 + k], u * P[4 * 19 + k]) / d3;
    }
    nan_to_zero4(F0); nan_to_zero4(F1); nan_to_zero4(F2); nan_to_zero4(F3);

    float su = 1.0f - u, sv = 1.0f - v;
    float a[4], b[4];
    bern_u(u, su, a);
    bern_v(v, sv, b);
    const float* row[4][4] = {
        { P + 4 * 0, P + 4 * 4, P + 4 * 9,  P + 4 * 1 },
        { P + 4 * 8, F0,        F1,         P + 4 * 5 },
        { P + 4 * 7, F3,        F2,         P + 4 * 10 },
        { P + 4 * 3, P + 4 * 11, P + 4 * 6, P + 4 * 2 },
    };
    p[0] = p[1] = p[2] = p[3] = 0.0f;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            acc4(p, a[i] * b[j], row[i][j]);
}

void or_eval_patch(const float* cps, int type, int pid, float u, float v, float out[4])
{
    if (type == OR_GREGORY) eval_gregory(cps + (long)pid * 20 * 4, u, v, out);
    else                    eval_bezier(cps + (long)pid * 16 * 4, u, v, out);
}

/* FP64 closed form (test helper): sum B_i(u) B_j(v) P_ij */
void or_eval_patch_f64(const double* cps, int type, int pid, double u, double v, double out[4])
{
    double su = 1.0 - u, sv = 1.0 - v;
    double a[4] = { su * su * su, 3 * su * su * u, 3 * su * u * u, u * u * u };
    double b[4] = { sv * sv * sv, 3 * sv * sv * v, 3 * sv * v * v, v * v * v };
    out[0] = out[1] = out[2] = out[3] = 0.0;
    if (type == OR_GREGORY) {
        const double* P = cps + (long)pid * 20 * 4;
        double F[4][4];
        double dd[4] = { u + v, 1 - u + v, 2 - u - v, 1 + u - v };
        for (int k = 0; k < 4; ++k) {
            F[0][k] = dd[0] != 0 ? (u * P[4 * 12 + k] + v * P[4 * 16 + k]) / dd[0] : 0;
            F[1][k] = dd[1] != 0 ? ((1 - u) * P[4 * 17 + k] + v * P[4 * 13 + k]) / dd[1] : 0;
            F[2][k] = dd[2] != 0 ? ((1 - u) * P[4 * 14 + k] + (1 - v) * P[4 * 18 + k]) / dd[2] : 0;
            F[3][k] = dd[3] != 0 ? (u * P[4 * 19 + k] + (1 - v) * P[4 * 15 + k]) / dd[3] : 0;
        }
        const double* row[4][4] = {
            { P + 0, P + 16, P + 36, P + 4 },
            { P + 32, F[0], F[1], P + 20 },
            { P + 28, F[3], F[2], P + 40 },
            { P + 12, P + 44, P + 24, P + 8 },
        };
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) out[k] += a[i] * b[j] * row[i][j][k];
    } else {
        const double* P = cps + (long)pid * 16 * 4;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) out[k] += a[i] * b[j] * P[4 * (4 * i + j) + k];
    }
}

/* ------------------------------------------------------------------------ */
/* Value noise displacement.  Defined by this build (the reference only has a
 * native_sin hack, kernels/dice.cl:51-67).                                   */

static inline uint32_t noise_hash(int32_t x, int32_t y, int32_t z, uint32_t seed)
{
    uint32_t h = seed ^ 0x9e3779b9u;
    h ^= (uint32_t)x * 0x8da6b343u;
    h ^= (uint32_t)y * 0xd8163841u;
    h ^= (uint32_t)z * 0xcb1ab31fu;
    h ^= h >> 16; h *= 0x85ebca6bu;
    h ^= h >> 13; h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

static inline float lattice(int32_t x, int32_t y, int32_t z, uint32_t seed)
{
    return (float)(noise_hash(x, y, z, seed) >> 8) * (1.0f / 8388608.0f) - 1.0f;
}

static inline float fade(float t) { return (t * t) * (3.0f - 2.0f * t); }
static inline float lerpf_(float a, float b, float t) { return mad(b - a, t, a); }

static float vnoise(float x, float y, float z, uint32_t seed)
{
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    float wx = fade(x - fx), wy = fade(y - fy), wz = fade(z - fz);
    float c000 = lattice(ix, iy, iz, seed),         c100 = lattice(ix + 1, iy, iz, seed);
    float c010 = lattice(ix, iy + 1, iz, seed),     c110 = lattice(ix + 1, iy + 1, iz, seed);
    float c001 = lattice(ix, iy, iz + 1, seed),     c101 = lattice(ix + 1, iy, iz + 1, seed);
    float c011 = lattice(ix, iy + 1, iz + 1, seed), c111 = lattice(ix + 1, iy + 1, iz + 1, seed);
    float x00 = lerpf_(c000, c100, wx), x10 = lerpf_(c010, c110, wx);
    float x01 = lerpf_(c001, c101, wx), x11 = lerpf_(c011, c111, wx);
    float y0 = lerpf_(x00, x10, wy), y1 = lerpf_(x01, x11, wy);
    return lerpf_(y0, y1, wz);
}

float or_noise(const or_params* prm, float x, float y, float z)
{
    float acc = 0.0f, amp = 1.0f, norm = 0.0f, freq = prm->disp_frequency;
    for (int k = 0; k < prm->disp_octaves; ++k) {
        acc = mad(amp, vnoise(x * freq, y * freq, z * freq, prm->disp_seed + (uint32_t)k), acc);
        norm = norm + amp;
        amp = amp * 0.5f;
        freq = freq * 2.0f;
    }
    return acc / norm;
}

/* Bezier partials for the displacement normal.  dB/dt in this fixed order:
 * -3 s^2, 3 s (s - 2t), 3 t (2s - t), 3 t^2. */
static inline void dbern(float t, float s, float* d)
{
    d[0] = -3.0f * (s * s);
    d[1] = 3.0f * (s * (s - 2.0f * t));
    d[2] = 3.0f * (t * (2.0f * s - t));
    d[3] = 3.0f * (t * t);
}

/* Displaced object-space point: P + amp * noise(P) * n, n = unit
 * cross(dP/dv, dP/du) (the orientation shade uses, reyes.cl:177); no
 * displacement where that cross product is zero.  Bezier only. */
void or_displaced_point(const float* cps, int type, int pid, float u, float v,
                        const or_params* prm, float out[4])
{
    or_eval_patch(cps, type, pid, u, v, out);
    if (!prm->displacement || type != OR_BEZIER) return;
    const float* cp = cps + (long)pid * 16 * 4;
    /* tangents of the displacement normal, summed over i first (column sums
     * S_j = sum_i a_i(u) P_ij, D_j = sum_i a_i'(u) P_ij, i ascending), then
     * dP/du = sum_j b_j(v) D_j and dP/dv = sum_j b_j'(v) S_j, j ascending */
    float su = 1.0f - u, sv = 1.0f - v;
    float a[4], b[4], da[4], db[4];
    bern_u(u, su, a);
    bern_v(v, sv, b);
    dbern(u, su, da);
    dbern(v, sv, db);
    float S[4][3], D[4][3];
    for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 3; ++k) {
            float sacc = 0.0f, dacc = 0.0f;
            for (int i = 0; i < 4; ++i) {
                const float P = cp[4 * (4 * i + j) + k];
                sacc = mad(a[i], P, sacc);
                dacc = mad(da[i], P, dacc);
            }
            S[j][k] = sacc;
            D[j][k] = dacc;
        }
    float pu[3] = { 0, 0, 0 }, pv[3] = { 0, 0, 0 };
    for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 3; ++k) {
            pu[k] = mad(b[j], D[j][k], pu[k]);
            pv[k] = mad(db[j], S[j][k], pv[k]);
        }
    float n[3];
    cross3(pv, pu, n);
    float len2 = dot3(n, n);
    float nh[3] = { 0.0f, 0.0f, 0.0f };
    if (len2 > 0.0f) {
        float inv = 1.0f / sqrtf(len2);
        nh[0] = n[0] * inv; nh[1] = n[1] * inv; nh[2] = n[2] * inv;
    }
    float d = prm->disp_amplitude * or_noise(prm, out[0], out[1], out[2]);
    out[0] = mad(d, nh[0], out[0]);
    out[1] = mad(d, nh[1], out[1]);
    out[2] = mad(d, nh[2], out[2]);
}

/* ------------------------------------------------------------------------ */
/* bound() and outside_frustum() (kernels/bound_n_split.h:10-100).           */

static int outside_frustum(const float* pmin, const float* pmax, const or_view* P, float ribbon)
{
    float n = -pmax[2];
    float f = -pmin[2];
    if (f < P->near_ || n > P->far_) return 1;
    n = cl_maxf(P->near_, n);
    float hw = (float)P->width * 0.5f, hh = (float)P->height * 0.5f;
    float sminx = mad(pmin[0] / ((pmin[0] < 0) ? n : f), P->f[0], hw);
    float sminy = mad(pmin[1] / ((pmin[1] < 0) ? n : f), P->f[1], hh);
    float smaxx = mad(pmax[0] / ((pmax[0] > 0) ? n : f), P->f[0], hw);
    float smaxy = mad(pmax[1] / ((pmax[1] > 0) ? n : f), P->f[1], hh);
    return (sminx > (float)(P->width - 1) + ribbon || smaxx < -ribbon ||
            sminy > (float)(P->height - 1) + ribbon || smaxy < -ribbon);
}

int or_bound(const float* cps, int type, int pid, int depth,
             float umin, float vmin, float umax, float vmax,
             const float mv[16], const or_view* P, const or_params* prm, float split_limit)
{
    enum { MAXRES = 8 };
    const int RES = prm->bound_sample_rate;
    float ppos[MAXRES][MAXRES][2];
    float bmin[3] = { INFINITY, INFINITY, INFINITY };
    float bmax[3] = { -INFINITY, -INFINITY, -INFINITY };

    if (depth >= prm->max_split_depth - 1) return OR_CULL;
    if (RES < 2 || RES > MAXRES) return OR_CULL;

    const float step = 1.0f / (float)(RES - 1);
    for (int iu = 0; iu < RES; ++iu) {
        for (int iv = 0; iv < RES; ++iv) {
            float tu = (float)iu * step, tv = (float)iv * step;
            float u = umin + (umax - umin) * tu;
            float v = vmin + (vmax - vmin) * tv;
            float p[4], e[4], q[4];
            or_eval_patch(cps, type, pid, u, v, p);
            mat4_mul(mv, p, e);
            for (int k = 0; k < 3; ++k) {
                bmin[k] = cl_minf(bmin[k], e[k]);
                bmax[k] = cl_maxf(bmax[k], e[k]);
            }
            mat4_mul(P->proj, e, q);
            float rx = q[0] / q[3], ry = q[1] / q[3];
            /* mul_cm2v2: m[0]*x + m[1]*y with m[0]=(s0,s1), m[1]=(s2,s3) */
            ppos[iv][iu][0] = mad(P->screen[2], ry, P->screen[0] * rx);
            ppos[iv][iu][1] = mad(P->screen[3], ry, P->screen[1] * rx);
        }
    }

    if (outside_frustum(bmin, bmax, P, prm->cull_ribbon)) return OR_CULL;
    if (bmin[2] < P->eps && bmax[2] > P->near_)
        return ((umax - umin) < (vmax - vmin)) ? OR_VSPLIT : OR_HSPLIT;

    float hlen = 0.0f, vlen = 0.0f;
    for (int i = 0; i < RES; ++i) {
        float h = 0.0f, v = 0.0f;
        for (int j = 0; j < RES - 1; ++j) {
            float dx = ppos[i][j][0] - ppos[i][j + 1][0];
            float dy = ppos[i][j][1] - ppos[i][j + 1][1];
            h = h + sqrtf(mad(dy, dy, dx * dx));
            dx = ppos[j][i][0] - ppos[j + 1][i][0];
            dy = ppos[j][i][1] - ppos[j + 1][i][1];
            v = v + sqrtf(mad(dy, dy, dx * dx));
        }
        hlen = cl_maxf(h, hlen);
        vlen = cl_maxf(v, vlen);
    }
    if (hlen <= split_limit && vlen <= split_limit) return OR_DRAW;
    return (hlen > vlen) ? OR_VSPLIT : OR_HSPLIT;
}

/* ------------------------------------------------------------------------ */
/* CPU bound&split method (src/Reyes/BoundNSplitCLCPU.cpp, src/base/Patch.cpp,
 * src/base/Projection.cpp).  Host code built on glm; glm is not vendored in
 * the reference, so the vector operations follow glm 0.9.9's published
 * definitions: mat4 * vec4 = (c0*x + c1*y) + (c2*z + c3*w) per row,
 * dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, length = sqrt(dot(v,v)),
 * distance(a,b) = length(b - a), scalar * vec and vec / scalar per component. */

/* eval_spline (Patch.cpp:107-113): s*s*s*p0 + 3*s*s*t*p1 + 3*s*t*t*p2 + t*t*t*p3 */
static void cpu_spline(const float* p0, const float* p1, const float* p2, const float* p3,
                       float t, float* dst)
{
    float s = 1.0f - t;
    float a = (s * s) * s, b = ((3.0f * s) * s) * t, c = ((3.0f * s) * t) * t, d = (t * t) * t;
    for (int k = 0; k < 3; ++k)
        dst[k] = ((a * p0[k] + b * p1[k]) + c * p2[k]) + d * p3[k];
}

/* eval_patch (Patch.cpp:115-129): rows P[i][0..3] along t, then along s.
 * cp holds 16 points of 4 floats (P[i][j] at 4i+j). */
static void cpu_eval_bezier(const float* cp, float t, float s, float* dst)
{
    float p[4][3];
    for (int i = 0; i < 4; ++i)
        cpu_spline(cp + 4 * (4 * i + 0), cp + 4 * (4 * i + 1), cp + 4 * (4 * i + 2),
                   cp + 4 * (4 * i + 3), t, p[i]);
    cpu_spline(p[0], p[1], p[2], p[3], s, dst);
}

/* eval_gregory_patch (Patch.cpp:186-211): u and v swapped on entry; the F
 * terms have no NaN guard, so every patch corner evaluates to NaN. */
static void cpu_eval_gregory(const float* P, float u, float v, float* dst)
{
    float tmp = u; u = v; v = tmp;
    const float *p = P, *ep = P + 16, *em = P + 32, *fp = P + 48, *fm = P + 64;
    float F0[3], F1[3], F2[3], F3[3];
    for (int k = 0; k < 3; ++k) {
        F0[k] = (u * fp[0 + k] + v * fm[0 + k]) / (u + v);
        F1[k] = ((1.0f - u) * fm[4 + k] + v * fp[4 + k]) / ((1.0f - u) + v);
        F2[k] = ((1.0f - u) * fm[8 + k] + (1.0f - v) * fp[8 + k]) / ((2.0f - u) - v);
        F3[k] = (u * fm[12 + k] + (1.0f - v) * fp[12 + k]) / ((1.0f + u) - v);
    }
    float R[4][3];
    cpu_spline(p + 0,  em + 0, ep + 12, p + 12, u, R[0]);
    cpu_spline(ep + 0, F0,     F3,      em + 12, u, R[1]);
    cpu_spline(em + 4, F1,     F2,      ep + 8, u, R[2]);
    cpu_spline(p + 4,  ep + 4, em + 8,  p + 8,  u, R[3]);
    cpu_spline(R[0], R[1], R[2], R[3], v, dst);
}

void or_eval_patch_cpu(const float* cps, int type, int pid, float u, float v, float out[3])
{
    if (type == OR_GREGORY) cpu_eval_gregory(cps + (long)pid * 20 * 4, u, v, out);
    else                    cpu_eval_bezier(cps + (long)pid * 16 * 4, u, v, out);
}

static float glm_distance(const float* a, const float* b)
{
    float d[3] = { b[0] - a[0], b[1] - a[1], b[2] - a[2] };
    return sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
}

/* Projection::bound (Projection.cpp:57-106).  Returns 1 when culled. */
int or_projection_bound(const float bmin[3], const float bmax[3], const or_view* P,
                        float cull_ribbon, float size[2])
{
    float n = -bmax[2];
    float f = -bmin[2];
    if (f < P->near_) return 1;
    n = (P->near_ > n) ? P->near_ : n;          /* maximum(_near, n) */
    const float fx = P->f[0], fy = P->f[1];
    const float vpx = (float)P->width / 2.0f, vpy = (float)P->height / 2.0f;
    float mnx = (bmin[0] < 0) ? bmin[0] / n * fx + vpx : bmin[0] / f * fx + vpx;
    float mny = (bmin[1] < 0) ? bmin[1] / n * fy + vpy : bmin[1] / f * fy + vpy;
    float mxx = (bmax[0] > 0) ? bmax[0] / n * fx + vpx : bmax[0] / f * fx + vpx;
    float mxy = (bmax[1] > 0) ? bmax[1] / n * fy + vpy : bmax[1] / f * fy + vpy;
    if (mnx > (float)(P->width - 1) + cull_ribbon || mxx < -cull_ribbon ||
        mny > (float)(P->height - 1) + cull_ribbon || mxy < -cull_ribbon)
        return 1;
    size[0] = mxx - mnx;
    size[1] = mxy - mny;
    return 0;
}

/* One step of BoundNSplitCLCPU::do_bound_n_split (:102-148) for one range:
 * bound_patch_range / bound_gregory_patch_range (:249-336) with RES = 3,
 * Projection::bound, then the DRAW / drop / split decision.  Returns OR_CULL,
 * OR_DRAW, OR_CPU_DROP (depth > max_split_depth), OR_HSPLIT (halve the
 * range's second coordinate: vsplit_range) or OR_VSPLIT (halve the first:
 * hsplit_range). */
int or_bound_cpu(const float* cps, int type, int pid, int depth,
                 float umin, float vmin, float umax, float vmax,
                 const float mv[16], const or_view* view, const or_params* prm, float split_limit)
{
    float ps[3][3][3];
    float bmin[3] = { INFINITY, INFINITY, INFINITY };
    float bmax[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (int iu = 0; iu < 3; ++iu) {
        for (int iv = 0; iv < 3; ++iv) {
            /* (max - min) * (size_t)i * (1.0f / (RES-1)) */
            float v = umin + ((umax - umin) * (float)iu) * 0.5f;
            float u = vmin + ((vmax - vmin) * (float)iv) * 0.5f;
            float pos[3];
            if (type == OR_GREGORY) cpu_eval_gregory(cps + (long)pid * 20 * 4, u, v, pos);
            else                    cpu_eval_bezier(cps + (long)pid * 16 * 4, u, v, pos);
            float* pt = ps[iu][iv];
            for (int r = 0; r < 3; ++r)
                pt[r] = (mv[r] * pos[0] + mv[4 + r] * pos[1]) + (mv[8 + r] * pos[2] + mv[12 + r] * 1.0f);
            for (int k = 0; k < 3; ++k) {            /* BBox::add_point */
                bmin[k] = (pt[k] < bmin[k]) ? pt[k] : bmin[k];
                bmax[k] = (pt[k] > bmax[k]) ? pt[k] : bmax[k];
            }
        }
    }
    float vlen = 0.0f, hlen = 0.0f;
    for (int i = 0; i < 3; ++i) {
        float h = 0.0f, v = 0.0f;
        for (int j = 0; j < 2; ++j) {
            v = v + glm_distance(ps[j][i], ps[j + 1][i]);
            h = h + glm_distance(ps[i][j], ps[i][j + 1]);
        }
        vlen = (v > vlen) ? v : vlen;
        hlen = (h > hlen) ? h : hlen;
    }
    float size[2];
    if (or_projection_bound(bmin, bmax, view, prm->cull_ribbon, size)) return OR_CULL;
    if (bmin[2] < 0 && size[0] < split_limit && size[1] < split_limit) return OR_DRAW;
    if (depth > prm->max_split_depth) return OR_CPU_DROP;
    return (vlen < hlen) ? OR_HSPLIT : OR_VSPLIT;
}

/* Children of a split (bound_n_split_breadthfirst.cl:99-113). */
static void split_range(const or_range* r, int flags, or_range* c0, or_range* c1)
{
    float cx = (r->umin + r->umax) * 0.5f;
    float cy = (r->vmin + r->vmax) * 0.5f;
    *c0 = *r; *c1 = *r;
    c0->depth = r->depth + 1;
    c1->depth = r->depth + 1;
    if (flags & 4) { c0->umax = cx; c1->umin = cx; }
    else           { c0->vmax = cy; c1->vmin = cy; }
}

/* ------------------------------------------------------------------------ */
/* Split schedules.                                                         */

typedef struct rvec { or_range* d; long n, cap; } rvec;
static void rpush(rvec* v, or_range r)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 1024;
        v->d = (or_range*)realloc(v->d, (size_t)v->cap * sizeof(or_range));
    }
    v->d[v->n++] = r;
}

typedef struct split_ctx {
    const float* cps; int type; const float* mv; const or_view* view;
    const or_params* prm; float limit;
} split_ctx;

static int sbound(const split_ctx* c, const or_range* r)
{
    return or_bound(c->cps, c->type, r->pid, r->depth, r->umin, r->vmin, r->umax, r->vmax,
                    c->mv, c->view, c->prm, c->limit);
}

long or_split(int method, const float* cps, int type, int n_patches,
              const float mv[16], const or_view* view, const or_params* prm,
              float split_limit, or_range* out, long out_cap,
              int32_t* pass_counts, int max_passes, int32_t* n_passes,
              long* n_bounds, long* n_dropped)
{
    split_ctx c = { cps, type, mv, view, prm, split_limit };
    long nout = 0, nb = 0, dropped = 0;
    int passes = 0;
    const long ppp = prm->patches_per_pass;

    if (method == OR_METHOD_BREADTH) {
        /* Level-synchronous; stable compaction (bound_n_split_breadthfirst.cl);
         * the renderer dices the first reyes_patches_per_pass DRAW ranges of
         * each level (RendererCL.cpp:209). */
        rvec cur = { 0 }, nxt = { 0 };
        for (int i = 0; i < n_patches; ++i) {
            or_range r = { i, 0, 0.0f, 0.0f, 1.0f, 1.0f };
            rpush(&cur, r);
        }
        while (cur.n > 0) {
            long draws = 0;
            nxt.n = 0;
            for (long i = 0; i < cur.n; ++i) {
                int fl = sbound(&c, &cur.d[i]);
                ++nb;
                if (fl & 1) {
                    if (draws < ppp) {
                        if (nout >= out_cap) { free(cur.d); free(nxt.d); return -1; }
                        out[nout++] = cur.d[i];
                    } else {
                        ++dropped;
                    }
                    ++draws;
                } else if (fl & 2) {
                    or_range c0, c1;
                    split_range(&cur.d[i], fl, &c0, &c1);
                    rpush(&nxt, c0);
                    rpush(&nxt, c1);
                }
            }
            if (passes < max_passes && pass_counts) pass_counts[passes] = (int32_t)draws;
            ++passes;
            rvec t = cur; cur = nxt; nxt = t;
        }
        free(cur.d); free(nxt.d);
    } else if (method == OR_METHOD_BOUNDED) {
        /* LIFO stack processed in batches of reyes_patches_per_pass
         * (BoundNSplitCLBounded.cpp:144-214, bound_n_split_multipass.cl). */
        rvec st = { 0 };
        for (int i = 0; i < n_patches; ++i) {
            or_range r = { i, 0, 0.0f, 0.0f, 1.0f, 1.0f };
            rpush(&st, r);
        }
        long height = st.n;
        or_range* pad = (or_range*)malloc((size_t)(ppp > 0 ? ppp : 1) * sizeof(or_range));
        int* flg = (int*)malloc((size_t)(ppp > 0 ? ppp : 1) * sizeof(int));
        while (height > 0) {
            long b = height < ppp ? height : ppp;
            height -= b;
            memcpy(pad, st.d + height, (size_t)b * sizeof(or_range));
            long draws = 0, splits = 0;
            for (long i = 0; i < b; ++i) { flg[i] = sbound(&c, &pad[i]); ++nb; }
            for (long i = 0; i < b; ++i) {
                if (flg[i] & 1) {
                    if (nout >= out_cap) { free(st.d); free(pad); free(flg); return -1; }
                    out[nout++] = pad[i];
                    ++draws;
                } else if (flg[i] & 2) {
                    or_range c0, c1;
                    split_range(&pad[i], flg[i], &c0, &c1);
                    long pos0 = height + splits * 2;
                    while (st.n < pos0 + 2) rpush(&st, c0);
                    st.d[pos0] = c0;
                    st.d[pos0 + 1] = c1;
                    ++splits;
                }
            }
            height += splits * 2;
            if (st.n < height) st.n = height;
            if (passes < max_passes && pass_counts) pass_counts[passes] = (int32_t)draws;
            ++passes;
        }
        free(st.d); free(pad); free(flg);
    } else if (method == OR_METHOD_CPU) {
        /* Host DFS of BoundNSplitCLCPU::do_bound_n_split (:75-159): the stack
         * holds the roots in patch order and pops from the back; a batch ends
         * when it holds reyes_patches_per_pass DRAW ranges (:135-137). */
        rvec st = { 0 };
        for (int i = 0; i < n_patches; ++i) {
            or_range r = { i, 0, 0.0f, 0.0f, 1.0f, 1.0f };
            rpush(&st, r);
        }
        while (st.n > 0) {
            long draws = 0;
            while (st.n > 0) {
                or_range r = st.d[--st.n];
                int fl = or_bound_cpu(cps, type, r.pid, r.depth, r.umin, r.vmin, r.umax, r.vmax,
                                      mv, view, prm, split_limit);
                ++nb;
                if (fl == OR_CULL) continue;
                if (fl == OR_DRAW) {
                    if (nout >= out_cap) { free(st.d); return -1; }
                    out[nout++] = r;
                    if (++draws >= ppp) break;
                } else if (fl == OR_CPU_DROP) {
                    /* depth > max_split_depth: dropped (:139-141) */
                } else {
                    or_range c0, c1;
                    split_range(&r, fl, &c0, &c1);
                    rpush(&st, c0);   /* lower half first: the upper half */
                    rpush(&st, c1);   /* is popped next (:231-245)        */
                }
            }
            if (passes < max_passes && pass_counts) pass_counts[passes] = (int32_t)draws;
            ++passes;
        }
        free(st.d);
    } else {
        /* LOCAL / BALANCED: the DRAW multiset is the canonical set of the
         * predicate (order is scheduling-dependent in the reference).
         * Depth-first traversal; one pass. */
        rvec st = { 0 };
        for (int i = n_patches - 1; i >= 0; --i) {
            or_range r = { i, 0, 0.0f, 0.0f, 1.0f, 1.0f };
            rpush(&st, r);
        }
        while (st.n > 0) {
            or_range r = st.d[--st.n];
            int fl = sbound(&c, &r);
            ++nb;
            if (fl & 1) {
                if (nout >= out_cap) { free(st.d); return -1; }
                out[nout++] = r;
            } else if (fl & 2) {
                or_range c0, c1;
                split_range(&r, fl, &c0, &c1);
                rpush(&st, c1);
                rpush(&st, c0);
            }
        }
        free(st.d);
        if (pass_counts && max_passes > 0) pass_counts[0] = (int32_t)nout;
        passes = 1;
    }
    if (n_passes) *n_passes = passes;
    if (n_bounds) *n_bounds = nb;
    if (n_dropped) *n_dropped = dropped;
    return nout;
}

/* Literal emulation of kernels/bound_n_split_local.cl + the host loop of
 * BoundNSplitCLLocal.cpp:185-245, work-groups executed one after another. */
long or_split_local_emulated(const float* cps, int type, int n_patches,
                             const float mv[16], const or_view* view, const or_params* prm,
                             float split_limit, int wgs, int wgc, long batch,
                             or_range* out, long out_cap, int32_t* pass_counts, int max_passes,
                             int32_t* n_passes, long* n_bounds)
{
    split_ctx c = { cps, type, mv, view, prm, split_limit };
    const int msd = prm->max_split_depth;
    long items = ((long)n_patches + wgc - 1) / wgc * wgc + (long)msd * wgs * wgc;
    long stride = items / wgc;
    or_range* in = (or_range*)calloc((size_t)items, sizeof(or_range));
    int32_t* in_cnt = (int32_t*)calloc((size_t)wgc, sizeof(int32_t));
    or_range* stack = (or_range*)malloc((size_t)wgs * (msd + 2) * sizeof(or_range));
    or_range* lane = (or_range*)malloc((size_t)wgs * sizeof(or_range));
    int* fl = (int*)malloc((size_t)wgs * sizeof(int));
    or_range* obuf = (or_range*)malloc((size_t)batch * sizeof(or_range));
    long nout = 0, nb = 0;
    int passes = 0;

    for (int gid = 0; gid < n_patches; ++gid) {            /* init_range_buffers */
        int wid = gid % wgc, lid = gid / wgc;
        or_range r = { gid, 0, 0.0f, 0.0f, 1.0f, 1.0f };
        in[(long)lid + (long)wid * stride] = r;
    }
    for (int w = 0; w < wgc; ++w)                           /* init_count_buffers */
        in_cnt[w] = (n_patches + (wgc - w - 1)) / wgc;

    for (int iteration = 1;; ++iteration) {
        long out_cnt = 0;
        for (int w = 0; w < wgc; ++w) {
            long sh = 0;
            for (;;) {
                long cnt, start, scnt;
                if (sh < wgs) {
                    cnt = wgs - sh;
                    long top = in_cnt[w];
                    in_cnt[w] -= (int32_t)cnt;
                    cnt = cnt < (top > 0 ? top : 0) ? cnt : (top > 0 ? top : 0);
                    start = top - cnt;
                    scnt = sh < wgs - cnt ? sh : wgs - cnt;
                    sh -= scnt;
                } else {
                    start = 0; cnt = 0;
                    scnt = sh < wgs ? sh : wgs;
                    sh -= scnt;
                }
                if (cnt == 0 && scnt == 0) break;
                int occ = (int)(cnt + scnt);
                for (int l = 0; l < occ; ++l)
                    lane[l] = (l < cnt) ? in[(long)w * stride + start + l] : stack[sh + l - cnt];
                for (int l = 0; l < occ; ++l) { fl[l] = sbound(&c, &lane[l]); ++nb; }
                /* split compaction onto the stack */
                long s0 = sh, k = 0;
                for (int l = 0; l < occ; ++l)
                    if (fl[l] & 2) {
                        or_range c0, c1;
                        split_range(&lane[l], fl[l], &c0, &c1);
                        stack[s0 + 2 * k] = c0;
                        stack[s0 + 2 * k + 1] = c1;
                        ++k;
                    }
                sh += 2 * k;
                /* draw compaction into the shared output */
                long nd = 0;
                for (int l = 0; l < occ; ++l) if (fl[l] & 1) ++nd;
                long dstart = out_cnt;
                out_cnt += nd;
                int overflow = (nd + dstart >= batch);
                long offset = 0;
                if (overflow) {
                    offset = in_cnt[w] > 0 ? in_cnt[w] : 0;
                    long keep = batch - dstart > 0 ? batch - dstart : 0;
                    in_cnt[w] = (int32_t)(offset + sh + nd - keep);
                }
                long rank = 0;
                for (int l = 0; l < occ; ++l)
                    if (fl[l] & 1) {
                        if (dstart + rank < batch) obuf[dstart + rank] = lane[l];
                        else if (overflow) {
                            long keep = batch - dstart > 0 ? batch - dstart : 0;
                            in[(long)w * stride + offset + sh + rank - keep] = lane[l];
                        }
                        ++rank;
                    }
                if (overflow) {
                    for (long i = 0; i < sh; ++i) in[(long)w * stride + offset + i] = stack[i];
                    break;
                }
            }
        }
        long emitted = out_cnt < batch ? out_cnt : batch;
        for (long i = 0; i < emitted; ++i) {
            if (nout >= out_cap) { nout = -1; goto done; }
            out[nout++] = obuf[i];
        }
        if (out_cnt <= 0 || iteration > 200) break;
        if (passes < max_passes && pass_counts) pass_counts[passes] = (int32_t)emitted;
        ++passes;
    }
done:
    if (n_passes) *n_passes = passes;
    if (n_bounds) *n_bounds = nb;
    free(in); free(in_cnt); free(stack); free(lane); free(fl); free(obuf);
    return nout;
}

/* Canonical order of DRAW ranges: (pid, umin, vmin).  The split tree's
 * leaves of one patch are disjoint dyadic rectangles, so (umin, vmin) is
 * unique per leaf.
 * Key of a range inside its patch: umin and vmin are dyadic with at most 24
 * halvings per axis (max_split_depth <= 23, the CPU method draws at depth
 * max_split_depth + 1), so umin * 2^24 and vmin * 2^24 are exact integers
 * below 2^24. */
uint64_t or_canonical_key(const or_range* r)
{
    uint64_t u = (uint64_t)(uint32_t)(r->umin * 16777216.0f);
    uint64_t v = (uint64_t)(uint32_t)(r->vmin * 16777216.0f);
    return (u << 24) | v;
}

static int cmp_range(const void* a, const void* b)
{
    const or_range* ra = (const or_range*)a;
    const or_range* rb = (const or_range*)b;
    if (ra->pid != rb->pid) return ra->pid < rb->pid ? -1 : 1;
    uint64_t ka = or_canonical_key(ra), kb = or_canonical_key(rb);
    return ka < kb ? -1 : (ka > kb ? 1 : 0);
}

void or_canonical_sort(or_range* r, long n)
{
    qsort(r, (size_t)n, sizeof(or_range), cmp_range);
}

/* ------------------------------------------------------------------------ */
/* dice (kernels/dice.cl:28-80)                                              */

void or_dice(const float* cps, int type, const or_range* r, const float mv[16],
             const or_view* view, const or_params* prm,
             float* pos, int32_t* pxl, float* depth)
{
    const int N = prm->patch_size;
    const int32_t vsx = view->width << prm->subpixel_bits;
    const int32_t vsy = view->height << prm->subpixel_bits;
    for (int nv = 0; nv <= N; ++nv) {
        for (int nu = 0; nu <= N; ++nu) {
            float tu = (float)nu / (float)N, tv = (float)nv / (float)N;
            float u = r->umin + (r->umax - r->umin) * tu;
            float v = r->vmin + (r->vmax - r->vmin) * tv;
            float P[4], e[4], q[4];
            or_displaced_point(cps, type, r->pid, u, v, prm, P);
            mat4_mul(mv, P, e);
            mat4_mul(view->proj, e, q);
            int gi = nu + nv * (N + 1);
            pos[4 * gi + 0] = e[0]; pos[4 * gi + 1] = e[1];
            pos[4 * gi + 2] = e[2]; pos[4 * gi + 3] = e[3];
            pxl[2 * gi + 0] = f2i_sat((q[0] / q[3]) * (float)vsx / 2.0f + (float)(vsx / 2));
            pxl[2 * gi + 1] = f2i_sat((q[1] / q[3]) * (float)vsy / 2.0f + (float)(vsy / 2));
            depth[gi] = q[2] / q[3];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* shade (kernels/reyes.cl:54-199).  block_index layout: block (bu,bv) at
 * 4*(bu + bv*N/8): (xmin, ymin, xmax, ymax) in sub-pixels.  quad_s[u + v*N]
 * is the Lambert term max(dot(n,l),0) (NaN when the reference colour is NaN
 * through the zero-weighted specular term). */

static int front_facing(const int32_t* px, const int32_t* py, int culling)
{
    if (!culling) return 1;
    int32_t d1x = wsub(px[1], px[0]), d1y = wsub(py[1], py[0]);
    int32_t d2x = wsub(px[2], px[0]), d2y = wsub(py[2], py[0]);
    int32_t d3x = wsub(px[3], px[0]), d3y = wsub(py[3], py[0]);
    return (wsub(wmul(d1x, d3y), wmul(d3x, d1y)) < 0 || wsub(wmul(d3x, d2y), wmul(d2x, d3y)) < 0);
}

static float quad_lambert(const float* p0, const float* p1, const float* p2, const float* p3)
{
    float du[3], dv[3], m[3], n[3], l[3], sum[3], vv[3], lh[3], h[3];
    for (int k = 0; k < 3; ++k) {
        du[k] = ((p1[k] - p0[k]) + (p3[k] - p2[k])) * 0.5f;
        dv[k] = ((p2[k] - p0[k]) + (p3[k] - p1[k])) * 0.5f;
    }
    cross3(dv, du, m);
    normalize3(m, n);
    const float l0[3] = { 4.0f, 3.0f, 8.0f };
    normalize3(l0, l);
    for (int k = 0; k < 3; ++k) sum[k] = ((p0[k] + p1[k]) + p2[k]) + p3[k];
    normalize3(sum, vv);
    for (int k = 0; k < 3; ++k) { vv[k] = -vv[k]; lh[k] = l[k] + vv[k]; }
    normalize3(lh, h);
    float spec = cl_maxf(dot3(n, h), 0.0f);
    float s = cl_maxf(dot3(n, l), 0.0f);
    /* pow(spec, 60) * 0 is 0 for finite spec >= 0 and NaN for NaN spec */
    return (spec != spec) ? spec : s;
}

void or_shade(const float* pos, const int32_t* pxl, const or_view* view, const or_params* prm,
              int32_t* block_index, float* quad_s)
{
    const int N = prm->patch_size, B = N / 8;
    const int sh = prm->subpixel_bits;
    const int32_t vmaxx = (view->width << sh) - 1, vmaxy = (view->height << sh) - 1;
    for (int bv = 0; bv < B; ++bv) {
        for (int bu = 0; bu < B; ++bu) {
            int32_t xmin = vmaxx, ymin = vmaxy, xmax = 0, ymax = 0;
            int allnormal = 1;
            for (int qv = 0; qv < 8; ++qv) {
                for (int qu = 0; qu < 8; ++qu) {
                    int u = bu * 8 + qu, v = bv * 8 + qv;
                    int32_t px[4], py[4];
                    int32_t pminx = vmaxx, pminy = vmaxy, pmaxx = 0, pmaxy = 0;
                    const float* p[4];
                    for (int i = 0; i < 4; ++i) {
                        int gi = (u + (i & 1)) + (v + (i >> 1)) * (N + 1);
                        p[i] = pos + 4 * gi;
                        if (p[i][2] == 0.0f) allnormal = 0;
                        px[i] = pxl[2 * gi]; py[i] = pxl[2 * gi + 1];
                        pminx = imin(pminx, px[i]); pminy = imin(pminy, py[i]);
                        pmaxx = imax(pmaxx, px[i]); pmaxy = imax(pmaxy, py[i]);
                    }
                    if (front_facing(px, py, prm->backface_culling)) {
                        xmin = imin(xmin, pminx); ymin = imin(ymin, pminy);
                        xmax = imax(xmax, pmaxx); ymax = imax(ymax, pmaxy);
                    }
                    quad_s[u + v * N] = quad_lambert(p[0], p[1], p[2], p[3]);
                }
            }
            xmin = imax(0, xmin); ymin = imax(0, ymin);
            xmax = imin(vmaxx, xmax); ymax = imin(vmaxy, ymax);
            if (!allnormal) { xmin = 1; ymin = 1; xmax = -1; ymax = -1; }
            int32_t* bi = block_index + 4 * (bu + bv * B);
            bi[0] = xmin; bi[1] = ymin; bi[2] = xmax; bi[3] = ymax;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* sample (kernels/reyes.cl:202-398) with a deterministic visibility key:
 * key = (monotone_bits(depth) << 32) | quad_id, depth being the value the
 * reference quantises (reyes.cl:359-364).  The smallest key wins; its depth
 * order refines the reference's quantised idepth order, so the winner is
 * always one the reference's race could pick. */

static inline uint32_t mono_bits(float d)
{
    uint32_t b;
    memcpy(&b, &d, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

static inline float unmono_bits(uint32_t m)
{
    uint32_t b = (m & 0x80000000u) ? (m & 0x7fffffffu) : ~m;
    float d;
    memcpy(&d, &b, 4);
    return d;
}

int or_framebuffer_index(int x, int y, int width)
{
    int gw = (width + 7) / 8;
    int gx = x / 8, gy = y / 8;
    return (gx + gw * gy) * 64 + (x - gx * 8) + 8 * (y - gy * 8);
}

long or_framebuffer_size(int width, int height)
{
    return (long)((width + 7) / 8) * 8 * (long)((height + 7) / 8) * 8;
}

/* inside_triangle (reyes.cl:240-265), wrapping int32 arithmetic */
static int inside_tri(const int32_t* Px, const int32_t* Py, const float* dv,
                      int32_t tx, int32_t ty, float* depth)
{
    int32_t Dx[3], Dy[3], O[3], C[3], V[3];
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3;
        Dx[i] = wsub(Py[j], Py[i]);
        Dy[i] = wsub(Px[i], Px[j]);
    }
    for (int i = 0; i < 3; ++i) {
        O[i] = wadd(wmul(Dx[i], Px[i]), wmul(Dy[i], Py[i]));
        C[i] = (Dx[i] > 0 || (Dx[i] == 0 && Dy[i] > 0)) ? -1 : 0;
        V[i] = wsub(wadd(wmul(Dx[i], tx), wmul(Dy[i], ty)), O[i]);
    }
    int success = (V[0] > C[0]) && (V[1] > C[1]) && (V[2] > C[2]);
    float w[3];
    for (int k = 0; k < 3; ++k) {
        int j = (k + 1) % 3;
        int32_t den = wsub(wadd(wmul(Dx[j], Px[k]), wmul(Dy[j], Py[k])), O[j]);
        w[k] = (float)V[j] / (float)den;
    }
    float d = mad(dv[2], w[2], mad(dv[1], w[1], dv[0] * w[0]));
    if (success && d < *depth) *depth = d;
    return success;
}

void or_sample(const int32_t* block_index, const int32_t* pxl, const float* depth,
               const or_view* view, const or_params* prm, uint32_t id_base,
               uint64_t* keybuf)
{
    const int N = prm->patch_size, B = N / 8;
    const int sh = prm->subpixel_bits;
    const int32_t vmaxx = (view->width << sh) - 1, vmaxy = (view->height << sh) - 1;
    const int32_t maxl = (8 << sh) - 1;
    for (int bv = 0; bv < B; ++bv) {
        for (int bu = 0; bu < B; ++bu) {
            const int32_t* bb = block_index + 4 * (bu + bv * B);
            int32_t mintx = bb[0] >> (sh + 3), minty = bb[1] >> (sh + 3);
            int32_t maxtx = bb[2] >> (sh + 3), maxty = bb[3] >> (sh + 3);
            for (int32_t ty = minty; ty <= maxty; ++ty) {
                for (int32_t tx = mintx; tx <= maxtx; ++tx) {
                    int32_t osx = (tx * 8) << sh, osy = (ty * 8) << sh;
                    for (int qv = 0; qv < 8; ++qv) {
                        for (int qu = 0; qu < 8; ++qu) {
                            int u = bu * 8 + qu, v = bv * 8 + qv;
                            int32_t Px[4], Py[4];
                            float dv[4];
                            int32_t mingx = vmaxx + 1, mingy = vmaxy + 1, maxgx = -1, maxgy = -1;
                            for (int i = 0; i < 4; ++i) {
                                int gi = (u + (i & 1)) + (v + (i >> 1)) * (N + 1);
                                Px[i] = pxl[2 * gi]; Py[i] = pxl[2 * gi + 1];
                                dv[i] = depth[gi];
                                mingx = imin(mingx, Px[i]); mingy = imin(mingy, Py[i]);
                                maxgx = imax(maxgx, Px[i]); maxgy = imax(maxgy, Py[i]);
                            }
                            int32_t minpx = iclamp(wsub(mingx, osx), 0, maxl) >> sh;
                            int32_t minpy = iclamp(wsub(mingy, osy), 0, maxl) >> sh;
                            int32_t maxpx = iclamp(wsub(maxgx, osx), 0, maxl) >> sh;
                            int32_t maxpy = iclamp(wsub(maxgy, osy), 0, maxl) >> sh;
                            const int32_t t1x[3] = { Px[0], Px[1], Px[3] }, t1y[3] = { Py[0], Py[1], Py[3] };
                            const int32_t t2x[3] = { Px[0], Px[3], Px[2] }, t2y[3] = { Py[0], Py[3], Py[2] };
                            const float d1[3] = { dv[0], dv[1], dv[3] }, d2[3] = { dv[0], dv[3], dv[2] };
                            uint32_t id = id_base + (uint32_t)(u + v * N);
                            for (int32_t y = minpy; y <= maxpy; ++y) {
                                for (int32_t x = minpx; x <= maxpx; ++x) {
                                    int32_t tpx = (x << sh) + osx, tpy = (y << sh) + osy;
                                    float dd = 1.0f;
                                    int in1 = inside_tri(t1x, t1y, d1, tpx, tpy, &dd);
                                    int in2 = inside_tri(t2x, t2y, d2, tpx, tpy, &dd);
                                    if (in1 || in2) {
                                        uint64_t key = ((uint64_t)mono_bits(dd) << 32) | id;
                                        int fi = or_framebuffer_index(tx * 8 + x, ty * 8 + y, view->width);
                                        if (key < keybuf[fi]) keybuf[fi] = key;
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
    }
}

void or_render(const float* cps, int type, const or_range* ranges, long n,
               uint32_t range_base, const float mv[16], const or_view* view,
               const or_params* prm, uint64_t* keybuf, float* quad_store)
{
    const int N = prm->patch_size, G = (N + 1) * (N + 1), B = N / 8;
    float* pos = (float*)malloc((size_t)G * 4 * sizeof(float));
    int32_t* pxl = (int32_t*)malloc((size_t)G * 2 * sizeof(int32_t));
    float* dep = (float*)malloc((size_t)G * sizeof(float));
    int32_t* bidx = (int32_t*)malloc((size_t)B * B * 4 * sizeof(int32_t));
    for (long i = 0; i < n; ++i) {
        uint32_t rid = range_base + (uint32_t)i;
        or_dice(cps, type, &ranges[i], mv, view, prm, pos, pxl, dep);
        or_shade(pos, pxl, view, prm, bidx, quad_store + (size_t)rid * N * N);
        or_sample(bidx, pxl, dep, view, prm, rid * (uint32_t)(N * N), keybuf);
    }
    free(pos); free(pxl); free(dep); free(bidx);
}

uint8_t or_to_srgb8(float c)
{
    if (!(c > 0.0f)) return 0;
    float g = powf(c, 0.454545f);
    if (!(g < 1.0f)) return 255;
    return (uint8_t)(g * 255.0f + 0.5f);
}

void or_resolve(const uint64_t* keybuf, const float* quad_store, const float* range_color,
                const or_view* view, const or_params* prm,
                uint8_t* rgba8, float* depth, uint32_t* ids, float* rgba32)
{
    const int N = prm->patch_size;
    const float ac[4] = { 0.015f, 0.015f, 0.015f, 1.0f };
    for (int y = 0; y < view->height; ++y) {
        for (int x = 0; x < view->width; ++x) {
            long o = (long)y * view->width + x;
            uint64_t key = keybuf[or_framebuffer_index(x, y, view->width)];
            float c[4];
            float d = 1.0f;
            uint32_t id = 0xFFFFFFFFu;
            if (key == OR_EMPTY_KEY) {
                memcpy(c, view->clear_linear, sizeof(c));
            } else {
                id = (uint32_t)key;
                d = unmono_bits((uint32_t)(key >> 32));
                uint32_t r = id / (uint32_t)(N * N);
                float s = quad_store[id];
                const float* dc = range_color + 4 * (size_t)r;
                for (int k = 0; k < 4; ++k) c[k] = mad(s, dc[k], ac[k] * dc[k]) + 0.0f;
            }
            if (rgba8) {
                rgba8[4 * o + 0] = or_to_srgb8(c[0]);
                rgba8[4 * o + 1] = or_to_srgb8(c[1]);
                rgba8[4 * o + 2] = or_to_srgb8(c[2]);
                rgba8[4 * o + 3] = 255;
            }
            if (depth) depth[o] = d;
            if (ids) ids[o] = id;
            if (rgba32) memcpy(rgba32 + 4 * o, c, sizeof(c));
        }
    }
}

/* ------------------------------------------------------------------------ */

int32_t or_prefix_sum(const int32_t* in, int32_t* out, long n)
{
    int32_t s = 0;
    for (long i = 0; i < n; ++i) { s = wadd(s, in[i]); out[i] = s; }
    return s;
}

/* glibc random() TYPE_3 (additive feedback, r[i] = r[i-3] + r[i-31]) */
void or_glibc_rand(uint32_t seed, int32_t* out, long n)
{
    long total = 344 + n;
    int32_t* r = (int32_t*)malloc((size_t)total * sizeof(int32_t));
    r[0] = (int32_t)(seed ? seed : 1);
    for (int i = 1; i < 31; ++i) {
        long long v = (16807LL * r[i - 1]) % 2147483647LL;
        if (v < 0) v += 2147483647LL;
        r[i] = (int32_t)v;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (long i = 34; i < total; ++i) r[i] = (int32_t)((uint32_t)r[i - 31] + (uint32_t)r[i - 3]);
    for (long k = 0; k < n; ++k) out[k] = (int32_t)(((uint32_t)r[k + 344]) >> 1);
    free(r);
}

void or_make_view(float fovy_deg, float near_, int width, int height,
                  const float clear_color[4], or_view* v)
{
    memset(v, 0, sizeof(*v));
    v->width = width;
    v->height = height;
    float aspect = (float)width / (float)height;
    /* glm::perspective<float>(fovy*pi/180, aspect, near, 10000) */
    float fovy = (float)((double)fovy_deg * 3.14159265358979323846 / 180.0);
    float zfar = 10000.0f;
    float th = tanf(fovy / 2.0f);
    v->proj[0] = 1.0f / (aspect * th);
    v->proj[5] = 1.0f / th;
    v->proj[10] = -(zfar + near_) / (zfar - near_);
    v->proj[11] = -1.0f;
    v->proj[14] = -(2.0f * zfar * near_) / (zfar - near_);
    /* Projection::Projection */
    float fy = 1.0f / (float)tan((double)fovy_deg * 3.14159265358979323846 / 360.0);
    float fx = fy / aspect;
    float vpx = (float)width / 2.0f, vpy = (float)height / 2.0f;
    fx *= vpx;
    fy *= vpy;
    v->f[0] = fx;
    v->f[1] = fy;
    v->screen[0] = vpx; v->screen[1] = 0.0f; v->screen[2] = 0.0f; v->screen[3] = vpy;
    v->near_ = near_;
    v->far_ = 1000.0f;
    v->eps = (float)((double)near_ * 0.1);
    for (int k = 0; k < 3; ++k) v->clear_linear[k] = powf(clear_color[k], 2.2f);
    v->clear_linear[3] = 1000.0f;
}
