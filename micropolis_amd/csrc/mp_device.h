// Device math of the REYES hot path for gfx950.
//
// Every floating-point expression follows the operation order written in the
// reference kernels (kernels/utility.h, kernels/bound_n_split.h,
// kernels/dice.cl, kernels/reyes.cl).  The reference builds its kernels with
// -cl-mad-enable (src/CL/Program.cpp:118), which lets a product and the sum
// it feeds be contracted; here that contraction is one fixed rule: s + p*q
// is fma(p, q, s) and p*q + r*s (or -) is fma(r, s, p*q), the left product
// rounded, each written out with mad() / mad_diff().  Nothing else is
// contracted (the library is built with -ffp-contract=off) and '/' and sqrt
// are the default correctly rounded ones, so every result equals the scalar
// evaluation with C99 fmaf at the same places.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mp_types.h"

#define MP_DEV __device__ __forceinline__

// Global-address-space view of a device pointer held in a table (object,
// view or queue records).  Through it the compiler emits global_* memory
// instructions instead of flat_* ones; flat instructions also count against
// the LDS/scalar wait counter, so every later wait for an LDS result would
// wait for them as well.
#define MP_GLOBAL __attribute__((address_space(1)))
template <typename T>
MP_DEV MP_GLOBAL T* gptr(T* p) { return (MP_GLOBAL T*)p; }

typedef float mp_f4v __attribute__((ext_vector_type(4)));
// 16-byte global load / store of a float4 held behind a table pointer.
MP_DEV float4 ldg4(const float4* p)
{
    const mp_f4v v = *(const MP_GLOBAL mp_f4v*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
MP_DEV void stg4(float4* p, const float4& v)
{
    mp_f4v w;
    w.x = v.x; w.y = v.y; w.z = v.z; w.w = v.w;
    *(MP_GLOBAL mp_f4v*)p = w;
}

// Visibility update: 64-bit atomic minimum, relaxed, device scope, no return.
MP_DEV void key_min(unsigned long long* p, unsigned long long v)
{
    __hip_atomic_fetch_min(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

namespace mp {

// Wave votes over the active lanes: the ballot's lane mask against exec.
// A ballot of one comparison is that comparison's lane mask; a compound
// predicate is voted on per comparison (wave_all2 / wave_all3) so the masks
// combine on the scalar unit instead of going through a vector select.
MP_DEV uint64_t lane_mask(bool p) { return __builtin_amdgcn_ballot_w64(p); }
MP_DEV bool wave_all(bool p) { return lane_mask(p) == __builtin_amdgcn_read_exec(); }
MP_DEV bool wave_all2(bool p, bool q) { return (lane_mask(p) & lane_mask(q)) == __builtin_amdgcn_read_exec(); }
MP_DEV bool wave_all3(bool p, bool q, bool r)
{
    return (lane_mask(p) & lane_mask(q) & lane_mask(r)) == __builtin_amdgcn_read_exec();
}
MP_DEV bool wave_any(bool p) { return lane_mask(p) != 0; }
// Lane mask of v_cmp_class_f32 (the compiler turns finiteness tests into
// class tests, whose ballots it would route through a vector select).
constexpr int kClassFinite = 0x1F8;          // +-normal, +-subnormal, +-0
constexpr int kClassFiniteNonZero = 0x198;   // +-normal, +-subnormal
MP_DEV uint64_t class_mask(float x, int cls)
{
    uint64_t m;
    asm("v_cmp_class_f32_e64 %0, %1, %2" : "=s"(m) : "v"(x), "s"(cls));
    return m;
}

// OpenCL min(x,y): y if y < x else x.  OpenCL max(x,y): y if x < y else x.
MP_DEV float cl_min(float x, float y) { return (y < x) ? y : x; }
MP_DEV float cl_max(float x, float y) { return (x < y) ? y : x; }

MP_DEV int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
MP_DEV int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
MP_DEV int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }

// float -> int32: NaN -> 0, saturating, else truncation toward zero.  That
// is v_cvt_i32_f32's own definition (out-of-range values and infinities
// saturate, NaN converts to 0), so one instruction does it; C++'s (int)x
// leaves those cases undefined and the compiler guards them with branches.
MP_DEV int32_t f2i_sat(float x)
{
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// s + p*q as one fused multiply-add (the contraction rule above)
MP_DEV float mad(float p, float q, float s) { return __builtin_fmaf(p, q, s); }
// p*q - r*s: the left product rounded, the right one fused
MP_DEV float mad_diff(float p, float q, float r, float s) { return __builtin_fmaf(-r, s, p * q); }

MP_DEV float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    return mad(az, bz, mad(ay, by, ax * bx));
}

// column-major 4x4 times vector, rows summed left to right
MP_DEV void mat4_mul(const float* __restrict__ m, const float v[4], float o[4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i)
        o[i] = mad(m[12 + i], v[3], mad(m[8 + i], v[2], mad(m[4 + i], v[1], m[0 + i] * v[0])));
}

// proj * e for the perspective matrices of every view (glm::perspective,
// Projection.cpp:39: only entries 0, 5, 10, 11 = -1 and 14 are non-zero).
// With e finite and e.z, e.w non-zero the full product's rows reduce to
// p0 e0, p5 e1, fma(p14, e3, p10 e2) and -e2 exactly (a fused zero product
// adds a signed zero to a non-zero sum), up to the sign of a zero x
// or y row (zero terms add signed zeros), which no pixel coordinate
// computed from q.x / q.w or q.y / q.w can tell apart; the wave takes the
// full product when any lane's point is not such a point.
MP_DEV void perspective_mul(const float* __restrict__ m, const float e[4], float q[4])
{
    const uint64_t plain = class_mask(e[0], kClassFinite) & class_mask(e[1], kClassFinite) &
                           class_mask(e[2], kClassFiniteNonZero) & class_mask(e[3], kClassFiniteNonZero);
    if (plain == __builtin_amdgcn_read_exec()) {
        q[0] = m[0] * e[0];
        q[1] = m[5] * e[1];
        q[2] = mad(m[14], e[3], m[10] * e[2]);
        q[3] = -e[2];
    } else {
        mat4_mul(m, e, q);
    }
}

MP_DEV void bern_u(float t, float s, float a[4])
{
    a[0] = (s * s) * s;
    a[1] = ((3.0f * s) * s) * t;
    a[2] = ((3.0f * s) * t) * t;
    a[3] = (t * t) * t;
}
MP_DEV void bern_v(float t, float s, float b[4])
{
    b[0] = (s * s) * s;
    b[1] = ((3.0f * t) * s) * s;
    b[2] = ((3.0f * t) * t) * s;
    b[3] = (t * t) * t;
}
MP_DEV void dbern(float t, float s, float d[4])
{
    d[0] = -3.0f * (s * s);
    d[1] = 3.0f * (s * (s - 2.0f * t));
    d[2] = 3.0f * (t * (2.0f * s - t));
    d[3] = 3.0f * (t * t);
}

MP_DEV void acc4(float p[4], float w, const float4 P)
{
    p[0] = mad(w, P.x, p[0]);
    p[1] = mad(w, P.y, p[1]);
    p[2] = mad(w, P.z, p[2]);
    p[3] = mad(w, P.w, p[3]);
}

// Bicubic Bezier from its Bernstein weights a (u) and b (v).
MP_DEV void eval_bezier_w(const float4* __restrict__ cp, const float a[4], const float b[4], float p[4])
{
    p[0] = p[1] = p[2] = p[3] = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            acc4(p, a[i] * b[j], cp[4 * i + j]);
}

// Bicubic Bezier (utility.h:41-72); cp = 16 control points (xyz, w=1).
MP_DEV void eval_bezier(const float4* __restrict__ cp, float u, float v, float p[4])
{
    float su = 1.0f - u, sv = 1.0f - v;
    float a[4], b[4];
    bern_u(u, su, a);
    bern_v(v, sv, b);
    eval_bezier_w(cp, a, b, p);
}

// Two bicubic Bezier points, p0 of the patch cp0 and p1 of cp1, each sum in
// eval_bezier's order of operations; with cp0 == cp1 each control point is
// read once for both.
MP_DEV void eval_bezier2(const float4* cp0, const float4* cp1, float u0, float v0, float u1, float v1,
                         float p0[4], float p1[4])
{
    float a0[4], b0[4], a1[4], b1[4];
    bern_u(u0, 1.0f - u0, a0);
    bern_v(v0, 1.0f - v0, b0);
    bern_u(u1, 1.0f - u1, a1);
    bern_v(v1, 1.0f - v1, b1);
    p0[0] = p0[1] = p0[2] = p0[3] = 0.0f;
    p1[0] = p1[1] = p1[2] = p1[3] = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            acc4(p0, a0[i] * b0[j], cp0[4 * i + j]);
            acc4(p1, a1[i] * b1[j], cp1[4 * i + j]);
        }
}

MP_DEV float nz(float x) { return (x != x) ? 0.0f : x; }

MP_DEV float4 greg_f(float nx, float ny, float nz_, float nw, float d)
{
    return make_float4(nz(nx / d), nz(ny / d), nz(nz_ / d), nz(nw / d));
}

// 20-point Gregory patch (utility.h:75-126).
MP_DEV void eval_gregory(const float4* __restrict__ P, float u, float v, float p[4])
{
    const float omu = 1.0f - u, omv = 1.0f - v;
    const float d0 = u + v, d1 = (1.0f - u) + v, d2 = (2.0f - u) - v, d3 = (1.0f + u) - v;
    const float4 a12 = P[12], a16 = P[16], a17 = P[17], a13 = P[13];
    const float4 a14 = P[14], a18 = P[18], a19 = P[19], a15 = P[15];
    float4 F0 = greg_f(mad(v, a16.x, u * a12.x), mad(v, a16.y, u * a12.y), mad(v, a16.z, u * a12.z),
                       mad(v, a16.w, u * a12.w), d0);
    float4 F1 = greg_f(mad(v, a13.x, omu * a17.x), mad(v, a13.y, omu * a17.y), mad(v, a13.z, omu * a17.z),
                       mad(v, a13.w, omu * a17.w), d1);
    float4 F2 = greg_f(mad(omv, a18.x, omu * a14.x), mad(omv, a18.y, omu * a14.y),
                       mad(omv, a18.z, omu * a14.z), mad(omv, a18.w, omu * a14.w), d2);
    float4 F3 = greg_f(mad(omv, a15.x, u * a19.x), mad(omv, a15.y, u * a19.y), mad(omv, a15.z, u * a19.z),
                       mad(omv, a15.w, u * a19.w), d3);
    float su = 1.0f - u, sv = 1.0f - v;
    float a[4], b[4];
    bern_u(u, su, a);
    bern_v(v, sv, b);
    p[0] = p[1] = p[2] = p[3] = 0.0f;
    acc4(p, a[0] * b[0], P[0]);  acc4(p, a[0] * b[1], P[4]);
    acc4(p, a[0] * b[2], P[9]);  acc4(p, a[0] * b[3], P[1]);
    acc4(p, a[1] * b[0], P[8]);  acc4(p, a[1] * b[1], F0);
    acc4(p, a[1] * b[2], F1);    acc4(p, a[1] * b[3], P[5]);
    acc4(p, a[2] * b[0], P[7]);  acc4(p, a[2] * b[1], F3);
    acc4(p, a[2] * b[2], F2);    acc4(p, a[2] * b[3], P[10]);
    acc4(p, a[3] * b[0], P[3]);  acc4(p, a[3] * b[1], P[11]);
    acc4(p, a[3] * b[2], P[6]);  acc4(p, a[3] * b[3], P[2]);
}

MP_DEV void eval_patch(const float4* __restrict__ cp, int type, float u, float v, float p[4])
{
    if (type == MP_TYPE_GREGORY) eval_gregory(cp, u, v, p);
    else                         eval_bezier(cp, u, v, p);
}

// ---- 3D value noise (defined by this library; see the oracle for the
// identical scalar statement) ----
MP_DEV uint32_t noise_hash(int32_t x, int32_t y, int32_t z, uint32_t seed)
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
MP_DEV float lattice(int32_t x, int32_t y, int32_t z, uint32_t seed)
{
    return (float)(noise_hash(x, y, z, seed) >> 8) * (1.0f / 8388608.0f) - 1.0f;
}
// Same value for wave-uniform arguments; the readfirstlane keeps the hash on
// the scalar unit (the compiler would otherwise merge it with the per-lane
// path and run it on the vector unit).
MP_DEV float lattice_wave(int32_t x, int32_t y, int32_t z, uint32_t seed)
{
    const uint32_t h = (uint32_t)__builtin_amdgcn_readfirstlane((int)noise_hash(x, y, z, seed));
    // m * 2^-23 - 1 for m = h >> 8 < 2^24, built on the scalar unit: the
    // float with mantissa bits (m mod 2^23) and exponent 0 is 1 + (m mod 2^23)
    // * 2^-23, so subtracting 2 (m < 2^23) or 1 (m >= 2^23) gives the value
    // exactly, with one vector instruction instead of three
    const uint32_t m = h >> 8;
    const float one_plus = __uint_as_float(0x3f800000u | (m & 0x7fffffu));
    return one_plus - ((m >> 23) ? 1.0f : 2.0f);
}
// (t * t) * (3 - 2t): 2t is exact, so 3 - 2t as one fma is the written
// difference
MP_DEV float fade(float t) { return (t * t) * __builtin_fmaf(-2.0f, t, 3.0f); }
MP_DEV float lerpf_(float a, float b, float t) { return mad(b - a, t, a); }

// Value of v held by lane byte_addr / 4 of the wave (ds_bpermute_b32).
MP_DEV float lane_value(int byte_addr, float v)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(byte_addr, __float_as_int(v)));
}

// Lattice schemes of vnoise: per-lane hashing of all eight corners; scalar
// hashing when the whole wave sits in one cell; or the lane window below.
enum { kNoisePerLane = 0, kNoiseWaveCell = 1, kNoiseLaneWindow = 2 };

// Lane window (whole wave active): the wave's cells lie within +-1 cell of
// lane 32's cell in each axis in practice, so the 4x4x4 lattice corners
// around that cell are exactly 64 values.  Lane k hashes corner
// (bx + (k & 3), by + (k >> 2 & 3), bz + (k >> 4)) once, and every lane
// reads its eight corners from the lanes that hashed them.  One hash per lane
// per octave instead of eight; the values are lattice()'s (coordinates wrap
// in 32 bits, as lattice()'s multiplications do).  Returns false (nothing
// read) when a lane's cell falls outside the window.
MP_DEV bool window_corners(int32_t ix, int32_t iy, int32_t iz, uint32_t seed, float c[8])
{
    const uint32_t bx = (uint32_t)__builtin_amdgcn_readlane(ix, 32) - 1u;
    const uint32_t by = (uint32_t)__builtin_amdgcn_readlane(iy, 32) - 1u;
    const uint32_t bz = (uint32_t)__builtin_amdgcn_readlane(iz, 32) - 1u;
    const uint32_t dx = (uint32_t)ix - bx, dy = (uint32_t)iy - by, dz = (uint32_t)iz - bz;
    if (!wave_all(dx <= 2u && dy <= 2u && dz <= 2u)) return false;
    const uint32_t k = threadIdx.x & 63u;
    const float hv = lattice((int32_t)(bx + (k & 3u)), (int32_t)(by + ((k >> 2) & 3u)),
                             (int32_t)(bz + (k >> 4)), seed);
    const int a = (int)((dx + 4u * dy + 16u * dz) << 2);
    c[0] = lane_value(a, hv);      c[1] = lane_value(a + 4, hv);    // z0 y0: x0 x1
    c[2] = lane_value(a + 16, hv); c[3] = lane_value(a + 20, hv);   // z0 y1
    c[4] = lane_value(a + 64, hv); c[5] = lane_value(a + 68, hv);   // z1 y0
    c[6] = lane_value(a + 80, hv); c[7] = lane_value(a + 84, hv);   // z1 y1
    return true;
}

MP_DEV float vnoise(float x, float y, float z, uint32_t seed, int scheme = kNoisePerLane)
{
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    float wx = fade(x - fx), wy = fade(y - fy), wz = fade(z - fz);
    float c000, c100, c010, c110, c001, c101, c011, c111;
    const int32_t jx = __builtin_amdgcn_readfirstlane(ix);
    const int32_t jy = __builtin_amdgcn_readfirstlane(iy);
    const int32_t jz = __builtin_amdgcn_readfirstlane(iz);
    float c[8];
    if (scheme == kNoiseLaneWindow && window_corners(ix, iy, iz, seed, c)) {
        c000 = c[0]; c100 = c[1]; c010 = c[2]; c110 = c[3];
        c001 = c[4]; c101 = c[5]; c011 = c[6]; c111 = c[7];
    } else if (scheme == kNoiseWaveCell && wave_all(ix == jx && iy == jy && iz == jz)) {
        // every active lane is in the same lattice cell: the eight corner
        // hashes are wave-uniform (scalar unit), the values are identical
        c000 = lattice_wave(jx, jy, jz, seed);         c100 = lattice_wave(jx + 1, jy, jz, seed);
        c010 = lattice_wave(jx, jy + 1, jz, seed);     c110 = lattice_wave(jx + 1, jy + 1, jz, seed);
        c001 = lattice_wave(jx, jy, jz + 1, seed);     c101 = lattice_wave(jx + 1, jy, jz + 1, seed);
        c011 = lattice_wave(jx, jy + 1, jz + 1, seed); c111 = lattice_wave(jx + 1, jy + 1, jz + 1, seed);
    } else {
        c000 = lattice(ix, iy, iz, seed);         c100 = lattice(ix + 1, iy, iz, seed);
        c010 = lattice(ix, iy + 1, iz, seed);     c110 = lattice(ix + 1, iy + 1, iz, seed);
        c001 = lattice(ix, iy, iz + 1, seed);     c101 = lattice(ix + 1, iy, iz + 1, seed);
        c011 = lattice(ix, iy + 1, iz + 1, seed); c111 = lattice(ix + 1, iy + 1, iz + 1, seed);
    }
    float x00 = lerpf_(c000, c100, wx), x10 = lerpf_(c010, c110, wx);
    float x01 = lerpf_(c001, c101, wx), x11 = lerpf_(c011, c111, wx);
    float y0 = lerpf_(x00, x10, wy), y1 = lerpf_(x01, x11, wy);
    return lerpf_(y0, y1, wz);
}

// Lane windows kept by one wave across calls (the raster's dice passes and
// the ranges after them, which mostly fall in the same lattice cells): per
// octave the window's base corner, kept in LDS (one uint4 per octave per
// wave: x, y, z, and w = 1 once hashed), and this lane's hashed corner value
// in a register.  A call whose cells all fit the kept window reads it with no
// hashing; otherwise a new window around lane 32's cell is hashed and kept
// (as window_corners does), and if even that does not fit, each lane hashes
// its own eight corners.  The values are lattice()'s either way.  Whole
// waves only; each wave owns its slots (LDS accesses of one wave are done in
// order, so no barrier is needed between its calls).
constexpr int kKeptOctaves = 4;
struct NoiseWindows {
    uint4* base;   // LDS, kKeptOctaves slots of the calling wave
    float hv[kKeptOctaves];
    MP_DEV void reset(uint4* slots)
    {
        base = slots;
#pragma unroll
        for (int k = 0; k < kKeptOctaves; ++k) {
            hv[k] = 0.0f;
            slots[k] = make_uint4(0u, 0u, 0u, 0u);
        }
    }
};

MP_DEV float vnoise_kept(float x, float y, float z, uint32_t seed, uint4* slot, float& hv)
{
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    float wx = fade(x - fx), wy = fade(y - fy), wz = fade(z - fz);
    const uint4 B = *slot;
    uint32_t dx = (uint32_t)ix - B.x, dy = (uint32_t)iy - B.y, dz = (uint32_t)iz - B.z;
    bool fits = wave_all2(B.w != 0u, max(max(dx, dy), dz) <= 2u);
    if (!fits) {
        const uint32_t nx = (uint32_t)__builtin_amdgcn_readlane(ix, 32) - 1u;
        const uint32_t ny = (uint32_t)__builtin_amdgcn_readlane(iy, 32) - 1u;
        const uint32_t nz = (uint32_t)__builtin_amdgcn_readlane(iz, 32) - 1u;
        dx = (uint32_t)ix - nx; dy = (uint32_t)iy - ny; dz = (uint32_t)iz - nz;
        if (wave_all(max(max(dx, dy), dz) <= 2u)) {
            const uint32_t k = threadIdx.x & 63u;
            hv = lattice((int32_t)(nx + (k & 3u)), (int32_t)(ny + ((k >> 2) & 3u)), (int32_t)(nz + (k >> 4)), seed);
            *slot = make_uint4(nx, ny, nz, 1u);   // every lane stores the same value
            fits = true;
        }
    }
    float c000, c100, c010, c110, c001, c101, c011, c111;
    if (fits) {
        const int a = (int)((dx + 4u * dy + 16u * dz) << 2);
        c000 = lane_value(a, hv);      c100 = lane_value(a + 4, hv);
        c010 = lane_value(a + 16, hv); c110 = lane_value(a + 20, hv);
        c001 = lane_value(a + 64, hv); c101 = lane_value(a + 68, hv);
        c011 = lane_value(a + 80, hv); c111 = lane_value(a + 84, hv);
    } else {
        c000 = lattice(ix, iy, iz, seed);         c100 = lattice(ix + 1, iy, iz, seed);
        c010 = lattice(ix, iy + 1, iz, seed);     c110 = lattice(ix + 1, iy + 1, iz, seed);
        c001 = lattice(ix, iy, iz + 1, seed);     c101 = lattice(ix + 1, iy, iz + 1, seed);
        c011 = lattice(ix, iy + 1, iz + 1, seed); c111 = lattice(ix + 1, iy + 1, iz + 1, seed);
    }
    float x00 = lerpf_(c000, c100, wx), x10 = lerpf_(c010, c110, wx);
    float x01 = lerpf_(c001, c101, wx), x11 = lerpf_(c011, c111, wx);
    float y0 = lerpf_(x00, x10, wy), y1 = lerpf_(x01, x11, wy);
    return lerpf_(y0, y1, wz);
}

// noise3 of a whole wave with kept windows for the first kKeptOctaves
// octaves and the lane window for the rest (the same value as noise3).
MP_DEV float noise3_kept(const DevParams& prm, float x, float y, float z, NoiseWindows& w)
{
    float acc = 0.0f, amp = 1.0f, norm = 0.0f, freq = prm.disp_frequency;
#pragma unroll
    for (int k = 0; k < kKeptOctaves; ++k) {
        if (k < prm.disp_octaves) {
            acc = mad(amp, vnoise_kept(x * freq, y * freq, z * freq, prm.disp_seed + (uint32_t)k, w.base + k,
                                       w.hv[k]), acc);
            norm = norm + amp;
            amp = amp * 0.5f;
            freq = freq * 2.0f;
        }
    }
    for (int k = kKeptOctaves; k < prm.disp_octaves; ++k) {
        acc = mad(amp, vnoise(x * freq, y * freq, z * freq, prm.disp_seed + (uint32_t)k, kNoiseLaneWindow), acc);
        norm = norm + amp;
        amp = amp * 0.5f;
        freq = freq * 2.0f;
    }
    return acc / norm;
}

// full_wave: every lane of the wave runs this call (the lane window needs
// all 64 lanes); otherwise the window scheme falls back to the scalar one.
MP_DEV float noise3(const DevParams& prm, float x, float y, float z, bool full_wave = false)
{
    float acc = 0.0f, amp = 1.0f, norm = 0.0f, freq = prm.disp_frequency;
    const int scheme = prm.noise_wave_cells == kNoiseLaneWindow && !full_wave ? kNoiseWaveCell
                                                                              : prm.noise_wave_cells;
    for (int k = 0; k < prm.disp_octaves; ++k) {
        acc = mad(amp, vnoise(x * freq, y * freq, z * freq, prm.disp_seed + (uint32_t)k, scheme), acc);
        norm = norm + amp;
        amp = amp * 0.5f;
        freq = freq * 2.0f;
    }
    return acc / norm;
}

// Column sums of the displacement normal at u, i ascending:
// col[j] = sum_i a_i(u) P_ij, col[4 + j] = sum_i a_i'(u) P_ij (xyz; w = 0).
// Every grid point of one u column shares them.
MP_DEV float4 normal_column(const float4* __restrict__ cp, float u, int k, float w[4])
{
    const float su = 1.0f - u;
    if (k < 4) bern_u(u, su, w);
    else       dbern(u, su, w);
    const int j = k & 3;
    float x = 0.0f, y = 0.0f, z = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float4 P = cp[4 * i + j];
        x = mad(w[i], P.x, x); y = mad(w[i], P.y, y); z = mad(w[i], P.z, z);
    }
    return make_float4(x, y, z, 0.0f);
}

// 1 / sqrtf(s), both correctly rounded.  Where every active lane's s lies
// in [2^-80, 2^80] this is the IEEE expansions' own steps for that range:
// v_sqrt_f32 with the neighbour test of the correctly rounded square root
// (its scaling of operands below 2^-96 and its fix-up of zeros and infinity
// are identities there), then the division's reciprocal refinement and
// correction steps on the unscaled operands (v_div_scale and v_div_fixup
// leave a numerator of 1 and a denominator in [2^-40, 2^40] unchanged).
// Otherwise the wave takes '/' and sqrtf.
MP_DEV float inv_sqrt_ieee(float s)
{
    if (wave_all2(s >= 0x1p-80f, s <= 0x1p80f)) {
        float r = __builtin_amdgcn_sqrtf(s);
        const float rd = __int_as_float(__float_as_int(r) - 1);
        const float ru = __int_as_float(__float_as_int(r) + 1);
        const float ed = __builtin_fmaf(-rd, r, s);
        const float eu = __builtin_fmaf(-ru, r, s);
        r = ed <= 0.0f ? rd : r;
        r = eu > 0.0f ? ru : r;
        const float y0 = __builtin_amdgcn_rcpf(r);
        const float y = __builtin_fmaf(__builtin_fmaf(-r, y0, 1.0f), y0, y0);
        const float q = __builtin_fmaf(__builtin_fmaf(-r, y, 1.0f), y, y);
        return __builtin_fmaf(__builtin_fmaf(-r, q, 1.0f), y, q);
    }
    return 1.0f / sqrtf(s);
}

// Unit displacement direction from the column sums at v:
// dP/du = sum_j b_j(v) col[4 + j], dP/dv = sum_j b_j'(v) col[j] (j
// ascending), n = unit(cross(dP/dv, dP/du)), zero where it vanishes.
MP_DEV void normal_from_columns(const float4* __restrict__ col, float v, float n[3])
{
    const float sv = 1.0f - v;
    float b[4], db[4];
    bern_v(v, sv, b);
    dbern(v, sv, db);
    float pux = 0.f, puy = 0.f, puz = 0.f, pvx = 0.f, pvy = 0.f, pvz = 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float4 S = col[j], D = col[4 + j];
        pux = mad(b[j], D.x, pux); puy = mad(b[j], D.y, puy); puz = mad(b[j], D.z, puz);
        pvx = mad(db[j], S.x, pvx); pvy = mad(db[j], S.y, pvy); pvz = mad(db[j], S.z, pvz);
    }
    float nx = mad_diff(pvy, puz, pvz, puy);
    float ny = mad_diff(pvz, pux, pvx, puz);
    float nz_ = mad_diff(pvx, puy, pvy, pux);
    float len2 = dot3(nx, ny, nz_, nx, ny, nz_);
    n[0] = n[1] = n[2] = 0.0f;
    if (len2 > 0.0f) {
        float inv = 1.0f / sqrtf(len2);
        n[0] = nx * inv; n[1] = ny * inv; n[2] = nz_ * inv;
    }
}

// normal_from_columns with the column sums in LDS, component-major:
// s_cols[(c * 8 + k) * G1 + nu] is component c (x, y, z) of col[k] of
// column nu.  The lanes of a wave read one (c, k) at up to G1 <= 33
// consecutive words, so no two columns of a read share an LDS bank (a
// float4 per column and k, 128 bytes apart, put every column on the same
// four banks).  b, db: the Bernstein weights bern_v and dbern at the
// point's v.
template <int G1>
MP_DEV void normal_from_lds_columns(const float* __restrict__ s_cols, int nu, const float b[4],
                                    const float db[4], float n[3])
{
    const float* cx = s_cols + nu;
    float pux = 0.f, puy = 0.f, puz = 0.f, pvx = 0.f, pvy = 0.f, pvz = 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float Sx = cx[(0 * 8 + j) * G1], Sy = cx[(1 * 8 + j) * G1], Sz = cx[(2 * 8 + j) * G1];
        const float Dx = cx[(0 * 8 + 4 + j) * G1], Dy = cx[(1 * 8 + 4 + j) * G1], Dz = cx[(2 * 8 + 4 + j) * G1];
        pux = mad(b[j], Dx, pux); puy = mad(b[j], Dy, puy); puz = mad(b[j], Dz, puz);
        pvx = mad(db[j], Sx, pvx); pvy = mad(db[j], Sy, pvy); pvz = mad(db[j], Sz, pvz);
    }
    float nx = mad_diff(pvy, puz, pvz, puy);
    float ny = mad_diff(pvz, pux, pvx, puz);
    float nz_ = mad_diff(pvx, puy, pvy, pux);
    float len2 = dot3(nx, ny, nz_, nx, ny, nz_);
    n[0] = n[1] = n[2] = 0.0f;
    if (len2 > 0.0f) {
        float inv = inv_sqrt_ieee(len2);
        n[0] = nx * inv; n[1] = ny * inv; n[2] = nz_ * inv;
    }
}

// Displaced point of a Bezier patch with the normal's column sums in LDS
// (normal_from_lds_columns) and the point's Bernstein weights given: a =
// bern_u at u, b = bern_v and db = dbern at v.
template <int G1>
MP_DEV void displaced_point_lds_cols(const float4* __restrict__ cp, const float* __restrict__ s_cols, int nu,
                                     const float a[4], const float b[4], const float db[4],
                                     const DevParams& prm, float amp, float out[4], NoiseWindows& kept)
{
    eval_bezier_w(cp, a, b, out);
    float h[3];
    normal_from_lds_columns<G1>(s_cols, nu, b, db, h);
    float d = amp * noise3_kept(prm, out[0], out[1], out[2], kept);
    out[0] = mad(d, h[0], out[0]);
    out[1] = mad(d, h[1], out[1]);
    out[2] = mad(d, h[2], out[2]);
}

// Unit displacement direction at (u,v) (Bezier patches).
MP_DEV void bezier_unit_normal(const float4* __restrict__ cp, float u, float v, float n[3])
{
    float4 col[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float w[4];
        col[k] = normal_column(cp, u, k, w);
    }
    normal_from_columns(col, v, n);
}

// Displaced object-space point P + amp*noise(P)*unit(cross(dP/dv, dP/du)).
MP_DEV void displaced_point(const float4* __restrict__ cp, int type, float u, float v,
                            const DevParams& prm, float amp, float out[4])
{
    eval_patch(cp, type, u, v, out);
    if (!prm.displacement || type != MP_TYPE_BEZIER) return;
    float h[3];
    bezier_unit_normal(cp, u, v, h);
    float d = amp * noise3(prm, out[0], out[1], out[2]);
    out[0] = mad(d, h[0], out[0]);
    out[1] = mad(d, h[1], out[1]);
    out[2] = mad(d, h[2], out[2]);
}

// ---- bound() (kernels/bound_n_split.h:43-100) with RES = 3 ----
// One of bound()'s 3x3 samples: eye-space point and screen position
// (bound_n_split.h:52-68).
// V: DevView or any table with its proj, screen, f, near_, far_, eps, width
// and height fields.
template <int NCP, typename V>
MP_DEV void bound_sample(const float4 (&C)[NCP], int iu, int iv, float umin, float vmin, float umax,
                         float vmax, const float* __restrict__ mv, const V& P, float e[3],
                         float& sx, float& sy)
{
    const float tu = (float)iu * 0.5f, tv = (float)iv * 0.5f;
    const float u = umin + (umax - umin) * tu;
    const float v = vmin + (vmax - vmin) * tv;
    float p[4], ee[4], q[4];
    if (NCP == 20) eval_gregory(C, u, v, p);
    else           eval_bezier(C, u, v, p);
    mat4_mul(mv, p, ee);
    mat4_mul(P.proj, ee, q);
    const float rx = q[0] / q[3], ry = q[1] / q[3];
    sx = mad(P.screen[2], ry, P.screen[0] * rx);
    sy = mad(P.screen[3], ry, P.screen[1] * rx);
    e[0] = ee[0]; e[1] = ee[1]; e[2] = ee[2];
}

// bound()'s decision from its nine samples, sample k = iu * 3 + iv
// (bound_n_split.h:10-31, 70-100), in three parts that bound_decide runs in
// order and the forest split runs on separate lanes: the box in sample order
// with the frustum cull and the eye split (bound_box_test: MP_CULL, a split,
// or -1 to go on), the polyline lengths of one row i (bound_row: h along
// samples (0, i), (1, i), (2, i), v along (i, 0), (i, 1), (i, 2)), and the
// decision from the three rows' lengths (bound_rows_decide).
template <typename V>
MP_DEV int bound_box_test(const float (&ex)[9], const float (&ey)[9], const float (&ez)[9], float umin,
                          float vmin, float umax, float vmax, const V& P, const DevParams& prm)
{
    float bminx = INFINITY, bminy = INFINITY, bminz = INFINITY;
    float bmaxx = -INFINITY, bmaxy = -INFINITY, bmaxz = -INFINITY;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        bminx = cl_min(bminx, ex[k]); bminy = cl_min(bminy, ey[k]); bminz = cl_min(bminz, ez[k]);
        bmaxx = cl_max(bmaxx, ex[k]); bmaxy = cl_max(bmaxy, ey[k]); bmaxz = cl_max(bmaxz, ez[k]);
    }
    // outside_frustum (bound_n_split.h:10-31)
    {
        float n = -bmaxz;
        float f = -bminz;
        if (f < P.near_ || n > P.far_) return MP_CULL;
        n = cl_max(P.near_, n);
        float hw = (float)P.width * 0.5f, hh = (float)P.height * 0.5f;
        float sminx = mad(bminx / ((bminx < 0) ? n : f), P.f[0], hw);
        float sminy = mad(bminy / ((bminy < 0) ? n : f), P.f[1], hh);
        float smaxx = mad(bmaxx / ((bmaxx > 0) ? n : f), P.f[0], hw);
        float smaxy = mad(bmaxy / ((bmaxy > 0) ? n : f), P.f[1], hh);
        const float rib = prm.cull_ribbon;
        if (sminx > (float)(P.width - 1) + rib || smaxx < -rib ||
            sminy > (float)(P.height - 1) + rib || smaxy < -rib)
            return MP_CULL;
    }
    if (bminz < P.eps && bmaxz > P.near_)
        return ((umax - umin) < (vmax - vmin)) ? MP_VSPLIT : MP_HSPLIT;
    return -1;
}

// a0..a2: screen positions of samples (0, i), (1, i), (2, i); b0..b2: of
// samples (i, 0), (i, 1), (i, 2)
MP_DEV void bound_row(float ax0, float ay0, float ax1, float ay1, float ax2, float ay2, float bx0, float by0,
                      float bx1, float by1, float bx2, float by2, float& h, float& v)
{
    const float ax[3] = { ax0, ax1, ax2 }, ay[3] = { ay0, ay1, ay2 };
    const float bx[3] = { bx0, bx1, bx2 }, by[3] = { by0, by1, by2 };
    h = 0.0f;
    v = 0.0f;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        float dx = ax[j] - ax[j + 1];
        float dy = ay[j] - ay[j + 1];
        h = h + sqrtf(mad(dy, dy, dx * dx));
        dx = bx[j] - bx[j + 1];
        dy = by[j] - by[j + 1];
        v = v + sqrtf(mad(dy, dy, dx * dx));
    }
}

MP_DEV int bound_rows_decide(float h0, float v0, float h1, float v1, float h2, float v2, float split_limit)
{
    float hlen = 0.0f, vlen = 0.0f;
    hlen = cl_max(h0, hlen); vlen = cl_max(v0, vlen);
    hlen = cl_max(h1, hlen); vlen = cl_max(v1, vlen);
    hlen = cl_max(h2, hlen); vlen = cl_max(v2, vlen);
    if (hlen <= split_limit && vlen <= split_limit) return MP_DRAW;
    return (hlen > vlen) ? MP_VSPLIT : MP_HSPLIT;
}

template <typename V>
MP_DEV int bound_decide(const float (&ex)[9], const float (&ey)[9], const float (&ez)[9],
                        const float (&sx)[9], const float (&sy)[9], float umin, float vmin,
                        float umax, float vmax, const V& P, const DevParams& prm,
                        float split_limit)
{
    const int r = bound_box_test(ex, ey, ez, umin, vmin, umax, vmax, P, prm);
    if (r >= 0) return r;
    // px[iv][iu] = sx[iu * 3 + iv]
    float h[3], v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        bound_row(sx[i], sy[i], sx[3 + i], sy[3 + i], sx[6 + i], sy[6 + i], sx[i * 3], sy[i * 3], sx[i * 3 + 1],
                  sy[i * 3 + 1], sx[i * 3 + 2], sy[i * 3 + 2], h[i], v[i]);
    return bound_rows_decide(h[0], v[0], h[1], v[1], h[2], v[2], split_limit);
}

// The control points are loaded into registers once per range (NCP = 16
// Bezier, 20 Gregory): the nine evaluations then cost one memory round trip.
template <int NCP>
MP_DEV int bound3_regs(const float4 (&C)[NCP], float umin, float vmin, float umax, float vmax,
                       const float* __restrict__ mv, const DevView& P, const DevParams& prm,
                       float split_limit)
{
    float ex[9], ey[9], ez[9], sx[9], sy[9];
#pragma unroll
    for (int iu = 0; iu < 3; ++iu) {
#pragma unroll
        for (int iv = 0; iv < 3; ++iv) {
            float e[3];
            const int k = iu * 3 + iv;
            bound_sample<NCP>(C, iu, iv, umin, vmin, umax, vmax, mv, P, e, sx[k], sy[k]);
            ex[k] = e[0]; ey[k] = e[1]; ez[k] = e[2];
        }
    }
    return bound_decide(ex, ey, ez, sx, sy, umin, vmin, umax, vmax, P, prm, split_limit);
}

MP_DEV int bound3(const float4* __restrict__ cp, int type, int depth, float umin, float vmin,
                  float umax, float vmax, const float* __restrict__ mv, const DevView& P,
                  const DevParams& prm, float split_limit)
{
    if (depth >= prm.max_split_depth - 1) return MP_CULL;
    if (type == MP_TYPE_GREGORY) {
        float4 C[20];
#pragma unroll
        for (int k = 0; k < 20; ++k) C[k] = ldg4(cp + k);
        return bound3_regs<20>(C, umin, vmin, umax, vmax, mv, P, prm, split_limit);
    }
    float4 C[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) C[k] = ldg4(cp + k);
    return bound3_regs<16>(C, umin, vmin, umax, vmax, mv, P, prm, split_limit);
}

// Children of a split range (bound_n_split_breadthfirst.cl:99-113)
MP_DEV void split_children(const RangeRec& r, int flags, RangeRec& c0, RangeRec& c1)
{
    float cx = (r.umin + r.umax) * 0.5f;
    float cy = (r.vmin + r.vmax) * 0.5f;
    c0 = r; c1 = r;
    uint32_t od = r.obj_depth + (1u << 24);
    c0.obj_depth = od; c1.obj_depth = od;
    if (flags & 4) { c0.umax = cx; c1.umin = cx; }
    else           { c0.vmax = cy; c1.vmin = cy; }
}

MP_DEV uint64_t canonical_key(const RangeRec& r)
{
    // umin, vmin are dyadic with at most 24 halvings per axis: exact below 2^24
    uint64_t u = (uint64_t)(uint32_t)(r.umin * 16777216.0f);
    uint64_t v = (uint64_t)(uint32_t)(r.vmin * 16777216.0f);
    return (u << 24) | v;
}

// monotone uint32 image of a float (total order consistent with <)
MP_DEV uint32_t mono_bits(float d)
{
    uint32_t b = __float_as_uint(d);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
MP_DEV float unmono_bits(uint32_t m)
{
    uint32_t b = (m & 0x80000000u) ? (m & 0x7fffffffu) : ~m;
    return __uint_as_float(b);
}

// inside_triangle (kernels/reyes.cl:240-265), wrapping int32 arithmetic.
MP_DEV bool inside_tri(int32_t x0, int32_t y0, int32_t x1, int32_t y1, int32_t x2, int32_t y2,
                       float d0, float d1, float d2, int32_t tx, int32_t ty, float& depth)
{
    const int32_t Px[3] = { x0, x1, x2 }, Py[3] = { y0, y1, y2 };
    int32_t Dx[3], Dy[3], O[3], V[3];
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3;
        Dx[i] = wsub(Py[j], Py[i]);
        Dy[i] = wsub(Px[i], Px[j]);
        O[i] = wadd(wmul(Dx[i], Px[i]), wmul(Dy[i], Py[i]));
        int32_t C = (Dx[i] > 0 || (Dx[i] == 0 && Dy[i] > 0)) ? -1 : 0;
        V[i] = wsub(wadd(wmul(Dx[i], tx), wmul(Dy[i], ty)), O[i]);
        ok = ok && (V[i] > C);
    }
    if (ok) {
        float w[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int j = (k + 1) % 3;
            int32_t den = wsub(wadd(wmul(Dx[j], Px[k]), wmul(Dy[j], Py[k])), O[j]);
            w[k] = (float)V[j] / (float)den;
        }
        float d = mad(d2, w[2], mad(d1, w[1], d0 * w[0]));
        if (d < depth) depth = d;
    }
    return ok;
}

// inside_tri in the exact-integer regime (every vertex and sample within
// +-2^18 sub-pixels, every edge shorter than 2^11): each product has 24-bit
// operands and a result below 2^30, so v_mul_i32_i24 and plain adds give the
// wrapping arithmetic's values, and the three barycentric denominators are
// one integer (twice the signed area: edge j at its opposite vertex).  Set
// up once per triangle, tested per pixel.
struct SmallTri {
    int32_t dx[3], dy[3], o[3], c[3];
    float den, rden;     // denominator and its Newton-refined reciprocal
};

MP_DEV void small_tri_setup(SmallTri& t, int32_t x0, int32_t y0, int32_t x1, int32_t y1, int32_t x2, int32_t y2)
{
    const int32_t Px[3] = { x0, x1, x2 }, Py[3] = { y0, y1, y2 };
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3;
        t.dx[i] = Py[j] - Py[i];
        t.dy[i] = Px[i] - Px[j];
        t.o[i] = __mul24(t.dx[i], Px[i]) + __mul24(t.dy[i], Py[i]);
        t.c[i] = (t.dx[i] > 0 || (t.dx[i] == 0 && t.dy[i] > 0)) ? -1 : 0;
    }
    t.den = (float)(__mul24(t.dx[1], Px[0]) + __mul24(t.dy[1], Py[0]) - t.o[1]);
    const float r0 = __builtin_amdgcn_rcpf(t.den);
    t.rden = __builtin_fmaf(__builtin_fmaf(-t.den, r0, 1.0f), r0, r0);
}

// a / den for integers 0 <= a <= den < 2^24: the correction steps of the
// IEEE f32 division sequence (v_div_scale / v_div_fmas / v_div_fixup leave
// operands in this range unscaled and the quotient unchanged), with the
// reciprocal shared by the three weights.
MP_DEV float small_div(float a, const SmallTri& t)
{
    float q = a * t.rden;
    q = __builtin_fmaf(__builtin_fmaf(-t.den, q, a), t.rden, q);
    return __builtin_fmaf(__builtin_fmaf(-t.den, q, a), t.rden, q);
}

MP_DEV bool small_tri_test(const SmallTri& t, int32_t tx, int32_t ty, float d0, float d1, float d2, float& depth)
{
    int32_t V[3];
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        V[i] = __mul24(t.dx[i], tx) + __mul24(t.dy[i], ty) - t.o[i];
        ok = ok && (V[i] > t.c[i]);
    }
    if (ok) {
        const float w0 = small_div((float)V[1], t), w1 = small_div((float)V[2], t), w2 = small_div((float)V[0], t);
        const float d = mad(d2, w2, mad(d1, w1, d0 * w0));
        if (d < depth) depth = d;
    }
    return ok;
}

MP_DEV int fb_index(int x, int y, int grid_w)
{
    return (((x >> 3) + grid_w * (y >> 3)) << 6) + (x & 7) + ((y & 7) << 3);
}

}  // namespace mp
