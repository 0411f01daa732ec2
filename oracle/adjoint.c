/*
 * FP64 reference of the differentiable extension -- TEST INFRASTRUCTURE ONLY.
 *
 * Loss with visibility held fixed to a forward visible-id buffer:
 *   L = sum over covered pixels p, channels c in {r,g,b} of (C_c(q_p) - T_c(p))^2
 * where q_p is the micropolygon visible at p and its flat colour is
 *   C_c(q) = (0.015 + s(q)) * albedo_c,  s(q) = max(n(q).l, 0)
 * (kernels/reyes.cl:172-191 with the zero-weighted specular term dropped),
 * n(q) = normalize(cross(dv, du)) from the quad's four eye-space corners, and
 * each corner e = MV * (P + amp * noise(P) * unit(cross(dP/dv, dP/du))) with
 * P the bicubic Bezier point (the same displacement as the forward dice).
 * Silhouette terms (changes of visibility) are ignored by construction.
 *
 * or_adjoint_f64 evaluates L in double precision and its exact gradient with
 * respect to every control-point coordinate, the amplitude and the albedo,
 * recomputing each range's grid (no stored forward state).  The gradient is
 * pinned by central finite differences of this same function
 * (tests/test_adjoint_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct { double x, y, z; } d3;

static inline d3 v3(double x, double y, double z) { d3 r = { x, y, z }; return r; }
static inline d3 add(d3 a, d3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 sub(d3 a, d3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 mul(d3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 crs(d3 a, d3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

/* value noise in double with its gradient (same lattice as or_noise) */
static uint32_t nhash(int32_t x, int32_t y, int32_t z, uint32_t seed)
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
static double lat(int32_t x, int32_t y, int32_t z, uint32_t seed)
{
    return (double)(float)((float)(nhash(x, y, z, seed) >> 8) * (1.0f / 8388608.0f) - 1.0f);
}
static double fade(double t) { return t * t * (3.0 - 2.0 * t); }
static double dfade(double t) { return 6.0 * t * (1.0 - t); }
static double lrp(double a, double b, double t) { return a + (b - a) * t; }

static double vnoise_g(double x, double y, double z, uint32_t seed, d3* grad)
{
    double fx = floor(x), fy = floor(y), fz = floor(z);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    double tx = x - fx, ty = y - fy, tz = z - fz;
    double wx = fade(tx), wy = fade(ty), wz = fade(tz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lat(ix + a, iy + b, iz + d, seed);
    double x00 = lrp(c[0][0][0], c[1][0][0], wx), x10 = lrp(c[0][1][0], c[1][1][0], wx);
    double x01 = lrp(c[0][0][1], c[1][0][1], wx), x11 = lrp(c[0][1][1], c[1][1][1], wx);
    double y0 = lrp(x00, x10, wy), y1 = lrp(x01, x11, wy);
    double val = lrp(y0, y1, wz);
    if (grad) {
        double gx00 = c[1][0][0] - c[0][0][0], gx10 = c[1][1][0] - c[0][1][0];
        double gx01 = c[1][0][1] - c[0][0][1], gx11 = c[1][1][1] - c[0][1][1];
        double dwx = lrp(lrp(gx00, gx10, wy), lrp(gx01, gx11, wy), wz);
        double dwy = lrp(x10 - x00, x11 - x01, wz);
        double dwz = y1 - y0;
        *grad = v3(dwx * dfade(tx), dwy * dfade(ty), dwz * dfade(tz));
    }
    return val;
}

static double noise_g(const or_params* prm, d3 p, d3* grad)
{
    double acc = 0.0, amp = 1.0, norm = 0.0, freq = prm->disp_frequency;
    d3 g = v3(0, 0, 0);
    for (int k = 0; k < prm->disp_octaves; ++k) {
        d3 gk;
        acc += amp * vnoise_g(p.x * freq, p.y * freq, p.z * freq, prm->disp_seed + (uint32_t)k, &gk);
        g = add(g, mul(gk, amp * freq));
        norm += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    if (grad) *grad = mul(g, 1.0 / norm);
    return acc / norm;
}

/* per grid point forward state kept for the reverse pass */
typedef struct {
    double w[16], wu[16], wv[16];   /* Bernstein weights and u/v derivatives  */
    d3 P, Pu, Pv, N, nh;            /* surface point, partials, normal        */
    double nlen, nz, d;             /* |N|, noise value, displacement          */
    d3 gnz;                         /* noise gradient                          */
    d3 e;                           /* eye-space displaced point (xyz)         */
} gpt;

static void bern(double t, double* b, double* db)
{
    double s = 1.0 - t;
    b[0] = s * s * s; b[1] = 3 * s * s * t; b[2] = 3 * s * t * t; b[3] = t * t * t;
    db[0] = -3 * s * s; db[1] = 3 * s * (s - 2 * t); db[2] = 3 * t * (2 * s - t); db[3] = 3 * t * t;
}

static void grid_point(const double* cp /* 16*3 */, double u, double v, const double* mv,
                       const or_params* prm, double amp, gpt* g)
{
    double a[4], da[4], b[4], db[4];
    bern(u, a, da);
    bern(v, b, db);
    g->P = g->Pu = g->Pv = v3(0, 0, 0);
    double w = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            int k = 4 * i + j;
            d3 C = v3(cp[3 * k], cp[3 * k + 1], cp[3 * k + 2]);
            g->w[k] = a[i] * b[j];
            g->wu[k] = da[i] * b[j];
            g->wv[k] = a[i] * db[j];
            g->P = add(g->P, mul(C, g->w[k]));
            g->Pu = add(g->Pu, mul(C, g->wu[k]));
            g->Pv = add(g->Pv, mul(C, g->wv[k]));
            w += g->w[k];
        }
    g->N = crs(g->Pv, g->Pu);
    g->nlen = sqrt(dot(g->N, g->N));
    g->nh = g->nlen > 0 ? mul(g->N, 1.0 / g->nlen) : v3(0, 0, 0);
    g->nz = 0.0;
    g->d = 0.0;
    g->gnz = v3(0, 0, 0);
    d3 Pd = g->P;
    if (prm->displacement) {
        g->nz = noise_g(prm, g->P, &g->gnz);
        g->d = amp * g->nz;
        Pd = add(g->P, mul(g->nh, g->d));
    }
    /* e = MV * (Pd, w), column-major */
    g->e = v3(mv[0] * Pd.x + mv[4] * Pd.y + mv[8] * Pd.z + mv[12] * w,
              mv[1] * Pd.x + mv[5] * Pd.y + mv[9] * Pd.z + mv[13] * w,
              mv[2] * Pd.x + mv[6] * Pd.y + mv[10] * Pd.z + mv[14] * w);
}

/* quad colour term s = max(n.l, 0) and d s / d corners (eye space) */
static double quad_s(d3 p0, d3 p1, d3 p2, d3 p3, d3 L, d3* g /* 4 corners, may be NULL */)
{
    d3 du = mul(add(sub(p1, p0), sub(p3, p2)), 0.5);
    d3 dv = mul(add(sub(p2, p0), sub(p3, p1)), 0.5);
    d3 m = crs(dv, du);
    double ml = sqrt(dot(m, m));
    if (!(ml > 0)) {
        if (g) g[0] = g[1] = g[2] = g[3] = v3(0, 0, 0);
        return 0.0;
    }
    d3 n = mul(m, 1.0 / ml);
    double nl = dot(n, L);
    if (g) {
        if (nl > 0) {
            d3 gn = L;
            d3 gm = mul(sub(gn, mul(n, dot(n, gn))), 1.0 / ml);
            d3 gdv = crs(du, gm);
            d3 gdu = crs(gm, dv);
            g[0] = mul(add(gdu, gdv), -0.5);
            g[1] = mul(sub(gdu, gdv), 0.5);
            g[2] = mul(sub(gdv, gdu), 0.5);
            g[3] = mul(add(gdu, gdv), 0.5);
        } else {
            g[0] = g[1] = g[2] = g[3] = v3(0, 0, 0);
        }
    }
    return nl > 0 ? nl : 0.0;
}

double or_adjoint_f64(const float* cps_xyz, int n_patches, const or_range* ranges, long n_ranges,
                      const float mv[16], const or_params* prm, double amplitude,
                      const double albedo[3], const uint32_t* ids, const float* target_rgb,
                      int W, int H, double* grad_cps, double* grad_albedo, double* grad_amp,
                      double* color_out)
{
    const int N = prm->patch_size, G1 = N + 1, Q = N * N;
    double mvd[16];
    for (int k = 0; k < 16; ++k) mvd[k] = mv[k];
    const d3 L = mul(v3(4, 3, 8), 1.0 / sqrt(89.0));
    const double ac = 0.015;

    /* dL/ds per quad and the albedo terms, from the fixed visibility */
    double* gs = (double*)calloc((size_t)n_ranges * Q, sizeof(double));
    unsigned char* vis = (unsigned char*)calloc((size_t)n_ranges, 1);
    /* colours need s per visible quad: first pass computes s for ranges with
     * visible quads */
    for (long p = 0; p < (long)W * H; ++p) {
        uint32_t id = ids[p];
        if (id == 0xFFFFFFFFu) continue;
        long r = id / Q;
        if (r < n_ranges) vis[r] = 1;
    }
    double* sq = (double*)calloc((size_t)n_ranges * Q, sizeof(double));
    gpt* grid = (gpt*)malloc(sizeof(gpt) * (size_t)G1 * G1);
    double cp[48];
    for (long r = 0; r < n_ranges; ++r) {
        if (!vis[r]) continue;
        const or_range* R = &ranges[r];
        for (int k = 0; k < 48; ++k) cp[k] = cps_xyz[(long)R->pid * 48 + k];
        for (int nv = 0; nv <= N; ++nv)
            for (int nu = 0; nu <= N; ++nu) {
                double u = R->umin + ((double)R->umax - R->umin) * ((double)nu / N);
                double v = R->vmin + ((double)R->vmax - R->vmin) * ((double)nv / N);
                grid_point(cp, u, v, mvd, prm, amplitude, &grid[nu + nv * G1]);
            }
        for (int q = 0; q < Q; ++q) {
            int u = q % N, v = q / N, g0 = u + v * G1;
            sq[r * Q + q] = quad_s(grid[g0].e, grid[g0 + 1].e, grid[g0 + G1].e, grid[g0 + G1 + 1].e, L, NULL);
        }
    }
    double loss = 0.0;
    double ga[3] = { 0, 0, 0 };
    for (long p = 0; p < (long)W * H; ++p) {
        uint32_t id = ids[p];
        if (id == 0xFFFFFFFFu) continue;
        double s = sq[id];
        double acc = 0.0;
        for (int c = 0; c < 3; ++c) {
            double col = (ac + s) * albedo[c];
            double diff = col - (double)target_rgb[3 * p + c];
            if (color_out) color_out[3 * p + c] = col;
            loss += diff * diff;
            acc += 2.0 * diff * albedo[c];
            ga[c] += 2.0 * diff * (ac + s);
        }
        gs[id] += acc;
    }
    if (grad_albedo)
        for (int c = 0; c < 3; ++c) grad_albedo[c] += ga[c];

    if (grad_cps || grad_amp) {
        d3* ge = (d3*)malloc(sizeof(d3) * (size_t)G1 * G1);
        for (long r = 0; r < n_ranges; ++r) {
            if (!vis[r]) continue;
            const or_range* R = &ranges[r];
            for (int k = 0; k < 48; ++k) cp[k] = cps_xyz[(long)R->pid * 48 + k];
            for (int nv = 0; nv <= N; ++nv)
                for (int nu = 0; nu <= N; ++nu) {
                    double u = R->umin + ((double)R->umax - R->umin) * ((double)nu / N);
                    double v = R->vmin + ((double)R->vmax - R->vmin) * ((double)nv / N);
                    grid_point(cp, u, v, mvd, prm, amplitude, &grid[nu + nv * G1]);
                    ge[nu + nv * G1] = v3(0, 0, 0);
                }
            for (int q = 0; q < Q; ++q) {
                double g = gs[r * Q + q];
                if (g == 0.0) continue;
                int u = q % N, v = q / N, g0 = u + v * G1;
                d3 gc[4];
                quad_s(grid[g0].e, grid[g0 + 1].e, grid[g0 + G1].e, grid[g0 + G1 + 1].e, L, gc);
                ge[g0] = add(ge[g0], mul(gc[0], g));
                ge[g0 + 1] = add(ge[g0 + 1], mul(gc[1], g));
                ge[g0 + G1] = add(ge[g0 + G1], mul(gc[2], g));
                ge[g0 + G1 + 1] = add(ge[g0 + G1 + 1], mul(gc[3], g));
            }
            for (int i = 0; i < G1 * G1; ++i) {
                const gpt* P = &grid[i];
                d3 e = ge[i];
                if (e.x == 0 && e.y == 0 && e.z == 0) continue;
                /* g_Pd = MV3^T g_e */
                d3 gPd = v3(mvd[0] * e.x + mvd[1] * e.y + mvd[2] * e.z,
                            mvd[4] * e.x + mvd[5] * e.y + mvd[6] * e.z,
                            mvd[8] * e.x + mvd[9] * e.y + mvd[10] * e.z);
                d3 gP = gPd, gPu = v3(0, 0, 0), gPv = v3(0, 0, 0);
                if (prm->displacement) {
                    double nd = dot(P->nh, gPd);
                    if (grad_amp) *grad_amp += P->nz * nd;
                    gP = add(gP, mul(P->gnz, amplitude * nd));
                    if (P->nlen > 0) {
                        d3 gnh = mul(gPd, P->d);
                        d3 gN = mul(sub(gnh, mul(P->nh, dot(P->nh, gnh))), 1.0 / P->nlen);
                        gPv = crs(P->Pu, gN);
                        gPu = crs(gN, P->Pv);
                    }
                }
                if (grad_cps) {
                    double* o = grad_cps + (long)R->pid * 48;
                    for (int k = 0; k < 16; ++k) {
                        o[3 * k + 0] += P->w[k] * gP.x + P->wu[k] * gPu.x + P->wv[k] * gPv.x;
                        o[3 * k + 1] += P->w[k] * gP.y + P->wu[k] * gPu.y + P->wv[k] * gPv.y;
                        o[3 * k + 2] += P->w[k] * gP.z + P->wu[k] * gPu.z + P->wv[k] * gPv.z;
                    }
                }
            }
        }
        free(ge);
    }
    (void)n_patches;
    free(grid);
    free(gs);
    free(sq);
    free(vis);
    return loss;
}
