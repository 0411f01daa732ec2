"""Independent pins of the oracle's shading, resolve, projection, display and
value-noise functions.  Each expected value below comes from FP64 mathematics
or a structural property, never from re-running the oracle's own formula:

  * or_shade's Lambert term (reyes.cl:172-191): planar parallelograms of known
    orientation give max(n . normalize(4,3,8), 0) with n the analytic plane
    normal on the cross(dv, du) side; back-lit quads clamp to 0; a quad whose
    corner sum is zero has an undefined view vector and yields NaN.
  * or_resolve's colour (reyes.cl:181-189): 0.015 * dc + s * dc, alpha
    dc.a + s * dc.a; empty pixels get the clear value (rgb^2.2, 1000) of
    Framebuffer.cpp:65-68, depth 1 and id 0xFFFFFFFF.
  * or_make_view (Projection.cpp:23-50 + glm::perspective): the published
    perspective entries, points on the frustum's top plane / near / far planes
    mapping to NDC 1 / -1 / 1, and Projection::f (fx == fy == H / (2 tan(fovy/2))).
  * or_to_srgb8 (tex_draw.frag:37): FP64 pow(c, 0.454545) * 255 + 0.5,
    truncated, away from byte boundaries.
  * or_noise: the seeded lattice value at integer points, a smoothstep fade
    (0 at 0, 1 at 1, flat ends), the trilinear mean at cell centres and the
    octave sum sum_k 0.5^k n_k(2^k x) / sum_k 0.5^k with seed + k per octave.
  * or_framebuffer_index / _size (reyes.cl:37-45, tex_draw.frag:25-31): a
    bijection of the padded window onto 8x8 tiles of 64 consecutive entries.
"""
import ctypes as C

import numpy as np
import pytest

N = 8
G1 = N + 1
EMPTY = 0xFFFFFFFFFFFFFFFF
LIGHT = np.array([4.0, 3.0, 8.0]) / np.sqrt(89.0)


# ---------------------------------------------------------------- shading

def _shade_grid(oracle, pos, W=256, H=192):
    """or_shade over one N=8 grid of eye-space points (81, 3) -> quad_s (64,)."""
    view = oracle.make_view(60.0, 0.01, W, H)
    prm = oracle.make_params(reyes_patch_size=N, backface_culling=0)
    p4 = np.ones((G1 * G1, 4), np.float32)
    p4[:, :3] = pos
    pxl = np.zeros((G1 * G1, 2), np.int32)
    bidx = np.zeros(4, np.int32)
    qs = np.zeros(N * N, np.float32)
    oracle.load().or_shade(p4.ctypes.data, pxl.ctypes.data, C.byref(view), C.byref(prm),
                           bidx.ctypes.data, qs.ctypes.data)
    return qs


def _parallelogram_grid(origin, e_u, e_v):
    """Grid point (nu, nv) at origin + nu * e_u + nv * e_v, index nu + nv * (N + 1)."""
    nu, nv = np.meshgrid(np.arange(G1), np.arange(G1))
    pts = origin[None, :] + nu.reshape(-1, 1) * e_u[None, :] + nv.reshape(-1, 1) * e_v[None, :]
    return pts.astype(np.float32)


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.diag(r))


def test_lambert_term_of_planar_quads_matches_the_plane_normal(oracle):
    rng = np.random.default_rng(21)
    lit = dark = 0
    for _ in range(60):
        R = _rotation(rng)
        e_u = R[:, 0] * rng.uniform(0.02, 0.2)
        e_v = R[:, 1] * rng.uniform(0.02, 0.2)
        origin = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-8, -3)])
        pos = _parallelogram_grid(origin, e_u, e_v)
        qs = _shade_grid(oracle, pos)
        # du ~ e_u, dv ~ e_v for a parallelogram; n = unit cross(dv, du)
        n = np.cross(e_v, e_u)
        n /= np.linalg.norm(n)
        want = max(float(n @ LIGHT), 0.0)
        assert np.all(np.isfinite(qs))
        assert np.allclose(qs, want, rtol=0, atol=2e-5), (want, qs[:4])
        lit += want > 0.05
        dark += want == 0.0
    assert lit > 10 and dark > 10   # both sides of the clamp were exercised


def test_lambert_term_of_axis_quads_is_exact_on_both_faces(oracle):
    # x along u, y along v: cross(dv, du) = (0, 0, -1) faces away from the light
    pos = _parallelogram_grid(np.array([0.0, 0.0, -5.0]), np.array([0.1, 0, 0]), np.array([0, 0.1, 0]))
    assert np.all(_shade_grid(oracle, pos) == 0.0)
    # swap u and v: normal (0, 0, 1), n . l = 8 / sqrt(89)
    pos = _parallelogram_grid(np.array([0.0, 0.0, -5.0]), np.array([0, 0.1, 0]), np.array([0.1, 0, 0]))
    assert np.allclose(_shade_grid(oracle, pos), 8.0 / np.sqrt(89.0), rtol=0, atol=1e-6)
    # normal along +x / +y: the light's x and y components
    # (cross(z, y) = -x, so u along z and v along y gives +x; u along x, v along z gives +y)
    pos = _parallelogram_grid(np.array([0.0, 0.0, -5.0]), np.array([0, 0, 0.1]), np.array([0, 0.1, 0]))
    assert np.allclose(_shade_grid(oracle, pos), 4.0 / np.sqrt(89.0), rtol=0, atol=1e-6)
    pos = _parallelogram_grid(np.array([0.0, 0.0, -5.0]), np.array([0.1, 0, 0]), np.array([0, 0, 0.1]))
    assert np.allclose(_shade_grid(oracle, pos), 3.0 / np.sqrt(89.0), rtol=0, atol=1e-6)


def test_quad_with_zero_corner_sum_is_nan(oracle):
    # the view vector -normalize(p0+p1+p2+p3) is 0/0: the zero-weighted specular
    # term pow(max(dot(n,h),0),60) * 0 is NaN, so the colour is NaN
    pos = _parallelogram_grid(np.array([0.0, 0.0, -5.0]), np.array([0, 0.1, 0]), np.array([0.1, 0, 0]))
    quad = np.array([[-1.0, -1.0, 1.0], [1.0, -1.0, -1.0], [-1.0, 1.0, -1.0], [1.0, 1.0, 1.0]], np.float32)
    for k, g in enumerate([0, 1, G1, G1 + 1]):
        pos[g] = quad[k]
    qs = _shade_grid(oracle, pos)
    assert np.isnan(qs[0])
    assert np.all(np.isfinite(qs[1:]))


# ---------------------------------------------------------------- resolve

def _mono(d):
    """Order-preserving map float -> uint32 (larger float -> larger integer)."""
    b = np.asarray(d, np.float32).view(np.uint32).astype(np.uint64)
    neg = (b & 0x80000000) != 0
    return np.where(neg, (~b) & 0xFFFFFFFF, b | 0x80000000).astype(np.uint64)


def test_monotone_depth_bits_preserve_order():
    rng = np.random.default_rng(3)
    d = np.concatenate([rng.normal(size=500).astype(np.float32), np.float32([0.0, 1.0, -1.0, 1e-30, -1e-30])])
    assert np.array_equal(np.argsort(d, kind="stable"), np.argsort(_mono(d), kind="stable"))


def test_resolve_colour_depth_id_and_clear(oracle):
    W, H = 24, 16
    clear = (0.1, 0.3, 0.7, 0.0)
    view = oracle.make_view(60.0, 0.01, W, H, clear)
    prm = oracle.make_params(reyes_patch_size=N)
    lib = oracle.load()
    rng = np.random.default_rng(8)
    n_ranges = 5
    quad = rng.uniform(0, 1, n_ranges * N * N).astype(np.float32)
    quad[7] = 0.0
    colors = rng.uniform(0, 1, (n_ranges, 4)).astype(np.float32)
    keys = np.full(lib.or_framebuffer_size(W, H), EMPTY, np.uint64)
    want_id = np.full((H, W), 0xFFFFFFFF, np.uint32)
    want_d = np.ones((H, W), np.float32)
    covered = rng.random((H, W)) < 0.6
    for y, x in zip(*np.nonzero(covered)):
        qid = int(rng.integers(0, n_ranges * N * N))
        d = np.float32(rng.uniform(-1, 1))
        keys[lib.or_framebuffer_index(int(x), int(y), W)] = (int(_mono(d)) << 32) | qid
        want_id[y, x] = qid
        want_d[y, x] = d
    out = dict(rgba8=np.zeros((H, W, 4), np.uint8), depth=np.zeros((H, W), np.float32),
               ids=np.zeros((H, W), np.uint32), rgba32=np.zeros((H, W, 4), np.float32))
    lib.or_resolve(keys.ctypes.data, quad.ctypes.data, colors.ctypes.data, C.byref(view), C.byref(prm),
                   out["rgba8"].ctypes.data, out["depth"].ctypes.data, out["ids"].ctypes.data,
                   out["rgba32"].ctypes.data)
    assert np.array_equal(out["ids"], want_id)
    assert np.array_equal(out["depth"].view(np.uint32), want_d.view(np.uint32))
    clear64 = np.array([clear[0] ** 2.2, clear[1] ** 2.2, clear[2] ** 2.2, 1000.0])
    for y in range(H):
        for x in range(W):
            c = out["rgba32"][y, x].astype(np.float64)
            if covered[y, x]:
                qid = int(want_id[y, x])
                s = float(quad[qid])
                dc = colors[qid // (N * N)].astype(np.float64)
                ref = np.array([0.015, 0.015, 0.015, 1.0]) * dc + s * dc
            else:
                ref = clear64
            assert np.allclose(c, ref, rtol=3e-7, atol=1e-9), (x, y, c, ref)
            assert out["rgba8"][y, x, 3] == 255


def test_clear_value_is_rgb_pow_2_2_and_1000(oracle):
    for clear in [(0.1, 0.1, 0.7, 0.0), (0.0, 1.0, 0.5, 0.3), (0.25, 0.9, 0.003, 1.0)]:
        v = oracle.make_view(45.0, 0.1, 64, 48, clear)
        cl = np.array(v.clear_linear, np.float64)
        assert np.allclose(cl[:3], np.array(clear[:3]) ** 2.2, rtol=3e-7, atol=0)
        assert cl[3] == 1000.0


# ---------------------------------------------------------------- projection

@pytest.mark.parametrize("fovy,near,W,H", [(60.0, 0.01, 1920, 1080), (45.0, 0.1, 1024, 768),
                                           (90.0, 1.0, 37, 23), (30.0, 0.05, 640, 480)])
def test_make_view_matches_glm_perspective_and_projection(oracle, fovy, near, W, H):
    v = oracle.make_view(fovy, near, W, H)
    P = np.array(v.proj, np.float64).reshape(4, 4).T      # column-major storage -> rows
    aspect = W / H
    t = np.tan(np.radians(fovy) / 2)
    far = 10000.0
    ref = np.zeros((4, 4))
    ref[0, 0] = 1 / (aspect * t)
    ref[1, 1] = 1 / t
    ref[2, 2] = -(far + near) / (far - near)
    ref[3, 2] = -1.0
    ref[2, 3] = -(2 * far * near) / (far - near)
    assert np.allclose(P, ref, rtol=1e-6, atol=0)
    # frustum planes: top edge at NDC y = +1, right edge at x = +1, near/far at z = -1/+1
    d = 7.0
    for pt, axis, want in [((0.0, d * t, -d), 1, 1.0), ((d * t * aspect, 0.0, -d), 0, 1.0),
                           ((0.0, 0.0, -near), 2, -1.0), ((0.0, 0.0, -far), 2, 1.0)]:
        c = P @ np.array(pt + (1.0,))
        assert abs(c[axis] / c[3] - want) < 2e-5
    # Projection::f: fx = fy = H / (2 tan(fovy/2)) (equal in exact arithmetic)
    fy = H / (2 * t)
    assert np.allclose([v.f[0], v.f[1]], [fy, fy], rtol=4e-7, atol=0)
    assert list(v.screen) == [W / 2.0, 0.0, 0.0, H / 2.0]
    assert v.near_ == np.float32(near) and v.far_ == 1000.0
    assert v.eps == np.float32(np.float64(np.float32(near)) * 0.1)


# ---------------------------------------------------------------- display byte

def test_display_byte_matches_fp64_gamma_away_from_boundaries(oracle):
    f = oracle.load().or_to_srgb8
    rng = np.random.default_rng(4)
    c = np.concatenate([rng.random(20000), 10.0 ** rng.uniform(-8, 0, 5000)]).astype(np.float32)
    g = c.astype(np.float64) ** 0.454545 * 255.0 + 0.5
    frac = g - np.floor(g)
    keep = (frac > 1e-4) & (frac < 1 - 1e-4)
    want = np.minimum(np.floor(g[keep]), 255).astype(np.int64)
    got = np.array([f(float(x)) for x in c[keep]], np.int64)
    assert keep.sum() > 20000
    assert np.array_equal(got, want)
    for x, b in [(0.0, 0), (-0.5, 0), (float("nan"), 0), (1.0, 255), (7.0, 255), (float("inf"), 255),
                 (1e-45, 0)]:
        assert f(x) == b, x


# ---------------------------------------------------------------- value noise

def _hash(x, y, z, seed):
    """The lattice hash the noise is defined with (32-bit multiply-xorshift)."""
    M = 0xFFFFFFFF
    h = (seed ^ 0x9E3779B9) & M
    h ^= (x * 0x8DA6B343) & M
    h ^= (y * 0xD8163841) & M
    h ^= (z * 0xCB1AB31F) & M
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M
    h ^= h >> 16
    return h


def _lattice(x, y, z, seed):
    return (_hash(x & 0xFFFFFFFF, y & 0xFFFFFFFF, z & 0xFFFFFFFF, seed) >> 8) / 8388608.0 - 1.0


def _noise(oracle, prm, p):
    return oracle.load().or_noise(C.byref(prm), float(p[0]), float(p[1]), float(p[2]))


def test_noise_equals_the_seeded_lattice_at_integer_points(oracle):
    prm = oracle.make_params(displacement=1, displacement_frequency=1.0, displacement_octaves=1,
                             displacement_seed=77)
    rng = np.random.default_rng(6)
    vals = []
    for p in rng.integers(-40, 40, (300, 3)):
        got = _noise(oracle, prm, p)
        want = _lattice(int(p[0]), int(p[1]), int(p[2]), 77)
        assert got == np.float32(want), (p, got, want)
        vals.append(got)
    vals = np.array(vals)
    assert -1.0 <= vals.min() and vals.max() < 1.0 and abs(vals.mean()) < 0.15


def test_noise_fade_is_smoothstep_with_flat_ends(oracle):
    prm = oracle.make_params(displacement=1, displacement_frequency=1.0, displacement_octaves=1,
                             displacement_seed=5)
    checked = 0
    for (iy, iz) in [(0, 0), (3, -2), (-7, 11), (5, 5), (1, -9)]:
        for ix in range(-3, 3):
            c0, c1 = _lattice(ix, iy, iz, 5), _lattice(ix + 1, iy, iz, 5)
            if abs(c1 - c0) < 0.2:
                continue
            for t in [0.0, 1e-3, 0.1, 0.25, 0.5, 0.75, 0.9, 1 - 1e-3]:
                n = _noise(oracle, prm, (ix + t, iy, iz))
                fade = (n - c0) / (c1 - c0)
                s = np.float64(np.float32(t))
                assert abs(fade - s * s * (3 - 2 * s)) < 3e-6 / abs(c1 - c0), (t, fade)
            # flat ends: the slope over the first / last 1e-3 is ~3e-3, not ~1 (a linear fade)
            lo = (_noise(oracle, prm, (ix + 1e-3, iy, iz)) - c0) / (c1 - c0) / 1e-3
            hi = (c1 - _noise(oracle, prm, (ix + 1 - 1e-3, iy, iz))) / (c1 - c0) / 1e-3
            assert abs(lo) < 0.05 and abs(hi) < 0.05
            checked += 1
    assert checked >= 5


def test_noise_at_cell_centres_is_the_corner_mean(oracle):
    prm = oracle.make_params(displacement=1, displacement_frequency=1.0, displacement_octaves=1,
                             displacement_seed=13)
    rng = np.random.default_rng(7)
    for i, j, k in rng.integers(-20, 20, (40, 3)):
        corners = [_lattice(int(i + a), int(j + b), int(k + c), 13) for a in (0, 1) for b in (0, 1) for c in (0, 1)]
        got = _noise(oracle, prm, (i + 0.5, j + 0.5, k + 0.5))
        assert abs(got - np.mean(corners)) < 1e-6


def test_noise_octaves_are_halved_normalised_and_reseeded(oracle):
    prm = oracle.make_params(displacement=1, displacement_frequency=2.0, displacement_octaves=4,
                             displacement_seed=100)
    rng = np.random.default_rng(9)
    for p in rng.integers(-30, 30, (60, 3)):
        x = p / 2.0              # octave k samples the lattice at p * 2^k: all integer points
        want = sum(0.5 ** k * _lattice(*(int(c) * 2 ** k for c in p), 100 + k) for k in range(4)) / 1.875
        got = _noise(oracle, prm, x)
        assert abs(got - want) < 1e-6, (p, got, want)


# ---------------------------------------------------------------- framebuffer layout

@pytest.mark.parametrize("W,H", [(37, 23), (64, 48), (1, 1), (65, 9)])
def test_framebuffer_index_is_a_tiled_bijection(oracle, W, H):
    lib = oracle.load()
    gw, gh = (W + 7) // 8, (H + 7) // 8
    size = lib.or_framebuffer_size(W, H)
    assert size == gw * gh * 64
    idx = np.array([[lib.or_framebuffer_index(x, y, W) for x in range(gw * 8)] for y in range(gh * 8)])
    assert sorted(idx.ravel().tolist()) == list(range(size))
    for ty in range(gh):
        for tx in range(gw):
            tile = idx[ty * 8:(ty + 1) * 8, tx * 8:(tx + 1) * 8]
            assert tile.min() == (tx + gw * ty) * 64
            assert np.array_equal(tile - tile.min(), np.arange(64).reshape(8, 8))
