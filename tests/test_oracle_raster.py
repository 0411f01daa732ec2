"""Pins dice / shade / sample of the oracle:
  * dice: closed-form projection of known points (FP64)
  * sample: coverage of hand-built micropolygon grids against an exact
    integer brute force (pixel-corner sampling, two triangles per quad, the
    top-left rule of reyes.cl:252), watertightness (every pixel corner
    strictly inside a proper grid's outline is covered exactly once),
    back-face rejection and depth ordering between overlapping grids."""
import ctypes as C

import numpy as np

W, H = 256, 192
N = 8
G1 = N + 1
EMPTY = 0xFFFFFFFFFFFFFFFF


def run_dice(oracle, P, mv, rng=(0.0, 0.0, 1.0, 1.0)):
    view = oracle.make_view(60.0, 0.01, W, H)
    prm = oracle.make_params(reyes_patch_size=N)
    r = np.zeros(1, oracle.RANGE_DTYPE)
    r["umin"], r["vmin"], r["umax"], r["vmax"] = rng
    pos = np.zeros((G1 * G1, 4), np.float32)
    pxl = np.zeros((G1 * G1, 2), np.int32)
    dep = np.zeros(G1 * G1, np.float32)
    oracle.load().or_dice(oracle.fptr(oracle.as_cps(P)), 0, r.ctypes.data, oracle.fptr(oracle.colmajor(mv)),
                          C.byref(view), C.byref(prm), pos.ctypes.data, pxl.ctypes.data, dep.ctypes.data)
    return view, pos, pxl, dep


def test_dice_matches_closed_form_projection(oracle):
    z = -5.0
    P = np.zeros((1, 16, 3), np.float32)
    for i in range(4):
        for j in range(4):
            P[0, 4 * i + j] = (-1 + 2 * i / 3.0, -0.75 + 1.5 * j / 3.0, z)
    view, pos, pxl, dep = run_dice(oracle, P, np.eye(4))
    proj = np.array(view.proj, np.float64).reshape(4, 4).T
    for nv in range(G1):
        for nu in range(G1):
            g = nu + nv * G1
            x = -1 + 2 * nu / N
            y = -0.75 + 1.5 * nv / N
            assert np.allclose(pos[g, :3], (x, y, z), atol=1e-6)
            clip = proj @ np.array([x, y, z, 1.0])
            sx = clip[0] / clip[3] * (W * 64) / 2 + (W * 64) / 2
            sy = clip[1] / clip[3] * (H * 64) / 2 + (H * 64) / 2
            assert abs(pxl[g, 0] - np.trunc(sx)) <= 1 and abs(pxl[g, 1] - np.trunc(sy)) <= 1
            assert abs(dep[g] - clip[2] / clip[3]) < 1e-6
    c = 4 + 4 * G1                              # patch centre (0, 0, z) -> screen centre
    assert tuple(pxl[c]) == (W * 32, H * 32)


def exact_inside(Px, Py, tx, ty):
    for i in range(3):
        j = (i + 1) % 3
        Dx = Py[j] - Py[i]
        Dy = Px[i] - Px[j]
        O = Dx * Px[i] + Dy * Py[i]
        Cb = -1 if (Dx > 0 or (Dx == 0 and Dy > 0)) else 0
        if not (Dx * tx + Dy * ty - O > Cb):
            return False
    return True


def brute_coverage(pxl):
    """pixel -> list of quad ids covering its corner, exact integers."""
    cov = {}
    for v in range(N):
        for u in range(N):
            g = [u + v * G1, u + 1 + v * G1, u + (v + 1) * G1, u + 1 + (v + 1) * G1]
            X = [int(pxl[k, 0]) for k in g]
            Y = [int(pxl[k, 1]) for k in g]
            for y in range(min(Y) // 64, max(Y) // 64 + 1):
                for x in range(min(X) // 64, max(X) // 64 + 1):
                    tx, ty = x * 64, y * 64
                    if (exact_inside([X[0], X[1], X[3]], [Y[0], Y[1], Y[3]], tx, ty) or
                            exact_inside([X[0], X[3], X[2]], [Y[0], Y[3], Y[2]], tx, ty)):
                        cov.setdefault((x, y), []).append(u + v * N)
    return cov


def raster(oracle, pxl, depth, zpos=-1.0, id_base=0, keys=None):
    view = oracle.make_view(60.0, 0.01, W, H)
    prm = oracle.make_params(reyes_patch_size=N)
    pos = np.zeros((G1 * G1, 4), np.float32)
    pos[:, 2] = zpos
    pos[:, 3] = 1.0
    pos[:, 0] = np.arange(G1 * G1) % G1
    pos[:, 1] = np.arange(G1 * G1) // G1
    bidx = np.zeros(4, np.int32)
    qs = np.zeros(N * N, np.float32)
    lib = oracle.load()
    lib.or_shade(pos.ctypes.data, np.ascontiguousarray(pxl).ctypes.data, C.byref(view), C.byref(prm),
                 bidx.ctypes.data, qs.ctypes.data)
    if keys is None:
        keys = np.full(lib.or_framebuffer_size(W, H), EMPTY, np.uint64)
    lib.or_sample(bidx.ctypes.data, np.ascontiguousarray(pxl).ctypes.data,
                  np.ascontiguousarray(depth, np.float32).ctypes.data, C.byref(view), C.byref(prm),
                  id_base, keys.ctypes.data)
    return keys, bidx


def decode(oracle, keys):
    lib = oracle.load()
    out = {}
    for y in range(H):
        for x in range(W):
            k = int(keys[lib.or_framebuffer_index(x, y, W)])
            if k != EMPTY:
                out[(x, y)] = k
    return out


def grid(seed, origin=(3000, 2500), step=(250, 230), jitter=60):
    rng = np.random.default_rng(seed)
    pxl = np.zeros((G1 * G1, 2), np.int32)
    for nv in range(G1):
        for nu in range(G1):
            # x grows with nu, y decreases with nv: cross(d1, d3) < 0 (front-facing)
            x = origin[0] + nu * step[0] + nv * 40 + (rng.integers(-jitter, jitter) if 0 < nu < N else 0)
            y = origin[1] + 2000 - nv * step[1] + nu * 30 + (rng.integers(-jitter, jitter) if 0 < nv < N else 0)
            pxl[nu + nv * G1] = (x | 1, y | 1)      # odd sub-pixel coords: no corner on the outline
    return pxl


def test_sample_coverage_matches_exact_brute_force(oracle):
    for seed in range(4):
        pxl = grid(seed)
        dep = np.full(G1 * G1, 0.5, np.float32)
        keys, bidx = raster(oracle, pxl, dep)
        got = decode(oracle, keys)
        cov = brute_coverage(pxl)
        assert all(len(q) == 1 for q in cov.values()), "top-left rule double coverage"
        assert {p: q[0] for p, q in cov.items()} == {p: k & 0xFFFFFFFF for p, k in got.items()}
        assert len(got) > 200


def _inside_polygon(poly, x, y):
    inside = False
    n = len(poly)
    for i in range(n):
        x0, y0 = poly[i]
        x1, y1 = poly[(i + 1) % n]
        if (y0 > y) != (y1 > y) and x < x0 + (y - y0) * (x1 - x0) / (y1 - y0):
            inside = not inside
    return inside


def test_sample_is_watertight(oracle):
    pxl = grid(11)
    dep = np.full(G1 * G1, 0.5, np.float32)
    keys, _ = raster(oracle, pxl, dep)
    got = decode(oracle, keys)
    outline = [tuple(pxl[nu]) for nu in range(G1)] + \
              [tuple(pxl[N + nv * G1]) for nv in range(1, G1)] + \
              [tuple(pxl[nu + N * G1]) for nu in range(N - 1, -1, -1)] + \
              [tuple(pxl[nv * G1]) for nv in range(N - 1, 0, -1)]
    inside = {(x, y) for y in range(H) for x in range(W) if _inside_polygon(outline, x * 64, y * 64)}
    assert set(got) == inside


def test_backfacing_grid_draws_nothing(oracle):
    pxl = grid(2)
    flipped = pxl.reshape(G1, G1, 2)[::-1].reshape(-1, 2).copy()   # reverse v: flips winding
    keys, bidx = raster(oracle, flipped, np.full(G1 * G1, 0.5, np.float32))
    assert decode(oracle, keys) == {}
    assert bidx[0] > bidx[2]                                        # empty block bound


def test_zero_eye_depth_culls_the_block(oracle):
    keys, bidx = raster(oracle, grid(3), np.full(G1 * G1, 0.5, np.float32), zpos=0.0)
    assert decode(oracle, keys) == {} and tuple(bidx) == (1, 1, -1, -1)


def test_nearer_grid_wins_regardless_of_id(oracle):
    pxl = grid(4)
    near = np.full(G1 * G1, 0.25, np.float32)
    far = np.full(G1 * G1, 0.75, np.float32)
    keys, _ = raster(oracle, pxl, far, id_base=0)
    keys, _ = raster(oracle, pxl, near, id_base=1000, keys=keys)
    got = decode(oracle, keys)
    assert got and all((k & 0xFFFFFFFF) >= 1000 for k in got.values())
    keys2, _ = raster(oracle, pxl, near, id_base=1000)
    keys2, _ = raster(oracle, pxl, far, id_base=0, keys=keys2)
    assert decode(oracle, keys2) == got
