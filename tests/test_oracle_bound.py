"""Pins or_bound (kernels/bound_n_split.h:10-100) with hand-built cases:
planar patches facing the camera have analytic screen lengths, frustum and
near-plane cases have known outcomes, and the depth cap culls."""
import ctypes as C

import numpy as np

W, H = 800, 600


def planar_patch(x0, x1, y0, y1, z):
    """Bicubic patch of the rectangle [x0,x1]x[y0,y1] at depth z (u along x)."""
    P = np.zeros((1, 16, 3), np.float32)
    for i in range(4):
        for j in range(4):
            P[0, 4 * i + j] = (x0 + (x1 - x0) * i / 3.0, y0 + (y1 - y0) * j / 3.0, z)
    return P


def bound(oracle, P, mv=None, rng=(0.0, 0.0, 1.0, 1.0), depth=0, limit=8.0, **kw):
    view = oracle.make_view(60.0, 0.01, W, H)
    prm = oracle.make_params(**kw)
    cps = oracle.as_cps(P)
    mvc = oracle.colmajor(np.eye(4, dtype=np.float32) if mv is None else mv)
    return oracle.load().or_bound(oracle.fptr(cps), 0, 0, depth, *rng, oracle.fptr(mvc),
                                  C.byref(view), C.byref(prm), limit), view


def screen_len(view, side, z):
    """Analytic screen length of a segment of length `side` at distance -z,
    parallel to the image plane: screen scale * proj scale * side / |z|."""
    return view.screen[0] * view.proj[0] * side / abs(z), view.screen[3] * view.proj[5] * side / abs(z)


def test_draw_threshold_matches_analytic_length(oracle):
    z = -5.0
    _, view = bound(oracle, planar_patch(-0.01, 0.01, -0.01, 0.01, z))
    lx, ly = screen_len(view, 1.0, z)
    # sides just below and just above 8 pixels
    for frac, expect in ((0.98, 1), (1.02, 6)):
        s = 8.0 * frac / lx
        f, _ = bound(oracle, planar_patch(-s / 2, s / 2, -0.001, 0.001, z))
        assert f == expect
    s = 8.0 * 1.02 / ly
    f, _ = bound(oracle, planar_patch(-0.001, 0.001, -s / 2, s / 2, z))
    assert f == 2               # vertical extent too long -> split v (HSPLIT)


def test_split_direction_follows_longer_screen_side(oracle):
    f, _ = bound(oracle, planar_patch(-1.0, 1.0, -0.2, 0.2, -5.0))
    assert f == 6               # hlen > vlen -> VSPLIT (split u)
    f, _ = bound(oracle, planar_patch(-0.2, 0.2, -1.0, 1.0, -5.0))
    assert f == 2


def test_frustum_culling(oracle):
    assert bound(oracle, planar_patch(40.0, 41.0, 0.0, 1.0, -5.0))[0] == 0   # far right
    assert bound(oracle, planar_patch(-41.0, -40.0, 0.0, 1.0, -5.0))[0] == 0  # far left
    assert bound(oracle, planar_patch(0.0, 1.0, 30.0, 31.0, -5.0))[0] == 0   # above
    assert bound(oracle, planar_patch(-1.0, 1.0, -1.0, 1.0, 5.0))[0] == 0    # behind
    assert bound(oracle, planar_patch(-1.0, 1.0, -1.0, 1.0, -2000.0))[0] == 0  # beyond far=1000
    # inside the cull ribbon of 32 pixels: not culled
    _, view = bound(oracle, planar_patch(0, 1, 0, 1, -5.0))
    lx, _ = screen_len(view, 1.0, -5.0)
    x_edge = (W / 2 + 10) / lx      # 10 px right of the right border
    assert bound(oracle, planar_patch(x_edge, x_edge + 0.001, 0.0, 0.001, -5.0))[0] != 0


def test_eye_split_for_geometry_crossing_the_camera_plane(oracle):
    # z runs from -1 (in front) to +1 (behind) along u+v, so both half ranges
    # cross the camera plane: bbox_min.z < eps, bbox_max.z > near
    P = np.zeros((1, 16, 3), np.float32)
    for i in range(4):
        for j in range(4):
            P[0, 4 * i + j] = (-0.5 + i / 3.0, -0.2 + 0.4 * j / 3.0, -1.0 + (i + j) / 3.0)
    # du < dv -> VSPLIT, du >= dv -> HSPLIT (bound_n_split.h:80-81)
    assert bound(oracle, P, rng=(0.0, 0.0, 0.5, 1.0))[0] == 6
    assert bound(oracle, P, rng=(0.0, 0.0, 1.0, 0.5))[0] == 2


def test_depth_cap_culls(oracle):
    P = planar_patch(-1.0, 1.0, -1.0, 1.0, -5.0)
    assert bound(oracle, P, depth=21, max_split_depth=23)[0] != 0
    assert bound(oracle, P, depth=22, max_split_depth=23)[0] == 0
    assert bound(oracle, P, depth=3, max_split_depth=4)[0] == 0
