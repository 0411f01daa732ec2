"""CPU bound&split method (src/Reyes/BoundNSplitCLCPU.cpp, src/base/Patch.cpp,
src/base/Projection.cpp).

The oracle's restatement is pinned by the FP64 Bernstein closed form (with
the reference's host argument order: eval_patch(p, u, v) is called with u
from the range's second coordinate), analytic screen sizes of boxes through
Projection::bound, hand-built cull / drop / split-direction cases, the brute
force recursion of the predicate and the batch recurrence.  The library's
multi-threaded host split (mp_cpu_bound_n_split, no device needed) must then
match the oracle bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

from micropolis_amd import scenes

W, H = 800, 600


def planar_patch(x0, x1, y0, y1, z):
    """Bicubic patch of the rectangle [x0,x1]x[y0,y1] at depth z (index i along x)."""
    P = np.zeros((1, 16, 3), np.float32)
    for i in range(4):
        for j in range(4):
            P[0, 4 * i + j] = (x0 + (x1 - x0) * i / 3.0, y0 + (y1 - y0) * j / 3.0, z)
    return P


def bound_cpu(oracle, P, rng=(0.0, 0.0, 1.0, 1.0), depth=0, limit=8.0, mv=None, ptype=0, **kw):
    view = oracle.make_view(60.0, 0.01, W, H)
    prm = oracle.make_params(**kw)
    cps = oracle.as_cps(P)
    mvc = oracle.colmajor(np.eye(4, dtype=np.float32) if mv is None else mv)
    return oracle.load().or_bound_cpu(oracle.fptr(cps), ptype, 0, depth, *map(float, rng), oracle.fptr(mvc),
                                      C.byref(view), C.byref(prm), float(limit)), view


def eval_cpu(oracle, cps, ptype, u, v):
    out = np.zeros(3, np.float32)
    oracle.load().or_eval_patch_cpu(oracle.fptr(cps), ptype, 0, u, v, oracle.fptr(out))
    return out


def eval_f64(oracle, cps, ptype, u, v):
    c64 = np.ascontiguousarray(cps.astype(np.float64))
    out = np.zeros(4, np.float64)
    oracle.load().or_eval_patch_f64(c64.ctypes.data, ptype, 0, u, v, out.ctypes.data)
    return out[:3]


def test_host_bezier_eval_is_the_transposed_closed_form(oracle):
    rng = np.random.default_rng(3)
    cps = oracle.as_cps(rng.uniform(-1, 1, (1, 16, 3)))
    for u, v in rng.uniform(0, 1, (50, 2)).astype(np.float32):
        got = eval_cpu(oracle, cps, 0, u, v)
        # eval_patch(patch, t, s): t runs along P[i][0..3], s along i
        want = eval_f64(oracle, cps, 0, float(v), float(u))
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-6)
    for (u, v), k in (((0, 0), 0), ((1, 0), 3), ((0, 1), 12), ((1, 1), 15)):
        np.testing.assert_array_equal(eval_cpu(oracle, cps, 0, u, v), cps[0, k, :3])


def test_host_gregory_eval_matches_bezier_and_is_nan_at_corners(oracle):
    rng = np.random.default_rng(4)
    bez = rng.uniform(-1, 1, (1, 16, 3)).astype(np.float32)
    greg = bez[:, scenes.BEZ_TO_GREGORY, :]
    cb, cg = oracle.as_cps(bez), oracle.as_cps(greg)
    for u, v in rng.uniform(0.05, 0.95, (40, 2)).astype(np.float32):
        np.testing.assert_allclose(eval_cpu(oracle, cg, 1, u, v), eval_f64(oracle, cb, 0, float(v), float(u)),
                                   rtol=0, atol=5e-6)
    # no NaN guard on the interior terms (Patch.cpp:197-200): 0/0 at every corner
    for u, v in ((0, 0), (1, 0), (0, 1), (1, 1)):
        assert np.isnan(eval_cpu(oracle, cg, 1, u, v)).all()


def projection_bound(oracle, bmin, bmax, ribbon=32.0):
    view = oracle.make_view(60.0, 0.01, W, H)
    size = np.zeros(2, np.float32)
    lo, hi = np.asarray(bmin, np.float32), np.asarray(bmax, np.float32)
    cull = oracle.load().or_projection_bound(oracle.fptr(lo), oracle.fptr(hi), C.byref(view), ribbon,
                                             oracle.fptr(size))
    return cull, size, view


def test_projection_bound_sizes_and_culling(oracle):
    # a flat box at distance 5: size = extent * f / 5 on each axis
    cull, size, view = projection_bound(oracle, (-0.3, -0.1, -5.0), (0.2, 0.4, -5.0))
    assert cull == 0
    np.testing.assert_allclose(size, [0.5 * view.f[0] / 5.0, 0.5 * view.f[1] / 5.0], rtol=1e-5)
    # depth range: the near side (n) magnifies the parts away from the axis
    cull, size, view = projection_bound(oracle, (0.1, 0.1, -4.0), (0.3, 0.2, -2.0))
    np.testing.assert_allclose(size[0], 0.3 / 2.0 * view.f[0] - 0.1 / 4.0 * view.f[0], rtol=1e-5)
    # a box straddling the eye uses the near plane for n
    cull, size, view = projection_bound(oracle, (-0.01, -0.01, -1.0), (0.01, 0.01, 1.0))
    assert cull == 0
    np.testing.assert_allclose(size[0], 0.02 / view.near_ * view.f[0], rtol=1e-5)
    assert projection_bound(oracle, (-1, -1, 0.5), (1, 1, 1.0))[0] == 1          # behind the eye
    assert projection_bound(oracle, (-1, -1, -0.005), (1, 1, -0.001))[0] == 1    # closer than near
    # right of the screen: culled beyond the 32-pixel ribbon, kept inside it
    f = projection_bound(oracle, (0, 0, -5), (0, 0, -5))[2].f[0]
    x_in = (W / 2 + 20) * 5.0 / f
    x_out = (W / 2 + 40) * 5.0 / f
    assert projection_bound(oracle, (x_in, 0, -5), (x_in + 1e-3, 1e-3, -5))[0] == 0
    assert projection_bound(oracle, (x_out, 0, -5), (x_out + 1e-3, 1e-3, -5))[0] == 1


def test_cpu_bound_decisions(oracle):
    z = -5.0
    _, view = bound_cpu(oracle, planar_patch(-0.01, 0.01, -0.01, 0.01, z))
    px = view.f[0] / 5.0                      # pixels per unit at distance 5
    for frac, expect in ((0.98, 1), (1.02, 6)):
        s = 8.0 * frac / px
        assert bound_cpu(oracle, planar_patch(-s / 2, s / 2, -1e-3, 1e-3, z))[0] == expect
    s = 8.0 * 1.02 / (view.f[1] / 5.0)
    # long along j (the range's second coordinate): vsplit_range halves it
    assert bound_cpu(oracle, planar_patch(-1e-3, 1e-3, -s / 2, s / 2, z))[0] == 2
    big = planar_patch(-1.0, 1.0, -0.2, 0.2, z)
    assert bound_cpu(oracle, big)[0] == 6
    assert bound_cpu(oracle, big, depth=23, max_split_depth=23)[0] == 6
    assert bound_cpu(oracle, big, depth=24, max_split_depth=23)[0] == 8   # dropped
    assert bound_cpu(oracle, planar_patch(-1, 1, -1, 1, 5.0))[0] == 0    # behind
    assert bound_cpu(oracle, planar_patch(40, 41, 0, 1, z))[0] == 0      # off screen


def sphere(oracle):
    o = scenes.c1_sphere().views[0].objects[0]
    view = oracle.make_view(60.0, 0.01, 1024, 768)
    return o.patches, o.matrix, view


def tuples(a):
    return sorted((int(r["pid"]), int(r["depth"]), float(r["umin"]), float(r["vmin"]),
                   float(r["umax"]), float(r["vmax"])) for r in a)


def test_cpu_split_is_the_predicate_recursion(oracle):
    P, mv, view = sphere(oracle)
    prm = oracle.make_params(reyes_patches_per_pass=1000)
    cps = oracle.as_cps(P)
    ranges, passes, nb, _ = oracle.split("CPU", cps, 0, mv, view, prm, 8.0)
    mvc = oracle.colmajor(mv)
    lib = oracle.load()
    want, count = [], [0]

    def rec(r):
        count[0] += 1
        f = lib.or_bound_cpu(oracle.fptr(cps), 0, r[0], r[1], *r[2:], oracle.fptr(mvc), C.byref(view),
                             C.byref(prm), 8.0)
        if f == 1:
            want.append(r)
        elif f in (2, 6):
            pid, d, u0, v0, u1, v1 = r
            if f == 6:
                c = float(np.float32((np.float32(u0) + np.float32(u1)) * np.float32(0.5)))
                rec((pid, d + 1, u0, v0, c, v1)); rec((pid, d + 1, c, v0, u1, v1))
            else:
                c = float(np.float32((np.float32(v0) + np.float32(v1)) * np.float32(0.5)))
                rec((pid, d + 1, u0, v0, u1, c)); rec((pid, d + 1, u0, c, u1, v1))
    for pid in range(P.shape[0]):
        rec((pid, 0, 0.0, 0.0, 1.0, 1.0))
    assert tuples(ranges) == sorted(want)
    assert nb == count[0]
    # DFS from the back of the root list: the last patch is emitted first
    assert ranges["pid"][0] == P.shape[0] - 1
    # batches of reyes_patches_per_pass, the remainder last
    n = len(ranges)
    assert passes.tolist() == [1000] * (n // 1000) + ([n % 1000] if n % 1000 else [])


def test_cpu_split_batches_close_with_an_empty_pass_when_work_remains(oracle):
    # two patches: patch 1 visible (split into 2 ranges), patch 0 behind the eye.
    # patch 1 is processed first, so after its last DRAW the stack still holds
    # patch 0: with reyes_patches_per_pass 2 a final empty batch follows.
    view = oracle.make_view(60.0, 0.01, W, H)
    s = 12.0 / (view.f[0] / 5.0)
    P = np.concatenate([planar_patch(-1, 1, -1, 1, 5.0), planar_patch(-s / 2, s / 2, -1e-3, 1e-3, -5.0)])
    cps = oracle.as_cps(P)
    eye = np.eye(4, dtype=np.float32)
    r, passes, nb, _ = oracle.split("CPU", cps, 0, eye, view, oracle.make_params(reyes_patches_per_pass=2), 8.0)
    assert len(r) == 2 and passes.tolist() == [2, 0] and nb == 4
    r, passes, nb, _ = oracle.split("CPU", cps, 0, eye, view, oracle.make_params(reyes_patches_per_pass=3), 8.0)
    assert passes.tolist() == [2]
    # swapped order: the visible patch is processed last, nothing remains
    r, passes, nb, _ = oracle.split("CPU", cps[::-1].copy(), 0, eye, view,
                                    oracle.make_params(reyes_patches_per_pass=2), 8.0)
    assert passes.tolist() == [2]


def library_split(w, obj, ppp, threads):
    from micropolis_amd.renderer import cpu_bound_n_split
    os.environ["MP_CPU_THREADS"] = str(threads)
    try:
        opts = dict(w.options, reyes_patches_per_pass=ppp)
        opts.pop("bound_n_split_method", None)
        return cpu_bound_n_split(obj.patches, obj.patch_type, obj.matrix, w.views[0].projection,
                                 bound_n_split_method="CPU", **opts)
    finally:
        del os.environ["MP_CPU_THREADS"]


def oracle_split(oracle, w, obj, ppp):
    v = w.views[0]
    view = oracle.make_view(v.projection.fovy, v.projection.near, *v.projection.viewport)
    opts = dict(w.options, reyes_patches_per_pass=ppp)
    prm = oracle.make_params(opts)
    r, passes, nb, _ = oracle.split("CPU", oracle.as_cps(obj.patches), obj.patch_type, obj.matrix, view, prm,
                                    float(opts.get("bound_n_split_limit", 8.0)))
    return oracle.canonical_sort(r), passes, nb


@pytest.mark.parametrize("workload,ppp", [("c1", 1024), ("c1", 333), ("c2", 4096), ("c4", 5000)])
@pytest.mark.parametrize("threads", [1, 8])
def test_library_cpu_split_matches_oracle(oracle, workload, ppp, threads):
    w = {"c1": scenes.c1_sphere, "c2": scenes.c2_torus, "c4": scenes.c4_gregory}[workload]()
    obj = w.views[0].objects[0]
    got, passes, nb = library_split(w, obj, ppp, threads)
    want, wpasses, wnb = oracle_split(oracle, w, obj, ppp)
    assert len(got) == len(want) > 0
    for f in ("pid", "depth", "umin", "vmin", "umax", "vmax"):
        np.testing.assert_array_equal(got[f], want[f], err_msg=f)
    assert passes.tolist() == wpasses.tolist()
    assert nb == wnb
