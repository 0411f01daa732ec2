"""GPU parity on edge cases against the oracle: an empty frame, a patch
through the camera plane (near-plane splits, culls behind the eye),
degenerate patches (a point, a line), ragged window sizes (partial 8x8
tiles), mixed Bezier / Gregory objects with a patch set drawn twice, the
split depth limit, and one renderer reused across frames of different patch
counts."""
import numpy as np
import pytest

from micropolis_amd import Renderer, scenes
from micropolis_amd.renderer import Projection
from gpu_util import assert_frame_equal, render_plain

pytestmark = pytest.mark.gpu

LOCAL = dict(bound_n_split_method="LOCAL", reyes_patch_size=16, bound_n_split_limit=16.0)


def _flat_patch(x0, x1, z0, z1, y, z_along_v=False):
    """Bicubic Bezier control net of the planar rectangle [x0,x1] x [z0,z1] at
    height y; z varies along u (control index i) or, with z_along_v, along v."""
    P = np.zeros((16, 3), np.float32)
    for i in range(4):
        for j in range(4):
            a, b = (j, i) if not z_along_v else (i, j)
            P[4 * i + j] = (x0 + (x1 - x0) * a / 3.0, y, z0 + (z1 - z0) * b / 3.0)
    return P


def _render_both(oracle, w, what, **overrides):
    print("case", what, flush=True)
    r, got = render_plain(w, write_rgba32=1, **overrides)
    ref = oracle.render(scenes.Workload(w.name, dict(w.options, **overrides), w.views))[0]
    assert_frame_equal(got, ref, what)
    return r, got, ref


def test_empty_frame(oracle):
    w = scenes.Workload("empty", dict(LOCAL), [scenes.View(Projection(60.0, 0.01, (64, 48)), [])])
    r, got, ref = _render_both(oracle, w, "empty")
    assert got["stats"]["diced_ranges"] == 0
    assert (got["ids"] == 0xFFFFFFFF).all() and (got["depth"] == 1.0).all()
    r.close()


@pytest.mark.parametrize("method", ["LOCAL", "BALANCED", "BOUNDED"])
def test_plane_through_the_camera(oracle, method):
    """A ground plane from behind the eye to far ahead, the eye plane crossed
    along v: ranges straddling it split by the near-plane rule (which splits
    the shorter parametric side, bound_n_split.h:76-77), ranges behind the eye
    are culled, the rest drawn."""
    mv = np.eye(4, dtype=np.float32)
    # x runs from +8 to -8 so that the transposed net keeps the winding the
    # sampler's edge test accepts (reyes.cl:240-265 rejects the other one)
    o = scenes._obj(1, _flat_patch(8.0, -8.0, 4.0, -40.0, -0.5, z_along_v=True)[None], scenes.BEZIER, mv,
                    (0.4, 0.8, 0.3, 1.0))
    w = scenes.Workload("plane", dict(LOCAL, backface_culling=0),
                        [scenes.View(Projection(60.0, 0.1, (320, 240)), [o])])
    r, got, ref = _render_both(oracle, w, "plane/" + method, bound_n_split_method=method)
    assert got["stats"]["diced_ranges"] > 0 and (got["ids"] != 0xFFFFFFFF).sum() > 10000
    r.close()


def test_plane_crossing_along_u_reaches_the_depth_limit(oracle):
    """The same plane crossed along u: the near-plane rule keeps splitting v,
    so every range down to max_split_depth - 1 is bounded (2^23 - 1) and all
    are culled there; nothing is drawn."""
    mv = np.eye(4, dtype=np.float32)
    o = scenes._obj(1, _flat_patch(-8.0, 8.0, 4.0, -40.0, -0.5)[None], scenes.BEZIER, mv, (0.4, 0.8, 0.3, 1.0))
    w = scenes.Workload("plane_u", dict(LOCAL, backface_culling=0),
                        [scenes.View(Projection(60.0, 0.1, (320, 240)), [o])])
    r, got, ref = _render_both(oracle, w, "plane along u")
    assert ref["n_bounds"] == (1 << 23) - 1 and len(ref["ranges"]) == 0
    assert got["stats"]["bounded_patches"] == ref["n_bounds"]
    r.close()


def test_degenerate_patches(oracle):
    """A patch collapsed to a point, one collapsed to a line and a normal one."""
    point = np.tile(np.array([0.1, 0.05, -3.0], np.float32), (16, 1))
    line = np.zeros((16, 3), np.float32)
    for k in range(16):
        line[k] = (-0.5 + k / 15.0, -0.2, -3.0)
    ok = _flat_patch(-0.6, 0.6, -2.5, -3.5, -0.3)
    mv = np.eye(4, dtype=np.float32)
    o = scenes._obj(1, np.stack([point, line, ok]), scenes.BEZIER, mv, (1.0, 0.5, 0.2, 1.0))
    w = scenes.Workload("degenerate", dict(LOCAL, backface_culling=0),
                        [scenes.View(Projection(60.0, 0.01, (200, 120)), [o])])
    r, got, ref = _render_both(oracle, w, "degenerate")
    r.close()


@pytest.mark.parametrize("size", [(37, 23), (65, 9), (8, 8), (1, 1)])
def test_ragged_windows(oracle, size):
    w = scenes.c2_torus(width=size[0], height=size[1])
    r, got, ref = _render_both(oracle, w, "torus %dx%d" % size, displacement=0)
    assert got["ids"].shape == (size[1], size[0])
    r.close()


@pytest.mark.parametrize("n", [16, 8])
def test_mixed_objects_and_instancing(oracle, n):
    """Bezier torus drawn twice (same patches, two matrices), Gregory planes;
    at N = 8 the raster dices two consecutive ranges together only when they
    share a Bezier object, checked under every range schedule."""
    t = scenes.c2_torus(width=480, height=270)
    g = scenes.c4_gregory(planes=2, quads=32, width=480, height=270)
    tor = t.views[0].objects[0]
    mv2 = tor.matrix.copy()
    mv2[:3, 3] += np.array([0.4, -0.2, -0.6], np.float32)
    twin = scenes.SceneObject(tor.handle, tor.patches, tor.patch_type, mv2,
                              np.array([0.2, 0.4, 0.9, 1.0], np.float32))
    # patch sets are identified by handle: the Gregory planes get their own
    greg = [scenes.SceneObject(100 + k, o.patches, o.patch_type, o.matrix, o.color)
            for k, o in enumerate(g.views[0].objects)]
    objs = [tor, twin] + greg
    opts = dict(LOCAL, displacement=0, reyes_patch_size=n, bound_n_split_limit=float(n))
    w = scenes.Workload("mixed", opts, [scenes.View(t.views[0].projection, objs)])
    r, got, ref = _render_both(oracle, w, "mixed N=%d" % n)
    assert got["stats"]["diced_ranges"] == len(ref["ranges"])
    for chunk in (0, -1, 5, -16):
        t = {k: v for k, v in r.tuning().items() if k != "frame_ms"}
        r.set_tuning(**dict(t, raster_chunk=chunk))
        r.replay(2, use_graph=True)
        again = r.read_framebuffer()
        assert np.array_equal(again["ids"], got["ids"]), chunk
        assert np.array_equal(again["depth"], got["depth"]), chunk
        assert np.array_equal(again["rgba8"], got["rgba8"]), chunk
    r.close()


@pytest.mark.parametrize("depth", [2, 5])
def test_split_depth_limit(oracle, depth):
    """Ranges reaching max_split_depth - 1 are culled without evaluation."""
    w = scenes.c2_torus(width=640, height=360)
    r, got, ref = _render_both(oracle, w, "depth limit %d" % depth, max_split_depth=depth, displacement=0)
    assert (ref["ranges"]["depth"] < depth).all() if len(ref["ranges"]) else True
    r.close()


def test_atomic_rate_measurement():
    """The measurement kernel of the frame roofline runs and reports a rate."""
    w = scenes.c2_torus(width=64, height=64)
    r, got = render_plain(w, displacement=0)
    rate = r.measure_atomic_rate(1 << 20)
    assert 1e8 < rate < 1e14
    r.close()


def test_one_renderer_across_patch_counts(oracle):
    """Frames of more and of at most 4096 patches on one renderer, in both
    orders: the canonical order's count scan runs in its own launch or inside
    the scatter launch, and each frame must start from zero per-patch counts
    and bucket cursors whichever ran before."""
    rng = np.random.default_rng(11)
    opts = dict(window_size=(320, 240), bound_n_split_method="LOCAL", bound_n_split_limit=8.0,
                reyes_patch_size=8, backface_culling=1)
    frames = []
    for k, n in enumerate((6000, 1500, 5000, 3000)):
        radius = 2.0 * n ** (1.0 / 3.0)
        pats = scenes.random_patches(rng, n, radius)
        eye = (radius * 2.2, 0.3 * radius, radius * 0.5)
        o = scenes._obj(10 + k, pats, scenes.BEZIER, scenes.look_at(eye, (0, 0, 0)), (0.8, 0.6, 0.3, 1.0))
        frames.append(scenes.Workload("reuse%d" % k, opts, [scenes.View(Projection(60.0, 0.01, (320, 240)), [o])]))
    r = Renderer(**opts)
    for w in frames:
        print("case", w.name, flush=True)
        v = w.views[0]
        o = v.objects[0]
        r.load_patches(o.handle, o.patches, o.patch_type)
        r.prepare()
        r.draw_patches(o.handle, o.matrix, v.projection, o.color)
        r.finish()
        got = r.read_framebuffer()
        got["ranges"] = r.ranges()
        ref = oracle.render(w)[0]
        assert_frame_equal(got, ref, w.name)
        r.unload_patches(o.handle)
    r.close()


def test_draws_of_one_frame_share_the_projection():
    """A later draw with another fovy or near plane is refused (one
    projection per frame, as Scene::draw passes the active camera's)."""
    from micropolis_amd.renderer import MicropolisError
    w = scenes.c1_sphere()
    o = w.views[0].objects[0]
    r = Renderer(**dict(w.options, window_size=(64, 48)))
    r.load_patches(o.handle, o.patches, o.patch_type)
    r.prepare()
    r.draw_patches(o.handle, o.matrix, Projection(60.0, 0.01, (64, 48)), o.color)
    r.draw_patches(o.handle, o.matrix, Projection(60.0, 0.01, (64, 48)), o.color)
    with pytest.raises(MicropolisError, match="fovy"):
        r.draw_patches(o.handle, o.matrix, Projection(45.0, 0.01, (64, 48)), o.color)
    with pytest.raises(MicropolisError, match="fovy"):
        r.draw_patches(o.handle, o.matrix, Projection(60.0, 0.1, (64, 48)), o.color)
    r.finish()
    assert r.stats()["diced_ranges"] > 0
    r.close()
