"""GPU parity of the batched path and of the Gregory / multi-view workloads:
a batch of views renders exactly what one plain frame per view renders, and
each checked view matches the oracle."""
import numpy as np
import pytest

from micropolis_amd import Renderer, scenes
from gpu_util import assert_frame_equal, render_plain

pytestmark = pytest.mark.gpu


def render_batch(workload, **overrides):
    opts = dict(workload.options)
    opts.update(overrides)
    r = Renderer(**opts)
    r.batch_begin([v.projection for v in workload.views])
    for vi, v in enumerate(workload.views):
        for o in v.objects:
            if not r.are_patches_loaded(o.handle):
                r.load_patches(o.handle, o.patches, o.patch_type)
            r.batch_draw(vi, o.handle, o.matrix, o.color)
    r.batch_finish()
    return r


def test_c3_batch_equals_64_plain_frames():
    w = scenes.c3_batch()
    print("case c3 batch of %d views" % len(w.views), flush=True)
    rb = render_batch(w)
    batched = [rb.read_framebuffer(view=i) for i in range(len(w.views))]
    rb.close()
    rp = Renderer(**w.options)
    for vi, v in enumerate(w.views):
        o = v.objects[0]
        rp.load_patches(o.handle, o.patches, o.patch_type)
        rp.prepare()
        rp.draw_patches(o.handle, o.matrix, v.projection, o.color)
        rp.finish()
        plain = rp.read_framebuffer()
        for k in ("ids", "depth", "rgba8"):
            assert np.array_equal(plain[k], batched[vi][k]), "view %d %s" % (vi, k)
        rp.unload_patches(o.handle)
    rp.close()


def test_c3_views_match_oracle(oracle):
    w = scenes.c3_batch()
    rb = render_batch(w)
    for vi in (0, 1, 4, 6):
        sub = scenes.Workload("v", w.options, [w.views[vi]])
        ref = oracle.render(sub)[0]
        got = rb.read_framebuffer(view=vi)
        got["ranges"] = ref["ranges"]           # per-view ranges compared below
        assert_frame_equal(got, ref, "c3 view %d" % vi)
    rb.close()


def test_c3_batch_ranges_are_per_view_canonical(oracle):
    w = scenes.c3_batch(n_scenes=6)
    rb = render_batch(w)
    allr = rb.ranges()
    rb.close()
    off = 0
    for vi, v in enumerate(w.views):
        ref = oracle.render(scenes.Workload("v", w.options, [v]))[0]["ranges"]
        got = allr[off:off + len(ref)]
        off += len(ref)
        for f in ("pid", "depth", "umin", "vmin", "umax", "vmax"):
            assert np.array_equal(got[f], ref[f]), (vi, f)
    assert off == len(allr)


def test_c4_gregory_balanced_full_frame(oracle):
    w = scenes.c4_gregory()
    print("case c4 gregory 16384 patches BALANCED", flush=True)
    r, got = render_plain(w)
    ref = oracle.render(w)[0]
    assert_frame_equal(got, ref, "c4")
    r.close()


def test_c5_views_batched_match_oracle(oracle):
    w = scenes.c5_views()
    rb = render_batch(w)
    for vi in (0, 13, 31):
        ref = oracle.render(scenes.Workload("v", w.options, [w.views[vi]]))[0]
        got = rb.read_framebuffer(view=vi)
        got["ranges"] = ref["ranges"]
        assert_frame_equal(got, ref, "c5 view %d" % vi)
    rb.close()


def test_work_sharing_off_gives_the_same_frame(monkeypatch):
    w = scenes.c2_torus(width=960, height=540)
    r1, a = render_plain(w)
    r1.close()
    monkeypatch.setenv("MP_WORK_SHARING", "0")
    r2, b = render_plain(w)
    r2.close()
    assert np.array_equal(a["ids"], b["ids"]) and np.array_equal(a["depth"], b["depth"])


def test_small_batch_views_match_oracle(oracle):
    """Three views of under 4096 patches in all (the canonical order's count
    scan then runs inside the scatter launch, which also sets each view's
    range base): ranges per view and images against the oracle."""
    rng = np.random.default_rng(21)
    opts = dict(window_size=(320, 240), bound_n_split_method="LOCAL", bound_n_split_limit=8.0,
                reyes_patch_size=8, backface_culling=1)
    views = []
    for k, n in enumerate((900, 400, 1300)):
        radius = 2.0 * n ** (1.0 / 3.0)
        pats = scenes.random_patches(rng, n, radius)
        eye = (radius * 2.0, 0.4 * radius, -radius * 0.8)
        o = scenes._obj(50 + k, pats, scenes.BEZIER, scenes.look_at(eye, (0, 0, 0)), (0.5, 0.7, 0.9, 1.0))
        views.append(scenes.View(scenes.Projection(60.0, 0.01, (320, 240)), [o]))
    w = scenes.Workload("small_batch", opts, views)
    rb = render_batch(w)
    allr = rb.ranges()
    off = 0
    for vi, v in enumerate(w.views):
        ref = oracle.render(scenes.Workload("v", opts, [v]))[0]
        got = rb.read_framebuffer(view=vi)
        got["ranges"] = allr[off:off + len(ref["ranges"])]
        off += len(ref["ranges"])
        assert_frame_equal(got, ref, "small batch view %d" % vi)
    assert off == len(allr)
    rb.close()
