"""Helpers shared by the GPU parity tests: render a scenes.Workload through
the HIP library and compare with the oracle."""
import numpy as np

from micropolis_amd import Renderer


def render_plain(workload, view=0, **overrides):
    """One plain frame (prepare / draw_patches / finish) of one view."""
    opts = dict(workload.options)
    opts.update(overrides)
    v = workload.views[view]
    opts["window_size"] = v.projection.viewport
    r = Renderer(**opts)
    r.prepare()
    for o in v.objects:
        if not r.are_patches_loaded(o.handle):
            r.load_patches(o.handle, o.patches, o.patch_type)
        r.draw_patches(o.handle, o.matrix, v.projection, o.color)
    r.finish()
    out = r.read_framebuffer(rgba32=bool(opts.get("write_rgba32", 0)))
    out["ranges"] = r.ranges()
    out["stats"] = r.stats()
    out["pass_counts"] = r.pass_counts()
    return r, out


def assert_frame_equal(got, ref, what=""):
    """Bit-exact visible ids and ranges, depth within 1 ulp (here exact),
    RGBA8 bit-exact (the device converts with the host's threshold table),
    linear colour bit-exact when present."""
    gr, rr = got["ranges"], ref["ranges"]
    assert len(gr) == len(rr), "%s: %d vs %d ranges" % (what, len(gr), len(rr))
    for f in ("pid", "depth", "umin", "vmin", "umax", "vmax"):
        assert np.array_equal(gr[f], rr[f]), "%s: range field %s differs" % (what, f)
    ids_equal = np.array_equal(got["ids"], ref["ids"])
    if not ids_equal:
        bad = np.argwhere(got["ids"] != ref["ids"])
        raise AssertionError("%s: %d visible ids differ, first at %s (gpu %x, oracle %x)" % (
            what, len(bad), bad[0], got["ids"][tuple(bad[0])], ref["ids"][tuple(bad[0])]))
    gd = got["depth"].view(np.int32).astype(np.int64)
    rd = ref["depth"].view(np.int32).astype(np.int64)
    assert np.abs(gd - rd).max() <= 1, "%s: depth differs by more than 1 ulp" % what
    dc = np.abs(got["rgba8"].astype(int) - ref["rgba8"].astype(int)).max()
    assert dc == 0, "%s: display colour differs by %d/255" % (what, dc)
    if "rgba32" in got:
        a, b = got["rgba32"], ref["rgba32"]
        same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
        assert same.all(), "%s: linear colour differs" % what
