"""GPU parity: the HIP library against the CPU oracle on the same seeded
inputs (bit-exact split ranges and visible ids; depth 1 ulp; RGBA8 exact)."""
import numpy as np
import pytest

from micropolis_amd import scenes
from gpu_util import assert_frame_equal, render_plain

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["BREADTH", "BOUNDED", "LOCAL", "BALANCED", "CPU"])
def test_c1_sphere_all_methods(oracle, method):
    w = scenes.c1_sphere()
    print("case c1_sphere", method, flush=True)
    r, got = render_plain(w, bound_n_split_method=method, write_rgba32=1)
    ref = oracle.render(w, method=method)[0]
    assert_frame_equal(got, ref, "c1/" + method)
    if method in ("BREADTH", "BOUNDED", "CPU"):
        assert np.array_equal(got["pass_counts"], ref["pass_counts"])
    if method == "CPU":
        assert got["stats"]["bounded_patches"] == ref["n_bounds"]
    assert got["stats"]["dropped_ranges"] == ref["n_dropped"]
    r.close()


def test_c1_untruncated_methods_agree(oracle):
    """Every method that does not drop ranges produces the same frame."""
    w = scenes.c1_sphere()
    outs = []
    for m in ("BOUNDED", "LOCAL", "BALANCED"):
        r, got = render_plain(w, bound_n_split_method=m)
        outs.append(got)
        r.close()
    for g in outs[1:]:
        assert np.array_equal(g["ids"], outs[0]["ids"])
        assert np.array_equal(g["depth"], outs[0]["depth"])


def test_c2_torus_full_frame(oracle):
    w = scenes.c2_torus()
    print("case c2_torus LOCAL 1920x1080", flush=True)
    r, got = render_plain(w, write_rgba32=1)
    ref = oracle.render(w)[0]
    assert_frame_equal(got, ref, "c2")
    assert got["stats"]["slow_blocks"] >= 0
    r.close()


def test_replay_and_graph_identical(oracle):
    w = scenes.c2_torus(width=640, height=360)
    r, got = render_plain(w)
    for use_graph in (False, True):
        r.replay(3, use_graph=use_graph)
        again = r.read_framebuffer()
        assert np.array_equal(again["ids"], got["ids"])
        assert np.array_equal(again["rgba8"], got["rgba8"])
    r.close()


def test_cpu_method_gregory_and_replay(oracle):
    """CPU method on Gregory patches (NaN corner samples of the host
    evaluation) through the device back end, and its host-driven replay."""
    w = scenes.c4_gregory(planes=4, quads=128, width=640, height=360)
    w.options["reyes_patches_per_pass"] = 2000
    print("case c4_gregory small CPU", flush=True)
    r, got = render_plain(w, bound_n_split_method="CPU")
    ref = oracle.render(w, method="CPU")[0]
    assert_frame_equal(got, ref, "c4-small/CPU")
    assert np.array_equal(got["pass_counts"], ref["pass_counts"])
    r.replay(2, use_graph=False)
    again = r.read_framebuffer()
    assert np.array_equal(again["ids"], got["ids"])
    r.close()


def _frame_of(r):
    out = r.read_framebuffer()
    out["ranges"] = r.ranges()
    return out


def test_every_scheduling_gives_the_same_frame(oracle):
    """Split engine, work sharing and the raster range schedule change only
    the schedule (BALANCED)."""
    w = scenes.c4_gregory(planes=8, quads=128, width=960, height=540)
    print("case c4_gregory small BALANCED tunings", flush=True)
    ref = oracle.render(w)[0]
    r, got = render_plain(w)
    assert_frame_equal(got, ref, "c4-small/default")
    for engine, sharing, chunk, levels in ((0, 0, 0, 0), (0, 1, -1, 0), (1, 0, 0, 0), (1, 1, -1, 0),
                                           (0, 0, 7, 0), (1, 0, 0, 2), (1, 1, 0, 3), (2, 0, 0, 0), (2, 0, -1, 0), (2, 0, -16, 0)):
        r.set_tuning(split_engine=engine, work_sharing=sharing, raster_chunk=chunk, split_levels=levels)
        r.replay(2, use_graph=True)
        assert_frame_equal(_frame_of(r), ref, "c4-small/engine%d-sharing%d-chunk%d-levels%d"
                           % (engine, sharing, chunk, levels))
    r.close()


def test_split_budgets_give_the_same_frame(oracle):
    """LOCAL persistent split with step budgets (the left ranges dealt to
    all groups in a second launch), with and without work sharing, in both
    register builds of the split kernel."""
    w = scenes.c2_torus(width=640, height=360)
    print("case c2_torus 640x360 split budgets", flush=True)
    ref = oracle.render(w)[0]
    r, got = render_plain(w)
    assert_frame_equal(got, ref, "c2-small/default")
    for sharing, budget, waves in ((0, 0, 2), (0, 1, 2), (0, 3, 2), (0, 8, 2), (1, 3, 2), (0, 0, 3), (0, 8, 3)):
        r.set_tuning(split_engine=0, work_sharing=sharing, split_budget=budget, split_waves=waves)
        r.replay(2, use_graph=True)
        assert_frame_equal(_frame_of(r), ref, "c2-small/sharing%d-budget%d-waves%d" % (sharing, budget, waves))
        t = r.tuning()
        assert t["split_budget"] == budget and t["split_waves"] == waves
    r.close()


def test_autotune_keeps_the_frame(oracle):
    w = scenes.c2_torus(width=640, height=360)
    print("case c2_torus 640x360 autotune", flush=True)
    ref = oracle.render(w)[0]
    r, got = render_plain(w)
    t = r.autotune(2)
    assert t["split_engine"] in (0, 1, 2) and t["frame_ms"] > 0.0
    assert r.tuning()["split_engine"] == t["split_engine"]
    r.replay(2, use_graph=True)
    assert_frame_equal(_frame_of(r), ref, "c2-small/autotuned")
    r.close()
