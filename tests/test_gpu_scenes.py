"""GPU parity on the reference's own scene files (tests/golden/mscene):
the Utah teapot as Bezier and Gregory patches, 20 transformed teapots drawn
through the Scene::draw mirror, and a Gregory patch crossing the eye plane."""
import os

import numpy as np
import pytest

from micropolis_amd import Renderer, mscene
from gpu_util import assert_frame_equal, render_plain

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(__file__), "golden", "mscene")


def scene(name):
    return mscene.load(os.path.join(HERE, name + ".mscene"))


@pytest.mark.parametrize("name,method", [("teapot", "LOCAL"), ("teapot", "BREADTH"), ("teapot", "CPU"),
                                         ("teapot_greg", "BALANCED"), ("teapot_greg", "BOUNDED"),
                                         ("eye_split", "LOCAL")])
def test_reference_scene_matches_oracle(oracle, name, method):
    print("case", name, method, flush=True)
    w = mscene.draw_workload(scene(name), window=(800, 600),
                             options=dict(bound_n_split_method=method, reyes_patch_size=8,
                                          bound_n_split_limit=8.0, reyes_patches_per_pass=1000,
                                          write_rgba32=1))
    r, got = render_plain(w)
    ref = oracle.render(w)[0]
    assert_frame_equal(got, ref, "%s/%s" % (name, method))
    if method in ("BREADTH", "BOUNDED", "CPU"):
        assert np.array_equal(got["pass_counts"], ref["pass_counts"])
    r.close()


def test_scene_draw_mirror_renders_teapots(oracle):
    print("case teapots via Scene::draw mirror", flush=True)
    sc = scene("teapots")
    opts = dict(bound_n_split_method="LOCAL", reyes_patch_size=16, bound_n_split_limit=16.0)
    r = Renderer(window_size=(640, 480), **opts)
    mscene.draw(r, sc)
    got = r.read_framebuffer()
    got["ranges"] = r.ranges()
    ref = oracle.render(mscene.draw_workload(sc, window=(640, 480), options=opts))[0]
    assert_frame_equal(got, ref, "teapots")
    assert len(np.unique(got["ids"][got["ids"] != 0xFFFFFFFF] // 256)) > 100
    r.close()
