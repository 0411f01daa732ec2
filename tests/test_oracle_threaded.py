"""The threaded oracle frame (the all-cores CPU baseline of bench.py) equals
the single-threaded one: same DRAW ranges, ids, depth and colours."""
import numpy as np

from micropolis_amd import scenes


def test_threaded_render_equals_single_thread(oracle):
    w = scenes.c1_sphere()
    w = scenes.Workload(w.name, dict(w.options, bound_n_split_method="LOCAL"), w.views)
    ref = oracle.render(w)[0]
    for T in (1, 3, 8):
        got = oracle.render_threaded(w, T)[0]
        assert np.array_equal(got["ranges"], ref["ranges"])
        assert np.array_equal(got["ids"], ref["ids"])
        assert np.array_equal(got["depth"].view(np.uint32), ref["depth"].view(np.uint32))
        assert np.array_equal(got["rgba8"], ref["rgba8"])
        assert got["n_bounds"] == ref["n_bounds"]
