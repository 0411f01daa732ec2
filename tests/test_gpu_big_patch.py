"""One patch that splits into 154,634 DRAW ranges: the canonical order of a
bucket that large comes from the rank kernel's radix sort (linear in the
bucket) instead of the per-range count (quadratic).  Ranges, visible ids,
depth and colour must equal the oracle's; the statistics report every range
as sorted."""
import numpy as np
import pytest

from micropolis_amd import scenes
from gpu_util import assert_frame_equal, render_plain

pytestmark = pytest.mark.gpu


def test_one_patch_of_150k_ranges(oracle):
    w = scenes.big_patch(2.83)
    print("case big patch (limit 2.83)", flush=True)
    r, got = render_plain(w, write_rgba32=1)
    ref = oracle.render(w)[0]
    n = len(ref["ranges"])
    assert n >= 100000
    assert got["stats"]["sorted_ranges"] == n
    assert_frame_equal(got, ref, "big patch")
    # the order stage replayed: still the same canonical list
    r.replay(2, use_graph=True)
    again = r.ranges()
    assert np.array_equal(again, got["ranges"])
    st = r.replay_stages(3)
    print("big patch: %d ranges, stages %s" % (n, st), flush=True)
    r.close()
