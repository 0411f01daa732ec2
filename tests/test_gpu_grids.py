"""The dice stage's grids, read through the mp_read_grids tap, against the
oracle's or_dice bit for bit: eye-space position (x, y, z, w), sub-pixel
coordinate and depth of every grid point (dice.cl:75-79).  The tap runs the
raster's own dice code, so a grid point that is wrong but never wins a pixel
is caught here.  Workloads: the 64-patch sphere (N 8, BREADTH) and the
displaced 1080p torus (N 16, LOCAL, value-noise displacement), on a sample
of ranges that includes the deepest and the first and last of the frame."""
import ctypes as C

import numpy as np
import pytest

from micropolis_amd import scenes
from gpu_util import render_plain

pytestmark = pytest.mark.gpu


def oracle_grid(oracle, w, rec):
    prm = oracle.make_params(w.options)
    N = prm.patch_size
    G = (N + 1) ** 2
    v = w.views[0]
    o = v.objects[int(rec["object"])]
    W, H = v.projection.viewport
    view = oracle.make_view(v.projection.fovy, v.projection.near, W, H)
    r = np.zeros(1, oracle.RANGE_DTYPE)
    for f in ("pid", "depth", "umin", "vmin", "umax", "vmax"):
        r[f] = rec[f]
    pos = np.zeros((G, 4), np.float32)
    pxl = np.zeros((G, 2), np.int32)
    dep = np.zeros(G, np.float32)
    cps = oracle.as_cps(o.patches)
    oracle.load().or_dice(oracle.fptr(cps), o.patch_type, r.ctypes.data, oracle.fptr(oracle.colmajor(o.matrix)),
                          C.byref(view), C.byref(prm), pos.ctypes.data, pxl.ctypes.data, dep.ctypes.data)
    return pos, pxl, dep


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_grid_tap_matches_oracle_dice(oracle, which):
    w = scenes.c1_sphere() if which == "c1" else scenes.c2_torus()
    print("case grids/%s" % which, flush=True)
    r, out = render_plain(w)
    ranges = out["ranges"]
    n = len(ranges)
    rng = np.random.default_rng(11)
    pick = set(rng.choice(n, size=min(n, 160), replace=False).tolist())
    pick |= {0, n - 1}
    pick |= set(np.argsort(ranges["depth"])[-8:].tolist())
    checked = 0
    for k in sorted(pick):
        pos, pxl, dep = r.grids(k, 1)
        opos, opxl, odep = oracle_grid(oracle, w, ranges[k])
        assert np.array_equal(pxl[0], opxl), (which, k, "pxlpos")
        assert np.array_equal(pos[0].view(np.uint32), opos.view(np.uint32)), (which, k, "pos")
        assert np.array_equal(dep[0].view(np.uint32), odep.view(np.uint32)), (which, k, "depth")
        checked += 1
    # a block of consecutive ranges in one call
    pos, pxl, dep = r.grids(n // 2, 16)
    for j in range(16):
        opos, opxl, odep = oracle_grid(oracle, w, ranges[n // 2 + j])
        assert np.array_equal(pxl[j], opxl) and np.array_equal(pos[j].view(np.uint32), opos.view(np.uint32))
        assert np.array_equal(dep[j].view(np.uint32), odep.view(np.uint32))
    assert checked > 100
    r.close()
