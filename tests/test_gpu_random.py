"""GPU parity on seeded random frames: clouds of random Bezier patches with a
few Gregory planes, one to three objects per frame (patch sets drawn twice
under other matrices), ragged window sizes, 8x8 and 16x16 grids, LOCAL /
BALANCED / BREADTH, displacement on some; each frame is then replayed under
every raster range schedule and must not change."""
import numpy as np
import pytest

from micropolis_amd import scenes
from micropolis_amd.renderer import Projection
from gpu_util import assert_frame_equal, render_plain

pytestmark = pytest.mark.gpu


def random_frame(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.choice([8, 16]))
    method = str(rng.choice(["LOCAL", "BALANCED", "BREADTH"]))
    disp = n == 16 and rng.uniform() < 0.5   # displacement is defined for Bezier patches only
    size = (int(rng.integers(200, 700)), int(rng.integers(150, 500)))
    objs = []
    handle = 1
    for k in range(int(rng.integers(1, 4))):
        npat = int(rng.integers(20, 200))
        radius = 1.5 * npat ** (1.0 / 3.0)
        pats = scenes.random_patches(rng, npat, radius)
        dist = radius * 1.8 + rng.uniform(1.0, 4.0)
        ang = rng.uniform(0, 2 * np.pi)
        mv = scenes.look_at((dist * np.cos(ang), rng.uniform(-1, 1) * dist * 0.3, dist * np.sin(ang)), (0, 0, 0))
        col = (rng.uniform(0.3, 1.0), rng.uniform(0.3, 1.0), rng.uniform(0.3, 1.0), 1.0)
        objs.append(scenes.SceneObject(handle, pats, scenes.BEZIER, mv, np.array(col, np.float32)))
        if rng.uniform() < 0.5:   # the same patch set again, moved
            mv2 = mv.copy()
            mv2[:3, 3] += rng.uniform(-1.0, 1.0, 3).astype(np.float32)
            objs.append(scenes.SceneObject(handle, pats, scenes.BEZIER, mv2, np.array(col[::-1], np.float32)))
        handle += 1
    if not disp and rng.uniform() < 0.5:
        g = scenes.gregory_planes(rng, n_planes=2, nu=4, nv=2)
        mv = scenes.look_at((0.0, 0.0, 3.0), (0.0, 0.0, -0.6))
        objs.append(scenes.SceneObject(100, g, scenes.GREGORY, mv, np.array([0.6, 0.8, 0.9, 1.0], np.float32)))
    opts = dict(window_size=size, bound_n_split_method=method, bound_n_split_limit=float(n),
                reyes_patch_size=n, reyes_patches_per_pass=512, local_bns_work_groups=0,
                backface_culling=int(rng.integers(0, 2)), write_rgba32=1)
    if disp:
        opts.update(displacement=1, displacement_amplitude=0.05, displacement_frequency=2.0,
                    displacement_seed=int(rng.integers(1, 100)), displacement_octaves=int(rng.integers(1, 5)))
    proj = Projection(60.0, 0.01, size)
    return scenes.Workload("random%d" % seed, opts, [scenes.View(proj, objs)])


@pytest.mark.parametrize("seed", range(16))
def test_random_frame_matches_oracle(oracle, seed):
    w = random_frame(seed)
    print("case random frame", seed, w.options["bound_n_split_method"], w.options["reyes_patch_size"],
          w.options["window_size"], len(w.views[0].objects), "objects", flush=True)
    r, got = render_plain(w)
    ref = oracle.render(w)[0]
    assert_frame_equal(got, ref, "random%d" % seed)
    if w.options["bound_n_split_method"] != "BREADTH":
        for chunk in (0, -1, 3, -2, -16):
            t = dict(r.tuning(), raster_chunk=chunk)
            r.set_tuning(**t)
            r.replay(2, use_graph=True)
            again = r.read_framebuffer()
            assert np.array_equal(again["ids"], got["ids"]), (seed, chunk)
            assert np.array_equal(again["rgba8"], got["rgba8"]), (seed, chunk)
    r.close()
