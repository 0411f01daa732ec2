"""The synthetic workloads have the properties their definitions state."""
import numpy as np

from micropolis_amd import scenes


def projected_spans(patches, mv, width, height, fovy=60.0):
    """Largest screen extent (pixels) of each patch's projected control points,
    and whether all of them lie in front of the eye."""
    f = (height / 2) / np.tan(np.radians(fovy / 2))
    pts = patches.reshape(-1, 3) @ mv[:3, :3].T + mv[:3, 3]
    pts = pts.reshape(patches.shape[0], -1, 3)
    z = -pts[..., 2]
    front = (z > 0.01).all(1)
    zz = np.maximum(z, 1e-3)
    sx = pts[..., 0] / zz * f
    sy = pts[..., 1] / zz * f
    span = np.maximum(sx.max(1) - sx.min(1), sy.max(1) - sy.min(1))
    return span, front


def test_torus_camera_has_large_near_patches_and_deep_splits(oracle):
    w = scenes.c2_torus()
    v = w.views[0]
    o = v.objects[0]
    span, front = projected_spans(o.patches, o.matrix, 1920, 1080)
    frac = np.mean(front & (span > 200))
    assert 0.29 <= frac <= 0.32
    view = oracle.make_view(60.0, 0.01, 1920, 1080)
    prm = oracle.make_params(w.options)
    r, _, _, _ = oracle.split("LOCAL", oracle.as_cps(o.patches), 0, o.matrix, view, prm, 16.0)
    hist = np.bincount(r["depth"])
    assert r["depth"].min() <= 1 and r["depth"].max() == 10
    assert (hist[1:11] > 0).all()


def test_batch_scenes_are_uneven(oracle):
    w = scenes.c3_batch()
    counts = [v.objects[0].patches.shape[0] for v in w.views]
    assert len(counts) == 64 and min(counts) >= 100 and max(counts) <= 20000
    assert max(counts) / min(counts) > 20


def test_gregory_planes_have_distinct_face_points():
    w = scenes.c4_gregory(planes=2, quads=32)
    g = w.views[0].objects[0].patches
    assert g.shape[1:] == (20, 3) and np.isfinite(g).all()
    # f- (16..19) is perturbed away from f+ (12..15), so the rational
    # interior terms of the Gregory evaluation matter
    assert np.abs(g[:, 16:20] - g[:, 12:16]).max() > 1e-3
