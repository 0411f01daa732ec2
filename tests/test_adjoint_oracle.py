"""Pins the FP64 adjoint (oracle/adjoint.c) by central finite differences
(eps = 1e-3) of the same fixed-visibility loss: control points, the
displacement amplitude and the albedo."""
import numpy as np
import pytest

from micropolis_amd import scenes


@pytest.fixture(scope="module")
def problem(oracle):
    w = scenes.c5_views(n_views=1, width=240, height=136)
    v = w.views[0]
    o = v.objects[0]
    fwd = oracle.render(w)[0]
    rng = np.random.default_rng(0)
    tgt_patches = o.patches + rng.normal(0, 0.01, o.patches.shape).astype(np.float32)
    tw = scenes.Workload("t", dict(w.options, displacement_amplitude=0.025), [
        scenes.View(v.projection, [scenes.SceneObject(1, tgt_patches, 0, o.matrix, o.color + np.float32([0.1, 0.1, 0.1, 0]))])])
    tgt = oracle.render(tw)[0]["rgba32"][..., :3].copy()
    prm = oracle.make_params(w.options)
    return dict(patches=o.patches.copy(), ranges=fwd["ranges"], mv=o.matrix, prm=prm, amp=0.02,
                albedo=o.color[:3].astype(np.float64), ids=fwd["ids"], target=tgt)


def loss_at(oracle, p, patches=None, amp=None, albedo=None):
    return oracle.adjoint_f64(p["patches"] if patches is None else patches, p["ranges"], p["mv"], p["prm"],
                              p["amp"] if amp is None else amp, p["albedo"] if albedo is None else albedo,
                              p["ids"], p["target"], grads=False)


def central(f, eps):
    return (f(eps) - f(-eps)) / (2 * eps)


def richardson(f, eps):
    """Central differences at eps and eps/2 combined to cancel the eps^2 term."""
    return (4.0 * central(f, eps / 2) - central(f, eps)) / 3.0


def test_fp64_gradient_matches_central_differences(oracle, problem):
    """Central differences with eps = 1e-3 agree to 1% (their own truncation
    error is ~1e-3 at micropolygon scale); Richardson-extrapolated from the
    same eps they agree to 1e-3."""
    p = problem
    loss, g_cps, g_alb, g_amp = oracle.adjoint_f64(p["patches"], p["ranges"], p["mv"], p["prm"], p["amp"],
                                                   p["albedo"], p["ids"], p["target"])
    assert loss > 0 and np.abs(g_cps).max() > 0
    eps = 1e-3
    base = p["patches"].astype(np.float64)
    flat = g_cps.ravel()
    for i in np.argsort(np.abs(flat))[-10:]:
        def f(h, i=i):
            x = base.copy().ravel()
            x[i] += h
            return loss_at(oracle, p, patches=x.reshape(base.shape).astype(np.float32))
        # float32 inputs: use steps representable around the coordinate
        x0 = np.float32(base.ravel()[i])
        e = float(np.float32(x0 + np.float32(eps)) - x0)
        assert abs(central(f, e) - flat[i]) <= 1e-2 * abs(flat[i])
        assert abs(richardson(f, e) - flat[i]) <= 1e-3 * abs(flat[i]), (i, richardson(f, e), flat[i])
    fa = lambda h: loss_at(oracle, p, amp=p["amp"] + h)
    assert abs(central(fa, eps) - g_amp) <= 1e-2 * abs(g_amp)
    assert abs(richardson(fa, eps) - g_amp) <= 1e-3 * abs(g_amp)
    for c in range(3):
        def fb(h, c=c):
            a = p["albedo"].copy()
            a[c] += h
            return loss_at(oracle, p, albedo=a)
        assert abs(central(fb, eps) - g_alb[c]) <= 1e-6 * abs(g_alb[c]) + 1e-6   # quadratic: exact


def test_fp64_loss_matches_fp32_forward(oracle, problem):
    """The FP64 colours of the visible quads agree with the forward frame."""
    p = problem
    fwd_loss = loss_at(oracle, p)
    w = scenes.c5_views(n_views=1, width=240, height=136)
    img = oracle.render(w)[0]["rgba32"][..., :3]
    mask = p["ids"] != 0xFFFFFFFF
    ref = float((((img - p["target"]) ** 2).sum(axis=-1) * mask).sum())
    assert abs(fwd_loss - ref) <= 1e-4 * ref
