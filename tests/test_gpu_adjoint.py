"""GPU adjoint: gradients of an image loss w.r.t. control points, the
displacement amplitude and the albedo (visibility fixed to the forward ids)
against the FP64 CPU reference, and against Richardson-extrapolated central
differences (eps = 1e-3) of the fixed-visibility loss.  Tolerance: 1e-3
relative error."""
import numpy as np
import pytest

from micropolis_amd import Renderer, scenes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def setup_problem(n_views=2, width=480, height=272):
    from micropolis_amd.autograd import render_views
    w = scenes.c5_views(n_views=n_views, width=width, height=height)
    views = [(v.projection, v.objects[0].matrix) for v in w.views]
    base = w.views[0].objects[0]
    opts = dict(w.options, write_rgba32=1)
    rng = np.random.default_rng(0)
    cps0 = torch.tensor(base.patches, device="cuda")
    albedo0 = torch.tensor(base.color[:3], device="cuda")
    tgt_r = Renderer(**opts)
    tcps = cps0 + torch.tensor(rng.normal(0, 0.01, base.patches.shape).astype(np.float32), device="cuda")
    target, _ = render_views(tgt_r, tcps, 0.025, albedo0 + 0.1, views)
    r = Renderer(**opts)
    cps = cps0.clone().requires_grad_(True)
    amp = torch.tensor(0.02, device="cuda", requires_grad=True)
    albedo = albedo0.clone().requires_grad_(True)
    rgb, ids = render_views(r, cps, amp, albedo, views)
    mask = (ids != -1).unsqueeze(-1).float()
    loss = (((rgb - target) ** 2) * mask).sum()
    loss.backward()
    return dict(w=w, r=r, tgt_r=tgt_r, cps=cps, amp=amp, albedo=albedo, ids=ids, target=target, loss=loss,
                opts=opts, views=views, rgb=rgb.detach(), mask=mask)


def fp64_reference(oracle, pb):
    w, r = pb["w"], pb["r"]
    ranges = r.ranges()
    prm = oracle.make_params(pb["opts"])
    P = pb["cps"].detach().cpu().numpy()
    g_cps = np.zeros(P.shape, np.float64)
    g_alb = np.zeros(3)
    g_amp = 0.0
    loss = 0.0
    ids = pb["ids"].cpu().numpy().view(np.uint32)
    tgt = pb["target"].cpu().numpy()
    for vi, v in enumerate(w.views):
        rv = ranges[ranges["object"] == vi]
        rr = np.zeros(len(rv), oracle.RANGE_DTYPE)
        for f in ("pid", "depth", "umin", "vmin", "umax", "vmax"):
            rr[f] = rv[f]
        l, gc, ga, gm = oracle.adjoint_f64(P, rr, v.objects[0].matrix, prm, float(pb["amp"]),
                                          pb["albedo"].detach().cpu().numpy().astype(np.float64),
                                          ids[vi], tgt[vi])
        loss += l
        g_cps += gc
        g_alb += ga
        g_amp += gm
    return loss, g_cps, g_alb, g_amp


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_adjoint_matches_fp64_reference(oracle):
    pb = setup_problem()
    loss64, g_cps, g_alb, g_amp = fp64_reference(oracle, pb)
    print("loss gpu %.6f fp64 %.6f" % (float(pb["loss"]), loss64), flush=True)
    assert abs(float(pb["loss"]) - loss64) <= 1e-4 * loss64
    assert np.abs(g_cps).max() > 0
    assert rel(pb["cps"].grad.cpu().numpy(), g_cps) <= 1e-3
    assert rel(pb["albedo"].grad.cpu().numpy(), g_alb) <= 1e-3
    assert abs(float(pb["amp"].grad) - g_amp) <= 1e-3 * abs(g_amp)


def test_adjoint_matches_central_differences(oracle):
    """Richardson-extrapolated central differences (eps = 1e-3) of the FP64
    fixed-visibility loss, on the coordinates with the largest gradients."""
    pb = setup_problem(n_views=1)
    w, r = pb["w"], pb["r"]
    ranges = r.ranges()
    rr = np.zeros(len(ranges), oracle.RANGE_DTYPE)
    for f in ("pid", "depth", "umin", "vmin", "umax", "vmax"):
        rr[f] = ranges[f]
    prm = oracle.make_params(pb["opts"])
    P = pb["cps"].detach().cpu().numpy().astype(np.float64)
    ids = pb["ids"].cpu().numpy().view(np.uint32)[0]
    tgt = pb["target"].cpu().numpy()[0]
    alb = pb["albedo"].detach().cpu().numpy().astype(np.float64)
    mv = w.views[0].objects[0].matrix

    def L(x):
        return oracle.adjoint_f64(x.astype(np.float32), rr, mv, prm, float(pb["amp"]), alb, ids, tgt, grads=False)

    g = pb["cps"].grad.cpu().numpy().ravel()
    for i in np.argsort(np.abs(g))[-6:]:
        x0 = np.float32(P.ravel()[i])
        e = float(np.float32(x0 + np.float32(1e-3)) - x0)

        def f(h, i=i):
            x = P.copy().ravel()
            x[i] += h
            return L(x.reshape(P.shape))
        c1 = (f(e) - f(-e)) / (2 * e)
        c2 = (f(e / 2) - f(-e / 2)) / e
        fd = (4 * c2 - c1) / 3
        assert abs(fd - g[i]) <= 1e-3 * abs(fd), (i, fd, g[i])


def test_backward_refuses_a_frame_rendered_since(oracle):
    """Two forwards on one renderer, then the first one's backward: the
    library refuses (the visibility and control points it would use are the
    second frame's), and the second forward's backward still works."""
    from micropolis_amd.autograd import render_views
    from micropolis_amd.renderer import MicropolisError
    w = scenes.c5_views(n_views=1, width=96, height=64)
    views = [(v.projection, v.objects[0].matrix) for v in w.views]
    base = w.views[0].objects[0]
    r = Renderer(**dict(w.options, write_rgba32=1))
    cps = torch.tensor(base.patches, device="cuda").requires_grad_(True)
    albedo = torch.tensor(base.color[:3], device="cuda")
    rgb1, _ = render_views(r, cps, 0.02, albedo, views)
    rgb2, _ = render_views(r, cps, 0.02, albedo, views)
    with pytest.raises(MicropolisError, match="another frame"):
        rgb1.sum().backward()
    rgb2.sum().backward()
    assert cps.grad is not None and torch.isfinite(cps.grad).all()
    r.close()


def test_backward_and_readback_check_buffer_sizes():
    w = scenes.c5_views(n_views=2, width=96, height=64)
    base = w.views[0].objects[0]
    r = Renderer(**dict(w.options, write_rgba32=1))
    r.load_patches(base.handle, base.patches, base.patch_type)
    r.batch_begin([v.projection for v in w.views])
    for vi, v in enumerate(w.views):
        r.batch_draw(vi, base.handle, v.objects[0].matrix, base.color)
    r.batch_finish()
    from micropolis_amd.renderer import MicropolisError
    out = torch.empty((64, 96, 4), device="cuda")
    with pytest.raises(MicropolisError, match="width"):
        r.read_device(0, out.data_ptr(), None, 64 * 95)
    r.read_device(0, out.data_ptr(), None, 64 * 96)
    s = r.frame_serial()
    dl = torch.zeros((2, 64, 96, 4), device="cuda")
    g = torch.empty((2,) + base.patches.shape, device="cuda")
    gc = torch.empty((2, 4), device="cuda")
    ga = torch.empty((2,), device="cuda")
    with pytest.raises(MicropolisError, match="dl_dcolor"):
        r.backward(s, dl.data_ptr(), dl.numel() - 4, g.data_ptr(), g.numel(), gc.data_ptr(), ga.data_ptr(), 2)
    with pytest.raises(MicropolisError, match="grad_cps"):
        r.backward(s, dl.data_ptr(), dl.numel(), g.data_ptr(), g.numel() // 2, gc.data_ptr(), ga.data_ptr(), 2)
    with pytest.raises(MicropolisError, match="n_objects"):
        r.backward(s, dl.data_ptr(), dl.numel(), g.data_ptr(), g.numel(), gc.data_ptr(), ga.data_ptr(), 1)
    r.backward(s, dl.data_ptr(), dl.numel(), g.data_ptr(), g.numel(), gc.data_ptr(), ga.data_ptr(), 2)
    torch.cuda.synchronize()
    assert float(g.abs().max()) == 0.0   # zero loss gradient -> zero control-point gradient
    r.close()


def test_adjoint_at_the_c5_size(oracle):
    """The forward + adjoint workload at its full size: 32 views of the
    displaced torus at 960x540, loss = L2 to a target rendered with control points
    + N(0, 0.01), amplitude 0.025 and albedo + 0.1.  The GPU gradients w.r.t.
    all 4096x16x3 control points, the amplitude and the albedo against the
    FP64 oracle over all 32 views (<= 1e-3 relative), and the per-view
    gradients of mp_backward on four views against Richardson central
    differences (eps 1e-3) of the FP64 fixed-visibility loss on the largest
    control-point coordinates, the amplitude and the three albedo channels."""
    pb = setup_problem(n_views=32, width=960, height=540)
    loss64, g_cps, g_alb, g_amp = fp64_reference(oracle, pb)
    print("c5: loss gpu %.6f fp64 %.6f" % (float(pb["loss"]), loss64), flush=True)
    assert abs(float(pb["loss"]) - loss64) <= 1e-4 * loss64
    assert rel(pb["cps"].grad.cpu().numpy(), g_cps) <= 1e-3
    assert rel(pb["albedo"].grad.cpu().numpy(), g_alb) <= 1e-3
    assert abs(float(pb["amp"].grad) - g_amp) <= 1e-3 * abs(g_amp)

    # per-view gradients straight from mp_backward (the autograd node sums them)
    w, r = pb["w"], pb["r"]
    V = len(w.views)
    dl = torch.zeros(pb["rgb"].shape[:-1] + (4,), device="cuda")
    dl[..., :3] = 2.0 * (pb["rgb"] - pb["target"]) * pb["mask"]
    P = pb["cps"].detach().cpu().numpy()
    g_v = torch.empty((V,) + P.shape, device="cuda")
    g_c = torch.empty((V, 4), device="cuda")
    g_a = torch.empty((V,), device="cuda")
    torch.cuda.synchronize()
    r.backward(r.frame_serial(), dl.data_ptr(), dl.numel(), g_v.data_ptr(), g_v.numel(), g_c.data_ptr(),
               g_a.data_ptr(), V)
    sub = [0, 9, 17, 26]
    gv = g_v[sub].sum(0).cpu().numpy().ravel()
    gc = g_c[sub, :3].sum(0).cpu().numpy()
    ga = float(g_a[sub].sum())
    # the per-view outputs sum to the autograd gradient
    assert rel(g_v.sum(0).cpu().numpy(), pb["cps"].grad.cpu().numpy()) <= 1e-5

    ranges = r.ranges()
    prm = oracle.make_params(pb["opts"])
    ids = pb["ids"].cpu().numpy().view(np.uint32)
    tgt = pb["target"].cpu().numpy()
    amp0 = float(pb["amp"])
    alb0 = pb["albedo"].detach().cpu().numpy().astype(np.float64)
    per_view = []
    for vi in sub:
        rv = ranges[ranges["object"] == vi]
        rr = np.zeros(len(rv), oracle.RANGE_DTYPE)
        for f in ("pid", "depth", "umin", "vmin", "umax", "vmax"):
            rr[f] = rv[f]
        per_view.append((rr, w.views[vi].objects[0].matrix, ids[vi], tgt[vi]))

    def L(x=P, amp=amp0, alb=alb0):
        return sum(oracle.adjoint_f64(x, rr, mv, prm, amp, alb, iv, tv, grads=False)
                   for rr, mv, iv, tv in per_view)

    def richardson(f, e):
        c1 = (f(e) - f(-e)) / (2 * e)
        c2 = (f(e / 2) - f(-e / 2)) / e
        return (4 * c2 - c1) / 3

    for i in np.argsort(np.abs(gv))[-4:]:
        x0 = np.float32(P.ravel()[i])
        e = float(np.float32(x0 + np.float32(1e-3)) - x0)

        def fx(h, i=i):
            x = P.astype(np.float64).ravel()
            x[i] += h
            return L(x=x.reshape(P.shape).astype(np.float32))
        fd = richardson(fx, e)
        print("c5 cd: cp %d fd %.6g gpu %.6g" % (i, fd, gv[i]), flush=True)
        assert abs(fd - gv[i]) <= 1e-3 * abs(fd), (i, fd, gv[i])
    fd = richardson(lambda h: L(amp=amp0 + h), 1e-3)
    print("c5 cd: amplitude fd %.6g gpu %.6g" % (fd, ga), flush=True)
    assert abs(fd - ga) <= 1e-3 * abs(fd)
    for c in range(3):
        def fa(h, c=c):
            a = alb0.copy()
            a[c] += h
            return L(alb=a)
        fd = richardson(fa, 1e-3)
        print("c5 cd: albedo[%d] fd %.6g gpu %.6g" % (c, fd, gc[c]), flush=True)
        assert abs(fd - gc[c]) <= 1e-3 * abs(fd)

