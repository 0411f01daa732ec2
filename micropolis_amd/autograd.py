"""torch.autograd binding of the differentiable REYES frame.

``render_views(renderer, cps, amplitude, albedo, views)`` renders one shared
Bezier patch set from several cameras in one batch and returns the linear
RGB images as a (V, H, W, 3) CUDA tensor.  Its backward runs the library's
hand-written reverse pass (mp_backward): gradients with respect to the
control points, the displacement amplitude and the albedo, with visibility
held fixed to the forward visible-id buffers (silhouette terms ignored) and
micropolygon grids recomputed per range.

Every view of one call has the same viewport (the outputs are one stacked
tensor).  The backward is taken for the frame its forward rendered: if the
renderer renders another frame (or its patches change) in between, the
library refuses (MicropolisError) instead of differentiating the wrong frame.
"""
import torch

from .renderer import BEZIER, MicropolisError

_HANDLE = 0x5EED


class _RenderViews(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cps, amplitude, albedo, renderer, views):
        assert cps.is_cuda and cps.dtype == torch.float32
        cps_c = cps.detach().contiguous()
        n_points = cps_c.numel() // 3
        if not renderer.are_patches_loaded(_HANDLE):
            renderer.load_patches(_HANDLE, cps_c.cpu().numpy(), BEZIER)
        else:
            torch.cuda.synchronize()
            renderer.update_patches_device(_HANDLE, cps_c.data_ptr(), n_points)
        renderer.set_displacement_amplitude(float(amplitude))
        col = [float(a) for a in albedo.detach().cpu()] + [1.0]
        renderer.batch_begin([p for p, _ in views])
        for i, (_, mv) in enumerate(views):
            renderer.batch_draw(i, _HANDLE, mv, col)
        renderer.batch_finish()
        w, h = views[0][0].viewport
        out = torch.empty((len(views), h, w, 4), dtype=torch.float32, device=cps.device)
        ids = torch.empty((len(views), h, w), dtype=torch.int32, device=cps.device)
        torch.cuda.synchronize()
        for i in range(len(views)):
            renderer.read_device(i, out[i].data_ptr(), ids[i].data_ptr(), h * w)
        ctx.renderer = renderer
        ctx.serial = renderer.frame_serial()
        ctx.shape = cps.shape
        ctx.n_views = len(views)
        ctx.mark_non_differentiable(ids)
        return out[..., :3].contiguous(), ids

    @staticmethod
    def backward(ctx, grad_rgb, grad_ids):
        V = ctx.n_views
        dl = torch.zeros(grad_rgb.shape[:-1] + (4,), dtype=torch.float32, device=grad_rgb.device)
        dl[..., :3] = grad_rgb
        g_cps = torch.empty((V,) + tuple(ctx.shape), dtype=torch.float32, device=grad_rgb.device)
        g_col = torch.empty((V, 4), dtype=torch.float32, device=grad_rgb.device)
        g_amp = torch.empty((V,), dtype=torch.float32, device=grad_rgb.device)
        torch.cuda.synchronize()
        ctx.renderer.backward(ctx.serial, dl.data_ptr(), dl.numel(), g_cps.data_ptr(), g_cps.numel(),
                              g_col.data_ptr(), g_amp.data_ptr(), V)
        return g_cps.sum(0), g_amp.sum(), g_col[:, :3].sum(0), None, None


def check_views(views):
    """Raise unless every (projection, modelview) view shares one viewport."""
    if not views:
        raise MicropolisError("render_views needs at least one view")
    sizes = {tuple(p.viewport) for p, _ in views}
    if len(sizes) != 1:
        raise MicropolisError("render_views needs one viewport for every view, got %s" % sorted(sizes))


def render_views(renderer, cps, amplitude, albedo, views):
    """-> (rgb (V,H,W,3) float32, ids (V,H,W) int32 visible micropolygon ids)."""
    check_views(views)
    if not torch.is_tensor(amplitude):
        amplitude = torch.tensor(float(amplitude))
    return _RenderViews.apply(cps, amplitude, albedo, renderer, views)
