"""mp_render_noalloc (the allocation-free frame, scratch in a device array of
the library) against the context path (mp_finish + mp_read_framebuffer) on
the same configuration and draw: visible ids, depth and RGBA8 bit-exact, on
two calls in a row -- the displaced 1080p torus (4096 patches, LOCAL, N 16),
then the smaller 64-patch sphere (1024x768, LOCAL, N 8) -- and a third,
Gregory BALANCED, call; frames that cannot fit the scratch array are refused
before any launch."""
import ctypes as C

import numpy as np
import pytest

from micropolis_amd import Renderer, _lib, scenes
from micropolis_amd.renderer import make_config, _colmajor, _fptr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def context_frame(w, **over):
    opts = dict(w.options, **over)
    v = w.views[0]
    opts["window_size"] = v.projection.viewport
    r = Renderer(**opts)
    r.prepare()
    for o in v.objects:
        r.load_patches(o.handle, o.patches, o.patch_type)
        r.draw_patches(o.handle, o.matrix, v.projection, o.color)
    r.finish()
    out = r.read_framebuffer()
    r.close()
    return out


def noalloc_frame(w, **over):
    lib = _lib.load()
    opts = dict(w.options, **over)
    v = w.views[0]
    opts["window_size"] = v.projection.viewport
    cfg = make_config(**opts)
    (o,) = v.objects
    W, H = v.projection.viewport
    xyz = torch.tensor(np.ascontiguousarray(o.patches.reshape(-1, 3)), device="cuda")
    rgba8 = torch.full((H, W, 4), 7, dtype=torch.uint8, device="cuda")
    depth = torch.full((H, W), -5.0, dtype=torch.float32, device="cuda")
    ids = torch.full((H, W), 12345, dtype=torch.int32, device="cuda")
    mv = _colmajor(o.matrix)
    col = np.ascontiguousarray(np.asarray(o.color, np.float32))
    pr = v.projection.as_c()
    torch.cuda.synchronize()
    rc = lib.mp_render_noalloc(C.byref(cfg), C.c_void_p(xyz.data_ptr()), xyz.shape[0], int(o.patch_type),
                               _fptr(mv), C.byref(pr), _fptr(col), C.c_void_p(rgba8.data_ptr()),
                               C.c_void_p(depth.data_ptr()), C.c_void_p(ids.data_ptr()), None)
    assert rc == 0, lib.mp_create_error().decode()
    return {"rgba8": rgba8.cpu().numpy(), "depth": depth.cpu().numpy(),
            "ids": ids.cpu().numpy().view(np.uint32)}


def assert_same(got, ref, what):
    assert np.array_equal(got["ids"], ref["ids"]), what + ": ids differ"
    assert np.array_equal(got["depth"].view(np.uint32), ref["depth"].view(np.uint32)), what + ": depth differs"
    assert np.array_equal(got["rgba8"], ref["rgba8"]), what + ": rgba8 differs"


def test_noalloc_equals_the_context_frame_on_consecutive_calls():
    torus = scenes.c2_torus()
    print("case noalloc/torus 1080p", flush=True)
    a = noalloc_frame(torus)
    assert_same(a, context_frame(torus), "torus")
    assert (a["ids"] != 0xFFFFFFFF).sum() > 100000
    sphere = scenes.c1_sphere()
    print("case noalloc/sphere (smaller input)", flush=True)
    b = noalloc_frame(sphere, bound_n_split_method="LOCAL")
    assert_same(b, context_frame(sphere, bound_n_split_method="LOCAL"), "sphere")
    assert (b["ids"] != 0xFFFFFFFF).sum() > 10000
    greg = scenes.c4_gregory(planes=4, quads=64, width=640, height=360)
    print("case noalloc/gregory balanced", flush=True)
    c = noalloc_frame(greg)
    assert_same(c, context_frame(greg), "gregory")


def test_noalloc_refuses_what_it_cannot_hold():
    lib = _lib.load()
    sphere = scenes.c1_sphere()
    o = sphere.views[0].objects[0]
    mv = _colmajor(o.matrix)
    col = np.ascontiguousarray(np.asarray(o.color, np.float32))
    xyz = torch.tensor(np.ascontiguousarray(o.patches.reshape(-1, 3)), device="cuda")
    out = torch.empty(16, dtype=torch.int32, device="cuda")
    # a window whose visibility keys alone exceed the scratch array
    cfg = make_config(**dict(sphere.options, bound_n_split_method="LOCAL", window_size=(8192, 8192)))
    pr = _lib.mp_projection(60.0, 0.01, 8192, 8192)
    p = C.c_void_p(out.data_ptr())
    rc = lib.mp_render_noalloc(C.byref(cfg), C.c_void_p(xyz.data_ptr()), xyz.shape[0], 0, _fptr(mv), C.byref(pr),
                               _fptr(col), p, p, p, None)
    assert rc == -1   # MP_E_INVALID
    assert b"scratch" in lib.mp_create_error()
    # host-driven methods are not part of the allocation-free frame
    cfg = make_config(**dict(sphere.options, bound_n_split_method="BREADTH"))
    pr = sphere.views[0].projection.as_c()
    rc = lib.mp_render_noalloc(C.byref(cfg), C.c_void_p(xyz.data_ptr()), xyz.shape[0], 0, _fptr(mv), C.byref(pr),
                               _fptr(col), p, p, p, None)
    assert rc == -1 and b"LOCAL" in lib.mp_create_error()
