"""Host-side mirror of the reference renderer interface over the HIP library.

``Renderer`` follows Reyes::Renderer (src/Reyes/Renderer.h:29-45):
prepare / are_patches_loaded / load_patches / draw_patches / finish, plus the
framebuffer readback of Reyes::Framebuffer (src/Reyes/Framebuffer.cpp:107-124).
Option names are the reference's reyes.options names.  Errors raise
``MicropolisError`` with the library's message, as CL::Exception does in the
reference (src/CL/Exception.h:28-29).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (MP_BEZIER, MP_GREGORY, METHODS, mp_projection, mp_range_record,
                   mp_stats)

BEZIER, GREGORY = MP_BEZIER, MP_GREGORY


class MicropolisError(RuntimeError):
    pass


class Projection:
    """Reyes::Projection(fovy, hither, viewport) (src/base/Projection.h:44)."""

    def __init__(self, fovy, near, viewport):
        self.fovy = float(fovy)
        self.near = float(near)
        self.viewport = (int(viewport[0]), int(viewport[1]))

    def as_c(self):
        return mp_projection(self.fovy, self.near, self.viewport[0], self.viewport[1])


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _colmajor(m):
    m = np.asarray(m, dtype=np.float32)
    if m.shape == (4, 4):
        return np.ascontiguousarray(m.T.reshape(16))
    return np.ascontiguousarray(m.reshape(16))


def make_config(**options):
    """mp_config from reyes.options names (unset options keep their defaults)."""
    cfg = _lib.default_config()
    for key, val in options.items():
        if key == "window_size":
            cfg.window_width, cfg.window_height = int(val[0]), int(val[1])
        elif key == "clear_color":
            for i in range(4):
                cfg.clear_color[i] = float(val[i])
        elif key == "bound_n_split_method":
            cfg.bound_n_split_method = METHODS[val] if isinstance(val, str) else int(val)
        elif hasattr(cfg, key):
            setattr(cfg, key, type(getattr(cfg, key))(val))
        else:
            raise MicropolisError("unknown option %r" % key)
    return cfg


RANGE_DTYPE = np.dtype([("pid", "<i4"), ("object", "<i4"), ("depth", "<i4"), ("umin", "<f4"),
                        ("vmin", "<f4"), ("umax", "<f4"), ("vmax", "<f4")])


def cpu_bound_n_split(patches, patch_type, matrix, projection, **options):
    """The CPU method's host split of one patch set (Reyes::BoundNSplitCLCPU),
    run to completion without a device.  Returns (DRAW ranges in canonical
    order as a RANGE_DTYPE array, per-batch DRAW counts, ranges bounded)."""
    lib = _lib.load()
    cfg = make_config(**options)
    xyz = np.ascontiguousarray(np.asarray(patches, np.float32).reshape(-1, 3))
    mv = _colmajor(matrix)
    pr = projection.as_c()
    n = C.c_int64()
    npass = C.c_int32()
    nb = C.c_int64()
    rc = lib.mp_cpu_bound_n_split(C.byref(cfg), _fptr(xyz), xyz.shape[0], int(patch_type), _fptr(mv),
                                  C.byref(pr), None, 0, C.byref(n), None, 0, C.byref(npass), None)
    if rc != 0:
        raise MicropolisError(lib.mp_create_error().decode())
    out = np.zeros(n.value, RANGE_DTYPE)
    passes = np.zeros(max(1, npass.value), np.int32)
    rc = lib.mp_cpu_bound_n_split(C.byref(cfg), _fptr(xyz), xyz.shape[0], int(patch_type), _fptr(mv),
                                  C.byref(pr), out.ctypes.data_as(C.POINTER(mp_range_record)), n.value,
                                  C.byref(n), passes.ctypes.data_as(C.POINTER(C.c_int32)), passes.size,
                                  C.byref(npass), C.byref(nb))
    if rc != 0:
        raise MicropolisError(lib.mp_create_error().decode())
    return out, passes[:npass.value].copy(), nb.value


class Renderer:
    """Reyes::RendererCL over libmicropolis_hip.so."""

    def __init__(self, **options):
        lib = _lib.load()
        cfg = make_config(**options)
        self.cfg = cfg
        self._lib = lib
        ctx = lib.mp_create(C.byref(cfg))
        if not ctx:
            raise MicropolisError(lib.mp_create_error().decode())
        self._ctx = ctx
        self._views = [(cfg.window_width, cfg.window_height)]

    # -- helpers --
    def _check(self, rc):
        if rc != 0:
            raise MicropolisError(self._lib.mp_last_error(self._ctx).decode())

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.mp_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- Reyes::Renderer --
    def prepare(self):
        self._check(self._lib.mp_prepare(self._ctx))
        self._views = [(self.cfg.window_width, self.cfg.window_height)]

    def finish(self):
        self._check(self._lib.mp_finish(self._ctx))

    def are_patches_loaded(self, handle):
        return bool(self._lib.mp_are_patches_loaded(self._ctx, int(handle)))

    def load_patches(self, handle, patch_data, patch_type=BEZIER):
        pts = np.ascontiguousarray(np.asarray(patch_data, dtype=np.float32).reshape(-1, 3))
        self._check(self._lib.mp_load_patches(self._ctx, int(handle), _fptr(pts), pts.shape[0],
                                              int(patch_type)))

    def unload_patches(self, handle):
        self._check(self._lib.mp_unload_patches(self._ctx, int(handle)))

    def draw_patches(self, handle, matrix, projection, color):
        mv = _colmajor(matrix)
        col = np.ascontiguousarray(np.asarray(color, dtype=np.float32).reshape(4))
        pr = projection.as_c()
        self._check(self._lib.mp_draw_patches(self._ctx, int(handle), _fptr(mv), C.byref(pr),
                                              _fptr(col)))

    def set_displacement_amplitude(self, amplitude):
        self._check(self._lib.mp_set_displacement_amplitude(self._ctx, float(amplitude)))

    # -- batching --
    def batch_begin(self, projections):
        arr = (mp_projection * len(projections))(*[p.as_c() for p in projections])
        self._check(self._lib.mp_batch_begin(self._ctx, len(projections), arr))
        self._views = [p.viewport for p in projections]

    def batch_draw(self, view, handle, matrix, color):
        mv = _colmajor(matrix)
        col = np.ascontiguousarray(np.asarray(color, dtype=np.float32).reshape(4))
        self._check(self._lib.mp_batch_draw(self._ctx, int(view), int(handle), _fptr(mv), _fptr(col)))

    def batch_finish(self):
        self._check(self._lib.mp_batch_finish(self._ctx))

    # -- readback --
    def read_framebuffer(self, view=0, rgba32=False):
        w, h = self._views[view]
        out = {
            "rgba8": np.empty((h, w, 4), np.uint8),
            "depth": np.empty((h, w), np.float32),
            "ids": np.empty((h, w), np.uint32),
        }
        p32 = None
        if rgba32:
            out["rgba32"] = np.empty((h, w, 4), np.float32)
            p32 = out["rgba32"].ctypes.data
        self._check(self._lib.mp_batch_read(self._ctx, int(view), out["rgba8"].ctypes.data,
                                            out["depth"].ctypes.data, out["ids"].ctypes.data, p32))
        return out

    def ranges(self):
        n = C.c_int64()
        self._check(self._lib.mp_read_ranges(self._ctx, None, 0, C.byref(n)))
        recs = (mp_range_record * max(1, n.value))()
        self._check(self._lib.mp_read_ranges(self._ctx, recs, n.value, C.byref(n)))
        a = np.frombuffer(recs, dtype=np.dtype([("pid", "<i4"), ("object", "<i4"), ("depth", "<i4"),
                                                ("umin", "<f4"), ("vmin", "<f4"),
                                                ("umax", "<f4"), ("vmax", "<f4")]))
        return a[: n.value].copy()

    def grids(self, first, count):
        """Diced grids of canonical ranges [first, first + count): eye-space
        pos (count, (N+1)^2, 4), sub-pixel pxlpos (count, (N+1)^2, 2) and
        depth (count, (N+1)^2)."""
        G = (self.cfg.reyes_patch_size + 1) ** 2
        pos = np.empty((count, G, 4), np.float32)
        pxl = np.empty((count, G, 2), np.int32)
        dep = np.empty((count, G), np.float32)
        self._check(self._lib.mp_read_grids(self._ctx, int(first), int(count), pos.ctypes.data, pxl.ctypes.data,
                                            dep.ctypes.data))
        return pos, pxl, dep

    def pass_counts(self):
        n = C.c_int32()
        self._check(self._lib.mp_read_pass_counts(self._ctx, None, 0, C.byref(n)))
        arr = (C.c_int32 * max(1, n.value))()
        self._check(self._lib.mp_read_pass_counts(self._ctx, arr, n.value, C.byref(n)))
        return np.array(arr[: n.value], dtype=np.int32)

    def stats(self):
        st = mp_stats()
        self._check(self._lib.mp_get_stats(self._ctx, C.byref(st)))
        return {name: getattr(st, name) for name, _ in mp_stats._fields_}

    # -- differentiable extension (device pointers; see include/micropolis.h) --
    def frame_serial(self):
        """Serial number of the last finished frame or batch (0: none)."""
        s = C.c_uint64()
        self._check(self._lib.mp_frame_serial(self._ctx, C.byref(s)))
        return s.value

    def backward(self, frame_serial, dl_dcolor_ptr, dl_floats, grad_cps_ptr, grad_cps_floats, grad_color_ptr,
                 grad_amp_ptr, n_objects):
        """mp_backward for the frame `frame_serial`; the sizes are the caller's
        buffer sizes in floats (grad_color: 4 per object, grad_amp: 1)."""
        self._check(self._lib.mp_backward(self._ctx, int(frame_serial), dl_dcolor_ptr, int(dl_floats),
                                          grad_cps_ptr, int(grad_cps_floats), grad_color_ptr, grad_amp_ptr,
                                          int(n_objects)))

    def read_device(self, view, rgba32_ptr=None, ids_ptr=None, n_pixels=0):
        """Device copy of a view's RGBA32F colour / ids into buffers of n_pixels pixels."""
        self._check(self._lib.mp_batch_read_device(self._ctx, int(view), rgba32_ptr, ids_ptr, int(n_pixels)))

    def update_patches_device(self, handle, xyz_ptr, n_points):
        self._check(self._lib.mp_update_patches_device(self._ctx, int(handle), xyz_ptr, int(n_points)))

    def measure_atomic_rate(self, n=1 << 25):
        """64-bit atomic-min updates per second to distinct keys (device)."""
        r = C.c_double()
        self._check(self._lib.mp_measure_atomic_rate(self._ctx, int(n), C.byref(r)))
        return r.value

    def replay_stages(self, n):
        ms = (C.c_double * 4)()
        self._check(self._lib.mp_replay_stages(self._ctx, int(n), ms))
        return dict(zip(("split", "order", "raster", "resolve"), list(ms)))

    def dump_trace(self, path):
        """Renderer::dump_trace: stages of the last frame in the reference's
        trace format."""
        self._check(self._lib.mp_dump_trace(self._ctx, str(path).encode()))

    def dump_stats(self, path):
        """Statistics::dump_stats format for the last frame."""
        self._check(self._lib.mp_dump_stats(self._ctx, str(path).encode()))

    def tuning(self):
        t = _lib.mp_tuning()
        self._check(self._lib.mp_get_tuning(self._ctx, C.byref(t)))
        return {"split_engine": t.split_engine, "work_sharing": t.work_sharing,
                "raster_per_cu": t.raster_per_cu, "raster_chunk": t.raster_chunk,
                "split_levels": t.split_levels, "split_budget": t.split_budget,
                "split_waves": t.split_waves}

    def set_tuning(self, split_engine=0, work_sharing=0, raster_per_cu=0, raster_chunk=0, split_levels=0,
                   split_budget=0, split_waves=2):
        t = _lib.mp_tuning(int(split_engine), int(work_sharing), int(raster_per_cu), int(raster_chunk),
                           int(split_levels), int(split_budget), int(split_waves), 0.0)
        self._check(self._lib.mp_set_tuning(self._ctx, C.byref(t)))

    def autotune(self, replays=5):
        """Pick the fastest scheduling of the last LOCAL/BALANCED frame or
        batch (the output does not change)."""
        t = _lib.mp_tuning()
        self._check(self._lib.mp_autotune(self._ctx, int(replays), C.byref(t)))
        return {"split_engine": t.split_engine, "work_sharing": t.work_sharing,
                "raster_per_cu": t.raster_per_cu, "raster_chunk": t.raster_chunk,
                "split_levels": t.split_levels, "split_budget": t.split_budget,
                "split_waves": t.split_waves, "frame_ms": t.frame_ms}

    def replay(self, n, use_graph=False):
        ms = C.c_double()
        self._check(self._lib.mp_replay(self._ctx, int(n), int(bool(use_graph)), C.byref(ms)))
        return ms.value
