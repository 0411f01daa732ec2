"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

``render(workload)`` composes the oracle's per-range functions into a frame
with the same frame semantics as the HIP library: objects of a view in draw
order, each object's DRAW ranges in canonical (pid, umin, vmin) order, quad
id = view-local range index * N*N + v*N + u, smallest (depth, id) key wins.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

CULL, DRAW, HSPLIT, VSPLIT, CPU_DROP = 0, 1, 2, 6, 8
METHODS = {"CPU": 0, "BOUNDED": 1, "LOCAL": 2, "BREADTH": 3, "BALANCED": 4}
EMPTY_KEY = 0xFFFFFFFFFFFFFFFF

RANGE_DTYPE = np.dtype([("pid", "<i4"), ("depth", "<i4"), ("umin", "<f4"), ("vmin", "<f4"),
                        ("umax", "<f4"), ("vmax", "<f4")])


class or_view(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("proj", C.c_float * 16),
                ("screen", C.c_float * 4), ("f", C.c_float * 2), ("near_", C.c_float),
                ("far_", C.c_float), ("eps", C.c_float), ("clear_linear", C.c_float * 4)]


class or_params(C.Structure):
    _fields_ = [("patch_size", C.c_int32), ("max_split_depth", C.c_int32),
                ("bound_sample_rate", C.c_int32), ("cull_ribbon", C.c_float),
                ("backface_culling", C.c_int32), ("subpixel_bits", C.c_int32),
                ("patches_per_pass", C.c_int32), ("displacement", C.c_int32),
                ("disp_amplitude", C.c_float), ("disp_frequency", C.c_float),
                ("disp_seed", C.c_uint32), ("disp_octaves", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


_lib = None
_P = C.c_void_p
_FP = C.POINTER(C.c_float)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "or_eval_patch": (None, [_FP, C.c_int, C.c_int, C.c_float, C.c_float, _FP]),
        "or_eval_patch_f64": (None, [_P, C.c_int, C.c_int, C.c_double, C.c_double, _P]),
        "or_noise": (C.c_float, [_P, C.c_float, C.c_float, C.c_float]),
        "or_displaced_point": (None, [_FP, C.c_int, C.c_int, C.c_float, C.c_float, _P, _FP]),
        "or_bound": (C.c_int, [_FP, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                               C.c_float, _FP, _P, _P, C.c_float]),
        "or_eval_patch_cpu": (None, [_FP, C.c_int, C.c_int, C.c_float, C.c_float, _FP]),
        "or_projection_bound": (C.c_int, [_FP, _FP, _P, C.c_float, _FP]),
        "or_bound_cpu": (C.c_int, [_FP, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                   C.c_float, _FP, _P, _P, C.c_float]),
        "or_split": (C.c_long, [C.c_int, _FP, C.c_int, C.c_int, _FP, _P, _P, C.c_float, _P,
                                C.c_long, _P, C.c_int, _P, _P, _P]),
        "or_split_local_emulated": (C.c_long, [_FP, C.c_int, C.c_int, _FP, _P, _P, C.c_float,
                                               C.c_int, C.c_int, C.c_long, _P, C.c_long, _P,
                                               C.c_int, _P, _P]),
        "or_canonical_sort": (None, [_P, C.c_long]),
        "or_canonical_key": (C.c_uint64, [_P]),
        "or_dice": (None, [_FP, C.c_int, _P, _FP, _P, _P, _P, _P, _P]),
        "or_shade": (None, [_P, _P, _P, _P, _P, _P]),
        "or_sample": (None, [_P, _P, _P, _P, _P, C.c_uint32, _P]),
        "or_render": (None, [_FP, C.c_int, _P, C.c_long, C.c_uint32, _FP, _P, _P, _P, _P]),
        "or_resolve": (None, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
        "or_make_view": (None, [C.c_float, C.c_float, C.c_int, C.c_int, _FP, _P]),
        "or_framebuffer_index": (C.c_int, [C.c_int, C.c_int, C.c_int]),
        "or_framebuffer_size": (C.c_long, [C.c_int, C.c_int]),
        "or_to_srgb8": (C.c_uint8, [C.c_float]),
        "or_prefix_sum": (C.c_int32, [_P, _P, C.c_long]),
        "or_glibc_rand": (None, [C.c_uint32, _P, C.c_long]),
        "or_adjoint_f64": (C.c_double, [_FP, C.c_int, _P, C.c_long, _FP, _P, C.c_double, _P, _P, _P,
                                        C.c_int, C.c_int, _P, _P, _P, _P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def fptr(a):
    return a.ctypes.data_as(_FP)


def as_cps(patches):
    """(P, C, 3) control points -> (P, C, 4) float32 with w = 1 (PatchIndex.cpp:86-91)."""
    p = np.asarray(patches, np.float32)
    out = np.ones(p.shape[:-1] + (4,), np.float32)
    out[..., :3] = p
    return np.ascontiguousarray(out)


def colmajor(m):
    m = np.asarray(m, np.float32)
    return np.ascontiguousarray(m.T.reshape(16)) if m.shape == (4, 4) else np.ascontiguousarray(m)


def make_view(fovy, near, width, height, clear=(0.1, 0.1, 0.7, 0.0)):
    v = or_view()
    cc = np.asarray(clear, np.float32)
    load().or_make_view(fovy, near, width, height, fptr(cc), C.byref(v))
    return v


def make_params(opts=None, **kw):
    """or_params from reyes.options names (defaults of reyes.options)."""
    o = dict(reyes_patch_size=8, max_split_depth=23, bound_sample_rate=3, cull_ribbon=32.0,
             backface_culling=1, subpixel_bits=6, reyes_patches_per_pass=100000, displacement=0,
             displacement_amplitude=0.02, displacement_frequency=4.0, displacement_seed=1,
             displacement_octaves=4)
    o.update(opts or {})
    o.update(kw)
    p = or_params()
    p.patch_size = int(o["reyes_patch_size"])
    p.max_split_depth = int(o["max_split_depth"])
    p.bound_sample_rate = int(o["bound_sample_rate"])
    p.cull_ribbon = float(o["cull_ribbon"])
    p.backface_culling = int(o["backface_culling"])
    p.subpixel_bits = int(o["subpixel_bits"])
    p.patches_per_pass = int(o["reyes_patches_per_pass"])
    p.displacement = int(o["displacement"])
    p.disp_amplitude = float(o["displacement_amplitude"])
    p.disp_frequency = float(o["displacement_frequency"])
    p.disp_seed = int(o["displacement_seed"])
    p.disp_octaves = int(o["displacement_octaves"])
    return p


def split(method, cps, ptype, mv, view, prm, limit, cap=None):
    """-> (ranges in emission order, pass_counts, n_bounds, n_dropped)"""
    lib = load()
    P = cps.shape[0]
    cap = cap or max(1 << 16, 64 * P)
    while True:
        out = np.zeros(cap, RANGE_DTYPE)
        pc = np.zeros(4096, np.int32)
        npass, nb, nd = C.c_int32(), C.c_long(), C.c_long()
        mvc = colmajor(mv)
        m = METHODS[method] if isinstance(method, str) else int(method)
        n = lib.or_split(m, fptr(cps), ptype, P, fptr(mvc), C.byref(view), C.byref(prm),
                         float(limit), out.ctypes.data, cap, pc.ctypes.data, 4096,
                         C.byref(npass), C.byref(nb), C.byref(nd))
        if n >= 0:
            return out[:n].copy(), pc[: min(npass.value, 4096)].copy(), nb.value, nd.value
        cap *= 4


def canonical_sort(r):
    r = np.ascontiguousarray(r.copy())
    load().or_canonical_sort(r.ctypes.data, r.shape[0])
    return r


def defaults_from_options(opts):
    d = dict(bound_n_split_method="BOUNDED", bound_n_split_limit=8.0, clear_color=(0.1, 0.1, 0.7, 0.0))
    d.update(opts)
    return d


def render(workload, method=None, want_keys=False):
    """Render every view of a scenes.Workload; returns a list of per-view dicts."""
    lib = load()
    opts = defaults_from_options(workload.options)
    prm = make_params(opts)
    N = prm.patch_size
    meth = method or opts["bound_n_split_method"]
    limit = float(opts["bound_n_split_limit"])
    results = []
    for v in workload.views:
        pr = v.projection
        W, H = pr.viewport
        view = make_view(pr.fovy, pr.near, W, H, opts["clear_color"])
        fb = np.full(lib.or_framebuffer_size(W, H), EMPTY_KEY, np.uint64)
        all_ranges, obj_of_range, colors, passes = [], [], [], []
        n_bounds = n_dropped = 0
        base = 0
        quad_chunks = []
        for oi, ob in enumerate(v.objects):
            cps = as_cps(ob.patches)
            r, pc, nb, nd = split(meth, cps, ob.patch_type, ob.matrix, view, prm, limit)
            n_bounds += nb
            n_dropped += nd
            passes.extend(pc.tolist())
            r = canonical_sort(r)
            quad = np.zeros(max(1, (base + len(r))) * N * N, np.float32)
            mvc = colmajor(ob.matrix)
            if len(r):
                lib.or_render(fptr(cps), ob.patch_type, r.ctypes.data, len(r), base, fptr(mvc),
                              C.byref(view), C.byref(prm), fb.ctypes.data, quad.ctypes.data)
            quad_chunks.append(quad[base * N * N:(base + len(r)) * N * N])
            all_ranges.append(r)
            obj_of_range.append(np.full(len(r), oi, np.int32))
            colors.append(np.repeat(ob.color.reshape(1, 4), len(r), axis=0))
            base += len(r)
        quad_store = np.concatenate(quad_chunks) if quad_chunks else np.zeros(1, np.float32)
        quad_store = np.ascontiguousarray(quad_store if quad_store.size else np.zeros(1, np.float32))
        rc = np.ascontiguousarray(np.concatenate(colors).astype(np.float32)) if base else np.zeros((1, 4), np.float32)
        out = {
            "rgba8": np.empty((H, W, 4), np.uint8), "depth": np.empty((H, W), np.float32),
            "ids": np.empty((H, W), np.uint32), "rgba32": np.empty((H, W, 4), np.float32),
        }
        lib.or_resolve(fb.ctypes.data, quad_store.ctypes.data, rc.ctypes.data, C.byref(view),
                       C.byref(prm), out["rgba8"].ctypes.data, out["depth"].ctypes.data,
                       out["ids"].ctypes.data, out["rgba32"].ctypes.data)
        out["ranges"] = np.concatenate(all_ranges) if all_ranges else np.zeros(0, RANGE_DTYPE)
        out["objects"] = np.concatenate(obj_of_range) if obj_of_range else np.zeros(0, np.int32)
        out["pass_counts"] = np.array(passes, np.int32)
        out["n_bounds"] = n_bounds
        out["n_dropped"] = n_dropped
        out["quad_store"] = quad_store[: base * N * N]
        if want_keys:
            out["keys"] = fb
        results.append(out)
    return results


def adjoint_f64(patches, ranges, mv, prm, amplitude, albedo, ids, target_rgb, grads=True):
    """FP64 loss (and gradients) with visibility fixed to `ids`.
    Returns (loss, grad_cps (P,16,3), grad_albedo (3,), grad_amp)."""
    lib = load()
    xyz = np.ascontiguousarray(np.asarray(patches, np.float32).reshape(-1, 16, 3))
    r = np.ascontiguousarray(ranges)
    ids = np.ascontiguousarray(ids, np.uint32)
    H, W = ids.shape
    tgt = np.ascontiguousarray(np.asarray(target_rgb, np.float32).reshape(H, W, 3))
    alb = np.ascontiguousarray(np.asarray(albedo, np.float64).reshape(3))
    g_cps = np.zeros(xyz.shape, np.float64) if grads else None
    g_alb = np.zeros(3, np.float64) if grads else None
    g_amp = np.zeros(1, np.float64) if grads else None
    mvc = colmajor(mv)
    loss = lib.or_adjoint_f64(fptr(xyz), xyz.shape[0], r.ctypes.data, len(r), fptr(mvc), C.byref(prm),
                              float(amplitude), alb.ctypes.data, ids.ctypes.data, tgt.ctypes.data, W, H,
                              g_cps.ctypes.data if grads else None, g_alb.ctypes.data if grads else None,
                              g_amp.ctypes.data if grads else None, None)
    if not grads:
        return loss
    return loss, g_cps, g_alb, float(g_amp[0])


def render_threaded(workload, threads):
    """render() on `threads` host threads (LOCAL / BALANCED workloads: the
    DRAW set is the union of the per-patch sets).  Each thread splits every
    T-th patch of an object and rasterises a contiguous slice
    of the canonical ranges into its own key buffer; the buffers are reduced
    with a minimum (the visibility keys are order independent) and resolved
    once.  Returns the same per-view dicts as render() (no pass counts)."""
    from concurrent.futures import ThreadPoolExecutor

    lib = load()
    opts = defaults_from_options(workload.options)
    meth = opts["bound_n_split_method"]
    if meth not in ("LOCAL", "BALANCED"):
        raise ValueError("render_threaded splits LOCAL / BALANCED workloads")
    prm = make_params(opts)
    N = prm.patch_size
    limit = float(opts["bound_n_split_limit"])
    T = max(1, int(threads))
    results = []
    with ThreadPoolExecutor(max_workers=T) as pool:
        for v in workload.views:
            pr = v.projection
            W, H = pr.viewport
            view = make_view(pr.fovy, pr.near, W, H, opts["clear_color"])
            all_ranges, colors, n_bounds = [], [], 0
            for ob in v.objects:
                cps = as_cps(ob.patches)
                P = cps.shape[0]

                def split_slice(t, cps=cps, ob=ob, P=P):
                    # patches t, t + T, ...: the deep split trees of neighbouring
                    # patches spread over the threads
                    idx = np.arange(t, P, T)
                    if idx.size == 0:
                        return np.zeros(0, RANGE_DTYPE), 0
                    sub = np.ascontiguousarray(cps[idx])
                    r, _, nb, _ = split(meth, sub, ob.patch_type, ob.matrix, view, prm, limit)
                    r["pid"] = idx[r["pid"]]
                    return r, nb
                parts = list(pool.map(split_slice, range(T)))
                r = canonical_sort(np.concatenate([p[0] for p in parts]))
                n_bounds += sum(p[1] for p in parts)
                all_ranges.append((ob, cps, r))
                colors.append(np.repeat(ob.color.reshape(1, 4), len(r), axis=0))
            total = sum(len(x[2]) for x in all_ranges)
            quad = np.zeros(max(1, total) * N * N, np.float32)
            size = lib.or_framebuffer_size(W, H)
            jobs = []
            base = 0
            for ob, cps, r in all_ranges:
                cuts = np.linspace(0, len(r), T + 1).astype(int)
                for a, b in zip(cuts[:-1], cuts[1:]):
                    if b > a:
                        jobs.append((ob, cps, r, a, b, base))
                base += len(r)

            def raster(job):
                ob, cps, r, a, b, base = job
                fb = np.full(size, EMPTY_KEY, np.uint64)
                chunk = np.ascontiguousarray(r[a:b])
                lib.or_render(fptr(cps), ob.patch_type, chunk.ctypes.data, b - a, base + a,
                              fptr(colmajor(ob.matrix)), C.byref(view), C.byref(prm), fb.ctypes.data,
                              quad.ctypes.data)
                return fb
            fbs = list(pool.map(raster, jobs))
            fb = np.minimum.reduce(fbs) if fbs else np.full(size, EMPTY_KEY, np.uint64)
            fb = np.ascontiguousarray(fb)
            rc = np.ascontiguousarray(np.concatenate(colors).astype(np.float32)) if total else \
                np.zeros((1, 4), np.float32)
            out = {"rgba8": np.empty((H, W, 4), np.uint8), "depth": np.empty((H, W), np.float32),
                   "ids": np.empty((H, W), np.uint32), "rgba32": np.empty((H, W, 4), np.float32)}
            lib.or_resolve(fb.ctypes.data, quad.ctypes.data, rc.ctypes.data, C.byref(view), C.byref(prm),
                           out["rgba8"].ctypes.data, out["depth"].ctypes.data, out["ids"].ctypes.data,
                           out["rgba32"].ctypes.data)
            out["ranges"] = np.concatenate([x[2] for x in all_ranges]) if all_ranges else np.zeros(0, RANGE_DTYPE)
            out["n_bounds"] = n_bounds
            results.append(out)
    return results
