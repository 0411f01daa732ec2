"""ctypes binding of ``libmicropolis_hip.so`` (declared in include/micropolis.h).

The library is loaded from this package directory only; a missing library is
an error (there is no CPU fallback on the product path).
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmicropolis_hip.so")

MP_OK = 0
MP_BNS_CPU, MP_BNS_BOUNDED, MP_BNS_LOCAL, MP_BNS_BREADTH, MP_BNS_BALANCED = range(5)
MP_BEZIER, MP_GREGORY = 0, 1

METHODS = {"CPU": MP_BNS_CPU, "BOUNDED": MP_BNS_BOUNDED, "LOCAL": MP_BNS_LOCAL,
           "BREADTH": MP_BNS_BREADTH, "BALANCED": MP_BNS_BALANCED}


class mp_config(C.Structure):
    _fields_ = [
        ("window_width", C.c_int32), ("window_height", C.c_int32),
        ("clear_color", C.c_float * 4),
        ("reyes_patch_size", C.c_int32),
        ("bound_n_split_limit", C.c_float),
        ("cull_ribbon", C.c_float),
        ("backface_culling", C.c_int32),
        ("bound_n_split_method", C.c_int32),
        ("bound_sample_rate", C.c_int32),
        ("max_split_depth", C.c_int32),
        ("reyes_patches_per_pass", C.c_int32),
        ("framebuffer_tile_size", C.c_int32),
        ("subpixel_bits", C.c_int32),
        ("dice_group_width", C.c_int32),
        ("bns_pipeline_length", C.c_int32),
        ("local_bns_work_groups", C.c_int32),
        ("displacement", C.c_int32),
        ("displacement_amplitude", C.c_float),
        ("displacement_frequency", C.c_float),
        ("displacement_seed", C.c_uint32),
        ("displacement_octaves", C.c_int32),
        ("dummy_render", C.c_int32),
        ("write_rgba32", C.c_int32),
    ]


class mp_projection(C.Structure):
    _fields_ = [("fovy_deg", C.c_float), ("near_plane", C.c_float),
                ("width", C.c_int32), ("height", C.c_int32)]


class mp_stats(C.Structure):
    _fields_ = [
        ("patches_per_frame", C.c_int64), ("diced_ranges", C.c_int64),
        ("dropped_ranges", C.c_int64), ("bounded_patches", C.c_int64),
        ("micropolygons", C.c_int64), ("pass_count", C.c_int64),
        ("max_patches", C.c_int64), ("slow_blocks", C.c_int64),
        ("overflow_relaunches", C.c_int64), ("coverage_samples", C.c_int64),
        ("backward_ms", C.c_double), ("frame_ms", C.c_double),
        ("sorted_ranges", C.c_int64),
    ]


class mp_tuning(C.Structure):
    _fields_ = [("split_engine", C.c_int32), ("work_sharing", C.c_int32),
                ("raster_per_cu", C.c_int32), ("raster_chunk", C.c_int32),
                ("split_levels", C.c_int32), ("split_budget", C.c_int32), ("split_waves", C.c_int32),
                ("frame_ms", C.c_double)]


class mp_range_record(C.Structure):
    _fields_ = [("pid", C.c_int32), ("object", C.c_int32), ("depth", C.c_int32),
                ("umin", C.c_float), ("vmin", C.c_float),
                ("umax", C.c_float), ("vmax", C.c_float)]


# (name, restype, argtypes) of every entry point in include/micropolis.h
_P = C.c_void_p
_FP = C.POINTER(C.c_float)
EXPORTS = [
    ("mp_config_defaults", C.c_int, [C.POINTER(mp_config)]),
    ("mp_create", _P, [C.POINTER(mp_config)]),
    ("mp_create_error", C.c_char_p, []),
    ("mp_destroy", None, [_P]),
    ("mp_last_error", C.c_char_p, [_P]),
    ("mp_are_patches_loaded", C.c_int, [_P, C.c_uint64]),
    ("mp_load_patches", C.c_int, [_P, C.c_uint64, _FP, C.c_size_t, C.c_int32]),
    ("mp_unload_patches", C.c_int, [_P, C.c_uint64]),
    ("mp_prepare", C.c_int, [_P]),
    ("mp_draw_patches", C.c_int, [_P, C.c_uint64, _FP, C.POINTER(mp_projection), _FP]),
    ("mp_finish", C.c_int, [_P]),
    ("mp_read_framebuffer", C.c_int, [_P, _P, _P, _P, _P]),
    ("mp_read_ranges", C.c_int, [_P, C.POINTER(mp_range_record), C.c_int64, C.POINTER(C.c_int64)]),
    ("mp_read_grids", C.c_int, [_P, C.c_int64, C.c_int64, _P, _P, _P]),
    ("mp_read_pass_counts", C.c_int, [_P, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32)]),
    ("mp_get_stats", C.c_int, [_P, C.POINTER(mp_stats)]),
    ("mp_set_displacement_amplitude", C.c_int, [_P, C.c_float]),
    ("mp_batch_begin", C.c_int, [_P, C.c_int32, C.POINTER(mp_projection)]),
    ("mp_batch_draw", C.c_int, [_P, C.c_int32, C.c_uint64, _FP, _FP]),
    ("mp_batch_finish", C.c_int, [_P]),
    ("mp_batch_read", C.c_int, [_P, C.c_int32, _P, _P, _P, _P]),
    ("mp_replay", C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    ("mp_replay_stages", C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
    ("mp_measure_atomic_rate", C.c_int, [_P, C.c_uint32, C.POINTER(C.c_double)]),
    ("mp_display_thresholds", C.c_int, [C.POINTER(C.c_float)]),
    ("mp_backward", C.c_int, [_P, C.c_uint64, _P, C.c_size_t, _P, C.c_size_t, _P, _P, C.c_int32]),
    ("mp_frame_serial", C.c_int, [_P, C.POINTER(C.c_uint64)]),
    ("mp_batch_read_device", C.c_int, [_P, C.c_int32, _P, _P, C.c_size_t]),
    ("mp_update_patches_device", C.c_int, [_P, C.c_uint64, _P, C.c_size_t]),
    ("mp_dump_trace", C.c_int, [_P, C.c_char_p]),
    ("mp_dump_stats", C.c_int, [_P, C.c_char_p]),
    ("mp_get_tuning", C.c_int, [_P, C.POINTER(mp_tuning)]),
    ("mp_set_tuning", C.c_int, [_P, C.POINTER(mp_tuning)]),
    ("mp_autotune", C.c_int, [_P, C.c_int32, C.POINTER(mp_tuning)]),
    ("mp_render_noalloc", C.c_int, [C.POINTER(mp_config), _P, C.c_size_t, C.c_int32, _FP,
                                    C.POINTER(mp_projection), _FP, _P, _P, _P, _P]),
    ("mp_cpu_bound_n_split", C.c_int, [C.POINTER(mp_config), _FP, C.c_size_t, C.c_int32, _FP,
                                       C.POINTER(mp_projection), C.POINTER(mp_range_record), C.c_int64,
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_int32,
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
]

_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libmicropolis_hip.so is not built; run micropolis_amd/build.py "
            "(or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def default_config():
    cfg = mp_config()
    load().mp_config_defaults(C.byref(cfg))
    return cfg
