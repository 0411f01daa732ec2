"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/micropolis.h declares, and its defaults are reyes.options."""
import ctypes as C
import os
import re

from micropolis_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "micropolis.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mp_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    names = header_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(n for n, _, _ in _lib.EXPORTS) == names


def test_struct_sizes_match_header_layout(tmp_path):
    """ctypes mirrors against the C compiler's view of the header."""
    import subprocess
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include "micropolis.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(mp_config), '
                   'sizeof(mp_projection), sizeof(mp_range_record), sizeof(mp_stats), '
                   'sizeof(mp_tuning));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert sizes == [C.sizeof(_lib.mp_config), C.sizeof(_lib.mp_projection),
                     C.sizeof(_lib.mp_range_record), C.sizeof(_lib.mp_stats),
                     C.sizeof(_lib.mp_tuning)]


def test_defaults_are_reyes_options():
    cfg = _lib.default_config()
    assert (cfg.window_width, cfg.window_height) == (1600, 1200)
    assert cfg.reyes_patch_size == 8 and cfg.bound_n_split_limit == 8.0
    assert cfg.cull_ribbon == 32.0 and cfg.backface_culling == 1
    assert cfg.bound_n_split_method == _lib.MP_BNS_BOUNDED
    assert cfg.bound_sample_rate == 3 and cfg.max_split_depth == 23
    assert cfg.reyes_patches_per_pass == 100000 and cfg.subpixel_bits == 6
    assert cfg.local_bns_work_groups == 128


def test_create_rejects_bad_config_without_touching_the_gpu():
    lib = _lib.load()
    cfg = _lib.default_config()
    cfg.reyes_patch_size = 12
    assert not lib.mp_create(C.byref(cfg))
    assert b"reyes_patch_size" in lib.mp_create_error()
    cfg = _lib.default_config()
    cfg.bound_n_split_method = _lib.MP_BNS_CPU
    assert not lib.mp_create(C.byref(cfg))
