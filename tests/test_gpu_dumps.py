"""Trace and statistics files of a rendered frame (mp_dump_trace,
mp_dump_stats) in the reference's formats."""
import re

import pytest

from micropolis_amd import scenes, trace
from gpu_util import render_plain

pytestmark = pytest.mark.gpu


def test_trace_and_stats_files(tmp_path):
    w = scenes.c2_torus(width=640, height=360)
    print("case c2_torus 640x360 dumps", flush=True)
    r, got = render_plain(w)
    tp, sp = tmp_path / "frame.trace", tmp_path / "frame.stats"
    r.dump_trace(tp)
    r.dump_stats(sp)
    ev = trace.load(tp)
    assert [e.name for e in ev] == ["bound & split", "accumulate", "sample", "clear framebuffer"]
    assert all(e.end >= e.start for e in ev)
    assert all(ev[k].start >= ev[k - 1].end for k in range(1, len(ev)))
    assert ev[1].deps == [0] and ev[3].deps == [2]
    cats = trace.by_category(ev)
    assert set(cats) == {"subdivision", "sample", "clear"} and cats["sample"] > 0
    st = r.stats()
    kv = dict(re.findall(r"^([\w@& ]+) = ([^;]*);", sp.read_text(), re.M))
    assert int(kv["patches_per_frame"]) == st["patches_per_frame"]
    assert int(kv["processed_patches_per_frame"]) == st["bounded_patches"]
    assert int(kv["pass_count"]) == st["pass_count"]
    assert int(kv["total_input_patches"]) == 4096
    parts = [int(v) for k, v in kv.items() if k.startswith("opencl_mem@")]
    assert int(kv["opencl_mem"]) == sum(parts) and int(kv["opencl_mem@patch data"]) == 4096 * 16 * 16
    r.close()
