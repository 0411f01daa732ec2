"""Trace reader against the reference's own committed trace: the per-stage
totals of one HD 7970 frame (tests/golden/traces/README.md)."""
import os

import pytest

from micropolis_amd import trace

HERE = os.path.join(os.path.dirname(__file__), "golden", "traces")


def test_reference_trace_totals():
    ev = trace.load(os.path.join(HERE, "HD7970-host-bns.trace"))
    summ, span = trace.summary(ev)
    assert summ["frame@host"]["total_ms"] == pytest.approx(5.075, abs=5e-4)
    assert summ["dice@rasterization"]["count"] == 20
    assert summ["dice@rasterization"]["total_ms"] == pytest.approx(0.760, abs=5e-4)
    assert summ["shade@rasterization"]["total_ms"] == pytest.approx(0.515, abs=5e-4)
    assert summ["sample@rasterization"]["total_ms"] == pytest.approx(1.698, abs=5e-4)
    host_bns = [v for k, v in summ.items() if k.startswith("CPU bound & split")]
    assert len(host_bns) == 1 and host_bns[0]["total_ms"] == pytest.approx(3.673, abs=5e-4)
    assert span >= 5.075


def test_new_format_with_ids_and_dependencies():
    text = ("bound & split@device:100:100:100:1100:0:\n"
            "accumulate@device:1100:1100:1100:1600:1:0\n"
            "sample@device:1600:1600:1600:9600:2:1|0\n")
    ev = trace.parse(text)
    assert [e.index for e in ev] == [0, 1, 2] and ev[2].deps == [1, 0]
    cats = trace.by_category(ev)
    assert cats["subdivision"] == pytest.approx(0.0015) and cats["sample"] == pytest.approx(0.008)
    with pytest.raises(ValueError):
        trace.parse("no colon here\n")
