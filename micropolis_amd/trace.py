"""Reader and summary of trace files in the reference's format.

One event per line, `name@queue:queued:submit:start:end[:id:dependencies]`
with nanosecond timestamps (src/CL/Device.cpp:275-289; older traces omit the
id and dependency fields).  `mp_dump_trace` writes this format for the
library's stages, named after the reference's stage categories so that its
tools/trace_summary.py groups them.
"""
import re
from collections import defaultdict
from dataclasses import dataclass, field

_LINE = re.compile(r"^([^@:]*)@([^@:]*):(\d+):(\d+):(\d+):(\d+)(?::(\d*):([0-9|]*))?\s*$")

# stage categories (the grouping of tools/trace_summary.py)
CATEGORY = {
    "sample": "sample", "dice": "dice", "shade": "shade",
    "bound & split": "subdivision", "bound patches": "subdivision", "split patches": "subdivision",
    "reduce": "subdivision", "accumulate": "subdivision", "init patch ranges": "subdivision",
    "initialize counter buffers": "subdivision", "initialize range buffers": "subdivision",
    "Clear out_range_cnt": "subdivision", "read range count": "subdivision",
    "read processed counts": "subdivision", "initialize projection buffer": "subdivision",
    "buffer map": "subdivision", "clear framebuffer": "clear", "clear depthbuffer": "clear",
}


@dataclass
class Event:
    name: str
    queue: str
    queued: int
    submit: int
    start: int
    end: int
    index: int = 0
    deps: list = field(default_factory=list)

    @property
    def duration_ms(self):
        return (self.end - self.start) / 1e6


def parse(text):
    events = []
    for n, line in enumerate(text.splitlines(), 1):
        if not line.strip():
            continue
        m = _LINE.match(line)
        if not m:
            raise ValueError("trace line %d does not match name@queue:queued:submit:start:end" % n)
        idx = int(m.group(7)) if m.group(7) else 0
        deps = [int(d) for d in m.group(8).split("|") if d] if m.group(8) else []
        events.append(Event(m.group(1), m.group(2), int(m.group(3)), int(m.group(4)), int(m.group(5)),
                            int(m.group(6)), idx, deps))
    return events


def load(path):
    with open(path) as f:
        return parse(f.read())


def summary(events):
    """Per event name: count, total and mean ms; plus the time range."""
    by = defaultdict(list)
    for e in events:
        by[(e.name, e.queue)].append(e.duration_ms)
    out = {"%s@%s" % k: {"count": len(v), "total_ms": sum(v), "mean_ms": sum(v) / len(v)}
           for k, v in by.items()}
    t0 = min(e.start for e in events)
    t1 = max(e.end for e in events)
    return out, (t1 - t0) / 1e6


def by_category(events):
    """Device time per stage category (unknown names under their own name)."""
    acc = defaultdict(float)
    for e in events:
        if e.queue == "host" and e.name == "frame":
            continue
        acc[CATEGORY.get(e.name, e.name)] += e.duration_ms
    return dict(acc)
