"""Benchmark: plain 1080p REYES frame of the 4096-patch displaced torus.

Headline workload: 4096 bicubic Bezier patches on a 64x64 torus (R=1,
r=0.35), 3D value-noise displacement (4 octaves, amplitude 0.02), 1920x1080,
LOCAL bound-and-split, bound_n_split_limit 16, reyes_patch_size 16.

A step is one whole frame with the patches resident in HBM: persistent
bound-and-split, canonical ordering, fused dice/shade/sample, resolve to
RGBA8 + FP32 depth + visible ids (the frame's kernels replayed as one hipGraph).

Beside the headline the same JSON line reports, each measured in this run:
  api_frame   the headline frame through the API (mp_prepare .. mp_finish,
              host sync and counter readback included), median of 20
  c1          the 64-patch sphere (1024x768, BREADTH: host loop, one sync per
              split level), median of 20 API frames
  c4          16384 Gregory patches, depth complexity ~32 (1920x1080): the
              reference's multipass form (BOUNDED batches of
              reyes_patches_per_pass, one host sync per batch) against the
              persistent BALANCED work queue replayed as one hipGraph, with
              identical output checked
  batch       64 uneven scenes (640x480): 64 plain frames against one batched
              frame, both autotuned and graph-replayed
  adjoint     32 views (960x540): forward against forward + hand-written adjoint
  cpu_baseline the oracle (a C port of the reference kernels) on the host's
              cores and on one core, full frames of the headline workload
Prints one JSON line.
"""
import argparse
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MI355X_CUS = 256          # compute units
MI355X_CLOCK_HZ = 2.4e9   # peak engine clock
TUNING_KEYS = ("split_engine", "work_sharing", "split_levels", "split_budget", "split_waves", "raster_per_cu",
               "raster_chunk")


def draw_frame(r, v):
    """One frame through the API: prepare, draw every object, finish."""
    r.prepare()
    for o in v.objects:
        if not r.are_patches_loaded(o.handle):
            r.load_patches(o.handle, o.patches, o.patch_type)
        r.draw_patches(o.handle, o.matrix, v.projection, o.color)
    r.finish()


def api_frames(r, v, n):
    """Median host wall time (ms) of n API frames (mp_prepare .. mp_finish,
    which waits for the frame) and median device time of the same frames
    (HIP events inside mp_finish, stats frame_ms)."""
    wall, dev = [], []
    for _ in range(n):
        t = time.perf_counter()
        draw_frame(r, v)
        wall.append((time.perf_counter() - t) * 1e3)
        dev.append(r.stats()["frame_ms"])
    return float(np.median(wall)), float(np.median(dev))


def frame_roofline(st, patches, ncp, W, H, ms, atomic_rate):
    """The whole frame against the HBM + 64-bit atomic roofline: compulsory
    bytes = control points once (ncp x 12 B per patch) + per pixel of the
    padded frame key clear 8 + key read 8 + RGBA8 4 + depth 4 + range records
    in and out of the split (2 x 20 B per bound); atomics = coverage samples
    at the measured rate of 64-bit atomic minimum to distinct keys."""
    H_pad = (H + 7) // 8 * 8
    W_pad = (W + 7) // 8 * 8
    frame_bytes = patches * ncp * 12 + W_pad * H_pad * (8 + 8) + W * H * (4 + 4) + 2 * 20 * st["bounded_patches"]
    t_roof = max(frame_bytes / (HBM_PEAK_GBS * 1e9), st["coverage_samples"] / atomic_rate)
    return {"alg_bytes": int(frame_bytes), "atomics": int(st["coverage_samples"]),
            "t_roof_us": round(t_roof * 1e6, 2), "t_frame_us": round(ms * 1e3, 2),
            "frac": round(t_roof / (ms * 1e-3), 5)}


def frame_valu(profile, ms):
    """The frame's compute floor: its kernels' VALU wave instructions
    (rocprofv3 SQ_INSTS_VALU of the profiling command, recorded in the profile
    as 'frame_valu_wave_instructions: N'; not measured in this run) at one per
    CU per cycle (256 CUs, 2.4 GHz), against a frame of ms milliseconds."""
    path = os.path.join(ROOT, profile)
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("frame_valu_wave_instructions:"):
            n = int(ln.split(":")[1])
            t = n / (MI355X_CUS * MI355X_CLOCK_HZ)
            return {"wave_instructions": n, "issue_rate_per_s": MI355X_CUS * MI355X_CLOCK_HZ,
                    "t_valu_us": round(t * 1e6, 2), "frac": round(t / (ms * 1e-3), 4),
                    "source": profile + " (rocprofv3 SQ_INSTS_VALU of the frame's kernels)"}
    return None


def measure_c1(atomic_rate):
    """The 64-patch sphere, BREADTH (host-driven level loop), API frames."""
    from micropolis_amd import Renderer, scenes

    w = scenes.c1_sphere()
    v = w.views[0]
    r = Renderer(**w.options)
    draw_frame(r, v)
    st = r.stats()
    wall, dev = api_frames(r, v, 20)
    r.close()
    print("c1: %.3f ms per API frame (device %.3f ms), %d ranges" % (wall, dev, st["diced_ranges"]),
          file=sys.stderr, flush=True)
    W, H = v.projection.viewport
    return {"workload": w.name, "method": "BREADTH", "ranges": int(st["diced_ranges"]),
            "dropped_ranges": int(st["dropped_ranges"]), "micropolygons": int(st["micropolygons"]),
            "api_frame_ms": round(wall, 4), "device_frame_ms": round(dev, 4),
            "mmicropolys_per_s": round(st["micropolygons"] / (wall * 1e-3) / 1e6, 3),
            "frame_roofline": frame_roofline(st, w.patch_count, 16, W, H, wall, atomic_rate)}


def measure_c4(steps, atomic_rate, autotune=True):
    """16384 Gregory patches on 32 stacked planes: the multipass BOUNDED form
    (batches of reyes_patches_per_pass, a host sync per batch) against the
    persistent BALANCED work queue replayed as one hipGraph."""
    from micropolis_amd import Renderer, scenes

    w = scenes.c4_gregory()
    v = w.views[0]
    W, H = v.projection.viewport
    r = Renderer(**w.options)
    draw_frame(r, v)
    st = r.stats()
    tuning = r.autotune(3) if autotune else r.tuning()
    r.replay(1, use_graph=True)
    graph_ms = r.replay(steps, use_graph=True)
    api_wall, api_dev = api_frames(r, v, 5)
    ids_b = r.read_framebuffer()["ids"]
    r.close()
    rb = Renderer(**dict(w.options, bound_n_split_method="BOUNDED"))
    draw_frame(rb, v)
    stb = rb.stats()
    bw, bd = api_frames(rb, v, 5)
    ids_m = rb.read_framebuffer()["ids"]
    rb.close()
    print("c4: BALANCED graph %.3f ms (API %.3f ms), BOUNDED multipass %.3f ms (%d passes)" % (
        graph_ms, api_wall, bw, stb["pass_count"]), file=sys.stderr, flush=True)
    return {"workload": w.name, "ranges": int(st["diced_ranges"]), "micropolygons": int(st["micropolygons"]),
            "coverage_samples": int(st["coverage_samples"]),
            "persistent_balanced_graph_ms": round(graph_ms, 4), "persistent_balanced_api_ms": round(api_wall, 4),
            "multipass_bounded_api_ms": round(bw, 4), "multipass_passes": int(stb["pass_count"]),
            "speedup": round(bw / graph_ms, 3),
            "mmicropolys_per_s": round(st["micropolygons"] / (graph_ms * 1e-3) / 1e6, 3),
            "same_output": bool(np.array_equal(ids_b, ids_m)) and int(st["micropolygons"]) == int(stb["micropolygons"]),
            "tuning": {k: tuning[k] for k in TUNING_KEYS},
            "frame_roofline": dict(frame_roofline(st, w.patch_count, 20, W, H, graph_ms, atomic_rate),
                                   valu=frame_valu("profiles/frame_kernels_c4_gregory_16384p_1080p.md", graph_ms))}


def measure_batch(steps, use_graph):
    """64 uneven scenes (640x480, log-uniform 100..20000 random patches): one
    batched frame against 64 plain frames, each autotuned, both replayed
    device-resident."""
    from micropolis_amd import Renderer, scenes

    w = scenes.c3_batch()
    r = Renderer(**w.options)
    for v in w.views:
        o = v.objects[0]
        r.load_patches(o.handle, o.patches, o.patch_type)
    plain_ms = 0.0
    mpolys = 0
    for v in w.views:
        o = v.objects[0]
        r.prepare()
        r.draw_patches(o.handle, o.matrix, v.projection, o.color)
        r.finish()
        mpolys += r.stats()["micropolygons"]
        if use_graph:
            r.autotune(3)
        r.replay(1, use_graph=use_graph)
        plain_ms += r.replay(steps, use_graph=use_graph)
    r.batch_begin([v.projection for v in w.views])
    for vi, v in enumerate(w.views):
        o = v.objects[0]
        r.batch_draw(vi, o.handle, o.matrix, o.color)
    r.batch_finish()
    st = r.stats()
    tuning = r.autotune(3) if use_graph else r.tuning()
    r.replay(1, use_graph=use_graph)
    batch_ms = r.replay(steps, use_graph=use_graph)
    stages = {k: round(v, 4) for k, v in r.replay_stages(3).items()}
    r.close()
    print("batch: 64 plain frames %.3f ms, one batched frame %.3f ms" % (plain_ms, batch_ms),
          file=sys.stderr, flush=True)
    return {"workload": w.name, "scenes": len(w.views), "patches": w.patch_count,
            "micropolygons": int(st["micropolygons"]), "plain_64_frames_ms": round(plain_ms, 4),
            "plain_frames_autotuned": bool(use_graph),
            "batched_ms": round(batch_ms, 4), "speedup": round(plain_ms / batch_ms, 3),
            "batched_tuning": {k: tuning[k] for k in TUNING_KEYS},
            "batched_stages_ms": stages, "batched_ranges": int(st["diced_ranges"]),
            "same_micropolygons": int(mpolys) == int(st["micropolygons"])}


def run_adjoint_child(steps):
    """The adjoint measurement uses torch buffers; it runs in a child process
    that initialises torch first."""
    import subprocess
    res = subprocess.run([sys.executable, os.path.abspath(__file__), "--adjoint-only", str(steps)],
                         capture_output=True, text=True, timeout=600)
    sys.stderr.write(res.stderr[-2000:])
    if res.returncode != 0:
        return {"error": "adjoint child exited with %d" % res.returncode}
    return json.loads(res.stdout.strip().splitlines()[-1])


def measure_adjoint(steps):
    """32 camera views of the displaced torus at 960x540 in one batch: forward
    frame vs forward + hand-written adjoint (gradients w.r.t. all control
    points, amplitude and albedo; visibility fixed; grids recomputed)."""
    import torch
    from micropolis_amd import Renderer, scenes

    w = scenes.c5_views()
    r = Renderer(**dict(w.options, write_rgba32=1))
    base = w.views[0].objects[0]
    r.load_patches(base.handle, base.patches, base.patch_type)
    r.batch_begin([v.projection for v in w.views])
    for vi, v in enumerate(w.views):
        r.batch_draw(vi, base.handle, v.objects[0].matrix, base.color)
    r.batch_finish()
    r.replay(1, use_graph=True)
    fwd_ms = r.replay(steps, use_graph=True)
    V = len(w.views)
    H, W = w.views[0].projection.viewport[1], w.views[0].projection.viewport[0]
    dl = torch.randn((V, H, W, 4), device="cuda") * 0.1
    g_cps = torch.empty((V,) + base.patches.shape, device="cuda")
    g_col = torch.empty((V, 4), device="cuda")
    g_amp = torch.empty((V,), device="cuda")
    torch.cuda.synchronize()
    bwd = []
    serial = r.frame_serial()
    for _ in range(steps + 1):
        r.backward(serial, dl.data_ptr(), dl.numel(), g_cps.data_ptr(), g_cps.numel(), g_col.data_ptr(),
                   g_amp.data_ptr(), V)
        bwd.append(r.stats()["backward_ms"])
    bwd_ms = float(np.median(bwd[1:]))
    r.close()
    print("adjoint: forward %.4f ms, backward %.4f ms for %d views (backward runs %s)" % (fwd_ms, bwd_ms, V, bwd),
          file=sys.stderr, flush=True)
    return {"workload": w.name, "views": V, "forward_ms": round(fwd_ms, 4), "backward_ms": round(bwd_ms, 4),
            "forward_ms_per_view": round(fwd_ms / V, 4),
            "fwd_plus_adjoint_ms_per_view": round((fwd_ms + bwd_ms) / V, 4),
            "ratio_vs_forward": round((fwd_ms + bwd_ms) / fwd_ms, 3)}


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Threads for the all-cores CPU baseline: the cores this process may use,
    at most 16 (the GPU box's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def measure_cpu_baseline(w, micropolys):
    """The oracle (C port of the reference kernels) on full frames of the
    headline workload: 2 warm-up frames, then the median of 5, on all the
    host threads and on one."""
    from oracle import pyoracle

    out = {}
    for threads in (host_threads(), 1):
        times = []
        for k in range(7):
            t = time.perf_counter()
            ref = pyoracle.render_threaded(w, threads)[0]
            if k >= 2:
                times.append(time.perf_counter() - t)
        assert len(ref["ranges"]) * w.options["reyes_patch_size"] ** 2 == micropolys
        out[threads] = float(np.median(times))
        print("cpu baseline: %d thread(s) %.3f s/frame" % (threads, out[threads]), file=sys.stderr, flush=True)
    T = host_threads()
    return {"value": round(micropolys / out[T] / 1e6, 4), "unit": "Mmicropolys/s", "cores": T, "kind": "port",
            "sample": "full frames of the headline workload through the C oracle on %d host thread(s) "
                      "(split by patch, raster by range, per-thread key buffers min-reduced): median of 5 "
                      "after 2 warm-ups, %.3f s/frame" % (T, out[T]),
            "frame_s": round(out[T], 4),
            "one_thread": {"value": round(micropolys / out[1] / 1e6, 4), "cores": 1, "frame_s": round(out[1], 4)},
            "cpu": cpu_model(), "host_cpus": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default scheduling (profiling runs: one kernel configuration)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint measurement")
    ap.add_argument("--no-configs", action="store_true", help="skip the sphere and Gregory configurations")
    ap.add_argument("--adjoint-only", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 256 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": a.numel() * 4,
                          "bytes_written": b.numel() * 4, "launches": 3}))
        return
    if args.adjoint_only:
        # torch (the gradient buffers) must initialise HIP before the library
        import torch
        torch.cuda.init()
        print(json.dumps(measure_adjoint(args.adjoint_only)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only; this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from micropolis_amd import Renderer, scenes

    w = scenes.c2_torus()
    v = w.views[0]
    r = Renderer(**w.options)
    draw_frame(r, v)
    st = r.stats()
    use_graph = not args.no_graph
    # pick the fastest scheduling of this frame (output unchanged)
    tuning = r.autotune(5) if use_graph and not args.no_autotune else r.tuning()
    print("autotuned scheduling: %s" % tuning, file=sys.stderr, flush=True)
    print("first frame: %d ranges, %d micropolygons, %d coverage samples" % (
        st["diced_ranges"], st["micropolygons"], st["coverage_samples"]), file=sys.stderr, flush=True)

    if args.warmup:
        r.replay(args.warmup, use_graph=use_graph)
    t0 = time.perf_counter()
    ms_dev = r.replay(args.steps, use_graph=use_graph)
    t1 = time.perf_counter()
    ms_wall = (t1 - t0) * 1e3 / args.steps
    stages = r.replay_stages(max(3, min(args.steps, 20)))
    print("frame %.3f ms (device %.3f ms); stages %s" % (ms_wall, ms_dev, stages),
          file=sys.stderr, flush=True)

    mpolys = st["micropolygons"]
    value = mpolys / (ms_wall * 1e-3) / 1e6

    # roofline of the dominant kernel (fused dice + shade + sample): algorithmic
    # bytes = range records (24 B) + patch control points once (16 x 16 B) +
    # per-micropolygon colour term (4 B) + one 8-byte key atomic per sample;
    # its duration from HIP events around it on the library's stream
    N = w.options["reyes_patch_size"]
    P = w.patch_count
    alg_bytes = st["diced_ranges"] * (24 + 4 * N * N) + P * 16 * 16 + st["coverage_samples"] * 8
    raster_ms = stages["raster"]
    achieved = alg_bytes / (raster_ms * 1e-3) / 1e9

    W, H = v.projection.viewport
    atomic_rate = r.measure_atomic_rate(1 << 25)
    fr = frame_roofline(st, P, 16, W, H, ms_wall, atomic_rate)
    fr["atomic_rate_per_s"] = round(atomic_rate, 1)
    api_wall, api_dev = api_frames(r, v, 20)
    print("api frame: %.3f ms (device %.3f ms)" % (api_wall, api_dev), file=sys.stderr, flush=True)
    r.close()

    # HBM traffic and VALU issue of the raster kernel: rocprofv3 counters of
    # the profiling command recorded in the profile (not measured in this run)
    traffic = None
    issue = None
    prof_name = "profiles/k_raster16_c2_torus_1080p.md"
    prof = os.path.join(ROOT, prof_name)
    valu = frame_valu("profiles/frame_kernels_c2_torus_1080p.md", ms_wall)
    if valu:
        fr["valu"] = valu
    if os.path.exists(prof):
        for ln in open(prof):
            if ln.startswith("traffic_bytes_per_launch:"):
                traffic = int(ln.split(":")[1])
            if ln.startswith("VALU issue:"):
                # "VALU issue: I wave instructions against S issue slots"
                f = ln.replace(",", "").split()
                issue = {"valu_wave_instructions": int(f[2]), "issue_slots": int(f[6])}
        if issue:
            issue["frac"] = round(issue["valu_wave_instructions"] / issue["issue_slots"], 3)

    configs = {}
    if not args.no_configs:
        configs["c1"] = measure_c1(atomic_rate)
        configs["c4"] = measure_c4(max(3, min(args.steps, 10)), atomic_rate, not args.no_autotune)
    batch = None if args.no_batch else measure_batch(max(3, min(args.steps, 10)), use_graph)
    adjoint = None if args.no_adjoint else run_adjoint_child(max(3, min(args.steps, 10)))
    cpu = None if args.no_cpu_baseline else measure_cpu_baseline(w, mpolys)

    line = {
        "metric": "1080p frame ms + Mmicropolys/s (plain LOCAL frame, displaced torus)",
        "value": round(value, 3),
        "unit": "Mmicropolys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_wall, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded torus patches, value-noise displacement)",
        "config": {"workload": w.name, "patches": P, "width": W, "height": H,
                   "reyes_patch_size": N, "bound_n_split_limit": 16, "method": "LOCAL",
                   "graph": use_graph, "tuning": {k: tuning[k] for k in TUNING_KEYS},
                   "ranges": st["diced_ranges"], "micropolygons": mpolys,
                   "coverage_samples": st["coverage_samples"]},
        "stages_ms": {k: round(x, 4) for k, x in stages.items()},
        "api_frame": {"ms": round(api_wall, 4), "device_ms": round(api_dev, 4),
                      "mmicropolys_per_s": round(mpolys / (api_wall * 1e-3) / 1e6, 3),
                      "what": "mp_prepare..mp_finish host wall time, median of 20 (device: HIP events in mp_finish)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": prof_name + " (rocprofv3 FETCH_SIZE / WRITE_SIZE of the profiling command)",
                     "kernel": "k_raster<16>", "alg_bytes": int(alg_bytes), "kernel_ms": round(raster_ms, 4),
                     # the kernel's actual limit: VALU issue (rocprofv3 counters of
                     # the same command, one wave64 instruction per CU per cycle)
                     "valu_issue": issue},
        "frame_roofline": fr,
        "cpu_baseline": cpu,
        "configs": configs,
        "batch": batch,
        "adjoint": adjoint,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
