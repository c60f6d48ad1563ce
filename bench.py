#!/usr/bin/env python3
"""Frame benchmark of the meshlet raster -> vis-buffer path on one MI355X.

One step = one frame of the seeded 1080p scene with 10,240 grid-patch
meshlets (1,003,520 triangles, 16 materials): clear, frustum cull, mesh +
setup, guard-band clip and raster into the 64-bit vis-buffer, then the
lit resolve. Inputs are resident in HBM before the timed region. The frame is
a batch of one view at the launch shape the batch autotuner picks, launched
as one hipGraph replay or kernel by kernel, whichever a short untimed trial
finds faster. A second measurement renders the 64-view batching workload as
one batch launch (graph and eager) and as per-view launches (views/s, same
vis-buffers).

Prints one JSON line:
 - value: Mtri/s of submitted triangles per frame time; value_processed:
   Mtri/s of the TrianglesProcessed counter (triangles of meshlets that pass
   the cull, Rasterizer.cpp:545); fps;
 - kernels: each kernel's average duration over the timed steps of the same
   frame at the same launch shape, from events the runtime stamps at the
   dispatch's own start and end (hipExtLaunchKernel), with its algorithmic
   bytes, and (from rocprofv3 counter passes in child processes, at the same
   shape) HBM bytes and vector-ALU instructions;
 - roofline: the dominant kernel's algorithmic bytes over its duration
   against the 8 TB/s HBM peak;
 - cpu_baseline: the CPU oracle on a bounded sample of the same frame, on one
   host core and on all the host threads the box grants (row bands).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Vector-ALU issue rate: 256 CUs x 4 SIMDs, a wave64 VALU instruction occupies
# a 16-lane SIMD for 4 cycles at the 2.4 GHz peak clock.
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4


def algorithmic_bytes(stats: dict) -> dict:
    """Bytes each kernel must move per launch at least: every input it needs
    read once and every output written once, from counted quantities of the
    frame (meshlets visible, records, clip entries, bin entries, pixels).
    Data that several work items share (a record in several bins, a vertex
    of several pixels) counts once."""
    m_all, m_vis = stats["meshlets"], stats["visible"]
    S, Cq, N = stats["records"], stats["clipped"], stats["pixels"]
    P = stats.get("bin_entries", 0)
    ncov = stats.get("covered_pixels", 0)
    F = stats["fragments"]
    # per visible meshlet: positions 768 B, indices 384 B, count + material 5 B
    meshlet_in = m_vis * (768 + 384 + 5)
    return {
        "clear": 8 * N,
        # bounds in, survivor list out
        "cull": 16 * m_all + 8 * m_vis,
        # meshlets in, 32-byte records and 64-byte clip entries out
        "mesh_setup": 8 * m_vis + meshlet_in + 32 * S + 64 * Cq,
        "clip": 64 * Cq,
        # records in, covered fragments update one 8-byte key each
        "raster_direct": 32 * S + 8 * F,
        # records in, one 4-byte entry per (record, bin) pair out; the
        # queued clip triangles in (the binned pipeline clips here)
        "bin_scatter": 32 * S + 4 * P + 64 * Cq,
        # entries and records in, every pixel's key written once
        "bin_raster": 4 * P + 32 * S + 8 * N,
        # keys in, RGBA8 out; vertex data (positions, indices, UVs, normals)
        # of the visible meshlets once; one RGBA8 texel per covered pixel
        # (the mip level is chosen for about one texel per pixel)
        "resolve_lit": 8 * N + 4 * N + m_vis * (768 + 384 + 256 + 256) + 4 * ncov,
    }


# profiler name -> kernel symbol substring in rocprofv3 output
KERNEL_SYMBOLS = {
    "clear": "k_clear(", "cull": "k_cull<", "mesh_setup": "k_mesh_setup<", "clip": "k_clip<",
    "bin_scatter": "k_bin_scatter<", "bin_order": "k_bin_order<", "bin_raster": "k_bin_raster<",
    "resolve_lit": "k_resolve<1>",
}


def _rocprof_counters(counters: list, child_args: list, tag: str) -> dict | None:
    """Runs this script as a child under rocprofv3 --pmc <counters> (one pass)
    and returns {kernel: {counter: per-dispatch average}}, or None."""
    import csv
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    out = tempfile.mkdtemp(prefix="glimp_pmc_", dir=os.environ.get("TMPDIR", "/tmp"))
    cmd = ["timeout", "-s", "KILL", "150", exe, "--pmc", *counters, "--output-format", "csv", "-d", out, "-o", tag,
           "--", sys.executable, os.path.abspath(__file__), *child_args]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=170)
    except subprocess.TimeoutExpired:
        return None
    sums, counts = {}, {}
    for root, _, files in os.walk(out):
        for f in files:
            if f.endswith("counter_collection.csv"):
                for row in csv.DictReader(open(os.path.join(root, f))):
                    key = (row["Kernel_Name"], row["Counter_Name"])
                    sums[key] = sums.get(key, 0.0) + float(row["Counter_Value"])
                    counts[key] = counts.get(key, 0) + 1
    shutil.rmtree(out, ignore_errors=True)
    if r.returncode != 0 and not sums:
        return None
    res: dict = {}
    for (k, c), v in sums.items():
        res.setdefault(k, {})[c] = v / counts[(k, c)]
    return res


def _rocprof_counter(counter: str, child_args: list, tag: str) -> dict | None:
    """One counter: {kernel: per-dispatch average}, or None."""
    r = _rocprof_counters([counter], child_args, tag)
    return None if r is None else {k: v[counter] for k, v in r.items() if counter in v}


def _pick(d, sub):
    vals = [v for k, v in d.items() if sub in k]
    return vals[0] if vals else None


# TCC read requests to memory by size; on gfx950 FETCH_SIZE counts every
# 128-byte request as 64 bytes (its formula takes the 128-byte count from
# TCC_BUBBLE, which stays zero), so a copy kernel reads it as half its bytes
READ_REQUESTS = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum")


def _request_bytes(c: dict) -> float:
    return (32 * c.get("TCC_EA0_RDREQ_32B_sum", 0.0) + 64 * c.get("TCC_EA0_RDREQ_64B_sum", 0.0)
            + 128 * c.get("TCC_EA0_RDREQ_128B_sum", 0.0))


def measure_traffic(shape: dict | None) -> dict | None:
    """HBM bytes per launch of each kernel of the frame: FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) from two rocprofv3 runs of a short frame
    loop at the headline launch shape, each scaled by what the counter reads
    for a device copy of known size. A third run counts the read requests by
    size (32, 64 and 128 bytes), which gives the bytes read without a scale;
    both readings are checked on the copy and on a read-only reduction of
    known size."""
    n_bytes = 1 << 28
    frame = ["--steps", "3", "--warmup", "1", "--no-cpu-baseline", "--no-batch", "--no-traffic", "--no-counts"]
    if shape:
        frame += ["--bin-size", str(shape["bin_size"]), "--bin-capacity", str(shape["bin_capacity"]),
                  "--raster-threads", str(shape["raster_threads"])]
    fetch = _rocprof_counter("FETCH_SIZE", frame, "frame")
    write = _rocprof_counter("WRITE_SIZE", frame, "frame")
    reqs = _rocprof_counters(list(READ_REQUESTS), frame, "frame")
    copy_args, read_args = ["--calibrate-copy", str(n_bytes)], ["--calibrate-read", str(n_bytes)]
    cf, cw = _rocprof_counter("FETCH_SIZE", copy_args, "copy"), _rocprof_counter("WRITE_SIZE", copy_args, "copy")
    creq = _rocprof_counters(list(READ_REQUESTS), copy_args, "copy")
    rf, rreq = _rocprof_counter("FETCH_SIZE", read_args, "read"), _rocprof_counters(list(READ_REQUESTS), read_args, "read")
    if not (fetch and write and cf and cw):
        return None
    kf, kw = _pick(cf, "copyBuffer"), _pick(cw, "copyBuffer")
    if not kf or not kw:
        return None
    f_scale = n_bytes / (kf * 1024.0)      # bytes per counted KiB
    w_scale = n_bytes / (kw * 1024.0)
    out = {"fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4)}
    # known bytes over each reading: 1.0 is exact
    checks = {}
    rfv = _pick(rf, "reduce_kernel") if rf else None
    if rfv:
        checks["reduce_fetch_scaled"] = round(n_bytes / (rfv * 1024.0 * f_scale), 4)
    if creq and _pick(creq, "copyBuffer"):
        checks["copy_read_requests"] = round(n_bytes / _request_bytes(_pick(creq, "copyBuffer")), 4)
    if rreq and _pick(rreq, "reduce_kernel"):
        checks["reduce_read_requests"] = round(n_bytes / _request_bytes(_pick(rreq, "reduce_kernel")), 4)
    out["scale_checks"] = checks
    by_req = {}
    for name, sub in KERNEL_SYMBOLS.items():
        fv, wv = _pick(fetch, sub), _pick(write, sub)
        if fv is not None and wv is not None:
            out[name] = int(fv * 1024.0 * f_scale + wv * 1024.0 * w_scale)
            rq = _pick(reqs, sub) if reqs else None
            if rq:
                by_req[name] = int(_request_bytes(rq) + wv * 1024.0)
    out["by_request_size"] = by_req
    # vector-ALU instructions per launch (wave instructions)
    valu = _rocprof_counter("SQ_INSTS_VALU", frame, "frame")
    if valu:
        out["valu"] = {name: int(_pick(valu, sub)) for name, sub in KERNEL_SYMBOLS.items() if _pick(valu, sub)}
    return out


def _valu_roofline(traffic, name: str, avg_us) -> dict | None:
    """The kernel's vector-ALU instructions per launch against the VALU issue
    rate: the time they need at best, and that time over the measured one."""
    if not traffic or not avg_us or name not in traffic.get("valu", {}):
        return None
    insts = traffic["valu"][name]
    bound_us = insts / VALU_PEAK_WAVE_INSTS * 1e6
    return {"wave_insts": insts, "peak_wave_insts_per_s": VALU_PEAK_WAVE_INSTS, "bound_us": round(bound_us, 2),
            "frac": round(bound_us / avg_us, 4)}


def host_threads() -> int:
    """Host threads this process may use: the CPUs it is allowed to run on,
    capped by OMP_NUM_THREADS when that is set (the GPU box grants 16)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(cap))) if cap and cap.isdigit() else n


def cpu_baseline(scene, view, seconds: float) -> dict:
    """The CPU oracle (a scalar C++ restatement of the reference's algorithm,
    -O2 without fast-math) on the same frame: clear + cull + draw + lit
    resolve, on one host core, then on all granted host threads with the
    frame split into row bands (bit-identical to the one-core frame), each for
    about `seconds`."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle
    from glimpsw_amd import scene as S

    b = oracle.SceneBinding(scene)
    lighting = S.default_lighting(view)
    tris = scene.triangles()

    def timed(fn):
        fn()                                   # warm (tables, page faults)
        frames, t0 = 0, time.perf_counter()
        while True:
            fn()
            frames += 1
            el = time.perf_counter() - t0
            if el >= seconds:
                return frames, el

    def one_core():
        fb, _, _ = oracle.render_frame(scene, view, b)
        oracle.resolve_lit(fb, b, view.object_to_clip, lighting)

    f1, e1 = timed(one_core)
    threads = host_threads()
    # one band per thread: every band sets up all visible meshlets, so more
    # bands cost more than the better balance gains
    bands = oracle.row_bands(view.height, threads)
    with ThreadPoolExecutor(threads) as pool:
        fn, en = timed(lambda: oracle.render_frame_threads(scene, view, b, lighting, pool, bands))
    return {
        "value": round(tris * f1 / e1 / 1e6, 3),
        "unit": "Mtri/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{f1} frames of the same scene (clear+cull+draw+lit resolve, {tris} submitted triangles each) "
                  f"in {e1:.1f} s on one host core, scalar C++ oracle",
        "fps": round(f1 / e1, 3),
        "all_cores": {"value": round(tris * fn / en / 1e6, 3), "unit": "Mtri/s", "cores": threads,
                      "fps": round(fn / en, 3),
                      "sample": f"{fn} frames in {en:.1f} s, {len(bands)} row bands over {threads} threads "
                                f"(cull once, then every band's draw and lit resolve; bit-identical to one core)"},
    }


def bench_batch(ctx, stream, steps: int, warmup: int, shape: dict | None = None) -> dict:
    """Batching workload: 64 views over 8 scenes of 1k-50k meshlets, one
    batch launch per step (hipGraph replay or eager, the faster; shape picked
    by the autotuner unless given) against the same views drawn one
    clear+draw call at a time."""
    import torch
    import glimpsw_amd as g
    from glimpsw_amd import scene as S

    scenes, views = S.batch_views()
    dscenes = [g.Scene(ctx, sc) for sc in scenes]
    fbs = [g.Framebuffer(ctx, v.width, v.height) for _, v in views]
    b = g.Batch(ctx, [(dscenes[si], fb, v) for (si, v), fb in zip(views, fbs)], shape=shape)
    t_tune = time.perf_counter()
    tune_ms = None
    if shape is None:
        shape, tune_ms = b.autotune(stream)
    t_tune = time.perf_counter() - t_tune

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    ms_graph = timed(lambda: b.launch(True, stream))
    want = [fb.read() for fb in fbs]
    ms_eager = timed(lambda: b.launch(False, stream))
    same_eager = all((wi == gi).all() for (wi, _), (gi, _) in zip(want, [fb.read() for fb in fbs]))
    ms_batch = min(ms_graph, ms_eager)
    # per-kernel times of the eager batch at the tuned shape
    ctx.profile(True)
    ctx.profile_read(reset=True)
    for _ in range(steps):
        b.launch(False, stream)
    kern = ctx.profile_read(reset=True)
    ctx.profile(False)
    ms_per_view = timed(lambda: b.launch_per_view(stream))
    got = [fb.read() for fb in fbs]
    identical = all((wi == gi).all() and (wd.view(np.uint32) == gd.view(np.uint32)).all()
                    for (wi, wd), (gi, gd) in zip(want, got))
    b.close()
    for fb in fbs:
        fb.close()
    for d in dscenes:
        d.close()
    n = len(views)
    tris = sum(scenes[si].triangles() for si, _ in views)
    return {
        "workload": f"{n} views over {len(scenes)} scenes of {min(s.count for s in scenes)}-"
                    f"{max(s.count for s in scenes)} meshlets, {sum(v.width * v.height for _, v in views)} pixels",
        "views_per_s": round(n / (ms_batch * 1e-3), 1),
        "views_per_s_per_view_launch": round(n / (ms_per_view * 1e-3), 1),
        "speedup": round(ms_per_view / ms_batch, 2),
        "ms_per_batch": round(ms_batch, 4),
        "ms_per_batch_graph": round(ms_graph, 4),
        "ms_per_batch_eager": round(ms_eager, 4),
        "ms_per_batch_per_view_launch": round(ms_per_view, 4),
        "kernels_eager_us": {k: round(t / max(1, c) * 1e3, 2) for k, (c, t) in kern.items()},
        "submitted_tris": tris,
        "Mtri_per_s": round(tris / (ms_batch * 1e-3) / 1e6, 1),
        "autotuned_shape": shape,
        "autotune_ms": round(tune_ms, 4) if tune_ms is not None else None,
        "autotune_wall_s": round(t_tune, 2),
        "identical_vis_buffers": bool(identical and same_eager),
    }


def bench_clip_frame(ctx, stream, shape: dict, steps: int, warmup: int) -> dict:
    """The 1080p frame with its near-plane crossers reaching into the frustum
    (frame_1080p(crossers_inside=True): about ten times the clipped triangles
    of the headline frame, each covering much of the screen), at the headline
    launch shape: the frame time, the clipped-triangle count and the bin
    scatter's duration, which includes the clipping."""
    import torch
    import glimpsw_amd as g
    from glimpsw_amd import scene as S

    scene, view = S.frame_1080p(crossers_inside=True)
    ds = g.Scene(ctx, scene)
    fb = g.Framebuffer(ctx, view.width, view.height)
    planes = view.frustum_planes()
    b = g.Batch(ctx, [(ds, fb, view, planes, 0)], shape=shape)
    b.set_resolve([(view.object_to_clip, g.GLIMP_RESOLVE_LIT, S.default_lighting(view))])
    for _ in range(warmup):
        b.launch(True, stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.launch(True, stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ctx.profile(True)
    ctx.profile_read(reset=True)
    for _ in range(steps):
        b.launch(False, stream)
    kern = ctx.profile_read(reset=True)
    ctx.profile(False)
    ctx.reset_counters()
    fb.clear(0, 0.0, stream)
    fb.draw(ds, view.object_to_clip, view.meshlet_count, 0, planes, g.GLIMP_DRAW_COUNTERS, stream,
            shape["bin_size"], shape["bin_capacity"])
    c = ctx.counters()
    b.close()
    fb.close()
    ds.close()
    return {"workload": "the 1080p frame with its near-plane crossers 0.5-0.7 of their size off the view axis",
            "ms_per_frame_graph": round(ms, 4), "Mtri_per_s": round(scene.triangles() / (ms * 1e-3) / 1e6, 1),
            "clipped": int(c["triangles_clipped"]), "dropped": int(c["dropped"]),
            "fragments_per_pixel": round(c["fragments_covered"] / (view.width * view.height), 2),
            "kernels_eager_us": {k: round(t / max(1, n) * 1e3, 2) for k, (n, t) in kern.items()}}


def _calibrate(kind: str, n_bytes: int) -> None:
    """Known-bytes kernels for the HBM counter calibration (10 launches each):
    copy (a device copy that reads and writes n_bytes) and read (an int64 sum
    that reads n_bytes)."""
    import torch

    if kind == "copy":
        src = torch.empty(n_bytes // 4, dtype=torch.int32, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        fn = lambda: dst.copy_(src)                                    # noqa: E731
    else:
        src = torch.empty(n_bytes // 8, dtype=torch.int64, device="cuda").fill_(1)
        fn = lambda: src.sum()                                         # noqa: E731
    torch.cuda.synchronize()
    for _ in range(10):
        fn()
    torch.cuda.synchronize()
    print(json.dumps({"calibrate": kind, "bytes": n_bytes, "launches": 10}))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="seconds of each CPU baseline leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--direct", action="store_true", help="global-atomic raster instead of tile bins")
    ap.add_argument("--bin-size", type=int, default=0, help="screen bin side (8/16/32/64), 0 = autotuned")
    ap.add_argument("--bin-capacity", type=int, default=0, help="records per bin, 0 = autotuned")
    ap.add_argument("--raster-threads", type=int, default=0, help="bin raster workgroup (128/256/512/1024)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-view batching measurement")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--no-counts", action="store_true", help="skip the counting pass (frame kernels only)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 64-view batching measurement (its JSON object)")
    ap.add_argument("--batch-shape", default="", metavar="BIN,CAPACITY,THREADS",
                    help="launch shape of the batching measurement instead of the autotuner's")
    for kind in ("copy", "read"):
        ap.add_argument(f"--calibrate-{kind}", type=int, default=0, metavar="BYTES",
                        help=f"only run the {kind} calibration kernel over BYTES bytes 10 times")
    args = ap.parse_args()
    for kind in ("copy", "read"):
        if getattr(args, f"calibrate_{kind}"):
            _calibrate(kind, getattr(args, f"calibrate_{kind}"))
            return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; the path has no multi-GPU mode "
                                   "(views or meshlets would shard without exchange except a min-reduce "
                                   "of vis-buffer keys)", "n_gpus": args.gpus}))
        return 2

    if args.batch_only:
        import torch
        import glimpsw_amd as g

        torch.cuda.init()
        ctx = g.Context(0)
        bshape = None
        if args.batch_shape:
            bs, cap, th = (int(x) for x in args.batch_shape.split(","))
            bshape = {"bin_size": bs, "bin_capacity": cap, "raster_threads": th}
        print(json.dumps(bench_batch(ctx, torch.cuda.current_stream(), max(5, args.steps // 5), 3, bshape)))
        ctx.close()
        return 0

    import torch
    import glimpsw_amd as g
    from glimpsw_amd import scene as S

    t_gen = time.perf_counter()
    scene, view = S.frame_1080p()
    t_gen = time.perf_counter() - t_gen
    torch.cuda.init()
    stream = torch.cuda.current_stream()
    ctx = g.Context(0)
    ds = g.Scene(ctx, scene)
    fb = g.Framebuffer(ctx, view.width, view.height)
    planes = view.frustum_planes()
    flags = g.GLIMP_DRAW_DIRECT if args.direct else 0
    lighting = S.default_lighting(view)

    def run(fn, steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    # The frame (clear, cull, draw, lit resolve) as a batch of one view with
    # its resolve: replayed as one captured graph or launched kernel by
    # kernel. Vis-buffer and colour equal the separate API calls
    # (tests/test_gpu_batch.py).
    batch, shape = None, None
    if not args.direct:
        shape = {"bin_size": args.bin_size or 16, "bin_capacity": args.bin_capacity or 1024,
                 "raster_threads": args.raster_threads or 256}
        batch = g.Batch(ctx, [(ds, fb, view, planes, flags)], shape=shape)
        # the autotuner times the whole frame, resolve included
        batch.set_resolve([(view.object_to_clip, g.GLIMP_RESOLVE_LIT, lighting)])
        if not args.bin_size and not args.bin_capacity and not args.raster_threads:
            shape, _ = batch.autotune(stream)

    def frame_calls():
        # the reference's call sequence: Clear, DrawMeshlets (with the cull), Resolve
        fb.clear(0, 0.0, stream)
        fb.draw(ds, view.object_to_clip, view.meshlet_count, 0, planes, flags, stream,
                shape["bin_size"] if shape else args.bin_size, shape["bin_capacity"] if shape else args.bin_capacity)
        fb.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_LIT, lighting, stream)

    frame, frame_mode, trial = frame_calls, "separate calls", {}
    if batch is not None:
        modes = {"graph": lambda: batch.launch(True, stream), "eager": lambda: batch.launch(False, stream)}
        for name, fn in modes.items():
            for _ in range(args.warmup):
                fn()
            trial[name] = round(run(fn, 20), 4)
        best = min(trial, key=trial.get)
        frame = modes[best]
        frame_mode = f"one {'graph replay' if best == 'graph' else 'eager batch launch'} of clear+cull+draw+resolve"
    for _ in range(args.warmup):
        frame()
    ms = run(frame, args.steps)                    # the headline: no events, no counters

    # per-kernel durations of the same frame at the same shape: the eager
    # launch of the same kernels with dispatch-stamped events
    per_kernel_fn = (lambda: batch.launch(False, stream)) if batch is not None else frame_calls
    ctx.profile(True)
    ctx.profile_read(reset=True)
    ms_profiled = run(per_kernel_fn, args.steps)
    kern = ctx.profile_read(reset=True)
    ctx.profile(False)
    ms_calls = run(frame_calls, args.steps)

    # The lit resolve alone: `steps` back-to-back launches over the vis-buffer
    # the frame left, between two events on the stream it runs on. The
    # per-dispatch events above stamp each dispatch with extra packets that
    # add a few microseconds to every kernel; this loop does not.
    resolve_us = None
    if not args.direct:
        resolve = lambda: fb.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_LIT, lighting, stream)  # noqa: E731
        for _ in range(args.warmup):
            resolve()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            resolve()
        e1.record(stream)
        e1.synchronize()
        resolve_us = e0.elapsed_time(e1) * 1e3 / args.steps

    if args.no_counts:
        print(json.dumps({"ms_per_step": round(ms, 4), "shape": shape,
                          "kernels_us": {k: round(t / max(1, c) * 1e3, 2) for k, (c, t) in kern.items()}}))
        if batch is not None:
            batch.close()
        fb.close()
        ds.close()
        ctx.close()
        return 0

    # counted quantities of one frame at the same shape (separate, untimed run)
    ctx.reset_counters()
    fb.clear(0, 0.0, stream)
    fb.draw(ds, view.object_to_clip, view.meshlet_count, 0, planes, flags | g.GLIMP_DRAW_COUNTERS, stream,
            shape["bin_size"] if shape else 0, shape["bin_capacity"] if shape else 0)
    c = ctx.counters()
    _, visible = ds.cull(planes, 0, view.meshlet_count)
    stats = {
        "meshlets": scene.count, "visible": visible, "tris_visible": int(c["triangles_processed"]),
        "records": int(c["triangles_rasterized"]), "clipped": int(c["triangles_clipped"]),
        "fragments": int(c["fragments_covered"]), "pixels": view.width * view.height,
        "bin_entries": int(c["bin_entries"]), "bins_rasterized": int(c["bins_rasterized"]),
        "tile_lanes": int(c["tile_lanes"]), "dropped": int(c["dropped"]),
    }
    if stats["tile_lanes"]:
        stats["lane_utilisation"] = round(stats["fragments"] / stats["tile_lanes"], 4)
    _, depth = fb.read()
    stats["covered_pixels"] = int((depth > 0).sum())
    alg = algorithmic_bytes(stats)

    # counter passes in child processes at the headline shape (this process
    # keeps its context but is idle meanwhile)
    traffic = None if args.no_traffic else measure_traffic(shape)

    per_kernel = {}
    for name, (launches, total) in kern.items():
        avg_us = total / max(1, launches) * 1e3
        b = alg.get(name, 0)
        per_kernel[name] = {"avg_us": round(avg_us, 2), "launches": launches, "alg_bytes": b,
                            "GBps": round(b / (avg_us * 1e-6) / 1e9, 1) if avg_us > 0 and b else None,
                            "hbm_bytes": traffic.get(name) if traffic else None,
                            "hbm_bytes_by_request": traffic["by_request_size"].get(name) if traffic else None,
                            "valu": _valu_roofline(traffic, name, avg_us)}
    dom = max(kern.items(), key=lambda kv: kv[1][1])[0]
    dk = per_kernel[dom]
    # the roofline's duration: the dominant kernel's own back-to-back loop
    # where it has one (the lit resolve), else its dispatch-stamped average
    roof_us, roof_timing = dk["avg_us"], ("dispatch-stamped events (hipExtLaunchKernel) over the timed steps of "
                                          "the eager frame at the headline shape")
    if dom == "resolve_lit" and resolve_us:
        roof_us = resolve_us
        roof_timing = (f"{args.steps} back-to-back launches of the frame's lit resolve between two events on its "
                       "stream")
    roof_gbps = dk["alg_bytes"] / (roof_us * 1e-6) / 1e9 if dk["alg_bytes"] else None
    tris = scene.triangles()
    result = {
        "metric": "Mtri/s at 1080p/1M tris (frame: clear+cull+mesh/setup+clip+bin+raster into vis-buffer+lit resolve)",
        "value": round(tris / (ms * 1e-3) / 1e6, 2),
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "fps": round(1e3 / ms, 1),
        "value_processed": round(stats["tris_visible"] / (ms * 1e-3) / 1e6, 2),
        "value_definitions": {"value": "submitted triangles / frame time",
                              "value_processed": "TrianglesProcessed (triangles of meshlets passing the cull, "
                                                 "Rasterizer.cpp:545) / frame time"},
        "frame": frame_mode,
        "frame_shape": shape,
        "frame_trial_ms": trial,
        "ms_per_step_profiled": round(ms_profiled, 4),
        "ms_per_step_separate_calls": round(ms_calls, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+fp32",
        "data": "synthetic",
        "config": {"workload": "10,240 grid-patch meshlets (1,003,520 tris), 1920x1080, depth complexity ~3, "
                               "16 materials 2048^2 RGBA8",
                   "raster": "direct" if args.direct else (
                       f"binned {shape['bin_size']}x{shape['bin_size']}, {shape['bin_capacity']} "
                       f"records per bin, {shape['raster_threads']}-thread bin raster"),
                   "submitted_tris": tris},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(roof_gbps, 1) if roof_gbps else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(roof_gbps / HBM_PEAK_GBS, 4) if roof_gbps else None,
                     "traffic": traffic.get(dom) if traffic else None,
                     "traffic_by_request": traffic["by_request_size"].get(dom) if traffic else None,
                     "valu": _valu_roofline(traffic, dom, roof_us),
                     "alg_bytes_per_launch": dk["alg_bytes"], "avg_us": round(roof_us, 2),
                     "avg_us_dispatch_events": dk["avg_us"], "timing": roof_timing},
        "traffic_calibration": ({k: traffic[k] for k in ("fetch_scale", "write_scale", "scale_checks")}
                                if traffic else None),
        "kernels": per_kernel,
        "frame_counts": stats,
        "scene_gen_s": round(t_gen, 2),
    }
    if shape is not None:
        result["clip_frame"] = bench_clip_frame(ctx, stream, shape, max(5, args.steps // 5), 3)
    if not args.no_batch:
        result["batch"] = bench_batch(ctx, stream, max(5, args.steps // 5), 3)
    result["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline(scene, view, args.cpu_seconds)
    print(json.dumps(result))
    if batch is not None:
        batch.close()
    fb.close()
    ds.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
