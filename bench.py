#!/usr/bin/env python3
"""Frame benchmark of the meshlet raster -> vis-buffer path on one MI355X.

One step = one frame of the seeded 1080p scene with 10,240 grid-patch
meshlets (1,003,520 triangles, 16 materials): clear, frustum cull, mesh +
setup, guard-band clip and raster into the 64-bit vis-buffer, then the
lit resolve. Inputs are resident in HBM before the timed region; the frame is
launched as a batch of one view, replayed as one hipGraph or launched kernel
by kernel, whichever a short untimed trial finds faster. A second measurement
renders the 64-view batching workload as one batch launch (graph and eager)
and as per-view launches (views/s, same vis-buffers).

Prints one JSON line: Mtri/s (submitted triangles per second) and fps, the
roofline of the dominant kernel from HIP-event kernel times taken inside the
timed region, and the CPU oracle timed on a bounded sample of the same frame.

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
    m_all, m_vis, t_vis = stats["meshlets"], stats["visible"], stats["tris_visible"]
    S, Cq, F, N = stats["records"], stats["clipped"], stats["fragments"], stats["pixels"]
    P = stats.get("bin_entries", 0)
    ncov = stats.get("covered_pixels", 0)
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


# profiler scope name -> kernel symbol substring in rocprofv3 output
KERNEL_SYMBOLS = {
    "clear": "k_clear(", "cull": "k_cull<", "mesh_setup": "k_mesh_setup<", "clip": "k_clip<",
    "bin_scatter": "k_bin_scatter<", "bin_raster": "k_bin_raster<",
    "resolve_lit": "k_resolve<1>",
}


def _rocprof_counter(counter: str, child_args: list, tag: str) -> dict | None:
    """Runs this script as a child under rocprofv3 --pmc <counter> and returns
    the per-dispatch average of the counter for every kernel, or None."""
    import csv
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    out = tempfile.mkdtemp(prefix="glimp_pmc_", dir=os.environ.get("TMPDIR", "/tmp"))
    cmd = ["timeout", "-s", "KILL", "150", exe, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", tag,
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
                    k = row["Kernel_Name"]
                    sums[k] = sums.get(k, 0.0) + float(row["Counter_Value"])
                    counts[k] = counts.get(k, 0) + 1
    shutil.rmtree(out, ignore_errors=True)
    if r.returncode != 0 and not sums:
        return None
    return {k: sums[k] / counts[k] for k in sums}


def measure_traffic() -> dict | None:
    """HBM bytes per launch of each kernel of the frame: FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) from two rocprofv3 runs of a short frame
    loop, each scaled by what the counter reads for a device copy of known
    size."""
    copy_bytes = 1 << 28
    # the separate-call frames' launch shape (the library default) for the
    # batch frames too, so that each kernel has one instantiation
    frame = ["--steps", "3", "--warmup", "1", "--no-cpu-baseline", "--no-batch", "--no-traffic", "--no-counts",
             "--bin-size", "16", "--bin-capacity", "1024", "--raster-threads", "256"]
    copy = ["--calibrate-copy", str(copy_bytes)]
    fetch = _rocprof_counter("FETCH_SIZE", frame, "frame")
    write = _rocprof_counter("WRITE_SIZE", frame, "frame")
    cf = _rocprof_counter("FETCH_SIZE", copy, "copy")
    cw = _rocprof_counter("WRITE_SIZE", copy, "copy")
    if not (fetch and write and cf and cw):
        return None

    def pick(d, sub):
        vals = [v for k, v in d.items() if sub in k]
        return vals[0] if vals else None

    kf, kw = pick(cf, "copyBuffer"), pick(cw, "copyBuffer")
    if not kf or not kw:
        return None
    f_scale = copy_bytes / (kf * 1024.0)      # bytes per counted KiB
    w_scale = copy_bytes / (kw * 1024.0)
    out = {"fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4)}
    for name, sub in KERNEL_SYMBOLS.items():
        fv, wv = pick(fetch, sub), pick(write, sub)
        if fv is not None and wv is not None:
            out[name] = int(fv * 1024.0 * f_scale + wv * 1024.0 * w_scale)
    # vector-ALU instructions per launch (wave instructions): the bound these
    # kernels actually meet
    valu = _rocprof_counter("SQ_INSTS_VALU", frame, "frame")
    if valu:
        out["valu"] = {name: int(pick(valu, sub)) for name, sub in KERNEL_SYMBOLS.items() if pick(valu, sub)}
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


def cpu_baseline(scene, view, seconds: float) -> dict:
    import oracle
    from glimpsw_amd import scene as S

    b = oracle.SceneBinding(scene)
    frames = 0
    t0 = time.perf_counter()
    lighting = S.default_lighting(view)
    while True:
        fb, ctr, _ = oracle.render_frame(scene, view, b)
        oracle.resolve_lit(fb, b, view.object_to_clip, lighting)
        frames += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    tris = scene.triangles()
    return {
        "value": round(tris * frames / el / 1e6, 3),
        "unit": "Mtri/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{frames} frames of the same scene (clear+cull+draw+lit resolve, {tris} submitted triangles each) "
                  f"in {el:.1f} s on one host core, scalar C++ oracle",
        "fps": round(frames / el, 3),
    }


def bench_batch(ctx, stream, steps: int, warmup: int) -> dict:
    """Batching workload: 64 views over 8 scenes of 1k-50k meshlets, one
    batch launch per step (hipGraph replay or eager, the faster; shape picked
    by the autotuner) against the same views drawn one clear+draw call at a
    time."""
    import torch
    import glimpsw_amd as g
    from glimpsw_amd import scene as S

    scenes, views = S.batch_views()
    dscenes = [g.Scene(ctx, sc) for sc in scenes]
    fbs = [g.Framebuffer(ctx, v.width, v.height) for _, v in views]
    b = g.Batch(ctx, [(dscenes[si], fb, v) for (si, v), fb in zip(views, fbs)])
    shape, tune_ms = b.autotune(stream)

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
        "submitted_tris": tris,
        "Mtri_per_s": round(tris / (ms_batch * 1e-3) / 1e6, 1),
        "autotuned_shape": shape,
        "autotune_ms": round(tune_ms, 4),
        "identical_vis_buffers": bool(identical and same_eager),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--direct", action="store_true", help="global-atomic raster instead of tile bins")
    ap.add_argument("--bin-size", type=int, default=0, help="screen bin side (8/16/32/64), 0 = library default")
    ap.add_argument("--bin-capacity", type=int, default=0, help="records per bin, 0 = library default")
    ap.add_argument("--raster-threads", type=int, default=0, help="bin raster workgroup (128/256/512/1024), 0 = 256")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-view batching measurement")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM traffic passes")
    ap.add_argument("--no-counts", action="store_true", help="skip the counting pass (frame kernels only)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 64-view batching measurement (its JSON object)")
    ap.add_argument("--frame-calls", action="store_true",
                    help="time the frame as separate clear/draw/resolve calls instead of a graph replay")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES bytes 10 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        src = torch.empty(args.calibrate_copy, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.calibrate_copy, "launches": 10}))
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
        print(json.dumps(bench_batch(ctx, torch.cuda.current_stream(), max(5, args.steps // 5), 3)))
        ctx.close()
        return 0

    # HBM traffic from counter passes in child processes, before this
    # process touches the GPU
    traffic = None if args.no_traffic else measure_traffic()

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

    def frame_calls():
        # the reference's call sequence: Clear, DrawMeshlets (with the cull), Resolve
        fb.clear(0, 0.0, stream)
        fb.draw(ds, view.object_to_clip, view.meshlet_count, 0, planes, flags, stream, args.bin_size,
                args.bin_capacity)
        fb.resolve(ds, view.object_to_clip, g.GLIMP_RESOLVE_LIT, lighting, stream)

    # the same frame (clear, cull, draw, lit resolve) replayed as one captured
    # graph: a batch of one view with its resolve; vis-buffer and colour are
    # identical to the separate calls (tests/test_gpu_batch.py)
    graph = None
    frame_shape = None
    if not args.direct and not args.frame_calls:
        shape = {"bin_size": args.bin_size or 16, "bin_capacity": args.bin_capacity or 1024,
                 "raster_threads": args.raster_threads or 256}
        graph = g.Batch(ctx, [(ds, fb, view, planes, flags)], shape=shape)
        frame_shape = shape
        if not args.bin_size and not args.bin_capacity and not args.raster_threads:
            frame_shape, _ = graph.autotune(stream)     # the launch shape, as for any batch
        graph.set_resolve([(view.object_to_clip, g.GLIMP_RESOLVE_LIT, lighting)])

    def frame_graph():
        graph.launch(True, stream)

    def frame_batch_eager():
        # the same batch launched kernel by kernel: the queue stays full
        # without the graph's per-launch boundary
        graph.launch(False, stream)

    def run(fn, steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    # per-kernel times: HIP events around every launch of the separate-call
    # frame, on the launch stream, over the timed steps
    for _ in range(args.warmup):
        frame_calls()
    ctx.profile(True)
    ctx.profile_read(reset=True)
    ms_calls = run(frame_calls, args.steps)
    kern = ctx.profile_read(reset=True)
    ctx.profile(False)
    # the headline frame time, without per-kernel events: the faster of the
    # graph replay and the eager batch launch (a short untimed trial of
    # each), then exactly `steps` timed frames of it
    frame, frame_mode, trial = frame_calls, "separate calls", {}
    if graph is not None:
        for name, fn in (("graph", frame_graph), ("eager", frame_batch_eager)):
            for _ in range(args.warmup):
                fn()
            trial[name] = round(run(fn, 20), 4)
        if trial["graph"] <= trial["eager"]:
            frame, frame_mode = frame_graph, "one graph replay of clear+cull+draw+resolve"
        else:
            frame, frame_mode = frame_batch_eager, "one eager batch launch of clear+cull+draw+resolve"
    for _ in range(args.warmup):
        frame()
    ms = run(frame, args.steps)

    if args.no_counts:
        print(json.dumps({"ms_per_step": round(ms, 4), "kernels": {k: v for k, v in kern.items()}}))
        if graph is not None:
            graph.close()
        fb.close()
        ds.close()
        ctx.close()
        return 0

    # counted quantities of one frame (separate, untimed run)
    ctx.reset_counters()
    fb.clear(0, 0.0, stream)
    fb.draw(ds, view.object_to_clip, view.meshlet_count, 0, planes, flags | g.GLIMP_DRAW_COUNTERS, stream,
            args.bin_size, args.bin_capacity)
    c = ctx.counters()
    _, visible = ds.cull(planes, 0, view.meshlet_count)
    stats = {
        "meshlets": scene.count, "visible": visible, "tris_visible": int(c["triangles_processed"]),
        "records": int(c["triangles_rasterized"]), "clipped": int(c["triangles_clipped"]),
        "fragments": int(c["fragments_covered"]), "pixels": view.width * view.height,
        "bin_entries": int(c["bin_entries"]), "bins_rasterized": int(c["bins_rasterized"]),
    }
    _, depth = fb.read()
    stats["covered_pixels"] = int((depth > 0).sum())
    alg = algorithmic_bytes(stats)
    per_kernel = {}
    for name, (launches, total) in kern.items():
        avg_ms = total / max(1, launches)
        b = alg.get(name, 0)
        per_kernel[name] = {"avg_us": round(avg_ms * 1e3, 2), "launches": launches, "alg_bytes": b,
                            "GBps": round(b / (avg_ms * 1e-3) / 1e9, 1) if avg_ms > 0 else None,
                            "hbm_bytes": traffic.get(name) if traffic else None,
                            "valu": _valu_roofline(traffic, name, round(avg_ms * 1e3, 2))}
    dom = max(kern.items(), key=lambda kv: kv[1][1])[0]
    dk = per_kernel[dom]
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
        "frame": frame_mode,
        "frame_shape": frame_shape,
        "frame_trial_ms": trial,
        "ms_per_step_separate_calls": round(ms_calls, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+fp32",
        "data": "synthetic",
        "config": {"workload": "10,240 grid-patch meshlets (1,003,520 tris), 1920x1080, depth complexity ~3, "
                               "16 materials 2048^2 RGBA8",
                   "raster": "direct" if args.direct else (
                       f"binned {frame_shape['bin_size']}x{frame_shape['bin_size']}, {frame_shape['bin_capacity']} "
                       f"records per bin, {frame_shape['raster_threads']}-thread bin raster" if frame_shape else
                       f"binned {args.bin_size or 16}x{args.bin_size or 16}"),
                   "submitted_tris": tris},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": dk["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dk["GBps"] / HBM_PEAK_GBS, 4) if dk["GBps"] else None,
                     "traffic": traffic.get(dom) if traffic else None,
                     "valu": _valu_roofline(traffic, dom, dk["avg_us"]),
                     "alg_bytes_per_launch": dk["alg_bytes"], "avg_us": dk["avg_us"]},
        "traffic_calibration": ({"fetch_scale": traffic["fetch_scale"], "write_scale": traffic["write_scale"]}
                                if traffic else None),
        "kernels": per_kernel,
        "frame_counts": stats,
        "scene_gen_s": round(t_gen, 2),
    }
    if not args.no_batch:
        result["batch"] = bench_batch(ctx, stream, max(5, args.steps // 5), 3)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(scene, view, args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))
    if graph is not None:
        graph.close()
    fb.close()
    ds.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
