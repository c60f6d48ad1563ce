"""Property test over seeded random tiny scenes (tests/random_scenes.py):
for every case the HIP vis-buffer (ids and depth) and the perf counters equal
the oracle's bit for bit, through the binned route with a random bin shape
(including capacities small enough to spill) or the direct route, with one to
three draws into the same framebuffer, with or without the cull, clipping
and guard band; and the same cases as views of batches. Seed 8 holds a
clipped triangle whose fan keeps a triangle after a culled one: the
binning's pair search has to step over the culled triangle's empty range."""
import numpy as np
import pytest

import oracle
from tests import random_scenes as R

pytestmark = pytest.mark.gpu

SEEDS = range(120)


def _oracle(case):
    sc, view = case.scene, case.view
    b = oracle.SceneBinding(sc)
    fb = oracle.OracleFramebuffer(view.width, view.height)
    fb.clear(0x0BADF00D, case.clear_depth)
    total = {}
    for off, cnt in case.draws:
        bm = oracle.cull(b, view.frustum_planes(), off, cnt)[0] if case.cull else None
        ctr, _ = oracle.draw(fb, b, view.object_to_clip, cnt, off, bm, clipping=not (case.flags & 1),
                             guardband=not (case.flags & 2))
        for k, v in ctr.items():
            total[k] = total.get(k, 0) + v
    return fb.ids(), fb.depth(), total


def test_random_scenes(gpu_ctx):
    import glimpsw_amd as g

    for seed in SEEDS:
        case = R.make_case(seed)
        sc, view = case.scene, case.view
        print(f"[random] seed {seed}: {sc.count} meshlets {view.width}x{view.height} draws {case.draws} "
              f"cull {case.cull} flags {case.flags} bins {case.bin_size}/{case.bin_capacity} direct {case.direct}",
              flush=True)
        want_ids, want_depth, want_ctr = _oracle(case)
        ds = g.Scene(gpu_ctx, sc)
        fb = g.Framebuffer(gpu_ctx, view.width, view.height)
        try:
            fb.clear(0x0BADF00D, case.clear_depth)
            gpu_ctx.reset_counters()
            flags = case.flags | g.GLIMP_DRAW_COUNTERS | (g.GLIMP_DRAW_DIRECT if case.direct else 0)
            for off, cnt in case.draws:
                fb.draw(ds, view.object_to_clip, cnt, off, view.frustum_planes() if case.cull else None, flags,
                        bin_size=case.bin_size, bin_capacity=case.bin_capacity)
            ids, depth = fb.read()
            ctr = gpu_ctx.counters()
        finally:
            fb.close()
            ds.close()
        bad = (ids != want_ids) | (depth.view(np.uint32) != want_depth.view(np.uint32))
        assert not bad.any(), f"seed {seed}: {bad.sum()} pixels differ, first {np.argwhere(bad)[:3].tolist()}"
        for k in ("triangles_processed", "triangles_rasterized", "triangles_clipped", "fragments_covered"):
            assert ctr[k] == want_ctr[k], (seed, k, ctr[k], want_ctr[k])
        assert ctr["dropped"] == 0


@pytest.mark.parametrize("shape", [None, {"bin_size": 16, "bin_capacity": 3, "raster_threads": 256},
                                   {"bin_size": 64, "bin_capacity": 2048, "raster_threads": 1024}])
def test_random_scenes_in_batches(gpu_ctx, shape):
    """The same random cases as views of batches (one draw each, every view
    its own scene, framebuffer size, cull planes and clipping flags),
    launched eagerly and replayed as a graph: first all of them in one batch
    (together they queue far more clipped triangles than the 4,096 of one
    small draw's queue; a batch's shared queue is as large as its views'
    queues drawn one at a time), then 12 batches of every twelfth view."""
    import dataclasses

    import glimpsw_amd as g

    cases = [dataclasses.replace(c, draws=[(0, c.scene.count)], clear_depth=0.0)
             for c in (R.make_case(s) for s in SEEDS)]
    wants = [_oracle(c)[:2] for c in cases]
    for sel in [list(range(len(cases)))] + [list(range(k, len(cases), 12)) for k in range(12)]:
        items, scenes, fbs = [], [], []
        try:
            for i in sel:
                c = cases[i]
                ds = g.Scene(gpu_ctx, c.scene)
                fb = g.Framebuffer(gpu_ctx, c.view.width, c.view.height)
                scenes.append(ds)
                fbs.append(fb)
                items.append((ds, fb, c.view, c.view.frustum_planes() if c.cull else None, c.flags & 3))
            b = g.Batch(gpu_ctx, items, shape=shape, clear_color=0x0BADF00D, clear_depth=0.0)
            try:
                for graph in (False, True):
                    print(f"[random batch] views {sel} shape {shape} graph {graph}", flush=True)
                    gpu_ctx.reset_counters()
                    b.launch(graph=graph)
                    assert gpu_ctx.counters()["dropped"] == 0
                    for i, fb in zip(sel, fbs):
                        want_ids, want_depth = wants[i]
                        ids, depth = fb.read()
                        bad = (ids != want_ids) | (depth.view(np.uint32) != want_depth.view(np.uint32))
                        assert not bad.any(), f"case {i} (graph {graph}): {bad.sum()} pixels differ"
                        fb.clear(0, 0.0)
            finally:
                b.close()
        finally:
            for fb in fbs:
                fb.close()
            for ds in scenes:
                ds.close()


def test_random_scenes_deferred_and_overdraw(gpu_ctx):
    """Every third random case through the deferred pipeline (three G-buffer
    layers) and the overdraw pipeline (counts, depth maximum, counters and
    both overdraw resolves), against the oracle."""
    from tests import test_gpu_deferred as D
    from tests import test_gpu_overdraw as O

    for seed in range(0, len(SEEDS), 3):
        case = R.make_case(seed)
        gpu_ctx.reset_counters()
        D._check(f"random seed {seed}", gpu_ctx, case.scene, case.view, frames=[(0x0BADF00D, case.draws)])
        O._check(f"random seed {seed}", gpu_ctx, case.scene, case.view, 0, tuple(case.draws))
