"""The batching workload at full size: 64 views over 8 seeded scenes of
1.5k-45k meshlets at log-uniform resolutions from 256x256 to 1920x1080, the
views the reference renders one DrawMeshlets call at a time (Main.cpp:216-245
draws each view with its own Clear + DrawMeshlets + Resolve). At the launch
shape the autotuner picks, one batched launch (eager, and replayed as a
captured graph) must give every view exactly the vis-buffer of its own clear
+ draw, nothing may be dropped, and the smallest and the largest view must
equal the CPU oracle's frames."""
import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S

pytestmark = pytest.mark.gpu


def _same(a, b, label):
    for i, ((ia, da), (ib, db)) in enumerate(zip(a, b)):
        bad = (ia != ib) | (da.view(np.uint32) != db.view(np.uint32))
        assert not bad.any(), f"{label}: view {i} differs at {int(bad.sum())} pixels"


def test_full_batch_workload(gpu_ctx):
    import glimpsw_amd as g

    scenes, views = S.batch_views()
    assert len(views) == 64 and len(scenes) == 8
    print(f"[batch-full] scenes of {[s.count for s in scenes]} meshlets", flush=True)
    dscenes = [g.Scene(gpu_ctx, sc) for sc in scenes]
    fbs = [g.Framebuffer(gpu_ctx, v.width, v.height) for _, v in views]
    b = g.Batch(gpu_ctx, [(dscenes[si], fb, v) for (si, v), fb in zip(views, fbs)])
    try:
        shape, ms = b.autotune()
        print(f"[batch-full] autotuned shape {shape}, {ms:.3f} ms per batch", flush=True)
        b.launch_per_view()
        want = [fb.read() for fb in fbs]
        assert sum(int((ids != 0).sum()) for ids, _ in want) > 0, "nothing drawn"
        gpu_ctx.reset_counters()
        b.launch(graph=False)
        _same(want, [fb.read() for fb in fbs], "eager batch")
        b.launch(graph=True)
        _same(want, [fb.read() for fb in fbs], "graph batch")
        assert gpu_ctx.counters()["dropped"] == 0
        print("[batch-full] batch == per-view for all 64 views", flush=True)
        px = [v.width * v.height for _, v in views]
        for i in (int(np.argmin(px)), int(np.argmax(px))):
            si, v = views[i]
            print(f"[batch-full] oracle view {i}: {v.width}x{v.height}, scene {si} ({scenes[si].count} meshlets)",
                  flush=True)
            ofb, _, _ = oracle.render_frame(scenes[si], v)
            _same([(ofb.ids(), ofb.depth())], [want[i]], f"oracle view {i}")
    finally:
        b.close()
        for fb in fbs:
            fb.close()
        for d in dscenes:
            d.close()
