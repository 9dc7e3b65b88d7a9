"""Batched multi-camera launches (bhs_render_batch) and their HIP-graph replay, against the oracle.

Bar: every ray of every camera bit-identical to the oracle's law S (FP32 with the FP64 band) for a
mixed batch (sizes, step budgets and presets, jitter, one camera face-on, one far away); for the
64-camera batch, a strided sample of every camera. Graph replays (several, with eager work and
synchronisations between them) equal the eager batch bit for bit, also after a pose update.
"""
import numpy as np
import pytest

import oracle as O
from helpers import bit_mismatches, oracle_to_dict

pytestmark = pytest.mark.gpu


def _mixed_batch(bhs):
    specs = [(30.0, 85.0, 60.0, 96, 64, 450), (15.0, 80.0, 75.0, 33, 17, 700),
             (8.0, 45.0, 90.0, 64, 64, 200), (1000.0, 60.0, 20.0, 40, 72, 200),
             (40.0, 0.0, 60.0, 48, 48, 300), (120.0, 89.0, 30.0, 128, 8, 450),
             (25.0, 30.0, 50.0, 1, 1, 450), (60.0, 70.0, 40.0, 72, 40, 700)]
    cams, params = [], []
    for k, (r, i, f, w, h, st) in enumerate(specs):
        cams.append(bhs.orbit_camera(r, i, f, w, h, jitter_hashed=1, seed=77, camera_id=k))
        params.append(bhs.march_params(max_steps=st, root64=k % 2))
    return cams, params


def _oracle_cam(cam, p, stride=1):
    ss, dstep = p.step_scale, p.disk_step
    op = O.lawS_params(max_steps=p.max_steps, step_scale=ss, disk_step=dstep,
                       disk_inner=p.disk_inner, disk_outer=p.disk_outer,
                       disk_thickness=p.disk_thickness, root64=p.root64, band_rel=p.band_rel)
    return oracle_to_dict(O.render(cam.as_dict(), op, mode="band", stride=stride, threads=8))


def _slice(g, a, b, stride=1):
    return {k: v[a:b:stride] for k, v in g.items()}


def test_mixed_batch_bit_exact(tracer):
    import torch
    import blackholerender_amd as bhs
    cams, params = _mixed_batch(bhs)
    print("[batch] mixed 8 cameras", flush=True)
    out = tracer.render_batch(cams, params)
    torch.cuda.synchronize()
    g = out.numpy()
    offs = tracer.batch_offsets(cams)
    total = 0
    for k, (c, p) in enumerate(zip(cams, params)):
        ref = _oracle_cam(c, p)
        mm = bit_mismatches(_slice(g, offs[k], offs[k + 1]), ref)
        assert all(v == 0 for v in mm.values()), (k, mm)
        total += int(ref["steps"].sum())
    st = tracer.last_stats()
    assert st["rays"] == offs[-1] and st["steps"] == total


@pytest.mark.parametrize("waves", [8, 4])
def test_segmented_queue_migration_bit_exact(tracer, waves):
    """Whole-tile claims from the segmented work queue (refill threshold 64) when the resident
    waves far outnumber the tiles and the tile ranges differ widely in cost: a single-tile camera,
    a near camera whose tiles orbit for hundreds of steps, a far one whose rays escape at once and
    a 1x1 camera (padding items), so most waves find their first range used up and move on. Every
    ray is handed out exactly once (the launch counters count each ray and step once) and equals
    the oracle bit for bit."""
    import torch
    import blackholerender_amd as bhs
    specs = [(30.0, 85.0, 60.0, 16, 4, 450), (8.0, 80.0, 75.0, 192, 128, 700),
             (900.0, 40.0, 20.0, 96, 64, 200), (25.0, 30.0, 50.0, 1, 1, 450)]
    cams = [bhs.orbit_camera(r, i, f, w, h, jitter_hashed=1, seed=13, camera_id=k)
            for k, (r, i, f, w, h, _) in enumerate(specs)]
    params = [bhs.march_params(max_steps=st) for (*_, st) in specs]
    print(f"[batch] segmented queue, {waves} waves/SIMD", flush=True)
    prev = tracer.get_tuning()
    tracer.set_tuning(1, 256, waves, 64)
    try:
        out = tracer.render_batch(cams, params)
        torch.cuda.synchronize()
        st = tracer.last_stats()
    finally:
        tracer.set_tuning(prev["mode"], prev["block"], prev["waves_per_simd"], prev["refill_min"])
    g = out.numpy()
    offs = tracer.batch_offsets(cams)
    total = 0
    for k, (c, p) in enumerate(zip(cams, params)):
        ref = _oracle_cam(c, p)
        mm = bit_mismatches(_slice(g, offs[k], offs[k + 1]), ref)
        assert all(v == 0 for v in mm.values()), (k, mm)
        total += int(ref["steps"].sum())
    assert st["rays"] == offs[-1] and st["steps"] == total


def test_c3_batch_sampled_bit_exact(tracer):
    import torch
    from blackholerender_amd import workload as W
    cams, params = W.batch_c3()
    print("[batch] 64 cameras", flush=True)
    out = tracer.render_batch(cams, params)
    torch.cuda.synchronize()
    g = out.numpy()
    offs = tracer.batch_offsets(cams)
    for k, (c, p) in enumerate(zip(cams, params)):
        stride = 211
        ref = _oracle_cam(c, p, stride=stride)
        mm = bit_mismatches(_slice(g, offs[k], offs[k + 1], stride), ref)
        assert all(v == 0 for v in mm.values()), (k, mm)


def test_long_band_list_bit_exact(tracer):
    """A band list longer than the one-ray-per-lane band kernel covers (262,144 rays; the rest is
    marched by the refilling queue kernel): two 1024x1024 cameras at r_obs = 15 with a 25% band,
    one with FP64 crossing refinement, every ray bit-identical to the oracle."""
    import torch
    import blackholerender_amd as bhs
    cams = [bhs.orbit_camera(15.0, 80.0, 60.0, 1024, 1024, jitter_hashed=1, seed=9, camera_id=k,
                             azimuth_deg=30.0 * k) for k in range(2)]
    params = [bhs.march_params(max_steps=700, disk_outer=30.0, band_rel=0.25, root64=k)
              for k in range(2)]
    print("[batch] 2 x 1024x1024, 25% band", flush=True)
    out = tracer.render_batch(cams, params)
    torch.cuda.synchronize()
    st = tracer.last_stats()
    assert st["band64"] > 262144, st
    g = out.numpy()
    offs = tracer.batch_offsets(cams)
    for k, (c, p) in enumerate(zip(cams, params)):
        mm = bit_mismatches(_slice(g, offs[k], offs[k + 1]), _oracle_cam(c, p))
        assert all(v == 0 for v in mm.values()), (k, mm)


def test_graph_replays_equal_eager(tracer):
    import torch
    import blackholerender_amd as bhs
    cams, params = _mixed_batch(bhs)
    eager = tracer.render_batch(cams, params)
    torch.cuda.synchronize()
    ref = eager.numpy()
    gr = tracer.graph(cams, params)
    try:
        for rep in range(3):
            gr.out.cls.fill_(7)           # eager work between replays
            gr.out.esc.fill_(-1.0)
            torch.cuda.synchronize()
            gr.launch()
            torch.cuda.synchronize()
            mm = bit_mismatches(gr.out.numpy(), ref)
            assert all(v == 0 for v in mm.values()), (rep, mm)
            assert gr.stats()["rays"] == ref["cls"].size
        moved = [bhs.orbit_camera(c.meta["r_obs"] * 1.1, c.meta["inclination_deg"] * 0.9,
                                  c.meta["fov_deg"], c.width, c.height, jitter_hashed=1, seed=5,
                                  camera_id=c.camera_id) for c in cams]
        gr.update(moved, params)
        eager2 = tracer.render_batch(moved, params)
        torch.cuda.synchronize()
        ref2 = eager2.numpy()
        for rep in range(2):
            gr.launch()
            torch.cuda.synchronize()
            assert all(v == 0 for v in bit_mismatches(gr.out.numpy(), ref2).values()), rep
    finally:
        gr.close()


def _planes_equal(a, b):
    """Per plane of two RayBuffers, whether they are equal bit for bit (compared on the device)."""
    import torch
    return {k: bool(torch.equal(getattr(a, k).view(torch.uint8) if getattr(a, k).dtype !=
                                torch.uint8 else getattr(a, k),
                                getattr(b, k).view(torch.uint8) if getattr(b, k).dtype !=
                                torch.uint8 else getattr(b, k)))
            for k in ("cls", "flags", "n_cross", "steps", "esc", "disk", "b")}


def test_c3_graph_replays_equal_eager(tracer):
    """The 64-camera batch itself captured in one HIP graph: three replays, with eager work and
    synchronisations between them, each equal to the eager batch bit for bit (every ray of every
    camera), and the replay's counters equal to the eager call's."""
    import torch
    from blackholerender_amd import workload as W
    cams, params = W.batch_c3()
    print("[batch] 64-camera graph", flush=True)
    eager = tracer.render_batch(cams, params)
    torch.cuda.synchronize()
    est = tracer.last_stats()
    gr = tracer.graph(cams, params)
    try:
        for rep in range(3):
            gr.out.cls.fill_(7)
            gr.out.disk.fill_(-1.0)
            gr.out.steps.fill_(-1)
            torch.cuda.synchronize()
            gr.launch()
            torch.cuda.synchronize()
            eq = _planes_equal(gr.out, eager)
            assert all(eq.values()), (rep, eq)
            gst = gr.stats()
            for k in ("rays", "steps", "crossings", "escapes", "band64"):
                assert gst[k] == est[k], (rep, k, gst[k], est[k])
    finally:
        gr.close()


def test_batch_rejects_bad_arguments(tracer):
    import blackholerender_amd as bhs
    cams, params = _mixed_batch(bhs)
    with pytest.raises(bhs.BhsError):
        tracer.render_batch(cams * 9, params * 9)           # 72 > 64 cameras
    bad = list(params)
    bad[0] = bhs.march_params(precision=bhs.PREC_F64)
    with pytest.raises(bhs.BhsError):
        tracer.render_batch(cams, bad)
