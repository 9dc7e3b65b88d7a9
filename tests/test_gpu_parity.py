"""GPU parity of the HIP kernels against the CPU oracle, through the C-ABI.

Bar: bit-exact for every integer output (class, flags, steps, crossing count) and, because both
sides follow one pinned IEEE operation order, for every float output too (escape direction, disk
r/phi/g, impact parameter). Cases: the 256x256 reference camera (r_obs=30, i=85 deg, FOV 60, 450
steps, disk [6,20]) in FP32, FP32+FP64 band, FP64 and with FP64 root refinement; hashed jitter;
odd and degenerate image sizes; explicit rays; law T against the oracle's law T.
"""
import numpy as np
import pytest

import oracle as O
from helpers import bit_mismatches, dir_for_impact_parameter, oracle_to_dict

pytestmark = pytest.mark.gpu

PREC = {"f32": 0, "band": 1, "f64": 2}


def _params(bhs, mode, root64=0, **kw):
    return bhs.march_params(precision=PREC[mode], root64=root64, **kw)


def _oracle_params(root64=0, **kw):
    base = dict(max_steps=450, step_scale=0.05, disk_step=0.18, disk_inner=6.0, disk_outer=20.0,
                disk_thickness=0.035, disk_mode=1, band_rel=0.02)
    base.update(kw)
    return O.lawS_params(root64=root64, **base)


def _render_both(tracer, cam, mode, root64=0, **kw):
    import torch
    import blackholerender_amd as bhs
    out = tracer.render(cam, _params(bhs, mode, root64, **kw))
    torch.cuda.synchronize()
    g = out.numpy()
    ref = oracle_to_dict(O.render(cam.as_dict(), _oracle_params(root64, **kw), mode=mode, threads=8))
    return g, ref, tracer.last_stats()


@pytest.mark.parametrize("mode", ["f32", "band", "f64"])
def test_render_c1_bit_exact(tracer, mode):
    import blackholerender_amd as bhs
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, 256, 256)
    print(f"[parity] render 256x256 mode={mode}", flush=True)
    g, ref, st = _render_both(tracer, cam, mode)
    mm = bit_mismatches(g, ref)
    assert all(v == 0 for v in mm.values()), mm
    assert st["rays"] == 256 * 256
    assert st["steps"] == int(ref["steps"].sum())
    assert st["crossings"] == int(np.minimum(ref["n_cross"], 255).sum())
    if mode == "band":
        assert st["band64"] == int(((ref["flags"] & 4) > 0).sum()) > 0


@pytest.mark.parametrize("mode", ["f32", "band"])
def test_render_root64_bit_exact(tracer, mode):
    import blackholerender_amd as bhs
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, 192, 128)
    print(f"[parity] root64 192x128 mode={mode}", flush=True)
    g, ref, _ = _render_both(tracer, cam, mode, root64=1)
    mm = bit_mismatches(g, ref)
    assert all(v == 0 for v in mm.values()), mm
    assert ((g["flags"] & 8) > 0).sum() == (g["n_cross"] > 0).sum() > 0


@pytest.mark.parametrize("shape", [(1, 1), (17, 5), (33, 130)])
def test_render_ragged_sizes(tracer, shape):
    import blackholerender_amd as bhs
    w, h = shape
    cam = bhs.orbit_camera(15.0, 80.0, 75.0, w, h)
    print(f"[parity] ragged {w}x{h}", flush=True)
    g, ref, _ = _render_both(tracer, cam, "band")
    assert all(v == 0 for v in bit_mismatches(g, ref).values())


def test_render_hashed_jitter(tracer):
    import blackholerender_amd as bhs
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, 160, 90, jitter_hashed=1, seed=0x1234_5678_9ABC,
                           camera_id=3, sample=7)
    print("[parity] hashed jitter 160x90", flush=True)
    g, ref, _ = _render_both(tracer, cam, "band")
    assert all(v == 0 for v in bit_mismatches(g, ref).values())


@pytest.mark.parametrize("steps", [200, 300, 700])
def test_render_presets(tracer, steps):
    import blackholerender_amd as bhs
    ss, dstep = bhs.PRESETS[steps]
    cam = bhs.orbit_camera(60.0, 60.0, 40.0, 96, 64)
    print(f"[parity] preset {steps}", flush=True)
    g, ref, _ = _render_both(tracer, cam, "band", max_steps=steps, step_scale=ss, disk_step=dstep)
    assert all(v == 0 for v in bit_mismatches(g, ref).values())


def test_max_steps_zero_and_no_disk(tracer):
    import blackholerender_amd as bhs
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, 64, 64)
    g, ref, _ = _render_both(tracer, cam, "f32", max_steps=0)
    assert all(v == 0 for v in bit_mismatches(g, ref).values())
    assert (g["flags"] & 1).all() and (g["cls"] == 0).all()
    g, ref, _ = _render_both(tracer, cam, "f32", disk_mode=0)
    assert all(v == 0 for v in bit_mismatches(g, ref).values())
    assert (g["cls"] != 1).all()


def test_march_rays_bit_exact(tracer):
    import torch
    import blackholerender_amd as bhs
    rng = np.random.default_rng(7)
    n = 4096
    r_obs = np.exp(rng.uniform(np.log(8.0), np.log(1000.0), n))
    th = rng.uniform(0.0, np.pi, n)
    ph = rng.uniform(-np.pi, np.pi, n)
    pos = np.stack([r_obs * np.sin(th) * np.cos(ph), r_obs * np.cos(th),
                    r_obs * np.sin(th) * np.sin(ph)], 1).astype(np.float32)
    d = -pos / np.linalg.norm(pos, axis=1, keepdims=True) + rng.normal(0, 0.08, (n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    print("[parity] march_rays 4096", flush=True)
    for mode in ("f32", "band"):
        out = tracer.march_rays(torch.from_numpy(pos).cuda(), torch.from_numpy(d).cuda(),
                                _params(bhs, mode, root64=1))
        torch.cuda.synchronize()
        g = out.numpy()
        p = _oracle_params(root64=1)
        ref = oracle_to_dict(np.array([tuple(_ray_tuple(O.march(pos[i], d[i], p, mode)))
                                       for i in range(n)], dtype=O.LAWS_RAY_DTYPE))
        assert all(v == 0 for v in bit_mismatches(g, ref).values()), mode


def _ray_tuple(r):
    return (r.cls, r.flags, r.steps, r.n_cross, tuple(r.esc), tuple(r.disk_r), tuple(r.disk_phi),
            tuple(r.disk_g), r.b)


def test_trace_T_bit_exact(tracer):
    rng = np.random.default_rng(11)
    pos, dirs = [], []
    for D in (10.0, 25.0, 2000.0):
        for b in (1.0, 5.15, 5.196, 5.25, 8.0, 20.0):
            pos.append((0.0, 0.0, D))
            dirs.append(dir_for_impact_parameter(D, b))
    for _ in range(200):
        r = float(np.exp(rng.uniform(np.log(4.0), np.log(500.0))))
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        d = -u + rng.normal(0, 0.3, 3)
        pos.append(tuple(r * u))
        dirs.append(d / np.linalg.norm(d))
    pos, dirs = np.array(pos), np.array(dirs)
    print("[parity] law T", len(pos), flush=True)
    g = tracer.trace_batch(pos, dirs, 20000, 0.005, 1e-4, 30.0)
    ref = O.trace_batch(pos, dirs, 20000, 0.005, 1e-4, 30.0)
    for f in ("status", "steps", "equator_crossings"):
        assert np.array_equal(g[f], ref[f]), f
    assert np.array_equal(g["dir"].view(np.uint64), ref["dir"].view(np.uint64))
    assert np.array_equal(g["min_r"].view(np.uint64), ref["min_r"].view(np.uint64))


TUNINGS = [(0, 256, 8, 64), (0, 64, 8, 64), (1, 256, 8, 64), (1, 256, 8, 1), (1, 256, 7, 32),
           (1, 64, 6, 16), (1, 256, 4, 8)]


@pytest.mark.parametrize("tuning", TUNINGS)
def test_launch_shapes_bit_identical(tracer, tuning):
    """Every launch shape (static tiles, persistent work queue with any refill threshold and
    occupancy) gives the oracle's bits, root refinement and FP64 band included."""
    import blackholerender_amd as bhs
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, 200, 120, jitter_hashed=1, seed=99)
    print(f"[parity] tuning {tuning}", flush=True)
    tracer.set_tuning(*tuning)
    try:
        for root64 in (0, 1):
            g, ref, st = _render_both(tracer, cam, "band", root64=root64)
            mm = bit_mismatches(g, ref)
            assert all(v == 0 for v in mm.values()), (tuning, root64, mm)
            assert st["rays"] == cam.n_rays
    finally:
        tracer.set_tuning(1, 256, 6, 64)   # the context default


def test_autotune_picks_a_valid_shape(tracer):
    import blackholerender_amd as bhs
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, 320, 180)
    best = tracer.autotune(cam, bhs.march_params(), reps=2)
    assert best["mode"] in (0, 1) and best["block"] in (64, 256) and best["ms"] > 0
    assert tracer.get_tuning() == {k: best[k] for k in ("mode", "block", "waves_per_simd",
                                                        "refill_min")}
    g, ref, _ = _render_both(tracer, cam, "band")
    assert all(v == 0 for v in bit_mismatches(g, ref).values())
