"""GPU parity of the volumetric disk render (bhs_render_volumetric[_batch]) against the oracle.

Bar: bit-exact for every output (colour, sky weight, class, flags, steps, escape direction): the
HIP march and the oracle share one IEEE operation order, including the Hermite sub-samples, the
ceil of the sub-sample count and the opaque stop. Cases: the High preset on the 1080p camera's view
(FP32 with the FP64 band), the Ultra preset close to the hole with the physical look (Doppler,
inflow, disk clock), a dense disk (opaque stops), a single FP64 camera, a three-camera batch, a
whole 1920x1080 frame on a strided sample.
"""
import numpy as np
import pytest

import oracle as O
from helpers import oracle_threads

pytestmark = pytest.mark.gpu


def _op(p):
    return O.lawS_params(max_steps=p.max_steps, step_scale=p.step_scale, disk_step=p.disk_step,
                         disk_inner=p.disk_inner, disk_outer=p.disk_outer,
                         disk_thickness=p.disk_thickness, disk_mode=2, band_rel=p.band_rel)


def _osh(sh, p):
    return O.shading(sh.density, sh.opacity, sh.peak_temp, sh.brightness, sh.noise_scale,
                     sh.noise_strength, sh.warmth, sh.doppler, sh.time, sh.inflow, p.disk_inner,
                     p.disk_outer)


def _compare(g, r, what, sel=slice(None)):
    np.testing.assert_array_equal(g["rgb"][sel], r["rgb"], err_msg=f"{what}: rgb")
    np.testing.assert_array_equal(g["sky_weight"][sel], r["sky_weight"], err_msg=f"{what}: sky")
    np.testing.assert_array_equal(g["cls"][sel], r["cls"], err_msg=f"{what}: cls")
    np.testing.assert_array_equal(g["flags"][sel], r["flags"], err_msg=f"{what}: flags")
    np.testing.assert_array_equal(g["steps"][sel], r["steps"], err_msg=f"{what}: steps")
    np.testing.assert_array_equal(g["esc"][sel], r["esc"], err_msg=f"{what}: esc")


def _gpu(tracer, cams, params, sh):
    import torch
    import blackholerender_amd as bhs
    n = sum(c.width * c.height for c in cams)
    rays = bhs.RayBuffers(n, f"cuda:{tracer.device}")
    if len(cams) == 1:
        out, _ = tracer.render_volumetric(cams[0], params[0], sh, rays=rays)
    else:
        out, _ = tracer.render_volumetric_batch(cams, params, sh, rays=rays)
    torch.cuda.synchronize()
    r, o = rays.numpy(), out.numpy()
    return dict(rgb=o["rgb"], sky_weight=o["sky_weight"], cls=r["cls"], flags=r["flags"],
                steps=r["steps"], esc=r["esc"], n_cross=r["n_cross"])


CASES = [
    # r_obs, incl, fov, w, h, steps, disk, shading overrides
    (30.0, 85.0, 60.0, 96, 54, 450, (6.0, 20.0), {}),
    (15.0, 80.0, 60.0, 64, 48, 700, (6.0, 30.0), dict(doppler=1.0, inflow=0.08, time=17.0)),
    (30.0, 60.0, 45.0, 48, 48, 300, (6.0, 20.0), dict(density=8.0, opacity=6.0)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_volumetric_bit_exact(tracer, case):
    import blackholerender_amd as bhs
    r_obs, inc, fov, w, h, steps, (di, do), over = CASES[case]
    cam = bhs.orbit_camera(r_obs, inc, fov, w, h, jitter_hashed=1, seed=31 + case)
    p = bhs.march_params(max_steps=steps, disk_inner=di, disk_outer=do,
                         disk_mode=bhs.DISK_VOLUMETRIC)
    sh = bhs.disk_shading(**over)
    print(f"[volume] case {case}: {w}x{h} r_obs {r_obs}", flush=True)
    g = _gpu(tracer, [cam], [p], sh)
    r = O.volumetric(cam.as_dict(), _op(p), _osh(sh, p), mode="band", threads=oracle_threads())
    _compare(g, r, f"case {case}")
    assert np.all(g["n_cross"] == 0)
    assert g["rgb"].sum(1).max() > 0.5 and (g["cls"] == bhs.CLASS_ESCAPE).any()
    if over.get("density", 1.0) > 4:
        assert (g["cls"] == bhs.CLASS_DISK).sum() > 20        # absorbed (opaque) rays
    st = tracer.last_stats()
    assert st["rays"] == w * h and st["steps"] == int(g["steps"].sum())


def test_volumetric_fp64_camera(tracer):
    import blackholerender_amd as bhs
    cam = bhs.orbit_camera(20.0, 75.0, 50.0, 40, 30, jitter_hashed=1, seed=5)
    p = bhs.march_params(disk_mode=bhs.DISK_VOLUMETRIC, precision=bhs.PREC_F64)
    sh = bhs.disk_shading(warmth=0.0)
    g = _gpu(tracer, [cam], [p], sh)
    r = O.volumetric(cam.as_dict(), _op(p), _osh(sh, p), mode="f64", threads=oracle_threads())
    _compare(g, r, "f64")
    assert np.all((g["flags"] & bhs.FLAG_BAND64) > 0)


def test_volumetric_batch(tracer):
    import blackholerender_amd as bhs
    cams = [bhs.orbit_camera(30.0, 85.0, 60.0, 40, 24, jitter_hashed=1, seed=9, camera_id=0),
            bhs.orbit_camera(12.0, 45.0, 75.0, 17, 33, jitter_hashed=1, seed=9, camera_id=1),
            bhs.orbit_camera(200.0, 88.0, 10.0, 32, 32, camera_id=2)]
    params = [bhs.march_params(max_steps=st, disk_mode=bhs.DISK_VOLUMETRIC)
              for st in (450, 700, 200)]
    sh = bhs.disk_shading(doppler=0.5)
    print("[volume] batch", flush=True)
    g = _gpu(tracer, cams, params, sh)
    offs = tracer.batch_offsets(cams)
    for k, (c, p) in enumerate(zip(cams, params)):
        r = O.volumetric(c.as_dict(), _op(p), _osh(sh, p), mode="band", threads=oracle_threads())
        _compare({f: v[offs[k]:offs[k + 1]] for f, v in g.items()}, r, f"camera {k}")


def test_volumetric_1080p_frame_sampled(tracer):
    """A whole 1920x1080 frame of the benchmark camera at the High preset, compared on every
    97th pixel."""
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    cam = W.camera_for(W.C2, sample=0)
    p = bhs.march_params(disk_mode=bhs.DISK_VOLUMETRIC)
    sh = bhs.disk_shading()
    print("[volume] 1920x1080", flush=True)
    g = _gpu(tracer, [cam], [p], sh)
    r = O.volumetric(cam.as_dict(), _op(p), _osh(sh, p), mode="band", stride=97,
                     threads=oracle_threads())
    _compare(g, r, "1080p", sel=slice(0, None, 97))
    assert (g["flags"] & bhs.FLAG_BAND64).any()
