"""GPU parity of the thin-disk shading pass (bhs_shade_thin_disk[_batch]), its tables
(bhs_disk_textures) and the accumulation (bhs_accumulate) against the oracle.

Bar: bit-exact. The HIP pass and the oracle follow one IEEE operation order (elementary functions
built from fma, frexp and ldexp), so every colour, transmittance and opaque stop agrees bit for bit.
Cases: the 256x256 reference camera; a view with rays crossing the disk three or more times (the
completion march, FP32 and FP64 band rays) at the default and at a high opacity (opaque stops); the
physical look (Doppler on, inflow, disk clock); FP64 crossing refinement; a three-camera batch; a
surrogate render; a whole 1920x1080 frame of the benchmark; accumulation over several samples.
"""
import numpy as np
import pytest

import oracle as O
from helpers import oracle_params_for, oracle_threads

pytestmark = pytest.mark.gpu


def _oracle_params(p):
    return O.lawS_params(max_steps=p.max_steps, step_scale=p.step_scale, disk_step=p.disk_step,
                         disk_inner=p.disk_inner, disk_outer=p.disk_outer,
                         disk_thickness=p.disk_thickness, root64=p.root64, band_rel=p.band_rel)


def _oracle_shading(sh, p):
    return O.shading(sh.density, sh.opacity, sh.peak_temp, sh.brightness, sh.noise_scale,
                     sh.noise_strength, sh.warmth, sh.doppler, sh.time, sh.inflow, p.disk_inner,
                     p.disk_outer)


def _check(tracer, cam, p, sh, rays, out, what):
    import torch
    torch.cuda.synchronize()
    rec = rays.numpy() if hasattr(rays, "numpy") else rays
    g = out.numpy()
    rgb, sky = O.shade_rays(cam.as_dict(), _oracle_params(p), _oracle_shading(sh, p), rec)
    np.testing.assert_array_equal(g["rgb"], rgb, err_msg=f"{what}: rgb")
    np.testing.assert_array_equal(g["sky_weight"], sky, err_msg=f"{what}: sky weight")
    return rec, g


def test_disk_tables_equal_oracle(tracer):
    vol, lut = tracer.disk_textures()
    np.testing.assert_array_equal(vol, O.noise_volume(threads=oracle_threads()))
    np.testing.assert_array_equal(lut, O.blackbody_lut())


def test_shade_reference_camera(tracer):
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    cam = W.camera_for(W.C1)
    p = W.params_for(W.C1)
    sh = bhs.disk_shading()
    print("[shade] 256x256", flush=True)
    rays = tracer.render(cam, p)
    out = tracer.shade_thin_disk(cam, rays, p, sh)
    rec, g = _check(tracer, cam, p, sh, rays, out, "c1")
    disk = rec["n_cross"] > 0
    lum = g["rgb"][disk].sum(1)
    assert disk.sum() > 1000 and np.median(lum) > 0.05 and (lum > 0).mean() > 0.99
    assert np.all(g["rgb"][~disk] == 0)


NEAR = dict(r=15.0, incl=80.0, fov=45.0, w=128, h=96)


@pytest.mark.parametrize("opacity,doppler,root64", [(2.0, 0.0, 0), (40.0, 0.0, 0), (2.0, 1.0, 1)])
def test_shade_multiply_crossing_rays(tracer, opacity, doppler, root64):
    """Disk [2.5, 30]: rays winding round the photon sphere cross the annulus three or more times
    and go through the completion march, in FP32 and (FP64 band) in double."""
    import blackholerender_amd as bhs
    cam = bhs.orbit_camera(NEAR["r"], NEAR["incl"], NEAR["fov"], NEAR["w"], NEAR["h"],
                           jitter_hashed=1, seed=21)
    p = bhs.march_params(max_steps=700, disk_inner=2.5, disk_outer=30.0, root64=root64)
    sh = bhs.disk_shading(opacity=opacity, doppler=doppler, time=37.5, inflow=0.08 * doppler,
                          warmth=0.7)
    print(f"[shade] near view opacity={opacity} doppler={doppler} root64={root64}", flush=True)
    rays = tracer.render(cam, p)
    out = tracer.shade_thin_disk(cam, rays, p, sh)
    rec, g = _check(tracer, cam, p, sh, rays, out, "near")
    many = rec["n_cross"] > 2
    assert many.sum() > 50 and ((rec["flags"][many] & bhs.FLAG_BAND64) > 0).any()
    if opacity > 10:
        assert ((g["sky_weight"] == 0) & ((rec["flags"] & bhs.FLAG_ESCAPED) > 0)).sum() > 50


def test_shade_batch(tracer):
    import blackholerender_amd as bhs
    cams = [bhs.orbit_camera(30.0, 85.0, 60.0, 64, 48, jitter_hashed=1, seed=3, camera_id=0),
            bhs.orbit_camera(12.0, 40.0, 75.0, 33, 17, jitter_hashed=1, seed=3, camera_id=1),
            bhs.orbit_camera(15.0, 80.0, 45.0, 96, 64, jitter_hashed=1, seed=3, camera_id=2)]
    params = [bhs.march_params(max_steps=st, disk_inner=3.0, disk_outer=30.0)
              for st in (450, 300, 700)]
    sh = bhs.disk_shading(doppler=0.5, time=5.0)
    print("[shade] batch of 3", flush=True)
    rays = tracer.render_batch(cams, params)
    out = tracer.shade_thin_disk_batch(cams, rays, params, sh)
    import torch
    torch.cuda.synchronize()
    r, g = rays.numpy(), out.numpy()
    offs = tracer.batch_offsets(cams)
    for k, (c, p) in enumerate(zip(cams, params)):
        rec = {f: v[offs[k]:offs[k + 1]] for f, v in r.items()}
        rgb, sky = O.shade_rays(c.as_dict(), _oracle_params(p), _oracle_shading(sh, p), rec)
        np.testing.assert_array_equal(g["rgb"][offs[k]:offs[k + 1]], rgb, err_msg=f"cam {k}")
        np.testing.assert_array_equal(g["sky_weight"][offs[k]:offs[k + 1]], sky, err_msg=f"cam {k}")


def test_shade_surrogate_frame(tracer):
    import os
    import blackholerender_amd as bhs
    from blackholerender_amd.surrogate import Surrogate
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sur = Surrogate.load(tracer, os.path.join(root, "blackholerender_amd", "models",
                                              "lawS_450_disk6-20.bhsmlp"))
    sur.set_table(1024, 128)
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, 320, 180, jitter_hashed=1, seed=8)
    p = bhs.march_params()
    sh = bhs.disk_shading()
    print("[shade] surrogate 320x180", flush=True)
    rays = sur.render(cam, p)
    out = tracer.shade_thin_disk(cam, rays, p, sh)
    rec, _ = _check(tracer, cam, p, sh, rays, out, "surrogate")
    assert ((rec["flags"] & bhs.FLAG_SURROGATE) > 0).mean() > 0.9
    sur.close()


def test_shade_1080p_frame_and_accumulation(tracer):
    """A whole 1920x1080 frame of the benchmark camera, then four Halton-jittered samples of a
    64x48 view accumulated on the GPU and by the oracle."""
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    cam = W.camera_for(W.C2, sample=0)
    p = W.params_for(W.C2)
    sh = bhs.disk_shading(doppler=1.0, time=3.25)
    print("[shade] 1920x1080", flush=True)
    rays = tracer.render(cam, p)
    out = tracer.shade_thin_disk(cam, rays, p, sh)
    _check(tracer, cam, p, sh, rays, out, "1080p")
    base = bhs.orbit_camera(30.0, 85.0, 60.0, 64, 48)
    acc = torch.zeros((3, 64 * 48), dtype=torch.float32, device=f"cuda:{tracer.device}")
    ref = np.zeros((3, 64 * 48), np.float32)
    sh2 = bhs.disk_shading()
    for s in range(4):
        cam_s = bhs.Camera(**{**base.__dict__, "jitter": bhs.halton_jitter(s)})
        rays = tracer.render(cam_s, bhs.march_params())
        cur = tracer.shade_thin_disk(cam_s, rays, bhs.march_params(), sh2)
        tracer.accumulate(cur.rgb, acc, s, cap=256)
        torch.cuda.synchronize()
        rgb, _ = O.shade_rays(cam_s.as_dict(), _oracle_params(bhs.march_params()),
                              _oracle_shading(sh2, bhs.march_params()), rays.numpy())
        ref = O.accumulate(rgb.T.copy(), ref, s, 256)
        np.testing.assert_array_equal(acc.cpu().numpy(), ref, err_msg=f"sample {s}")
