"""GPU parity of the sky pass (bhs_sample_sky[_batch]) against the oracle, bit for bit.

The HIP pass and the oracle share one operation order: the quad derivatives and footprint of the
footprint pass, the star layers with the IEEE-only exp/log/pow of the shading and the correctly
rounded reciprocal square root, and the cube map lookup with the exact level of detail, trilinear
float weights and seamless edges. Cases: the 256x256 reference camera shaded by the thin disk
(stars only, weights from the shading); an odd-sized view (helper pixels) with a mip-mapped cube
map, the physical look and coarse derivatives; a three-camera batch; a whole 1920x1080 frame of the
benchmark on a strided sample.
"""
import numpy as np
import pytest

import oracle as O
from helpers import oracle_threads

pytestmark = pytest.mark.gpu


def _osky(k):
    return O.sky_params(k.sky_brightness, k.star_density, k.star_brightness, k.doppler)


def _cube(device, res=32, seed=0):
    import blackholerender_amd as bhs
    rng = np.random.default_rng(seed)
    faces = (rng.uniform(0.0, 0.05, (6, res, res, 3)) +
             0.2 * np.sin(np.linspace(0, 6, res))[None, :, None, None]).astype(np.float32)
    return bhs.SkyCube.from_faces(faces, device)


def _render_shaded(tracer, cam, p):
    import torch
    import blackholerender_amd as bhs
    rays = tracer.render(cam, p)
    sh = tracer.shade_thin_disk(cam, rays, p, bhs.disk_shading())
    torch.cuda.synchronize()
    return rays, sh


def test_sky_reference_camera_stars(tracer):
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    cam = W.camera_for(W.C1, sample=0)
    p = W.params_for(W.C1)
    rays, sh = _render_shaded(tracer, cam, p)
    k = bhs.sky_params(star_density=2.0, star_brightness=4.0)
    base = sh.rgb.clone()
    out = tracer.sample_sky(cam, rays, k, weight=sh.sky_weight, rgb=sh.rgb)
    torch.cuda.synchronize()
    r = rays.numpy()
    want = O.sky_pass(cam.as_dict(), r["flags"], r["esc"], _osky(k),
                      weight=sh.sky_weight.cpu().numpy(), rgb=base.cpu().numpy(),
                      threads=oracle_threads())
    np.testing.assert_array_equal(out.cpu().numpy(), want)
    assert (out.cpu().numpy() != base.cpu().numpy()).any()


def test_sky_odd_view_cube_physical_coarse(tracer):
    import torch
    import blackholerender_amd as bhs
    cam = bhs.orbit_camera(12.0, 70.0, 75.0, 97, 61, jitter_hashed=1, seed=8)
    p = bhs.march_params(max_steps=700)
    rays, sh = _render_shaded(tracer, cam, p)
    cube = _cube(f"cuda:{tracer.device}")
    k = bhs.sky_params(sky_brightness=2.0, star_density=1.5, star_brightness=2.0, doppler=1.0)
    for coarse in (False, True):
        out = tracer.sample_sky(cam, rays, k, cube=cube, weight=sh.sky_weight, coarse=coarse)
        torch.cuda.synchronize()
        r = rays.numpy()
        want = O.sky_pass(cam.as_dict(), r["flags"], r["esc"], _osky(k), cube=cube.numpy(),
                          weight=sh.sky_weight.cpu().numpy(), coarse=coarse,
                          threads=oracle_threads())
        np.testing.assert_array_equal(out.cpu().numpy(), want, err_msg=f"coarse={coarse}")


def test_sky_batch(tracer):
    import torch
    import blackholerender_amd as bhs
    cams = [bhs.orbit_camera(30.0, 85.0, 60.0, 40, 24, jitter_hashed=1, seed=9, camera_id=0),
            bhs.orbit_camera(12.0, 45.0, 75.0, 17, 33, jitter_hashed=1, seed=9, camera_id=1),
            bhs.orbit_camera(200.0, 88.0, 10.0, 32, 32, camera_id=2)]
    params = [bhs.march_params(max_steps=st) for st in (450, 700, 200)]
    n = sum(c.width * c.height for c in cams)
    rays = bhs.RayBuffers(n, f"cuda:{tracer.device}")
    tracer.render_batch(cams, params, rays)
    cube = _cube(f"cuda:{tracer.device}", res=16, seed=3)
    k = bhs.sky_params(star_density=3.0)
    out = tracer.sample_sky_batch(cams, rays, k, cube=cube)
    torch.cuda.synchronize()
    r, o = rays.numpy(), out.cpu().numpy()
    offs = tracer.batch_offsets(cams)
    for j, c in enumerate(cams):
        sl = slice(offs[j], offs[j + 1])
        want = O.sky_pass(c.as_dict(), r["flags"][sl], r["esc"][sl], _osky(k), cube=cube.numpy(),
                          threads=oracle_threads())
        np.testing.assert_array_equal(o[:, sl], want, err_msg=f"camera {j}")


def test_sky_1080p_frame_sampled(tracer):
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    cam = W.camera_for(W.C2, sample=0)
    p = W.params_for(W.C2)
    print("[sky] 1920x1080", flush=True)
    rays, sh = _render_shaded(tracer, cam, p)
    cube = _cube(f"cuda:{tracer.device}", res=256, seed=5)
    k = bhs.sky_params()
    base = sh.rgb.clone()
    out = tracer.sample_sky(cam, rays, k, cube=cube, weight=sh.sky_weight, rgb=sh.rgb)
    torch.cuda.synchronize()
    r = rays.numpy()
    want = O.sky_pass(cam.as_dict(), r["flags"], r["esc"], _osky(k), cube=cube.numpy(),
                      weight=sh.sky_weight.cpu().numpy(), rgb=base.cpu().numpy(), stride=101,
                      threads=oracle_threads())
    np.testing.assert_array_equal(out.cpu().numpy()[:, ::101], want[:, ::101])
