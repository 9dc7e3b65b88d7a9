"""GPU parity of the sky footprint post-pass (bhs_sky_footprint / bhs_sky_footprint_batch).

The HIP pass runs on the ray planes of a HIP render; the oracle (oracle_sky_footprint, the CPU
restatement of shaders/blackhole.frag:297-306 and shaders/sky.glsl:27-51) runs on host copies of
the same planes. Bar: bit-exact for every plane (both sides follow one IEEE operation order, with
correctly rounded divides and square roots). Cases: the 256x256 reference camera, hashed jitter at
odd sizes (missing quad partners; an even width with an odd height takes the paired path's
missing-row branch), 1x1 and 2x1 images, fine and coarse derivatives, a batch of
cameras of different sizes, NULL planes, and the planes of a surrogate render.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

PLANES = ("ddx", "ddy", "e1", "e2", "rinv", "area", "mu", "fallback")


def _check(g, ref, what):
    for k in PLANES:
        if k in g:
            np.testing.assert_array_equal(g[k], ref[k], err_msg=f"{what}: {k}")


@pytest.mark.parametrize("coarse", [False, True])
@pytest.mark.parametrize("w,h,hashed", [(256, 256, 0), (37, 29, 1), (38, 29, 1), (1, 1, 0),
                                         (2, 1, 1)])
def test_footprint_bit_exact(tracer, coarse, w, h, hashed):
    import torch
    import blackholerender_amd as bhs
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, w, h, jitter_hashed=hashed, seed=11)
    print(f"[footprint] {w}x{h} hashed={hashed} coarse={coarse}", flush=True)
    rays = tracer.render(cam, bhs.march_params())
    fp = tracer.sky_footprint(cam, rays, coarse=coarse)
    torch.cuda.synchronize()
    r = rays.numpy()
    ref = O.sky_footprint(cam.as_dict(), r["flags"], r["esc"], coarse=coarse)
    _check(fp.numpy(), ref, f"{w}x{h}")
    if w * h > 1000:
        esc = (r["flags"] & 2) > 0
        assert esc.any() and (~esc).any() and ref["fallback"].any()


def test_footprint_batch_equals_oracle_per_camera(tracer):
    import torch
    import blackholerender_amd as bhs
    cams = [bhs.orbit_camera(30.0, 85.0, 60.0, 64, 48, jitter_hashed=1, seed=3, camera_id=0),
            bhs.orbit_camera(12.0, 40.0, 75.0, 33, 17, jitter_hashed=1, seed=3, camera_id=1),
            bhs.orbit_camera(200.0, 89.0, 10.0, 128, 96, camera_id=2)]
    print("[footprint] batch of 3 cameras", flush=True)
    rays = tracer.render_batch(cams, bhs.march_params())
    fp = tracer.sky_footprint_batch(cams, rays)
    part = tracer.sky_footprint_batch(cams, rays, out=bhs.FootprintBuffers(
        rays.n, f"cuda:{tracer.device}", planes=("mu", "fallback")))
    torch.cuda.synchronize()
    r, g, gp = rays.numpy(), fp.numpy(), part.numpy()
    assert set(gp) == {"mu", "fallback"}
    offs = tracer.batch_offsets(cams)
    for k, cam in enumerate(cams):
        s = slice(offs[k], offs[k + 1])
        ref = O.sky_footprint(cam.as_dict(), r["flags"][s], r["esc"][s])
        _check({p: v[s] for p, v in g.items()}, ref, f"camera {k}")
        _check({p: v[s] for p, v in gp.items()}, ref, f"camera {k} (mu, fallback only)")


def test_footprint_of_a_surrogate_render(tracer):
    """The pass reads only flags and escape directions, so it serves surrogate renders alike."""
    import os
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd.surrogate import Surrogate
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sur = Surrogate.load(tracer, os.path.join(root, "blackholerender_amd", "models",
                                              "lawS_450_disk6-20.bhsmlp"))
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, 320, 180, jitter_hashed=1, seed=5)
    params = bhs.march_params()
    print("[footprint] surrogate render 320x180", flush=True)
    rays = sur.render(cam, params)
    fp = tracer.sky_footprint(cam, rays)
    torch.cuda.synchronize()
    r = rays.numpy()
    assert ((r["flags"] & bhs.FLAG_SURROGATE) > 0).mean() > 0.9
    _check(fp.numpy(), O.sky_footprint(cam.as_dict(), r["flags"], r["esc"]), "surrogate")


def test_footprint_rejects_missing_planes(tracer):
    import ctypes as C
    import blackholerender_amd as bhs
    from blackholerender_amd import _camera_struct
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, 8, 8)
    rays = bhs.RayBuffers(64, f"cuda:{tracer.device}")
    out = bhs.FootprintBuffers(64, f"cuda:{tracer.device}")
    cs, fs = _camera_struct(cam), out.soa()
    for plane in ("flags_dev", "esc_dev"):
        rs = rays.soa()
        setattr(rs, plane, None)
        assert tracer._lib.bhs_sky_footprint(tracer._ctx, C.byref(cs), C.byref(rs), 0,
                                             C.byref(fs), None) == -1
        assert tracer._lib.bhs_sky_footprint_batch(tracer._ctx, C.byref(cs), 1, C.byref(rs), 0,
                                                   C.byref(fs), None) == -1
    rs = rays.soa()
    assert tracer._lib.bhs_sky_footprint_batch(tracer._ctx, C.byref(cs), 0, C.byref(rs), 0,
                                               C.byref(fs), None) == -1


@pytest.mark.parametrize("coarse", [False, True])
def test_footprint_batch_paired_path_equals_oracle(tracer, coarse):
    """Every width even: the batch takes the paired-access kernel (the one the 1080p batches use);
    one camera has an odd height."""
    import torch
    import blackholerender_amd as bhs
    cams = [bhs.orbit_camera(30.0, 85.0, 60.0, 96, 54, jitter_hashed=1, seed=9, camera_id=0),
            bhs.orbit_camera(15.0, 60.0, 50.0, 34, 17, jitter_hashed=1, seed=9, camera_id=1),
            bhs.orbit_camera(30.0, 85.0, 60.0, 96, 54, jitter_hashed=1, seed=9, camera_id=2)]
    print(f"[footprint] paired batch coarse={coarse}", flush=True)
    rays = tracer.render_batch(cams, bhs.march_params())
    fp = tracer.sky_footprint_batch(cams, rays, coarse=coarse)
    torch.cuda.synchronize()
    r, g = rays.numpy(), fp.numpy()
    offs = tracer.batch_offsets(cams)
    for k, cam in enumerate(cams):
        s = slice(offs[k], offs[k + 1])
        ref = O.sky_footprint(cam.as_dict(), r["flags"][s], r["esc"][s], coarse=coarse)
        _check({p: v[s] for p, v in g.items()}, ref, f"camera {k}")


def test_footprint_1080p_single_and_batch(tracer):
    """The benchmarked shape: 1920x1080 frames of the 1080p benchmark (the paired kernel of even
    widths and its int64 plane offsets), one camera and a two-camera batch, every plane
    bit-identical to the oracle."""
    import torch
    from blackholerender_amd import workload as W
    cams = [W.camera_for(W.C2, sample=k) for k in range(2)]
    params = W.params_for(W.C2)
    print("[footprint] 1920x1080 single + 2-camera batch", flush=True)
    one = tracer.render(cams[0], params)
    fp1 = tracer.sky_footprint(cams[0], one)
    rays = tracer.render_batch(cams, params)
    fpb = tracer.sky_footprint_batch(cams, rays)
    torch.cuda.synchronize()
    r1 = one.numpy()
    _check(fp1.numpy(), O.sky_footprint(cams[0].as_dict(), r1["flags"], r1["esc"]), "1080p")
    r, g = rays.numpy(), fpb.numpy()
    n = 1920 * 1080
    for k in range(2):
        ref = O.sky_footprint(cams[k].as_dict(), r["flags"][k * n:(k + 1) * n],
                              r["esc"][k * n:(k + 1) * n])
        _check({p: v[k * n:(k + 1) * n] for p, v in g.items()}, ref, f"1080p batch camera {k}")
