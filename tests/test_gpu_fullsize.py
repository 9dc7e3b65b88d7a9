"""The benchmark workloads at their full sizes, against the CPU oracle.

* One 1920x1080 frame of the 1080p benchmark (camera 0 of its step: r_obs 30, inclination 85,
  FOV 60, 450 steps, disk [6, 20], seeded per-pixel jitter): the plain render (FP32 with the FP64
  band) bit-identical to the oracle on every ray; the surrogate render of the same frame (table and
  per-ray modes): served rays within 1e-4 rad of the oracle's FP64 march, with its class, escape
  flag and crossing count, crossing radii within 1e-3 (relative) of its refined FP64 crossings;
  every unserved ray bit-identical to the oracle.
* The full form (3840x2160, 8 jittered samples, 4 poses = 32 cameras, 265,420,800 rays, 700 steps,
  disk [6, 30]) in one batch call: the plain batch bit-identical to the oracle on a strided sample
  of every camera (about 8,300 rays each); the surrogate batch with the same bars as the frame on
  the same sample.
"""
import numpy as np
import pytest

import oracle as O
from helpers import (bit_mismatches, escape_angle_error, gather_rays, oracle_params_for,
                     oracle_threads, oracle_to_dict)

pytestmark = pytest.mark.gpu

MODELS = {450: "lawS_450_disk6-20.bhsmlp", 700: "lawS_700_disk6-30.bhsmlp"}


def _model_path(name):
    import os
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "blackholerender_amd", "models", name)


def _check_surrogate(S, band, f64, f64_root, tag):
    """Served rays against the oracle's FP64 march, unserved rays bit for bit against its band
    march (S, band, f64, f64_root: dicts in the RayBuffers.numpy() layout, same rays)."""
    import blackholerender_amd as bhs
    sv = (S["flags"] & bhs.FLAG_SURROGATE) > 0
    assert sv.mean() > 0.95, (tag, sv.mean())
    assert np.array_equal(S["cls"][sv], f64["cls"][sv]), tag
    assert np.array_equal(S["n_cross"][sv], f64["n_cross"][sv]), tag
    assert np.array_equal(S["flags"][sv] & bhs.FLAG_ESCAPED, f64["flags"][sv] & bhs.FLAG_ESCAPED)
    esc = sv & ((S["flags"] & bhs.FLAG_ESCAPED) > 0)
    err = escape_angle_error(S["esc"][esc], f64["esc"][esc])
    assert err.max() <= 1e-4, (tag, err.max())
    for k in range(2):
        ck = sv & (S["n_cross"] > k)
        if ck.any():
            rel = np.abs(S["disk_r"][ck, k].astype(np.float64) / f64_root["disk_r"][ck, k] - 1.0)
            assert rel.max() < 1e-3, (tag, k, rel.max())
    mm = bit_mismatches({k: v[~sv] for k, v in S.items()}, {k: v[~sv] for k, v in band.items()})
    assert all(v == 0 for v in mm.values()), (tag, mm)
    return int(sv.sum()), float(err.max()) if err.size else 0.0


@pytest.fixture(scope="module")
def c2_frame(tracer):
    from blackholerender_amd import workload as W
    cam = W.camera_for(W.C2, sample=0)
    th = oracle_threads()
    op = oracle_params_for(W.C2)
    print(f"[fullsize] oracle 1920x1080 band/f64/f64-root on {th} threads", flush=True)
    band = oracle_to_dict(O.render(cam.as_dict(), op, mode="band", threads=th))
    f64 = oracle_to_dict(O.render(cam.as_dict(), op, mode="f64", threads=th))
    f64r = oracle_to_dict(O.render(cam.as_dict(), oracle_params_for(W.C2, root64=1), mode="f64",
                                   threads=th))
    return cam, band, f64, f64r


def test_c2_frame_plain_bit_exact(tracer, c2_frame):
    import torch
    from blackholerender_amd import workload as W
    cam, band, _, _ = c2_frame
    print("[fullsize] plain 1920x1080 frame", flush=True)
    out = tracer.render(cam, W.params_for(W.C2))
    torch.cuda.synchronize()
    mm = bit_mismatches(out.numpy(), band)
    assert all(v == 0 for v in mm.values()), mm
    st = tracer.last_stats()
    assert st["rays"] == 1920 * 1080 and st["steps"] == int(band["steps"].sum())
    assert st["band64"] == int(((band["flags"] & 4) > 0).sum()) > 0


@pytest.mark.parametrize("table", [1, 0])
def test_c2_frame_surrogate(tracer, c2_frame, table):
    import torch
    from blackholerender_amd import workload as W
    from blackholerender_amd.surrogate import Surrogate
    cam, band, f64, f64r = c2_frame
    sur = Surrogate.load(tracer, _model_path(MODELS[450]))
    if table:
        sur.set_table(1024, 128)
    print(f"[fullsize] surrogate 1920x1080 frame table={table}", flush=True)
    S = sur.render(cam, W.params_for(W.C2)).numpy()
    torch.cuda.synchronize()
    served, err = _check_surrogate(S, band, f64, f64r, f"c2 table={table}")
    print(f"[fullsize] served {served}, max escape error {err:.2e} rad", flush=True)
    sur.close()


def _c5_sample(cams, stride=997):
    """Per camera, pixels 0, stride, 2 stride, ... (global ray indices and per-camera counts)."""
    n = cams[0].width * cams[0].height
    local = np.arange(0, n, stride, dtype=np.int64)
    return np.concatenate([k * n + local for k in range(len(cams))]), local.size, stride


@pytest.fixture(scope="module")
def c5_oracle(tracer):
    from blackholerender_amd import workload as W
    cams = W.cameras_c5()
    idx, per, stride = _c5_sample(cams)
    th = oracle_threads()
    print(f"[fullsize] oracle full-form sample ({idx.size} rays) on {th} threads", flush=True)
    refs = {}
    for mode, root in (("band", 0), ("f64", 0), ("f64", 1)):
        op = oracle_params_for(W.C5, root64=root)
        parts = [oracle_to_dict(O.render(c.as_dict(), op, mode=mode, stride=stride, threads=th))
                 for c in cams]
        refs[(mode, root)] = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
        assert refs[(mode, root)]["cls"].size == idx.size
    return cams, idx, refs


def test_c5_full_batch_plain_sampled(tracer, c5_oracle):
    import torch
    from blackholerender_amd import workload as W
    cams, idx, refs = c5_oracle
    print("[fullsize] plain full-form batch (265,420,800 rays)", flush=True)
    out = tracer.render_batch(cams, W.params_for(W.C5))
    torch.cuda.synchronize()
    st = tracer.last_stats()
    assert st["rays"] == 3840 * 2160 * 32
    g = gather_rays(out, idx)
    del out
    mm = bit_mismatches(g, refs[("band", 0)])
    assert all(v == 0 for v in mm.values()), mm


def test_c5_full_batch_surrogate_sampled(tracer, c5_oracle):
    import torch
    from blackholerender_amd import workload as W
    from blackholerender_amd.surrogate import Surrogate
    cams, idx, refs = c5_oracle
    sur = Surrogate.load(tracer, _model_path(MODELS[700]))
    sur.set_table(1024, 128)
    print("[fullsize] surrogate full-form batch", flush=True)
    out = sur.render_batch(cams, W.params_for(W.C5))
    torch.cuda.synchronize()
    assert tracer.last_stats()["surrogate"] > 0.95 * 3840 * 2160 * 32
    g = gather_rays(out, idx)
    del out
    served, err = _check_surrogate(g, refs[("band", 0)], refs[("f64", 0)], refs[("f64", 1)], "c5")
    print(f"[fullsize] full form: served {served} of {idx.size} sampled, max escape error "
          f"{err:.2e} rad", flush=True)
    sur.close()
