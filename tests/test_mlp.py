"""The learned surrogate: network oracle pins (CPU), GPU inference and gradients against the
oracle, training, and the surrogate render's fallback path against the plain integrator.

Bars: the FP64 oracle gradient agrees with central finite differences to 1e-4 relative; GPU
inference (bf16 MFMA, FP32 accumulation) agrees with the oracle's bf16-emulating forward pass to
2e-3 absolute (bf16 rounding flips of single activations) and 2e-5 at the median; GPU gradients
agree with the oracle's bf16-emulating gradient to 1e-2 relative (L2 norm per parameter block).
Rays the surrogate does not serve are marched by the integrator's own kernels and must be
bit-identical to the plain render.
"""
import numpy as np
import pytest

from oracle import mlp as OM

BLOCKS = [("W1", OM.W1, OM.B1), ("b1", OM.B1, OM.W2), ("W2", OM.W2, OM.B2), ("b2", OM.B2, OM.W3),
          ("W3", OM.W3, OM.B3), ("b3", OM.B3, OM.W4), ("W4", OM.W4, OM.B4), ("b4", OM.B4, OM.NPARAMS)]


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    t = rng.normal(0, 0.5, (n, 5))
    t[:, 3:] = rng.integers(0, 2, (n, 2))
    m = (rng.random((n, 5)) > 0.25).astype(np.float64)
    return x, t, m


# ------------------------------------------------------------------------------------------- CPU
def test_oracle_gradient_matches_finite_differences():
    p = OM.init_params(3)
    x, t, m = _batch(16, 0)
    _, g = OM.loss_grad(p, x, t, m, "exact")
    rng = np.random.default_rng(1)
    h = 1e-6
    for name, a, b in BLOCKS:
        for i in rng.choice(np.arange(a, b), min(5, b - a), replace=False):
            pp = p.copy()
            pp[i] += h
            lp, _ = OM.loss_grad(pp, x, t, m, "exact")
            pp[i] -= 2 * h
            lm, _ = OM.loss_grad(pp, x, t, m, "exact")
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-4 * max(abs(g[i]), 1e-3), (name, i, fd, g[i])


def test_oracle_closed_forms():
    assert OM.silu(0.0) == 0.0 and OM.dsilu(0.0) == 0.5
    z = np.array([-30.0, -2.0, 0.5, 40.0])
    assert np.allclose(OM.softplus(z), np.log1p(np.exp(z)))
    # bf16 rounding: exact values stay, ties go to even, 1 + 3 * 2^-9 rounds up
    assert OM.bf16(np.array([1.0, -2.5, 0.0]))[0] == 1.0
    assert OM.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    assert OM.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    assert OM.bf16(np.array([1.0 + 3 * 2.0 ** -9]))[0] == 1.0 + 2.0 ** -7


def test_oracle_bf16_mode_is_close_to_exact():
    p = OM.init_params(5)
    x, t, m = _batch(64, 2)
    _, ge = OM.loss_grad(p, x, t, m, "exact")
    _, gb = OM.loss_grad(p, x, t, m, "bf16")
    rel = np.linalg.norm(gb - ge) / np.linalg.norm(ge)
    assert 0 < rel < 0.05


# ------------------------------------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def model(tracer):
    from blackholerender_amd.surrogate import Surrogate
    s = Surrogate(tracer, seed=7)
    yield s
    s.close()


@pytest.mark.gpu
def test_gpu_forward_matches_oracle(model):
    p = OM.init_params(11).astype(np.float32)
    model.params = p
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (1000, 3)).astype(np.float32)   # not a multiple of 64: ragged tail
    print("[mlp] forward 1000 rows", flush=True)
    y = model.forward(x)
    ref, _ = OM.forward(p.astype(np.float64), x.astype(np.float64), "bf16")
    d = np.abs(y - ref)
    assert np.all(np.isfinite(y))
    assert d.max() <= 2e-3 * max(1.0, np.abs(ref).max()), d.max()
    assert np.median(d) <= 2e-5, np.median(d)


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle(model):
    p = OM.init_params(12).astype(np.float32)
    model.params = p
    x, t, m = _batch(777, 9)
    print("[mlp] gradient 777 rows", flush=True)
    g, loss = model.grads(x, t, m)
    lo, go = OM.loss_grad(p.astype(np.float64), x.astype(np.float32).astype(np.float64),
                          t.astype(np.float32), m.astype(np.float32), "bf16")
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    for name, a, b in BLOCKS:
        rel = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-30)
        assert rel <= 1e-2, (name, rel)
    # parameters untouched by a gradient-only call
    assert np.array_equal(model.params, p)


@pytest.mark.gpu
def test_gpu_training_reduces_loss(tracer):
    from blackholerender_amd.surrogate import Surrogate
    s = Surrogate(tracer, seed=3)
    print("[mlp] dataset 65536 rays", flush=True)
    s.make_dataset(1 << 16, seed=5)
    info = s.info()
    assert info["dataset_rays"] == 1 << 16
    assert 0 < max(info["steps_max"]) < 4096
    print("[mlp] train 400 steps", flush=True)
    loss = s.train(400, batch=2048, lr_start=2e-3, lr_end=2e-4)
    assert np.all(np.isfinite(loss))
    assert loss[-50:].mean() < 0.3 * loss[:10].mean(), (loss[:10].mean(), loss[-50:].mean())
    s.close()


@pytest.mark.gpu
def test_unvalidated_model_serves_nothing_and_falls_back_bit_exact(tracer, model):
    """An empty error map sends every ray through the fallback lists: the result must equal the
    plain render bit for bit (FP32 list march + FP64 band list + root refinement)."""
    import torch
    import blackholerender_amd as bhs
    from helpers import bit_mismatches
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, 200, 120)
    for root64 in (0, 1):
        params = bhs.march_params(root64=root64)
        print(f"[mlp] fallback-only render root64={root64}", flush=True)
        a = tracer.render(cam, params)
        torch.cuda.synchronize()
        ref = a.numpy()
        b = model.render(cam, params)
        torch.cuda.synchronize()
        st = tracer.last_stats()
        got = b.numpy()
        mm = bit_mismatches(got, ref)
        assert all(v == 0 for v in mm.values()), mm
        # FP64 band rays are marched from the start of the call; every other ray is a fallback
        assert st["surrogate"] == 0 and st["fallback"] + st["band64"] == 200 * 120
        assert st["rays"] == 200 * 120


@pytest.mark.gpu
def test_render_rejects_other_march_parameters(tracer, model):
    import blackholerender_amd as bhs
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, 64, 64)
    with pytest.raises(bhs.BhsError):
        model.render(cam, bhs.march_params(disk_outer=30.0))
    with pytest.raises(bhs.BhsError):
        model.render(cam, bhs.march_params(max_steps=700))   # other preset step control


# ------------------------------------------------------------------------- trained models (GPU)
MODEL_HIGH = "lawS_450_disk6-20.bhsmlp"


def _model_path(name):
    import os
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "blackholerender_amd", "models", name)


def _angle(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), (a * b).sum(1))


@pytest.mark.gpu
@pytest.mark.parametrize("table", [0, 1])
def test_trained_model_frame_accuracy(tracer, table):
    """480x270 frame of the 1080p camera (r_obs 30, i 85, FOV 60, High preset, disk [6, 20]):
    served escape directions within 1e-4 rad of FP64, classes and crossing counts equal to the
    FP64 reference march, crossing radii within 1e-3 (relative) of the FP64 march's refined
    crossings, every unserved ray bit-identical to the plain render. Per-ray inference and the
    1024 x 256 table mode."""
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd.surrogate import Surrogate
    from helpers import bit_mismatches
    sur = Surrogate.load(tracer, _model_path(MODEL_HIGH))
    if table:
        sur.set_table(1024, 256)
    cam = bhs.orbit_camera(30.0, 85.0, 60.0, 480, 270, jitter_hashed=1, seed=11)
    params = bhs.march_params()
    print(f"[mlp] trained model frame table={table}", flush=True)
    s = sur.render(cam, params)
    torch.cuda.synchronize()
    st = tracer.last_stats()
    S = s.numpy()
    P = tracer.render(cam, params).numpy()
    R = tracer.render(cam, bhs.march_params(precision=bhs.PREC_F64)).numpy()
    sv = (S["flags"] & bhs.FLAG_SURROGATE) > 0
    assert sv.mean() > 0.95, sv.mean()
    assert st["surrogate"] == sv.sum()
    esc = sv & ((S["flags"] & 2) > 0)
    assert np.array_equal(esc, sv & ((R["flags"] & 2) > 0))
    assert _angle(S["esc"][esc], R["esc"][esc]).max() <= 1e-4
    assert np.array_equal(S["cls"][sv], R["cls"][sv])
    assert np.array_equal(S["n_cross"][sv], R["n_cross"][sv])
    cr = sv & (S["n_cross"] > 0)
    rel = np.abs(S["disk_r"][cr, 0] / R["disk_r"][cr, 0] - 1.0)
    assert rel.max() < 5e-3, rel.max()
    RR = tracer.render(cam, bhs.march_params(precision=bhs.PREC_F64, root64=1)).numpy()
    for k in range(2):
        ck = sv & (S["n_cross"] > k)
        rel = np.abs(S["disk_r"][ck, k] / RR["disk_r"][ck, k] - 1.0)
        assert rel.max() < 1e-3, (k, rel.max())
    mm = bit_mismatches({k: v[~sv] for k, v in S.items()}, {k: v[~sv] for k, v in P.items()})
    assert all(v == 0 for v in mm.values()), mm
    sur.close()


@pytest.mark.gpu
def test_near_hole_frame_plunging_and_band_rays(tracer):
    """320x180 frame of the full-form camera (r_obs 15, i 80, FOV 60, 700 steps, disk [6, 30]; the
    shadow fills much of the view): plunging rays (b < 0.3 b_crit, decided by the exact orbit
    alone) and near-critical rays of the FP64 band that the model serves render the FP64 class and
    crossing count, served escape directions stay within 1e-4 rad of FP64, every unserved ray is
    bit-identical to the plain render."""
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd.surrogate import Surrogate
    from helpers import bit_mismatches
    sur = Surrogate.load(tracer, _model_path("lawS_700_disk6-30.bhsmlp"))
    sur.set_table(1024, 256)
    cam = bhs.orbit_camera(15.0, 80.0, 60.0, 320, 180, jitter_hashed=1, seed=21)
    params = bhs.march_params(max_steps=700, disk_outer=30.0)
    print("[mlp] near-hole frame", flush=True)
    S = sur.render(cam, params).numpy()
    torch.cuda.synchronize()
    P = tracer.render(cam, params).numpy()
    R = tracer.render(cam, bhs.march_params(max_steps=700, disk_outer=30.0,
                                            precision=bhs.PREC_F64)).numpy()
    sv = (S["flags"] & bhs.FLAG_SURROGATE) > 0
    beta = P["b"] / (3.0 * np.sqrt(3.0))
    plunge = beta < 0.3
    near = np.abs(beta - 1.0) < 0.02
    assert plunge.sum() > 300 and (sv & plunge).sum() > 0.9 * plunge.sum()
    assert (sv & near).sum() > 0.5 * near.sum(), ((sv & near).sum(), near.sum())
    assert np.array_equal(S["cls"][sv], R["cls"][sv])
    assert np.array_equal(S["n_cross"][sv], R["n_cross"][sv])
    assert np.array_equal(S["flags"][sv] & 2, R["flags"][sv] & 2)
    esc = sv & ((S["flags"] & 2) > 0)
    assert _angle(S["esc"][esc], R["esc"][esc]).max() <= 1e-4
    mm = bit_mismatches({k: v[~sv] for k, v in S.items()}, {k: v[~sv] for k, v in P.items()})
    assert all(v == 0 for v in mm.values()), mm
    sur.close()


@pytest.mark.gpu
def test_batch_equals_single_calls(tracer):
    """bhs_render_surrogate_batch over four cameras (three sizes and poses, the fourth a jittered
    copy of the first's pose, which shares its table) equals the single-camera calls bit for bit
    (table mode)."""
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd.surrogate import Surrogate
    from helpers import bit_mismatches
    sur = Surrogate.load(tracer, _model_path(MODEL_HIGH))
    sur.set_table(256, 64)
    cams = [bhs.orbit_camera(30.0, 85.0, 60.0, 200, 120, jitter_hashed=1, seed=3, camera_id=0),
            bhs.orbit_camera(60.0, 70.0, 40.0, 96, 160, jitter_hashed=1, seed=3, camera_id=1,
                             azimuth_deg=90.0),
            bhs.orbit_camera(12.0, 80.0, 75.0, 128, 64, jitter_hashed=1, seed=3, camera_id=2),
            bhs.orbit_camera(30.0, 85.0, 60.0, 200, 120, jitter_hashed=1, seed=4, camera_id=3)]
    params = bhs.march_params(root64=1)
    print("[mlp] surrogate batch of 4 cameras", flush=True)
    b = sur.render_batch(cams, params)
    torch.cuda.synchronize()
    B = b.numpy()
    off = tracer.batch_offsets(cams)
    for k, cam in enumerate(cams):
        one = sur.render(cam, params)
        torch.cuda.synchronize()
        O = one.numpy()
        part = {key: v[off[k]:off[k + 1]] for key, v in B.items()}
        mm = bit_mismatches(part, O)
        assert all(v == 0 for v in mm.values()), (k, mm)
    sur.close()


@pytest.mark.gpu
def test_network_gain_over_prior_alone(tracer):
    """What the trained network adds to the exact-orbit prior: on the same validation rays, the
    share of in-domain rays whose (r_obs, b) cell meets the 1e-4 rad escape tolerance, with the
    trained weights and with all weights zero (outputs 0: the prior alone). The network must serve
    over 85% and gain more than 12 points over the prior."""
    from blackholerender_amd.surrogate import Surrogate, MLP_PARAMS
    net = Surrogate.load(tracer, _model_path(MODEL_HIGH))
    prior = Surrogate.load(tracer, _model_path(MODEL_HIGH))
    prior.params = np.zeros(MLP_PARAMS, dtype=np.float32)
    print("[mlp] validation with and without the network", flush=True)
    vn = net.validate(1 << 20, seed=77)
    vp = prior.validate(1 << 20, seed=77)
    share_n = vn["served"] / vn["in_domain"]
    share_p = vp["served"] / vp["in_domain"]
    print(f"served share: network {share_n:.4f}, prior alone {share_p:.4f}")
    assert vn["mismatches"] == 0 and vp["mismatches"] == 0
    assert vn["phi_err_max"] <= 1e-4
    # measured on the committed model: 0.870 against 0.706 (seeded validation rays)
    assert share_n > 0.85 and share_n > share_p + 0.12, (share_n, share_p)
    net.close()
    prior.close()


@pytest.mark.gpu
def test_throughput_mode_batch_equals_single_calls(tracer):
    """Above 2^22 work items a table-mode call runs the two-phase pass (near-critical chunks
    first, no pre-pass); two 2048x1536 cameras (6.3 M items) must equal the two single-camera
    calls (3.1 M items each: pre-pass mode) bit for bit."""
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd.surrogate import Surrogate
    sur = Surrogate.load(tracer, _model_path(MODEL_HIGH))
    sur.set_table(256, 64)
    cams = [bhs.orbit_camera(30.0, 85.0, 60.0, 2048, 1536, jitter_hashed=1, seed=5, camera_id=k,
                             azimuth_deg=45.0 * k) for k in range(2)]
    params = bhs.march_params()
    print("[mlp] two-phase batch 2 x 2048x1536", flush=True)
    b = sur.render_batch(cams, params)
    torch.cuda.synchronize()
    st = tracer.last_stats()
    assert st["rays"] == 2 * 2048 * 1536 and st["surrogate"] > 0.9 * st["rays"]
    off = tracer.batch_offsets(cams)
    for k, cam in enumerate(cams):
        one = sur.render(cam, params)
        torch.cuda.synchronize()
        for name in ("cls", "flags", "n_cross", "steps", "b"):
            assert torch.equal(getattr(b, name)[off[k]:off[k + 1]], getattr(one, name)), (k, name)
        for name in ("esc", "disk"):
            assert torch.equal(getattr(b, name)[:, off[k]:off[k + 1]], getattr(one, name)), (k, name)
    sur.close()


@pytest.mark.gpu
def test_long_band_fallback_list_equals_plain(tracer):
    """A surrogate batch whose FP64 band list exceeds 262,144 rays (marched by the five-wave band
    variant): two 1024x1024 cameras at r_obs = 1200 (outside the served domain, so every ray falls
    back) with a 1 degree field on the shadow and a 25% band; the output equals the plain batch
    render (the other band variant) bit for bit."""
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd.surrogate import Surrogate
    from helpers import bit_mismatches
    sur = Surrogate.load(tracer, _model_path(MODEL_HIGH))
    sur.set_table(256, 64)
    cams = [bhs.orbit_camera(1200.0, 80.0, 1.0, 1024, 1024, jitter_hashed=1, seed=11, camera_id=k,
                             azimuth_deg=20.0 * k) for k in range(2)]
    params = bhs.march_params(band_rel=0.25)
    print("[mlp] surrogate batch 2 x 1024x1024 at r_obs 1200, 25% band", flush=True)
    b = sur.render_batch(cams, params)
    torch.cuda.synchronize()
    st = tracer.last_stats()
    assert st["surrogate"] == 0 and st["band64"] > 262144, st
    p = tracer.render_batch(cams, params)
    torch.cuda.synchronize()
    mm = bit_mismatches(b.numpy(), p.numpy())
    assert all(v == 0 for v in mm.values()), mm
    sur.close()


@pytest.mark.gpu
def test_large_single_camera_two_phase_equals_batch(tracer):
    """A single-camera call above 2^22 work items runs the two-phase pass in its one-camera form;
    a 2560x2048 camera alone must equal its slice of a batch with a second, small camera (the
    camera-table form of the same pass) bit for bit."""
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd.surrogate import Surrogate
    from helpers import bit_mismatches
    sur = Surrogate.load(tracer, _model_path(MODEL_HIGH))
    sur.set_table(1024, 128)
    big = bhs.orbit_camera(30.0, 85.0, 60.0, 2560, 2048, jitter_hashed=1, seed=7, camera_id=0)
    small = bhs.orbit_camera(40.0, 60.0, 50.0, 128, 96, jitter_hashed=1, seed=7, camera_id=1)
    params = bhs.march_params()
    print("[mlp] one-camera two-phase call 2560x2048", flush=True)
    one = sur.render(big, params)
    torch.cuda.synchronize()
    st = tracer.last_stats()
    assert st["rays"] == 2560 * 2048 and st["surrogate"] > 0.9 * st["rays"], st
    O = one.numpy()
    b = sur.render_batch([big, small], params)
    torch.cuda.synchronize()
    B = b.numpy()
    off = tracer.batch_offsets([big, small])
    part = {key: v[off[0]:off[1]] for key, v in B.items()}
    mm = bit_mismatches(part, O)
    assert all(v == 0 for v in mm.values()), mm
    sur.close()
