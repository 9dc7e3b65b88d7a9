"""bhs_render_noalloc (the render whose scratch is the library's static device array) against
bhs_render: two calls in a row on the same GPU, the second with a smaller image (the array keeps the
first call's band list and counters), every plane and the counters equal bit for bit; renders whose
scratch would exceed the static maximum are rejected before any launch."""
import pytest

pytestmark = pytest.mark.gpu


def _equal_planes(a, b):
    import torch
    out = {}
    for k in ("cls", "flags", "n_cross", "steps", "esc", "disk", "b"):
        x, y = getattr(a, k), getattr(b, k)
        if x.dtype != torch.uint8:
            x, y = x.view(torch.uint8), y.view(torch.uint8)
        out[k] = bool(torch.equal(x, y))
    return out


def test_noalloc_equals_render_on_two_calls(tracer):
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    cases = [(W.camera_for(W.C2, sample=3), W.params_for(W.C2)),
             (bhs.orbit_camera(15.0, 80.0, 60.0, 333, 197, jitter_hashed=1, seed=5),
              bhs.march_params(max_steps=700, disk_outer=30.0, root64=1))]
    for k, (cam, p) in enumerate(cases):
        print(f"[noalloc] call {k}: {cam.width}x{cam.height} root64={p.root64}", flush=True)
        n = cam.width * cam.height
        got = bhs.RayBuffers(n, f"cuda:{tracer.device}")
        got.cls.fill_(9)
        got.disk.fill_(-2.0)
        tracer.render_noalloc(cam, p, got)
        torch.cuda.synchronize()
        st_noalloc = tracer.last_stats()
        ref = tracer.render(cam, p)
        torch.cuda.synchronize()
        st_ref = tracer.last_stats()
        eq = _equal_planes(got, ref)
        assert all(eq.values()), (k, eq)
        assert st_noalloc == st_ref, (k, st_noalloc, st_ref)
        assert st_ref["rays"] == n and st_ref["band64"] > 0


def test_noalloc_rejects_oversized_scratch(tracer):
    """4096x4096 needs 64 MiB of band list, 2048x2048 with root64 235 MiB of root states: both are
    rejected by the host before any launch (full-size output buffers, so nothing could overrun)."""
    import blackholerender_amd as bhs
    dev = f"cuda:{tracer.device}"
    for w, root in ((4096, 0), (2048, 1)):
        cam = bhs.orbit_camera(30.0, 85.0, 60.0, w, w)
        with pytest.raises(bhs.BhsError, match="static maximum"):
            tracer.render_noalloc(cam, bhs.march_params(root64=root), bhs.RayBuffers(w * w, dev))
