"""Host-side argument checks of the Python layer: a buffer whose size, dtype or device does not
match the call raises ValueError before any C-ABI function runs (the C side takes plane strides
from the camera sizes, so a short buffer would be read and written out of bounds on the GPU).

Runs on the CPU: the Tracer is built around a recording stand-in of the library, and the buffers
are CPU tensors, so no kernel can launch; every case asserts that the library was never called.
"""
import pytest
import torch

import blackholerender_amd as bhs
from blackholerender_amd.surrogate import Surrogate


class _RecordingLib:
    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def f(*a):
            self.calls.append(name)
            return 0
        return f


def _tracer():
    t = object.__new__(bhs.Tracer)
    t._lib = _RecordingLib()
    t._ctx = 1
    t.device = 0
    return t


def _cuda_like(n, **planes):
    """RayBuffers whose tensors are CPU tensors (the device check must reject them)."""
    b = object.__new__(bhs.RayBuffers)
    b.n = n
    b.cls = torch.empty(n, dtype=torch.uint8)
    b.flags = torch.empty(n, dtype=torch.uint8)
    b.n_cross = torch.empty(n, dtype=torch.uint8)
    b.steps = torch.empty(n, dtype=torch.int16)
    b.esc = torch.empty((3, n), dtype=torch.float32)
    b.disk = torch.empty((6, n), dtype=torch.float32)
    b.b = torch.empty(n, dtype=torch.float32)
    for k, v in planes.items():
        setattr(b, k, v)
    return b


CAM = bhs.orbit_camera(30.0, 85.0, 60.0, 16, 8)


def test_render_rejects_wrong_ray_count():
    t = _tracer()
    with pytest.raises(ValueError, match="hold 100 rays, the call needs 128"):
        t.render(CAM, bhs.march_params(), out=_cuda_like(100))
    assert t._lib.calls == []


def test_render_rejects_host_tensors():
    t = _tracer()
    with pytest.raises(ValueError, match="expected cuda:0"):
        t.render(CAM, bhs.march_params(), out=_cuda_like(128))
    assert t._lib.calls == []


def test_render_rejects_wrong_plane_shape():
    t = _tracer()
    with pytest.raises(ValueError, match="shape"):
        t.render(CAM, bhs.march_params(), out=_cuda_like(128, esc=torch.empty((3, 127))))
    assert t._lib.calls == []


def test_render_batch_rejects_wrong_total():
    t = _tracer()
    cams = [CAM, bhs.orbit_camera(15.0, 80.0, 60.0, 4, 4)]
    with pytest.raises(ValueError, match="the call needs 144"):
        t.render_batch(cams, bhs.march_params(), out=_cuda_like(128))
    assert t._lib.calls == []


def test_footprint_rejects_mismatched_rays_and_planes():
    t = _tracer()
    with pytest.raises(ValueError, match="sky_footprint: rays"):
        t.sky_footprint(CAM, _cuda_like(64))
    fp = object.__new__(bhs.FootprintBuffers)
    fp.n = 64
    for k in bhs.FootprintBuffers.ALL:
        setattr(fp, k, None)
    with pytest.raises(ValueError, match="sky_footprint: rays"):
        t.sky_footprint(CAM, _cuda_like(100), out=fp)
    cams = [CAM, CAM]
    with pytest.raises(ValueError, match="sky_footprint_batch: rays"):
        t.sky_footprint_batch(cams, _cuda_like(128))
    assert t._lib.calls == []


def test_gas_lens_rejects_bad_table():
    t = _tracer()
    with pytest.raises(ValueError, match="shape"):
        t.gas_lens_table(30.0, 0.2, 64, 65, out=torch.empty((64, 64), dtype=torch.float32))
    with pytest.raises(ValueError, match="dtype"):
        t.gas_lens_table(30.0, 0.2, 64, 65, out=torch.empty((65, 64), dtype=torch.float64))
    with pytest.raises(ValueError, match="expected cuda:0"):
        t.gas_lens_table(30.0, 0.2, 64, 65, out=torch.empty((65, 64), dtype=torch.float32))
    with pytest.raises(ValueError, match="contiguous"):
        t.gas_lens_table(30.0, 0.2, 64, 65,
                         out=torch.empty((64, 65), dtype=torch.float32).t())
    assert t._lib.calls == []


def test_gas_lens_tables_rejects_bad_tables():
    t = _tracer()
    with pytest.raises(ValueError, match="shape"):
        t.gas_lens_tables([30.0, 8.0], [0.2, 0.5], 64, 65,
                          out=torch.empty((1, 65, 64), dtype=torch.float32))
    with pytest.raises(ValueError, match="1..64"):
        t.gas_lens_tables([], [], 64, 65)
    with pytest.raises(ValueError, match="one alpha_crit"):
        t.gas_lens_tables([30.0, 8.0], [0.2], 64, 65)
    assert t._lib.calls == []


def test_march_rays_rejects_host_rays():
    t = _tracer()
    p = torch.zeros((10, 3), dtype=torch.float32)
    with pytest.raises(ValueError, match="march_rays: pos3"):
        t.march_rays(p, p.clone(), bhs.march_params(), out=_cuda_like(10))
    assert t._lib.calls == []


def test_surrogate_render_rejects_wrong_ray_count():
    t = _tracer()
    s = object.__new__(Surrogate)
    s.tracer = t
    s._lib = t._lib
    s._m = 1
    with pytest.raises(ValueError, match="the call needs 128"):
        s.render(CAM, bhs.march_params(), out=_cuda_like(64))
    with pytest.raises(ValueError, match="the call needs 256"):
        s.render_batch([CAM, CAM], bhs.march_params(), out=_cuda_like(128))
    assert t._lib.calls == []
