"""Python view of the learned surrogate (bhs_mlp_* and bhs_render_surrogate of include/bhs_trace.h).

The network, its training and its inference run in the library's HIP kernels; this class only
passes arguments and buffers. No counterpart exists in the reference repository.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from .camera import Camera

MLP_PARAMS = _capi.MLP_PARAMS


def _struct_dict(s) -> dict:
    out = {}
    for name, _ in s._fields_:
        if name.endswith("_"):
            continue
        v = getattr(s, name)
        if isinstance(v, C.Array):
            v = list(v)
        elif isinstance(v, C.Structure):
            v = _struct_dict(v)
        out[name] = v
    return out


class Surrogate:
    """One bhs_mlp model on the device of `tracer` (a blackholerender_amd.Tracer)."""

    def __init__(self, tracer, trained_for: _capi.MarchParams | None = None, seed: int = 1,
                 _handle=None):
        from . import march_params
        self.tracer = tracer
        self._lib = tracer._lib
        if _handle is not None:
            self._m = _handle
            return
        if trained_for is None:
            trained_for = march_params()
        st = C.c_int(0)
        self._m = self._lib.bhs_mlp_create(tracer._ctx, C.byref(trained_for), seed, C.byref(st))
        if not self._m:
            tracer._check(st.value, "bhs_mlp_create")

    @classmethod
    def load(cls, tracer, path: str) -> "Surrogate":
        st = C.c_int(0)
        h = tracer._lib.bhs_mlp_load(tracer._ctx, path.encode(), C.byref(st))
        if not h:
            tracer._check(st.value, "bhs_mlp_load")
        return cls(tracer, _handle=h)

    def close(self):
        if getattr(self, "_m", None):
            self._lib.bhs_mlp_destroy(self._m)
            self._m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        self.tracer._check(rc, what)

    def make_dataset(self, n_rays: int, seed: int = 1):
        self._check(self._lib.bhs_mlp_make_dataset(self.tracer._ctx, self._m, n_rays, seed),
                    "bhs_mlp_make_dataset")

    def train(self, steps: int, batch: int = 8192, lr_start: float = 2e-3,
              lr_end: float = 2e-5) -> np.ndarray:
        loss = np.zeros(max(steps, 1), dtype=np.float32)
        self._check(self._lib.bhs_mlp_train(self.tracer._ctx, self._m, steps, batch, lr_start,
                                            lr_end, loss.ctypes.data), "bhs_mlp_train")
        return loss[:steps]

    def grads(self, x, y, mask):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1, 5)
        mask = np.ascontiguousarray(mask, dtype=np.float32).reshape(-1, 5)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = C.c_float()
        self._check(self._lib.bhs_mlp_grads(self.tracer._ctx, self._m, x.ctypes.data, y.ctypes.data,
                                            mask.ctypes.data, x.shape[0], g.ctypes.data,
                                            C.byref(loss)), "bhs_mlp_grads")
        return g, float(loss.value)

    def forward(self, x) -> np.ndarray:
        import torch
        xt = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3),
                             device=f"cuda:{self.tracer.device}")
        yt = torch.empty((xt.shape[0], 5), dtype=torch.float32, device=xt.device)
        stream = torch.cuda.current_stream(xt.device).cuda_stream
        self._check(self._lib.bhs_mlp_forward(self.tracer._ctx, self._m, xt.data_ptr(),
                                              xt.shape[0], yt.data_ptr(), C.c_void_p(stream)),
                    "bhs_mlp_forward")
        torch.cuda.synchronize(xt.device)
        return yt.cpu().numpy()

    def validate(self, n_rays: int, seed: int = 2, tol_phi: float = 1e-4) -> dict:
        v = _capi.MlpValidation()
        self._check(self._lib.bhs_mlp_validate(self.tracer._ctx, self._m, n_rays, seed, tol_phi,
                                               C.byref(v)), "bhs_mlp_validate")
        return _struct_dict(v)

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        self._check(self._lib.bhs_mlp_get_params(self.tracer._ctx, self._m, p.ctypes.data,
                                                 MLP_PARAMS), "bhs_mlp_get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).ravel()
        self._check(self._lib.bhs_mlp_set_params(self.tracer._ctx, self._m, p.ctypes.data, p.size),
                    "bhs_mlp_set_params")

    def info(self) -> dict:
        i = _capi.MlpInfo()
        self._check(self._lib.bhs_mlp_get_info(self.tracer._ctx, self._m, C.byref(i)),
                    "bhs_mlp_get_info")
        return _struct_dict(i)

    def set_table(self, n2: int, n3: int = 0):
        """Table mode on an n2 x n3 (b/b_crit, psi) grid per render; n2 = 0: per-ray inference."""
        self._check(self._lib.bhs_mlp_set_table(self.tracer._ctx, self._m, n2, n3),
                    "bhs_mlp_set_table")

    def save(self, path: str):
        self._check(self._lib.bhs_mlp_save(self.tracer._ctx, self._m, path.encode()),
                    "bhs_mlp_save")

    def render_batch(self, cams, params, out=None, stream=None):
        """Surrogate render of up to 64 cameras in one call (bhs_render_surrogate_batch), outputs
        laid out as Tracer.render_batch; returns the RayBuffers."""
        from . import RayBuffers
        t = self.tracer
        cs, ps = t._batch_structs(cams, params)
        n = sum(c.width * c.height for c in cams)
        if out is None:
            out = RayBuffers(n, f"cuda:{t.device}")
        out.check(n, t.device, "surrogate render_batch: out")
        soa = out.soa()
        self._check(self._lib.bhs_render_surrogate_batch(
            t._ctx, self._m, C.addressof(cs), C.addressof(ps), len(cams), C.byref(soa),
            C.c_void_p(t._stream_ptr(stream))), "bhs_render_surrogate_batch")
        return out

    def render(self, cam: Camera, params: _capi.MarchParams | None = None, out=None,
               stream=None):
        """Surrogate render with integrator fallback (stream-ordered); returns the RayBuffers."""
        from . import RayBuffers, _camera_struct, march_params
        t = self.tracer
        if params is None:
            params = march_params()
        if out is None:
            out = RayBuffers(cam.width * cam.height, f"cuda:{t.device}")
        out.check(cam.width * cam.height, t.device, "surrogate render: out")
        cs = _camera_struct(cam)
        soa = out.soa()
        self._check(self._lib.bhs_render_surrogate(t._ctx, self._m, C.byref(cs), C.byref(params),
                                                   C.byref(soa), C.c_void_p(t._stream_ptr(stream))),
                    "bhs_render_surrogate")
        return out
