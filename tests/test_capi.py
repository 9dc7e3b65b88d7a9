"""The C-ABI library loads without a GPU and exports every function include/bhs_trace.h declares;
argument errors come back as status codes (no compute call is made)."""
import ctypes as C

import pytest

from blackholerender_amd import _capi


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    names = _capi.header_symbols()
    assert "bhs_trace_rays" in names and "bhs_render" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version():
    assert _capi.load().bhs_version() == 1


def test_null_context_is_rejected():
    lib = _capi.load()
    assert lib.bhs_sync(None) == -1
    assert lib.bhs_render(None, None, None, None, None) == -1
    assert lib.bhs_trace_rays(None, None, None, 0, 0, 0.0, 0.0, 0.0, None) == -1
    assert lib.bhs_last_error(None) == b"no context"


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    st = C.c_int(0)
    ctx = _capi.load().bhs_create(0, C.byref(st))
    assert not ctx and st.value == -2
