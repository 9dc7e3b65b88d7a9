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


@pytest.mark.gpu
def test_cpp_dropin_reference_checks():
    """The reference's geodesic checks in C++ through include/bhs/physics/Geodesic.h (trace on the
    GPU via the C-ABI); the binary is built by __graft_entry__.build()."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "geodesic_dropin_tests")
    assert os.path.exists(exe), "build tests/cpp first (make -C tests/cpp)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
