"""GPU parity of the gas lens table (bhs_gas_lens_table; shaders/gas_lens.comp:11-44).

The HIP kernel and the oracle (oracle_gas_lens_table) follow one operation order: law S's float
setup and RK4, the IEEE-only atan2, log / cos / sin evaluated in double and rounded to float. Bar:
bit-exact tables. A double result within one double ulp of a float rounding boundary could round
apart on the two libraries (probability ~1e-9 per evaluation); the tested tables show none.
Cases: the reference's 1024 x 1024 table at r_cam = 30, a close camera (r_cam = 4, a third of the
fan inside the shadow), one inside the photon sphere (r_cam = 2.5), odd sizes, argument errors.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cam_r,n_rays,n_phis", [(30.0, 1024, 1024), (4.0, 257, 129),
                                                 (2.5, 100, 300), (1000.0, 33, 2)])
def test_gas_lens_table_bit_exact(tracer, cam_r, n_rays, n_phis):
    import torch
    import blackholerender_amd as bhs
    print(f"[gas lens] r={cam_r} {n_rays}x{n_phis}", flush=True)
    ac = bhs.shadow_angular_radius_f32(cam_r)
    g = tracer.gas_lens_table(cam_r, ac, n_rays, n_phis)
    torch.cuda.synchronize()
    ref = O.gas_lens_table(cam_r, ac, n_rays, n_phis)
    np.testing.assert_array_equal(g.cpu().numpy(), ref)


def test_gas_lens_table_rejects_bad_arguments(tracer):
    import ctypes as C
    import torch
    t = torch.empty(64, dtype=torch.float32, device=f"cuda:{tracer.device}")
    L, ctx = tracer._lib, tracer._ctx
    assert L.bhs_gas_lens_table(ctx, 2.0, 0.5, 8, 8, t.data_ptr(), None) == -1     # at the horizon
    assert L.bhs_gas_lens_table(ctx, 30.0, 0.0, 8, 8, t.data_ptr(), None) == -1    # alpha_crit
    assert L.bhs_gas_lens_table(ctx, 30.0, 0.2, 8, 1, t.data_ptr(), None) == -1    # n_phis < 2
    assert L.bhs_gas_lens_table(ctx, 30.0, 0.2, 8, 8, None, None) == -1
    assert L.bhs_gas_lens_table(ctx, float("nan"), 0.2, 8, 8, t.data_ptr(), None) == -1


def test_gas_lens_tables_batch_equals_single_tables(tracer):
    """bhs_gas_lens_tables: four radii (far, the reference's 30, close, inside the photon sphere)
    in one launch equal the single-table calls bit for bit, and the oracle on the smallest."""
    import torch
    import blackholerender_amd as bhs
    radii = [1000.0, 30.0, 4.0, 2.5]
    print(f"[gas lens batch] radii={radii} 257x129", flush=True)
    tabs = tracer.gas_lens_tables(radii, None, 257, 129)
    torch.cuda.synchronize()
    for k, r in enumerate(radii):
        one = tracer.gas_lens_table(r, bhs.shadow_angular_radius_f32(r), 257, 129)
        torch.cuda.synchronize()
        assert torch.equal(tabs[k], one), r
    ref = O.gas_lens_table(2.5, bhs.shadow_angular_radius_f32(2.5), 257, 129)
    np.testing.assert_array_equal(tabs[3].cpu().numpy(), ref)


def test_gas_lens_tables_rejects_bad_arguments(tracer):
    import ctypes as C
    import torch
    t = torch.empty(4 * 64, dtype=torch.float32, device=f"cuda:{tracer.device}")
    L, ctx = tracer._lib, tracer._ctx
    r = (C.c_float * 2)(30.0, 2.0)
    a = (C.c_float * 2)(0.2, 0.2)
    assert L.bhs_gas_lens_tables(ctx, 2, r, a, 8, 8, t.data_ptr(), None) == -1     # at the horizon
    assert L.bhs_gas_lens_tables(ctx, 0, r, a, 8, 8, t.data_ptr(), None) == -1     # no table
    assert L.bhs_gas_lens_tables(ctx, 65, r, a, 8, 8, t.data_ptr(), None) == -1    # over 64
    assert L.bhs_gas_lens_tables(ctx, 1, r, a, 8, 8, None, None) == -1
    assert L.bhs_gas_lens_tables(ctx, 1, None, a, 8, 8, t.data_ptr(), None) == -1
