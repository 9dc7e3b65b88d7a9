"""CPU pins of the oracle's sky footprint (shaders/blackhole.frag:297-306, shaders/sky.glsl:27-51).

The reference holds no tests for this pass, so the restatement is pinned by what the mathematics
and the construction fix:
  * makeFootprint is a QR factorisation of the 3x2 Jacobian J = [a b]: e1, e2 orthonormal,
    a = la e1, b = b1 e1 + lbp e2, the inverse R factor inverts R, Rinv Rinv^T inverts the Gram
    matrix of (a, b), area^2 is its determinant |a x b|^2, mu = omega0 / area; the clamps give the
    closed forms the reference states (axis floor 1e-3 sqrt(omega0), mu cap 1e4);
  * on an unlensed field (every pixel escaped along its own view direction) the footprint is the
    pixel's solid angle on a flat image plane, omega0 cos^3(theta), so mu = 1 / cos^3(theta) to the
    finite-difference error;
  * quads: the fallback pattern of status edges, fine against coarse derivatives, and odd image
    edges follow from the quad rule on images built for the purpose, and every derivative equals
    the difference of the inputs that rule names.
"""
import numpy as np
import pytest

import oracle as O


def _cam(w, h, fov=60.0, hashed=0, r=30.0, inc=85.0):
    import blackholerender_amd as bhs
    return bhs.orbit_camera(r, inc, fov, w, h, jitter_hashed=hashed, seed=7).as_dict()


def _dirs(cam, cols, rows):
    return np.array([[O.ray_dir_at(cam, c, r) for c in cols] for r in rows], np.float32)


def test_make_footprint_is_a_qr_factorisation():
    rng = np.random.default_rng(3)
    omega0 = np.float32(1e-5)
    for _ in range(200):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        a = rng.normal(size=3) * 10 ** rng.uniform(-3, -1)
        b = rng.normal(size=3) * 10 ** rng.uniform(-3, -1)
        f = O.make_footprint(d, a, b, omega0)
        e1, e2 = f["e1"].astype(np.float64), f["e2"].astype(np.float64)
        r11, r12, r22 = f["rinv"].astype(np.float64)
        a64, b64 = np.float32(a).astype(np.float64), np.float32(b).astype(np.float64)
        la, b1, lbp = e1 @ a64, e1 @ b64, e2 @ b64
        # float rounding of b's perpendicular part grows with kappa = |b| / lbp
        kappa = np.linalg.norm(b64) / lbp
        assert abs(e1 @ e1 - 1) < 1e-6 and abs(e2 @ e2 - 1) < 1e-6
        assert abs(e1 @ e2) < 1e-6 * kappa
        np.testing.assert_allclose(la * e1, a64, atol=2e-6 * np.linalg.norm(a64))
        np.testing.assert_allclose(b1 * e1 + lbp * e2, b64, atol=2e-6 * np.linalg.norm(b64))
        Rinv = np.array([[r11, r12], [0.0, r22]])
        R = np.array([[la, b1], [0.0, lbp]])
        P = Rinv @ R
        assert abs(P[0, 0] - 1) < 1e-6 and abs(P[1, 1] - 1) < 4e-6 * kappa
        assert abs(P[0, 1]) < 4e-6 * kappa * (1 + abs(b1) / la) and P[1, 0] == 0
        G = np.array([[a64 @ a64, a64 @ b64], [a64 @ b64, b64 @ b64]])
        np.testing.assert_allclose((Rinv @ Rinv.T) @ G, np.eye(2), atol=1e-5 * kappa ** 2)
        cross = np.cross(a64, b64)
        assert abs(float(f["area"]) ** 2 / (cross @ cross) - 1) < 1e-5 * kappa
        assert abs(float(f["mu"]) - min(float(omega0) / float(f["area"]), 1e4)) <= \
            1e-6 * float(f["mu"])


def test_make_footprint_clamps():
    omega0 = np.float32(4e-6)
    min_axis = 1e-3 * np.sqrt(4e-6)
    d = np.array([0.0, 0.0, 1.0], np.float32)
    # a = 0: e1 = normalize(cross(d, (1e-3, 1.001, 1e-3))), la floored at 1e-3 sqrt(omega0)
    f = O.make_footprint(d, [0, 0, 0], [0, 1e-3, 0], omega0)
    up = np.array([1e-3, 1.0 + 1e-3, 1e-3])
    e1 = np.cross(d.astype(np.float64), up)
    e1 /= np.linalg.norm(e1)
    np.testing.assert_allclose(f["e1"], e1, atol=1e-6)
    assert f["rinv"][0] == pytest.approx(1 / min_axis, rel=1e-6)
    # b parallel to a: e2 = normalize(cross(d, e1)), lbp floored
    f = O.make_footprint(d, [2e-3, 0, 0], [4e-3, 0, 0], omega0)
    np.testing.assert_allclose(f["e1"], [1, 0, 0], atol=0)
    np.testing.assert_allclose(f["e2"], np.cross([0, 0, 1.0], [1, 0, 0]), atol=1e-7)
    assert f["rinv"][2] == pytest.approx(1 / min_axis, rel=1e-6)
    assert f["area"] == pytest.approx(2e-3 * min_axis, rel=1e-6)
    # magnification cap
    f = O.make_footprint(d, [1e-9, 0, 0], [0, 1e-9, 0], omega0)
    assert f["mu"] == np.float32(1e4)


def test_unlensed_field_gives_the_flat_image_plane_solid_angle():
    """Every pixel escaped along its own view direction: mu = 1 / cos^3(theta)."""
    w, h = 256, 192
    cam = _cam(w, h, fov=60.0)
    D = _dirs(cam, range(w), range(h)).reshape(-1, 3)
    flags = np.full(w * h, 2, np.uint8)
    fp = O.sky_footprint(cam, flags, D)
    assert not fp["fallback"].any()
    back = np.asarray(cam["basis"], np.float64)[6:9]          # column 3: camera looks down -back
    cos_t = -(D.astype(np.float64) @ back)
    rel = fp["mu"] * cos_t ** 3 - 1.0
    pa = 2 * np.tan(np.radians(30.0)) / h
    assert np.abs(rel).max() < 3 * pa, np.abs(rel).max()      # forward differences: O(pixel angle)
    # and the derivatives are the forward differences of the view directions within each quad
    img = D.reshape(h, w, 3)
    ddx = fp["ddx"].reshape(h, w, 3)
    np.testing.assert_array_equal(ddx[:, 0::2], img[:, 1::2] - img[:, 0::2])
    np.testing.assert_array_equal(ddx[:, 1::2], img[:, 1::2] - img[:, 0::2])
    ddy = fp["ddy"].reshape(h, w, 3)
    np.testing.assert_array_equal(ddy[0::2], img[1::2] - img[0::2])
    np.testing.assert_array_equal(ddy[1::2], img[1::2] - img[0::2])


def _synthetic(cam, escaped):
    """Escape directions: the view direction turned by a smooth, visible deflection."""
    w, h = cam["width"], cam["height"]
    D = _dirs(cam, range(w), range(h)).reshape(-1, 3)
    rot = np.array([[np.cos(0.3), 0, np.sin(0.3)], [0, 1, 0], [-np.sin(0.3), 0, np.cos(0.3)]])
    E = (D.astype(np.float64) @ rot.T).astype(np.float32)
    flags = np.where(escaped.reshape(-1), 2, 0).astype(np.uint8)
    return D, E, flags


@pytest.mark.parametrize("coarse", [False, True])
def test_status_edges_fall_back_to_the_unlensed_derivatives(coarse):
    w, h = 10, 6
    cam = _cam(w, h, hashed=1)
    esc = np.zeros((h, w), bool)
    esc[:, :5] = True                 # vertical edge inside the quads of columns 4-5
    esc[4:, :] = False                # horizontal edge between quad rows (rows 3 | 4): no fallback
    esc[1, 8] = True                  # lone pixel at (8, 1): quad (8..9, 0..1)
    D, E, flags = _synthetic(cam, esc)
    fp = O.sky_footprint(cam, flags, E, coarse=coarse)
    assert set(np.unique(fp["fallback"])) <= {0, 1}   # even sizes: no helper pixels
    fb = fp["fallback"].reshape(h, w).astype(bool)
    want = np.zeros((h, w), bool)
    want[:4, 4:6] = True
    if coarse:
        # every pixel of the quad uses its first row and column: (8,0)-(8,1) mixes
        want[0:2, 8:10] = True
    else:
        want[1, 8:10] = True          # row pair (8,1)-(9,1) mixes
        want[0:2, 8] = True           # column pair (8,0)-(8,1) mixes
    np.testing.assert_array_equal(fb, want)
    Eimg = np.where(esc[..., None], E.reshape(h, w, 3), D.reshape(h, w, 3))
    Dimg = D.reshape(h, w, 3)
    ddx, ddy = fp["ddx"].reshape(h, w, 3), fp["ddy"].reshape(h, w, 3)
    for y in range(h):
        for x in range(w):
            x0, y0 = x & ~1, y & ~1
            ry, cx = (y0, x0) if coarse else (y, x)
            src = Dimg if fb[y, x] else Eimg
            np.testing.assert_array_equal(ddx[y, x], src[ry, x0 + 1] - src[ry, x0])
            np.testing.assert_array_equal(ddy[y, x], src[y0 + 1, cx] - src[y0, cx])
    # footprints exist for escaped pixels only
    assert (fp["area"].reshape(h, w)[~esc] == 0).all() and (fp["area"].reshape(h, w)[esc] > 0).all()


def test_odd_edges_count_as_not_escaped():
    w, h = 7, 5
    cam = _cam(w, h)
    esc = np.ones((h, w), bool)
    D, E, flags = _synthetic(cam, esc)
    fp = O.sky_footprint(cam, flags, E)
    fb = fp["fallback"].reshape(h, w)
    # the last column's row pairs and the last row's column pairs run through a helper pixel
    # beyond the edge: fallback 2, told apart from status edges (1)
    want = np.zeros((h, w), np.uint8)
    want[:, 6] = 2
    want[4, :] = 2
    np.testing.assert_array_equal(fb, want)
    # a status edge inside the image next to the odd column stays 1: pixel (5, 1) not escaped
    esc2 = esc.copy()
    esc2[1, 5] = False
    D2, E2, flags2 = _synthetic(cam, esc2)
    fb2 = O.sky_footprint(cam, flags2, E2)["fallback"].reshape(h, w)
    assert fb2[1, 4] == 1 and fb2[1, 5] == 1 and fb2[0, 5] == 1   # row pair (4,5), column pair
    assert fb2[1, 6] == 2 and fb2[0, 4] == 0
    # the missing partner's view direction is the ray generation continued past the edge
    ddx = fp["ddx"].reshape(h, w, 3)
    np.testing.assert_array_equal(ddx[2, 6], O.ray_dir_at(cam, 7, 2) - O.ray_dir_at(cam, 6, 2))
    ddy = fp["ddy"].reshape(h, w, 3)
    np.testing.assert_array_equal(ddy[4, 3], O.ray_dir_at(cam, 3, 5) - O.ray_dir_at(cam, 3, 4))


def test_lensed_render_magnifies_near_the_shadow():
    """On a real render the magnification exceeds the unlensed value near the photon ring."""
    cam = _cam(96, 96, fov=40.0, r=30.0, inc=85.0)
    rays = O.render(cam, O.lawS_params(disk_mode=0), threads=8)
    fp = O.sky_footprint(cam, rays["flags"].astype(np.uint8), rays["esc"])
    esc = (rays["flags"] & 2) > 0
    assert esc.mean() > 0.5
    mu = fp["mu"][esc & (fp["fallback"] == 0)]
    assert mu.max() > 5.0            # strong lensing close to the shadow edge
    assert np.isfinite(fp["mu"]).all() and (fp["mu"] <= 1e4).all()
