"""Closed forms of the camera, and the sampler's sine and cosine.

The reference's only view is glm::lookAt(eye = (2, 2, 2), at = 0, up = (0, 0, 1))
with a 45 degree vertical field of view (src/Renderer/Renderer.cpp:373-380).
For that camera the basis is known by hand:
  forward = -(1, 1, 1) / sqrt(3), right = (-1, 1, 0) / sqrt(2),
  up' = cross(right, forward) = (-1, -1, 2) / sqrt(6).
The checks below do not restate the ray formula; they test what any correct
pinhole camera with that convention must satisfy:
  * the image centre looks along forward;
  * the middle of the top edge lies fov/2 = 22.5 degrees above forward (toward
    up'), the middle of the bottom edge 22.5 degrees below: row 0 is the top,
    and the field of view is vertical;
  * the middle of the right edge lies atan(tan(22.5 deg) * aspect) to the
    right (toward `right`), the left edge as far to the left;
  * every direction is a unit vector in the plane spanned by the offsets.
A horizontal-for-vertical field of view, a flipped row order or a swapped
right vector each fails one of them.  They are applied to the oracle's
camera (orc_camera_rays), to the product's host basis (fnd_camera_basis) and
to the host ray sets of the traversal benchmark (rays.primary_rays).
"""
import numpy as np
import pytest

from foundation_amd import _abi as A
from foundation_amd import pt
from foundation_amd import rays as R

FWD = -np.ones(3) / np.sqrt(3.0)
RIGHT = np.array([-1.0, 1.0, 0.0]) / np.sqrt(2.0)
UP = np.array([-1.0, -1.0, 2.0]) / np.sqrt(6.0)
HALF = np.radians(22.5)


def reference_camera():
    c = A.Camera()
    c.eye[:] = [2.0, 2.0, 2.0]
    c.at[:] = [0.0, 0.0, 0.0]
    c.up[:] = [0.0, 0.0, 1.0]
    c.fov_y_deg = 45.0
    return c


def angle(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.arccos(np.clip(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b), -1.0, 1.0)))


def check_view(centre, top, bottom, left, right, aspect, tol):
    side = np.arctan(np.tan(HALF) * aspect)
    assert angle(centre, FWD) < tol
    assert abs(angle(top, FWD) - HALF) < tol and np.dot(top, UP) > 0 and abs(np.dot(top, RIGHT)) < tol
    assert abs(angle(bottom, FWD) - HALF) < tol and np.dot(bottom, UP) < 0 and abs(np.dot(bottom, RIGHT)) < tol
    assert abs(angle(right, FWD) - side) < tol and np.dot(right, RIGHT) > 0 and abs(np.dot(right, UP)) < tol
    assert abs(angle(left, FWD) - side) < tol and np.dot(left, RIGHT) < 0 and abs(np.dot(left, UP)) < tol


@pytest.mark.parametrize("w,h", [(64, 48), (48, 64), (256, 256), (1920, 1080)])
def test_oracle_camera_closed_form(oracle, w, h):
    cam = oracle.camera10(reference_camera())
    # image points: centre, top middle, bottom middle, left middle, right middle
    xs = np.array([w // 2, w // 2, w // 2, 0, w - 1], np.uint32)
    ys = np.array([h // 2, 0, h - 1, h // 2, h // 2], np.uint32)
    cx = 0.0 if w % 2 == 0 else 0.5
    cy = 0.0 if h % 2 == 0 else 0.5
    jxy = np.array([[cx, cy], [cx, 0.0], [cx, 1.0], [0.0, cy], [1.0, cy]], np.float32)
    rays = oracle.camera_rays(cam, w, h, points=(xs, ys, jxy))
    assert np.allclose(rays[:, 0:3], 2.0) and np.all(rays[:, 3] == 0)
    d = rays[:, 4:7].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    check_view(*d, aspect=w / h, tol=2e-6)


def test_oracle_camera_pixel_order_and_jitter(oracle):
    """Every pixel's ray, row-major with row 0 at the top, passes through its
    own pixel: its screen coordinates (the offsets along right and up' over
    the forward component) lie inside the pixel's square."""
    w, h = 40, 30
    cam = oracle.camera10(reference_camera())
    rays = oracle.camera_rays(cam, w, h, seed=3, sample=5)
    d = rays[:, 4:7].astype(np.float64)
    f = d @ FWD
    sx = (d @ RIGHT) / f / (np.tan(HALF) * w / h)
    sy = (d @ UP) / f / np.tan(HALF)
    px = (sx + 1.0) * 0.5 * w
    py = (1.0 - sy) * 0.5 * h
    ix = np.arange(w * h) % w
    iy = np.arange(w * h) // w
    assert np.all(px >= ix - 1e-4) and np.all(px <= ix + 1 + 1e-4)
    assert np.all(py >= iy - 1e-4) and np.all(py <= iy + 1 + 1e-4)
    # the jitter differs between samples and stays inside the pixel
    other = oracle.camera_rays(cam, w, h, seed=3, sample=6)
    assert np.mean(np.any(other != rays, axis=1)) > 0.99


def test_product_camera_basis_closed_form(lib):
    b = pt.camera_basis(reference_camera(), 64, 48)
    np.testing.assert_allclose(b["eye"], [2, 2, 2])
    np.testing.assert_allclose(b["forward"], FWD, atol=1e-7)
    np.testing.assert_allclose(b["right"], RIGHT, atol=1e-7)
    np.testing.assert_allclose(b["up"], UP, atol=1e-7)
    assert b["tan_half"] == pytest.approx(np.tan(HALF), rel=1e-7)
    assert b["aspect"] == pytest.approx(64 / 48, rel=1e-7)


def test_product_camera_rejects_degenerate_views(lib):
    def cam(eye, at, up, fov=45.0):
        c = A.Camera()
        c.eye[:], c.at[:], c.up[:] = eye, at, up
        c.fov_y_deg = fov
        return c

    bad = [cam([0, 0, 5], [0, 0, 0], [0, 0, 1]),           # looking straight down with up = z
           cam([1, 1, 1], [1, 1, 1], [0, 0, 1]),           # eye == at
           cam([2, 2, 2], [0, 0, 0], [0, 0, 0]),           # up of zero length
           cam([2, 2, 2], [0, 0, 0], [1, 1, 1]),           # up parallel to the view
           cam([2, 2, 2], [0, 0, 0], [0, 0, 1], 0.0),
           cam([2, 2, 2], [0, 0, 0], [0, 0, 1], 180.0),
           cam([np.nan, 2, 2], [0, 0, 0], [0, 0, 1]),
           cam([2, 2, 2], [0, 0, np.inf], [0, 0, 1])]
    for c in bad:
        with pytest.raises(pt.FndError) as e:
            pt.camera_basis(c, 8, 8)
        assert e.value.status == A.E_INVALID_ARG
    # nearly parallel but still a basis
    pt.camera_basis(cam([0, 0, 5], [0, 0, 0], [1e-3, 0, 1]), 8, 8)


def test_host_primary_rays_closed_form():
    """rays.primary_rays (pixel centres, no jitter): with odd sizes the middle
    pixel looks along forward, and the outermost pixel centres lie at the
    closed-form angles of their screen offsets (1 - 1/W) and (1 - 1/H)."""
    w, h = 63, 47
    rays = R.primary_rays(reference_camera(), w, h).reshape(h, w, 8)
    d = rays[..., 4:7].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=-1), 1.0, atol=1e-6)
    assert angle(d[h // 2, w // 2], FWD) < 1e-6
    ty = np.arctan((1 - 1 / h) * np.tan(HALF))
    tx = np.arctan((1 - 1 / w) * np.tan(HALF) * w / h)
    top, bottom = d[0, w // 2], d[h - 1, w // 2]
    left, right = d[h // 2, 0], d[h // 2, w - 1]
    assert abs(angle(top, FWD) - ty) < 1e-6 and np.dot(top, UP) > 0
    assert abs(angle(bottom, FWD) - ty) < 1e-6 and np.dot(bottom, UP) < 0
    assert abs(angle(right, FWD) - tx) < 1e-6 and np.dot(right, RIGHT) > 0
    assert abs(angle(left, FWD) - tx) < 1e-6 and np.dot(left, RIGHT) < 0


def test_sincos_2pi_against_libm(oracle):
    """The samplers' own sin/cos of 2*pi*u over every 2^-12 step and random
    multiples of 2^-24: within 4e-7 of the double-precision values, exact at
    quarter turns."""
    rng = np.random.default_rng(0)
    u = np.concatenate([np.arange(4096) / 4096.0,
                        rng.integers(0, 1 << 24, 200000) * 2.0 ** -24]).astype(np.float32)
    s, c = oracle.sincos_2pi(u)
    a = 2.0 * np.pi * u.astype(np.float64)
    assert np.abs(s - np.sin(a)).max() < 4e-7
    assert np.abs(c - np.cos(a)).max() < 4e-7
    q = np.array([0.0, 0.25, 0.5, 0.75], np.float32)
    s, c = oracle.sincos_2pi(q)
    assert list(s) == [0.0, 1.0, 0.0, -1.0] and list(c) == [1.0, 0.0, -1.0, 0.0]
