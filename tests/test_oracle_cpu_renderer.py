"""The oracle's restatement of RustGS's CPU renderer (render/tiled_renderer.rs,
used only as the CPU-path timing baseline), pinned by closed forms that
mirror the behaviours of render/tiled_renderer/tests.rs:
  * a single isotropic splat on the optical axis renders the analytic
    footprint o_eff * exp(-d^2/2) * colour with its depth (tests.rs:40-65);
  * splats at one pixel composite front to back whatever their input order
    (tests.rs:181-226 sort_by_depth);
  * an axis-aligned splat gives a mirror-symmetric image, a rotated
    elongated one does not (tests.rs:103-179 covariance cross term).
"""
import math

import numpy as np

from rustscan_amd import scene

SH0 = 0.28209479177387814


def splats_from(points):
    """points: list of (mean xyz, scale xyz, quat wxyz, opacity, rgb)."""
    n = len(points)
    tf = np.zeros((n, 10), np.float32)
    sh = np.zeros((n, 1, 3), np.float32)
    op = np.zeros(n, np.float32)
    for i, (m, s, q, o, rgb) in enumerate(points):
        tf[i, 0:3] = m
        tf[i, 3:7] = q
        tf[i, 7:10] = np.log(np.asarray(s, np.float64))
        op[i] = math.log(o / (1.0 - o))
        sh[i, 0, :] = (np.asarray(rgb, np.float64) - 0.5) / SH0
    return scene.Splats(tf, sh, op, 0)


def footprint(z, s, f, o, blur, dx, dy):
    """Analytic weight of an isotropic splat at depth z seen head-on."""
    var = f * f * s * s / (z * z)
    cxx = var + blur
    comp = math.sqrt(var * var / (cxx * cxx))
    d2 = (dx * dx + dy * dy) / cxx
    return o * comp * math.exp(-0.5 * d2), d2


def test_single_splat_closed_form(orc):
    f, z, s, o, rgb = 100.0, 2.0, 0.05, 0.6, (1.0, 0.5, 0.25)
    cam = scene.pinhole_identity(64, 64, f)
    sp = splats_from([((0.0, 0.0, z), (s, s, s), (1.0, 0.0, 0.0, 0.0), o, rgb)])
    color, depth = orc.tiled_renderer(sp, cam, cov_blur=0.3)
    for px, py in [(31, 31), (32, 30), (35, 33), (28, 36)]:
        w, d2 = footprint(z, s, f, o, 0.3, 32.0 - (px + 0.5), 32.0 - (py + 0.5))
        assert d2 < 9.0 and w > 0.001
        np.testing.assert_allclose(color[py, px], np.clip(np.array(rgb) * w, 0, 1), rtol=2e-5, atol=1e-6)
        assert depth[py, px] == np.float32(z)
    # outside the 3-sigma cutoff: background and no depth
    assert np.all(color[0, 0] == 0.0) and depth[0, 0] == np.float32(3.4028235e38)


def test_front_to_back_whatever_the_input_order(orc):
    f, s, blur = 100.0, 0.05, 0.3
    cam = scene.pinhole_identity(32, 32, f)
    pts = [((0.0, 0.0, 2.0), (s, s, s), (1, 0, 0, 0), 0.5, (1.0, 0.0, 0.0)),
           ((0.0, 0.0, 1.5), (s, s, s), (1, 0, 0, 0), 0.5, (0.0, 1.0, 0.0)),
           ((0.0, 0.0, 3.0), (s, s, s), (1, 0, 0, 0), 0.5, (0.0, 0.0, 1.0))]
    px = py = 15
    expect = np.zeros(3)
    acc = 0.0
    for i in sorted(range(3), key=lambda i: pts[i][0][2]):
        w, _ = footprint(pts[i][0][2], s, f, pts[i][3], blur, 16.0 - (px + 0.5), 16.0 - (py + 0.5))
        a = w * (1.0 - acc)
        expect += np.array(pts[i][4]) * a
        acc += a
    for perm in ([0, 1, 2], [2, 1, 0], [1, 2, 0]):
        color, depth = orc.tiled_renderer(splats_from([pts[i] for i in perm]), cam, cov_blur=blur)
        np.testing.assert_allclose(color[py, px], expect, rtol=1e-5, atol=1e-6)
        assert depth[py, px] == np.float32(1.5)  # nearest splat with alpha > 0.01


def test_covariance_cross_term(orc):
    cam = scene.pinhole_identity(64, 64, 500.0)
    axis = splats_from([((0.0, 0.0, 2.0), (0.02, 0.008, 0.004), (1, 0, 0, 0), 0.8, (1.0, 0.0, 0.0))])
    color, _ = orc.tiled_renderer(axis, cam, cov_blur=0.0)
    # axis-aligned ellipse centred on the pixel grid's centre line: mirror symmetric
    np.testing.assert_allclose(color[:, :, 0], color[:, ::-1, 0], atol=1e-6)
    np.testing.assert_allclose(color[:, :, 0], color[::-1, :, 0], atol=1e-6)
    h = math.pi / 8  # 45 degrees about z
    rot = splats_from([((0.0, 0.0, 2.0), (0.02, 0.008, 0.004), (math.cos(h), 0, 0, math.sin(h)), 0.8,
                        (1.0, 0.0, 0.0))])
    color_r, _ = orc.tiled_renderer(rot, cam, cov_blur=0.0)
    assert np.abs(color_r[:, :, 0] - color_r[:, ::-1, 0]).max() > 1e-2
    # but symmetric under the 180-degree rotation of the image
    np.testing.assert_allclose(color_r[:, :, 0], color_r[::-1, ::-1, 0], atol=1e-6)
