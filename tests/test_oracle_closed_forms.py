"""Oracle forward/Adam/loss against closed forms, invariants and brute force.

Pins (each failing for a plausible slip):
  * projection + compositing of one isotropic Gaussian facing the camera:
    cov2d = (f s / z)^2 I, compensation var/(var+blur), conic 1/(var+blur+1e-6),
    alpha(pixel) = min(0.999, o exp(-sigma)) -> the whole image in float64
    (a transposed view matrix, a wrong blur or a dropped 1/2 fails);
  * two Gaussians on the optical axis: front-to-back over-compositing;
  * rigid-motion invariance: rotating scene and camera together leaves the
    image unchanged (catches row/column-major mix-ups of the view matrix);
  * SH colour: degree-0 round trip (RustGS/src/sh/tests.rs:4-8) and
    orthonormality of the degree-3 basis over the sphere (Gauss-Legendre x
    uniform quadrature, exact for these polynomials);
  * Adam one- and two-step closed forms (engine/optimizer.rs:122-162) and
    f32 powi bias corrections;
  * L1 gradient sign(pred - target) / (3HW) on RGB, 0 on alpha;
  * binning invariants on config 1: sorted tile keys, per-tile (depth, gid)
    order, ranges covering [0, M), counts summing to M;
  * brute force (every splat at every pixel in depth order, no tiles) equals
    the tiled render on config 1.
"""
import math

import numpy as np
import pytest

from rustscan_amd import scene

SH_C0 = 0.28209479177387814


def single_gaussian(z=2.0, s=0.05, rgb=(1.0, 0.5, 0.25), logit=0.0, pos_xy=(0.0, 0.0)):
    tf = np.array([[pos_xy[0], pos_xy[1], z, 1, 0, 0, 0, math.log(s), math.log(s), math.log(s)]], np.float32)
    sh = np.array([[[(c - 0.5) / SH_C0 for c in rgb]]], np.float32)
    return scene.Splats(tf, sh, np.array([logit], np.float32), 0)


def closed_form_image(gs, W, H, f, blur):
    """gs: list of (z, s, rgb, opacity_sigmoid) on the optical axis, any order."""
    gs = sorted(gs, key=lambda g: g[0])
    img = np.zeros((H, W, 3))
    T = np.ones((H, W))
    px = np.arange(W) + 0.5
    py = np.arange(H) + 0.5
    dx = W / 2.0 - px[None, :]
    dy = H / 2.0 - py[:, None]
    for z, s, rgb, o in gs:
        var = (f * s / z) ** 2
        comp = var / (var + blur)
        a = 1.0 / (var + blur + 1e-6)
        sigma = 0.5 * a * (dx * dx + dy * dy)
        alpha = np.minimum(0.999, o * comp * np.exp(-sigma))
        acc = alpha >= 1.0 / 255.0
        alpha = np.where(acc, alpha, 0.0)
        for c in range(3):
            img[..., c] += max(rgb[c], 0.0) * alpha * T
        T = T * (1.0 - alpha)
    return img, 1.0 - T


def test_single_gaussian_projection_closed_form(orc):
    W = H = 64
    f, blur, z, s = 500.0, 0.3, 2.0, 0.05
    sp = single_gaussian(z=z, s=s)
    r = orc.Render(sp, scene.pinhole_identity(W, H, f), cov_blur=blur)
    p = r.projected()
    assert p.shape == (1, 9)
    var = (f * s / z) ** 2
    np.testing.assert_allclose(p[0, :2], [32.0, 32.0], rtol=0, atol=1e-6)
    np.testing.assert_allclose(p[0, 2], 1.0 / (var + blur + 1e-6), rtol=2e-6)
    assert p[0, 3] == 0.0
    np.testing.assert_allclose(p[0, 4], 1.0 / (var + blur + 1e-6), rtol=2e-6)
    np.testing.assert_allclose(p[0, 5:8], [1.0, 0.5, 0.25], atol=2e-7)
    np.testing.assert_allclose(p[0, 8], 0.5 * var / (var + blur), rtol=2e-6)


@pytest.mark.parametrize("blur", [0.0, 0.3])
def test_single_gaussian_image_closed_form(orc, blur):
    W = H = 64
    f, z, s = 500.0, 2.0, 0.05
    r = orc.Render(single_gaussian(z=z, s=s), scene.pinhole_identity(W, H, f), cov_blur=blur)
    img = r.image()
    ref_rgb, ref_a = closed_form_image([(z, s, (1.0, 0.5, 0.25), 0.5)], W, H, f, blur)
    # exclude pixels whose alpha sits within float32 rounding of the 1/255 threshold
    var = (f * s / z) ** 2
    o = 0.5 * var / (var + blur)
    px = np.arange(W) + 0.5
    d2 = (W / 2 - px[None, :]) ** 2 + (H / 2 - px[:, None]) ** 2
    alpha = o * np.exp(-0.5 * d2 / (var + blur + 1e-6))
    ok = np.abs(alpha - 1.0 / 255.0) > 1e-5
    assert ok.mean() > 0.99
    np.testing.assert_allclose(img[..., :3][ok], ref_rgb[ok], atol=2e-6)
    np.testing.assert_allclose(img[..., 3][ok], ref_a[ok], atol=2e-6)


def test_two_gaussians_over_compositing(orc):
    W = H = 64
    f, blur = 500.0, 0.3
    a = single_gaussian(z=3.0, s=0.08, rgb=(0.2, 0.9, 0.4), logit=1.0)
    b = single_gaussian(z=2.0, s=0.05, rgb=(1.0, 0.5, 0.25), logit=-0.5)
    two = scene.Splats(np.concatenate([a.transforms, b.transforms]), np.concatenate([a.sh, b.sh]),
                       np.concatenate([a.raw_opac, b.raw_opac]), 0)
    r = orc.Render(two, scene.pinhole_identity(W, H, f), cov_blur=blur)
    assert r.depth_order().tolist() == [1, 0]
    sig = lambda x: 1.0 / (1.0 + math.exp(-x))
    ref_rgb, ref_a = closed_form_image([(3.0, 0.08, (0.2, 0.9, 0.4), sig(1.0)), (2.0, 0.05, (1.0, 0.5, 0.25), sig(-0.5))],
                                       W, H, f, blur)
    img = r.image()
    assert np.abs(img[..., :3] - ref_rgb).max() < 2e-3  # threshold pixels included
    core = np.zeros((H, W), bool)
    core[20:44, 20:44] = True
    np.testing.assert_allclose(img[..., :3][core], ref_rgb[core], atol=3e-6)
    np.testing.assert_allclose(img[..., 3][core], ref_a[core], atol=3e-6)


def rot(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def quat_of(R):
    w = math.sqrt(max(1 + R[0, 0] + R[1, 1] + R[2, 2], 1e-12)) / 2
    return np.array([w, (R[2, 1] - R[1, 2]) / (4 * w), (R[0, 2] - R[2, 0]) / (4 * w), (R[1, 0] - R[0, 1]) / (4 * w)])


def quat_mul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def test_rigid_motion_invariance(orc):
    sc = scene.gaussian_scene(400, seed=21, sh_degree=0, center=(0.0, 0.0, 4.0))
    cam = scene.pinhole_identity(96, 64, 120.0)
    base = orc.Render(sc, cam).image()
    R = rot([0.3, -0.5, 0.8], 0.7)
    t = np.array([0.4, -1.2, 2.0])
    tf = sc.transforms.astype(np.float64).copy()
    tf[:, 0:3] = tf[:, 0:3] @ R.T + t
    qR = quat_of(R)
    tf[:, 3:7] = np.stack([quat_mul(qR, q) for q in tf[:, 3:7]])
    moved = scene.Splats(tf.astype(np.float32), sc.sh, sc.raw_opac, 0)
    # camera pose: world->camera = old_w2c composed with the inverse motion
    M = np.eye(4)
    M[:3, :3] = R.T
    M[:3, 3] = -R.T @ t
    cam2 = scene.Camera((M.T.reshape(16)).astype(np.float32), cam.fx, cam.fy, cam.cx, cam.cy, cam.width, cam.height,
                        t.astype(np.float32))
    img = orc.Render(moved, cam2).image()
    assert np.abs(img - base).max() < 2e-3
    assert np.mean(np.abs(img - base)) < 2e-5


def test_sh_degree0_round_trip(orc):
    rgb = np.array([0.25, 0.5, 0.75], np.float32)
    coeffs = ((rgb - 0.5) / np.float32(SH_C0)).astype(np.float32).reshape(1, 3)
    out = orc.sh_color(0, coeffs, [0.0, 0.0, 1.0])
    np.testing.assert_allclose(out, rgb, atol=1e-7)


def test_sh_basis_orthonormal(orc):
    deg, K = 3, 16
    xg, wg = np.polynomial.legendre.leggauss(24)
    nphi = 48
    phis = 2 * np.pi * (np.arange(nphi) + 0.5) / nphi
    B = []
    wts = []
    for ct, w in zip(xg, wg):
        st = math.sqrt(1 - ct * ct)
        for ph in phis:
            v = np.array([st * math.cos(ph), st * math.sin(ph), ct], np.float32)
            row = []
            for k in range(K):
                co = np.zeros((K, 3), np.float32)
                co[k, 0] = 1.0
                row.append(orc.sh_color(deg, co, v)[0] - 0.5)
            B.append(row)
            wts.append(w * 2 * np.pi / nphi)
    B = np.array(B, np.float64)
    G = (B * np.array(wts)[:, None]).T @ B
    np.testing.assert_allclose(G, np.eye(K), atol=5e-6)


def test_adam_closed_form(orc):
    b1, b2, eps = np.float32(0.9), np.float32(0.999), np.float32(1e-8)
    p = np.array([[1.0, 2.0, 3.0]], np.float32)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    g1 = np.array([[1.0, -0.5, 2e-3]], np.float32)
    lr = np.array([0.1, 0.01, 1.0], np.float32)
    p0 = p.copy()
    orc.adam_group(p, g1.copy(), m, v, lr, b1, b2, eps, 1)
    # step 1: m_hat = g, v_hat = g^2 -> p -= lr * g / (|g| + eps)
    np.testing.assert_allclose(p, p0 - lr * g1 / (np.abs(g1) + 1e-8), rtol=1e-6)
    g2 = np.array([[-2.0, 0.25, 1e-3]], np.float32)
    p1 = p.copy()
    orc.adam_group(p, g2.copy(), m, v, lr, b1, b2, eps, 2)
    # the reference computes betas, 1 - beta and the bias corrections in f32
    f1, f2 = np.float64(b1), np.float64(b2)
    o1, o2 = np.float64(np.float32(1) - b1), np.float64(np.float32(1) - b2)
    bc1 = np.float64(np.float32(1) - np.float32(orc.powi(b1, 2)))
    bc2 = np.float64(np.float32(1) - np.float32(orc.powi(b2, 2)))
    g1d, g2d = g1.astype(np.float64), g2.astype(np.float64)
    m2 = f1 * o1 * g1d + o1 * g2d
    v2 = f2 * o2 * g1d ** 2 + o2 * g2d ** 2
    upd = (m2 / bc1) / (np.sqrt(v2 / bc2) + 1e-8)
    np.testing.assert_allclose(p, p1 - lr * upd, rtol=2e-6, atol=1e-9)
    # weight decay (optimizer.rs:128-130): the gradient becomes g + wd * p
    # before the moments, so the first step moves p by lr * g' / (|g'| + eps)
    wd = np.float32(0.5)
    q = np.array([[1.0, -2.0, 4.0]], np.float32)
    gq = np.array([[0.25, 3.0, -2.0]], np.float32)
    mq, vq = np.zeros_like(q), np.zeros_like(q)
    q0 = q.copy()
    orc.adam_group(q, gq.copy(), mq, vq, lr, b1, b2, eps, 1, weight_decay=wd)
    gd = gq.astype(np.float64) + 0.5 * q0.astype(np.float64)  # (0.75, 2.0, 0.0)
    np.testing.assert_allclose(q, q0 - lr * gd / (np.abs(gd) + 1e-8), rtol=1e-6, atol=1e-9)
    assert q[0, 2] == q0[0, 2]  # g' = 0 there: no move
    np.testing.assert_allclose(mq, (1 - 0.9) * gd, rtol=1e-6)
    # the f32 powi of the bias correction
    # binary exponentiation: repeated squaring doubles the relative error of
    # each power, so the bound grows like step * 2^-24 (and is 0 at step 1)
    base = float(np.float32(0.999))
    for step in [1, 2, 3, 10, 100, 1000]:
        exact = base ** step
        assert abs(orc.powi(0.999, step) - exact) <= (step - 1) * 1.2e-7 * exact


def test_l1_gradient(orc):
    rng = np.random.default_rng(5)
    pred = rng.uniform(size=(6, 5, 4)).astype(np.float32)
    tgt = rng.uniform(size=(6, 5, 3)).astype(np.float32)
    tgt[0, 0, 0] = pred[0, 0, 0]
    loss, v = orc.l1_loss_grad(pred, tgt, weight=0.8)
    d = (pred[..., :3] - tgt).astype(np.float64)  # the difference is taken in f32
    np.testing.assert_allclose(loss, float(np.float32(0.8)) * np.abs(d).mean(), rtol=1e-12)
    scale = np.float32(0.8) / np.float32(90)
    np.testing.assert_array_equal(v[..., :3], (np.sign(d) * scale).astype(np.float32))
    assert v[0, 0, 0] == 0.0
    assert (v[..., 3] == 0).all()


@pytest.fixture(scope="module")
def c1(orc):
    sc, cams = scene.config1()
    return sc, cams[0], orc.Render(sc, cams[0])


def test_binning_invariants_config1(orc, c1):
    sc, cam, r = c1
    tiles, gids = r.isect()
    M = r.stats["num_intersections"]
    assert M == r.counts().sum() == len(tiles)
    assert np.all(np.diff(tiles.astype(np.int64)) >= 0)
    depths = r.depths()
    key = depths[gids].view(np.uint32).astype(np.uint64) << np.uint64(32) | gids.astype(np.uint64)
    same = tiles[1:] == tiles[:-1]
    assert np.all(key[1:][same] > key[:-1][same]), "per-tile order must be (depth bits, gid)"
    off = r.tile_offsets().reshape(-1, 2)
    cover = np.zeros(M, np.int64)
    for t, (s, e) in enumerate(off):
        if e > s:
            assert np.all(tiles[s:e] == t)
            cover[s:e] += 1
    assert np.all(cover == 1)
    order = r.depth_order()
    dv = depths[order]
    assert np.all((dv[1:] > dv[:-1]) | ((dv[1:] == dv[:-1]) & (order[1:] > order[:-1])))
    vis = r.visible()
    assert set(np.nonzero(vis)[0]).issubset(set(order.tolist()))
    assert r.stats["k1_count_mismatch"] == 0


def test_bruteforce_equals_tiled_config1(orc, c1):
    _, _, r = c1
    img = r.image()
    bf = r.bruteforce()
    assert np.abs(img - bf).max() <= 1e-6
