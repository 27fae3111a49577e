"""Pins of the loss oracle (oracle/loss.py, a restatement of
RustGS/src/training/engine/loss.rs): the reference's own loss tests
(loss_tests.rs), closed forms, an independent matrix form of the
edge-padded blur, and float64 central differences of the forward for the
analytic gradient."""
import numpy as np
import pytest

from oracle import loss as OL


def test_gaussian_kernel():
    k = OL.gaussian_kernel_1d(11, 1.5)
    assert k.dtype == np.float32 and len(k) == 11
    np.testing.assert_allclose(k.sum(dtype=np.float64), 1.0, atol=1e-6)
    np.testing.assert_array_equal(k, k[::-1])
    e = np.exp(-np.arange(-5, 6) ** 2 / (2 * 1.5 ** 2))
    np.testing.assert_allclose(k, e / e.sum(), rtol=1e-6)


def blur_matrix(n, k):
    """Dense 1D edge-padded blur: out[i] = sum_t k[t] in[clamp(i + t - r)]."""
    r = len(k) // 2
    m = np.zeros((n, n))
    for i in range(n):
        for t in range(len(k)):
            m[i, min(max(i + t - r, 0), n - 1)] += k[t]
    return m


def test_separable_blur_matches_reference_vector():
    """loss_tests.rs:271-300: arange(120)/120 as [1,3,5,8], kernel (0.25, 0.5, 0.25)."""
    x = (np.arange(120, dtype=np.float32) / np.float32(120.0)).reshape(1, 3, 5, 8)
    k = np.array([0.25, 0.5, 0.25], np.float32)
    got = OL.separable_blur(x, k)
    ref = np.einsum("hi,ncij,wj->nchw", blur_matrix(5, k), x.astype(np.float64), blur_matrix(8, k))
    np.testing.assert_allclose(got, ref, atol=1e-6)


def test_blur_transpose_is_the_adjoint():
    rng = np.random.default_rng(1)
    k = OL.gaussian_kernel_1d().astype(np.float64)
    for h, w in [(7, 9), (13, 4), (1, 12), (20, 20)]:
        x = rng.normal(size=(3, h, w))
        g = rng.normal(size=(3, h, w))
        lhs = (OL.separable_blur(x, k) * g).sum()
        rhs = (x * OL.blur_transpose(g, k)).sum()
        assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(lhs))
        ref = np.einsum("hi,cij,wj->chw", blur_matrix(h, k).T, g, blur_matrix(w, k).T)
        np.testing.assert_allclose(OL.blur_transpose(g, k), ref, atol=1e-12)


def test_ssim_constant_images_closed_form():
    """Constant images a, b: with S the sum of the (f32-rounded) kernel, each
    blur scales a constant by S^2, so mu = a S^2, E[x^2] = a^2 S^2 and the
    variances are a^2 (S^2 - S^4); at S = 1 this is (2ab + c1) / (a^2 + b^2 + c1)."""
    cfg = OL.LossConfig()
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    S2 = float(OL.gaussian_kernel_1d().astype(np.float64).sum()) ** 2
    for a, b in [(0.3, 0.7), (0.0, 1.0), (0.5, 0.5)]:
        x = np.full((3, 12, 10), a, np.float64)
        y = np.full((3, 12, 10), b, np.float64)
        s = OL.ssim_map(x, y, cfg, np.float64)["s"]
        mx, my = a * S2, b * S2
        vxx, vyy, vxy = a * a * (S2 - S2 * S2), b * b * (S2 - S2 * S2), a * b * (S2 - S2 * S2)
        ref = (2 * mx * my + c1) * (2 * vxy + c2) / ((mx * mx + my * my + c1) * (vxx + vyy + c2))
        np.testing.assert_allclose(s, ref, rtol=1e-9)
        np.testing.assert_allclose(s, (2 * a * b + c1) / (a * a + b * b + c1), rtol=1e-4)


def test_reference_loss_tests():
    cfg = OL.LossConfig(l1_weight=0.0, ssim_weight=1.0)
    ones = np.ones((16, 16, 3), np.float32)
    zeros = np.zeros((16, 16, 3), np.float32)
    # identical images: near-zero SSIM loss (loss_tests.rs:61-76); zeros stay finite (:78-90)
    assert abs(OL.forward(ones, ones, cfg)) < 0.01
    assert np.isfinite(OL.forward(zeros, zeros, cfg))
    # combined l1 = 1, ssim = 1 of zeros against ones > 0.5 (:92-119); closed form 1 + 1 - c1 / (1 + c1)
    comb = OL.forward(zeros, ones, OL.LossConfig(l1_weight=1.0, ssim_weight=1.0))
    assert comb > 0.5
    np.testing.assert_allclose(comb, 2.0 - 1e-4 / (1.0 + 1e-4), rtol=1e-6)


def test_residual_variants():
    """reconstruction_residual_loss on the reference tests' 1x2 images."""
    pred = np.zeros((1, 2, 3), np.float32)
    tgt = np.array([0.1, 0.1, 0.1, 1.0, 1.0, 1.0], np.float32).reshape(1, 2, 3)
    exact = OL.residual_loss(pred, tgt, OL.LossConfig())
    np.testing.assert_allclose(exact, 0.55, rtol=1e-6)
    robust = OL.residual_loss(pred, tgt, OL.LossConfig(robust_delta=0.1))
    # saturating L1: a * d / (a + d) -> (0.01/0.2 + 0.1/1.1) / 2
    np.testing.assert_allclose(robust, (0.05 + 0.1 / 1.1) / 2, rtol=1e-6)
    assert robust < exact  # loss_tests.rs:121-150
    weighted = OL.residual_loss(pred, tgt, OL.LossConfig(outlier_threshold=0.25, outlier_weight=0.25))
    # a * (floor + (1 - floor) * thr / (a + thr)) with thr = 0.25, floor = 0.25
    per = lambda a: a * (0.25 + 0.75 * 0.25 / (a + 0.25))
    np.testing.assert_allclose(weighted, (per(0.1) + per(1.0)) / 2, rtol=1e-6)
    assert 0.0 < weighted < exact  # :152-185
    # the combined loss of a one-pixel-high image is NaN whatever the image-gradient
    # weight: that term's y mean is over an empty difference (loss.rs:213-227, 139-141)
    with np.errstate(invalid="ignore"):
        assert np.isnan(OL.forward(pred, tgt, OL.LossConfig(l1_weight=1.0, ssim_weight=0.0)))


def test_gradient_difference_closed_form():
    """loss_tests.rs:225-250: zeros against columns 2..3 set to one on a 4x4
    image: only the x difference at column 1 differs, 12 of 36 entries."""
    pred = np.zeros((4, 4, 3), np.float32)
    tgt = np.zeros((4, 4, 3), np.float32)
    tgt[:, 2:4, :] = 1.0
    g = OL.forward(pred, tgt, OL.LossConfig(l1_weight=0.0, ssim_weight=0.0, gradient_weight=1.0))
    np.testing.assert_allclose(g, 0.5 * (12.0 / 36.0), rtol=1e-6)


@pytest.mark.parametrize("cfg", [
    OL.LossConfig(),
    OL.LossConfig(l1_weight=0.3, ssim_weight=0.5, gradient_weight=0.2),
    OL.LossConfig(l1_weight=1.0, ssim_weight=0.2, robust_delta=0.2),
    OL.LossConfig(l1_weight=1.0, ssim_weight=0.0, outlier_threshold=0.3, outlier_weight=0.4),
])
def test_grad_matches_finite_differences(cfg):
    rng = np.random.default_rng(7)
    H, W = 9, 13
    pred = rng.uniform(0.0, 1.0, (H, W, 3))
    tgt = rng.uniform(0.0, 1.0, (H, W, 3))
    # keep residuals and neighbour differences away from the kinks of |.|
    d = pred - tgt
    pred = np.where(np.abs(d) < 0.05, pred + 0.1, pred)
    g = OL.grad(pred, tgt, cfg)
    eps = 1e-6
    picks = rng.integers(0, [H, W, 3], size=(40, 3))
    worst = 0.0
    checked = 0
    for i, j, c in picks:
        def f(delta):
            q = pred.copy()
            q[i, j, c] += delta
            return OL.forward(q, tgt, cfg, dtype=np.float64)
        fd = (f(eps) - f(-eps)) / (2 * eps)
        fd2 = (f(eps / 2) - f(-eps / 2)) / eps
        if abs(fd - fd2) > 1e-3 * max(abs(fd), 1e-9):
            continue  # a kink inside the stencil
        checked += 1
        worst = max(worst, abs(fd - g[i, j, c]) / max(abs(fd), 1e-6))
    assert checked >= 30
    assert worst <= 1e-5, worst
