"""Pins of the evaluation oracle (oracle/evaluation.py, the CPU renderer's SH
and its math switch in oracle/rgs_oracle.cpp) and of the host-side
evaluation bookkeeping (rustscan_amd/evaluation.py), by routes independent
of the restated code:

* sh.rs evaluate_sh_rgb (degrees 0-4): every basis function equals the real
  spherical harmonic from scipy's associated Legendre functions, and the
  reference's degree-0 test (sh/tests.rs: evaluate_degree_zero_matches_
  sh0_conversion, an exact round trip).
* PSNR: closed forms (uniform error d gives 20 log10(1/d); identical images
  give 100 dB; empty or mismatched inputs give 0).
* Sharpness: closed forms on ramps, parabolas and a checkerboard, and the
  reference's own test (core/tests.rs: sharpness_metrics_detect_edges).
* The two accumulation modes: the float32 sequential sum equals a plain
  Python loop in the reference's visiting order on small images, and differs
  from the float64 sum by no more than the float32 rounding bound.
* summarize_psnr_samples, summarize_training_metrics, worst_frame_metrics,
  select_evaluation_frames: the reference's tests (core/tests.rs).
* The CPU renderer with the library's exponential (gpu_math) against libm's:
  the same image up to the exponentials' ulps.
"""
import math

import numpy as np
import pytest

from oracle import evaluation as OE
from rustscan_amd import evaluation as E
from rustscan_amd import scene
from tests.test_oracle_sh_clip import real_sh

F = np.float32


def test_tr_sh_basis_equals_legendre_real_sh(orc):
    rng = scene.Rng(23)
    dirs = rng.normal(3 * 64).reshape(64, 3)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    dirs = dirs.astype(np.float32)
    worst = 0.0
    for deg in range(5):
        K = (deg + 1) ** 2
        for l in range(deg + 1):
            for m in range(-l, l + 1):
                k = l * l + l + m
                co = np.zeros((dirs.shape[0], K, 3), np.float32)
                co[:, k] = [1.0, -2.0, 0.5]
                got = orc.tr_sh_rgb(co, deg, dirs) - 0.5
                want = np.array([[real_sh(l, m, v.astype(np.float64))] for v in dirs]) * np.array([1.0, -2.0, 0.5])
                worst = max(worst, float(np.abs(got - want).max()))
    assert worst <= 2e-6, worst


def test_tr_sh_degree_zero_round_trip(orc):
    """sh/tests.rs evaluate_degree_zero_matches_sh0_conversion."""
    rgb = np.array([0.25, 0.5, 0.75], np.float32)
    co = ((rgb - F(0.5)) / F(0.2820948)).astype(np.float32).reshape(1, 1, 3)
    out = orc.tr_sh_rgb(co, 0, np.array([[0.0, 0.0, 1.0]], np.float32))
    assert np.array_equal(out[0], rgb)


def test_tr_sh_missing_coefficients_read_as_zero(orc):
    """A degree-4 evaluation of a degree-3 row: coeff() returns 0 past the row
    (sh.rs:114-121); here the row is padded explicitly and must match."""
    rng = np.random.default_rng(3)
    co3 = rng.normal(size=(8, 16, 3)).astype(np.float32)
    co4 = np.zeros((8, 25, 3), np.float32)
    co4[:, :16] = co3
    d = rng.normal(size=(8, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    np.testing.assert_allclose(orc.tr_sh_rgb(co4, 4, d), orc.tr_sh_rgb(co3, 3, d), rtol=0, atol=1e-6)


@pytest.mark.parametrize("acc", ["f32", "f64"])
def test_psnr_closed_forms(acc):
    t = np.random.default_rng(0).uniform(0.2, 0.8, (9, 7, 3)).astype(np.float32)
    for d in (0.5, 0.125, 1.0 / 64):
        r = (t + F(d)).astype(np.float32)
        diff = (r - t).astype(np.float64)  # d up to the rounding of t + d
        want = 10.0 * math.log10(1.0 / np.mean(diff * diff))
        assert abs(float(OE.psnr(r, t, acc)) - want) < 2e-4
    assert OE.psnr(t, t, acc) == F(100.0)
    assert OE.psnr(t[:0], t[:0], acc) == F(0.0)
    assert OE.psnr(t[:2], t, acc) == F(0.0)


def _lum_image(L):
    """An RGB image whose luminance is L (grey: R = G = B = L up to the
    luminance weights' float32 sum)."""
    return np.repeat(L[..., None], 3, axis=2).astype(np.float32)


@pytest.mark.parametrize("acc", ["f32", "f64"])
def test_sharpness_closed_forms(acc):
    H, W = 6, 9
    x = np.arange(W, dtype=np.float32)[None, :].repeat(H, 0)
    y = np.arange(H, dtype=np.float32)[:, None].repeat(W, 1)
    wsum = float(F(0.2126) + F(0.7152) + F(0.0722))
    ramp = _lum_image(F(0.05) * x)
    want = wsum * 0.05 * (W - 1) * H / ((W - 1) * H + W * (H - 1))
    assert abs(float(OE.gradient_sharpness(ramp, acc)) - want) < 1e-6
    assert abs(float(OE.laplacian_sharpness(ramp, acc))) < 1e-10
    para = _lum_image(F(0.01) * y * y)  # Laplacian 0.02 everywhere: variance 0
    assert abs(float(OE.laplacian_sharpness(para, acc))) < 1e-9
    cb = _lum_image(((x + y) % 2).astype(np.float32))  # Laplacian -4 wsum on ones, +4 wsum on zeros
    lap = np.where(((x + y) % 2)[1:-1, 1:-1] == 1, -4.0 * wsum, 4.0 * wsum)
    assert abs(float(OE.laplacian_sharpness(cb, acc)) - float(np.var(lap))) < 1e-4
    assert OE.gradient_sharpness(ramp[:1], acc) == 0.0 and OE.laplacian_sharpness(ramp[:2], acc) == 0.0


def test_sharpness_detects_edges():
    """core/tests.rs sharpness_metrics_detect_edges: a flat 4x4 image scores
    0, one with a vertical edge scores > 0 on both measures."""
    flat = np.full((4, 4, 3), 0.5, np.float32)
    edge = np.zeros((4, 4, 3), np.float32)
    edge[:, 2:] = 1.0
    for acc in ("f32", "f64"):
        assert OE.gradient_sharpness(flat, acc) == 0.0 and OE.gradient_sharpness(edge, acc) > 0.0
        assert OE.laplacian_sharpness(flat, acc) == 0.0 and OE.laplacian_sharpness(edge, acc) > 0.0


def _loop_metrics(r, t):
    """The reference's loops, literally, in float32 (core.rs:238-306)."""
    H, W = r.shape[:2]
    s = F(0.0)
    for a, b in zip(r.ravel(), t.ravel()):
        d = F(a) - F(b)
        s = F(s + d * d)
    mse = F(s / F(r.size))

    def lum(img, x, yy):
        p = img[yy, x]
        return F(F(F(0.2126) * p[0]) + F(F(0.7152) * p[1])) + F(F(0.0722) * p[2])

    def grad(img):
        s, n = F(0.0), 0
        for yy in range(H):
            for x in range(W):
                c = lum(img, x, yy)
                if x + 1 < W:
                    s = F(s + abs(F(lum(img, x + 1, yy) - c)))
                    n += 1
                if yy + 1 < H:
                    s = F(s + abs(F(lum(img, x, yy + 1) - c)))
                    n += 1
        return F(s / F(n))

    def lap(img):
        s, s2, n = F(0.0), F(0.0), 0
        for yy in range(1, H - 1):
            for x in range(1, W - 1):
                v = F(F(-4.0) * lum(img, x, yy))
                for xx, y2 in ((x - 1, yy), (x + 1, yy), (x, yy - 1), (x, yy + 1)):
                    v = F(v + lum(img, xx, y2))
                s = F(s + v)
                s2 = F(s2 + F(v * v))
                n += 1
        mean = F(s / F(n))
        return max(F(F(s2 / F(n)) - F(mean * mean)), F(0.0))

    return mse, grad(r), grad(t), lap(r), lap(t)


def test_f32_accumulation_is_the_reference_loop():
    rng = np.random.default_rng(5)
    r = rng.uniform(0, 1, (7, 11, 3)).astype(np.float32)
    t = rng.uniform(0, 1, (7, 11, 3)).astype(np.float32)
    mse, rg, tg, rl, tl = _loop_metrics(r, t)
    m = OE.frame_metrics(r, t, "f32")
    assert m["psnr_db"] == F(10.0) * F(np.log10(F(1.0) / mse))
    assert m["render_grad"] == rg and m["target_grad"] == tg
    assert m["render_lap"] == rl and m["target_lap"] == tl


def test_f32_and_f64_accumulation_within_the_rounding_bound():
    rng = np.random.default_rng(6)
    r = rng.uniform(0, 1, (240, 320, 3)).astype(np.float32)
    t = rng.uniform(0, 1, (240, 320, 3)).astype(np.float32)
    d = (r - t).astype(np.float32)
    terms = (d * d).ravel()
    bound = (terms.size - 1) * 2.0 ** -24 * float(np.sum(np.abs(terms.astype(np.float64))))
    s32 = float(np.cumsum(terms, dtype=np.float32)[-1])
    s64 = float(np.sum(terms.astype(np.float64)))
    assert abs(s32 - s64) <= bound
    m32, m64 = OE.frame_metrics(r, t, "f32"), OE.frame_metrics(r, t, "f64")
    # a relative MSE error e moves PSNR by 10 log10(1 + e) dB
    rel = bound / s64
    assert abs(float(m32["psnr_db"]) - float(m64["psnr_db"])) <= 10 * math.log10(1 + rel) + 1e-5
    for k in ("render_grad", "target_grad"):
        assert abs(float(m32[k]) - float(m64[k])) <= 1e-3 * float(m64[k])


def test_summarize_psnr_samples_tracks_distribution():
    """core/tests.rs summarize_psnr_samples_tracks_distribution."""
    s = E.summarize_psnr_samples([4.0, 1.0, 3.0, 2.0])
    assert abs(s.mean_db - 2.5) < 1e-6 and abs(s.median_db - 2.5) < 1e-6
    assert s.min_db == 1.0 and s.max_db == 4.0
    assert abs(s.stddev_db - 1.118034) < 1e-5
    odd = E.summarize_psnr_samples([5.0, 1.0, 3.0])
    assert odd.median_db == 3.0
    empty = E.summarize_psnr_samples([])
    assert (empty.mean_db, empty.median_db, empty.min_db, empty.max_db) == (0.0, 0.0, 0.0, 0.0)


def test_summarize_training_metrics_tracks_last_epoch():
    """core/tests.rs summarize_training_metrics_tracks_last_epoch_mean_and_last_step."""
    m = E.summarize_training_metrics([4.0, 3.0, 2.0, 1.0], 2)
    assert m == dict(final_loss=1.5, final_step_loss=1.0)
    assert E.summarize_training_metrics([], 3) == dict(final_loss=0.0, final_step_loss=0.0)
    assert E.summarize_training_metrics([2.0], 0)["final_loss"] == 2.0


def test_worst_frames_and_selection():
    """core/tests.rs worst_frame_metrics_returns_low_psnr_prefix and
    select_evaluation_frames_copies_initial_points_and_stride."""
    frames = [dict(frame_id=i, psnr_db=p) for i, p in enumerate([30.0, 10.0, 20.0, 40.0])]
    worst = E.worst_frame_metrics(frames, 2)
    assert [f["frame_id"] for f in worst] == [1, 2]
    assert len(E.worst_frame_metrics(frames, 10)) == 4
    assert E.select_evaluation_frames(list(range(6)), 0, 2) == [0, 2, 4]
    assert E.select_evaluation_frames(list(range(6)), 4, 3) == [0, 3]
    assert E.select_evaluation_frames(list(range(3)), 0, 0) == [0, 1, 2]


def test_scaled_dimensions_rounds_half_away():
    assert E.scaled_dimensions(640, 480, 0.5) == (320, 240)
    assert E.scaled_dimensions(5, 3, 0.5) == (3, 2)  # 2.5 -> 3, 1.5 -> 2
    assert E.scaled_dimensions(100, 100, 0.0) == (6, 6)  # scale clamped to 1/16: 6.25 -> 6
    assert E.scaled_dimensions(8, 8, 0.0) == (1, 1)  # 0.5 -> 1 (half away from zero)
    assert E.scaled_dimensions(64, 64, 4.0) == (64, 64)


def test_gpu_math_renderer_matches_libm_renderer(orc):
    sc = scene.gaussian_scene(2048, seed=31, sh_degree=3)
    cam = scene.ring_cameras([(96, 80)], seed=32)[0]
    c0, d0 = orc.tiled_renderer(sc, cam, cov_blur=0.3, gpu_math=False)
    c1, d1 = orc.tiled_renderer(sc, cam, cov_blur=0.3, gpu_math=True)
    assert np.abs(c0 - c1).max() < 1e-4
    assert np.mean(d0 != d1) < 1e-3


def test_renderer_empty_map_and_single_splat(orc):
    """renderer/tests.rs test_render_empty_map and test_render_depth_and_color:
    no splats give the background and zero depth; one splat in front of the
    camera gives some depth and some colour."""
    cam = scene.pinhole_identity(64, 64, 500.0)
    empty = scene.Splats(np.zeros((0, 10), np.float32), np.zeros((0, 1, 3), np.float32), np.zeros(0, np.float32), 0)
    c8, d = orc.gaussian_renderer(empty, cam)
    assert c8.shape == (64, 64, 3) and d.shape == (64, 64)
    assert not c8.any() and not d.any()
    tf = np.zeros((1, 10), np.float32)
    tf[0, 2] = 1.0
    tf[0, 3] = 1.0
    tf[0, 7:10] = np.log(0.01)
    sh = ((np.array([1.0, 128 / 255, 64 / 255], np.float32) - F(0.5)) / F(0.2820948)).reshape(1, 1, 3)
    one = scene.Splats(tf, sh.astype(np.float32), np.zeros(1, np.float32), 0)
    c8, d = orc.gaussian_renderer(one, cam)
    assert (d > 0).any() and c8.any()
