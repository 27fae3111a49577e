"""The evaluation renderer and the per-frame metrics through the C ABI,
against the oracle:

* rgs_tiled_render (TiledRenderer::render_camera_splats,
  render/tiled_renderer.rs:174-492) equals the oracle's restatement with
  libm's expf, the reference's own arithmetic (orc_tiled_renderer,
  gpu_math=0), BIT FOR BIT, colour and depth, for SH degrees 0-4, a
  non-black background, ragged image sizes (1x1, one-pixel rows and
  columns, non-multiples of 16) and edge rows (non-finite parameters, splats
  behind the camera, one covering every tile, equal depths, vanishing
  scales).  The device's exp (rgs_expf_libm) is checked against libm's
  expf on every float in tests/test_oracle_math.py.
* rgs_render_splats (GaussianRenderer::render_splats, renderer.rs:72-98):
  u8 colour and zero-mapped depth, bit for bit.
* rgs_image_metrics: PSNR and sharpness against the oracle's float64
  accumulation of the reference's float32 terms (rtol 1e-6), and against
  its float32 sequential accumulation (the reference's own) within the
  rounding bound of that sum.
* rgs_evaluate_frame and evaluate_splats: the render and the numbers of one
  evaluation frame and the summary of a few frames.
"""
import numpy as np
import pytest

from oracle import evaluation as OE
from rustscan_amd import scene
from tests.gpu_util import say

pytestmark = pytest.mark.gpu

F = np.float32


def _tiled(hip_ctx, sc, cam, bg=(0.0, 0.0, 0.0), blur=0.3):
    from rustscan_amd.evaluation import TiledRenderer
    from rustscan_amd.rasterizer import DeviceSplats

    ds = DeviceSplats(hip_ctx, sc)
    try:
        color, depth = TiledRenderer(cam.width, cam.height).with_raster_cov_blur(blur).render_camera_splats(
            hip_ctx, ds, cam, bg)
        return color.cpu().numpy(), depth.cpu().numpy()
    finally:
        ds.close()


def _assert_same(got_c, got_d, ref_c, ref_d, what):
    diff = got_c.view(np.uint32) != ref_c.view(np.uint32)
    ddiff = got_d.view(np.uint32) != ref_d.view(np.uint32)
    say(f"  {what}: {int(diff.sum())} colour / {int(ddiff.sum())} depth values differ")
    assert not diff.any() and not ddiff.any()


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
def test_tiled_render_bit_exact(hip_ctx, orc, deg):
    sc = scene.gaussian_scene(6144, seed=400 + deg, sh_degree=deg, rest_sigma=0.2)
    cams = scene.ring_cameras([(97, 61), (160, 128)], seed=410 + deg)
    bg = (0.1, 0.25, 0.7)
    for i, cam in enumerate(cams):
        say(f"tiled render deg={deg} view={i} {cam.width}x{cam.height}")
        c, d = _tiled(hip_ctx, sc, cam, bg)
        rc, rd = orc.tiled_renderer(sc, cam, background=bg, cov_blur=0.3, gpu_math=False)
        assert (rd < F(3.4028235e38)).mean() > 0.05  # the view is not empty
        _assert_same(c, d, rc, rd, f"deg {deg} view {i}")


@pytest.mark.parametrize("wh", [(1, 1), (17, 1), (1, 33), (16, 16), (31, 47), (200, 3)])
def test_tiled_render_ragged_sizes(hip_ctx, orc, wh):
    sc = scene.gaussian_scene(3000, seed=420, sh_degree=1, log_scale_mu=-2.5)
    cam = scene.ring_cameras([wh], radius=3.0, seed=421)[0]
    say(f"tiled render ragged {wh}")
    c, d = _tiled(hip_ctx, sc, cam, (0.3, 0.0, 1.0), blur=0.0)
    rc, rd = orc.tiled_renderer(sc, cam, background=(0.3, 0.0, 1.0), cov_blur=0.0, gpu_math=False)
    _assert_same(c, d, rc, rd, f"{wh}")


def _edge_scene():
    sc = scene.gaussian_scene(4096, seed=430, sh_degree=2)
    tf, sh, op = sc.transforms.copy(), sc.sh.copy(), sc.raw_opac.copy()
    nan, inf = np.float32(np.nan), np.float32(np.inf)
    tf[0, 0] = nan  # NaN mean
    tf[1, 3:7] = nan  # NaN rotation
    tf[2, 7] = inf  # infinite scale
    tf[3, 7:10] = -60.0  # vanishing scale
    op[4] = nan  # NaN opacity
    sh[5, 0, 0] = inf  # infinite colour: clamps to 1
    sh[6, 1, 2] = nan  # NaN colour: NaN clamps to NaN
    tf[7, 0:3] = [0.0, 0.0, 50.0]  # far behind / beyond the scene
    tf[8, 7:10] = 1.5  # covers the whole view
    tf[8, 0:3] = [0.0, 0.0, 0.0]
    op[8] = -3.0
    for k in range(9, 17):  # eight splats at one position (one depth): index order decides
        tf[k, 0:3] = [0.1, 0.2, 0.0]
        tf[k, 3:7] = [1.0, 0.0, 0.0, 0.0]
        tf[k, 7:10] = -2.5 + 0.1 * (k - 9)
        op[k] = 0.5
        sh[k, 0, :] = [(k % 3) - 1.0, ((k + 1) % 3) - 1.0, ((k + 2) % 3) - 1.0]
    tf[17, 0:3] = [1e30, 0.0, 0.0]  # huge mean
    return scene.Splats(tf, sh, op, sc.sh_degree)


def test_tiled_render_edge_rows(hip_ctx, orc):
    sc = _edge_scene()
    cams = scene.ring_cameras([(128, 96), (75, 131)], radius=3.0, jitter_deg=0.0, seed=431)
    for i, cam in enumerate(cams):
        say(f"tiled render edge rows view={i}")
        c, d = _tiled(hip_ctx, sc, cam, (0.2, 0.2, 0.2))
        rc, rd = orc.tiled_renderer(sc, cam, background=(0.2, 0.2, 0.2), cov_blur=0.3, gpu_math=False)
        same_nan = np.isnan(c) == np.isnan(rc)
        assert same_nan.all()
        _assert_same(np.nan_to_num(c, nan=-1.0), d, np.nan_to_num(rc, nan=-1.0), rd, f"edge view {i}")


def test_tiled_render_all_culled(hip_ctx, orc):
    """Every splat behind the camera: the background everywhere, no depth."""
    sc = scene.gaussian_scene(256, seed=440, sh_degree=0, center=(0.0, 0.0, -5.0))
    cam = scene.pinhole_identity(40, 24, 30.0)
    c, d = _tiled(hip_ctx, sc, cam, (0.5, 0.25, 0.125))
    assert np.all(c == np.array([0.5, 0.25, 0.125], np.float32)) and np.all(d == F(3.4028235e38))


def test_tiled_render_large_scene(hip_ctx, orc):
    sc = scene.gaussian_scene(65536, seed=450, sh_degree=3)
    cam = scene.ring_cameras([(640, 480)], seed=451)[0]
    say("tiled render 65536 splats 640x480")
    c, d = _tiled(hip_ctx, sc, cam)
    rc, rd = orc.tiled_renderer(sc, cam, cov_blur=0.3, gpu_math=False)
    _assert_same(c, d, rc, rd, "65536 splats")


def test_gaussian_renderer_u8(hip_ctx, orc):
    from rustscan_amd.evaluation import GaussianRenderer
    from rustscan_amd.rasterizer import DeviceSplats

    sc = scene.gaussian_scene(8192, seed=460, sh_degree=3)
    cam = scene.ring_cameras([(143, 101)], seed=461)[0]
    ds = DeviceSplats(hip_ctx, sc)
    out = GaussianRenderer(cam.width, cam.height).with_background(0.0, 0.5, 1.0).with_raster_cov_blur(0.2) \
        .render_splats(hip_ctx, ds, cam)
    ds.close()
    rc8, rd = orc.gaussian_renderer(sc, cam, background=(0.0, 0.5, 1.0), cov_blur=0.2, gpu_math=False)
    c8, d = out.color.cpu().numpy(), out.depth.cpu().numpy()
    say(f"gaussian renderer u8: {int((c8 != rc8).sum())} bytes differ")
    assert np.array_equal(c8, rc8) and np.array_equal(d.view(np.uint32), rd.view(np.uint32))
    assert (d == 0).any() and (d > 0).any()


@pytest.mark.parametrize("hw", [(1, 1), (2, 2), (3, 3), (1, 40), (37, 53), (480, 640)])
def test_image_metrics(hip_ctx, hw):
    import torch

    from rustscan_amd.evaluation import image_metrics

    H, W = hw
    rng = np.random.default_rng(H * 7919 + W)
    t = rng.uniform(0, 1, (H, W, 3)).astype(np.float32)
    r = np.clip(t + rng.normal(0, 0.05, t.shape), 0, 1).astype(np.float32)
    got = image_metrics(hip_ctx, torch.from_numpy(r).cuda(), torch.from_numpy(t).cuda())
    ref = OE.frame_metrics(r, t, "f64")
    ref32 = OE.frame_metrics(r, t, "f32")
    say(f"metrics {H}x{W}: psnr {got['psnr_db']:.6f} (f64 {ref['psnr_db']:.6f}, f32 {ref32['psnr_db']:.6f})")
    for k, v in ref.items():
        np.testing.assert_allclose(got[k], float(v), rtol=1e-6, atol=1e-12, err_msg=k)
    # the reference's float32 running sum of the squared errors
    d = (r - t).astype(np.float32)
    terms = (d * d).ravel().astype(np.float64)
    rel = (terms.size - 1) * 2.0 ** -24 + 1e-6
    assert abs(got["psnr_db"] - float(ref32["psnr_db"])) <= 10 * np.log10(1 + rel) + 1e-5


def test_image_metrics_identical_and_constant(hip_ctx):
    import torch

    from rustscan_amd.evaluation import image_metrics

    t = torch.full((20, 30, 3), 0.25, device="cuda")
    m = image_metrics(hip_ctx, t, t)
    assert m["psnr_db"] == 100.0 and m["render_grad"] == 0.0 and m["render_lap"] == 0.0
    assert m["sharpness_grad_ratio"] == 0.0 and m["sharpness_lap_ratio"] == 0.0
    m = image_metrics(hip_ctx, t + 0.125, t)
    np.testing.assert_allclose(m["psnr_db"], 20 * np.log10(8.0), rtol=1e-6)


def test_evaluate_frame_and_splats(hip_ctx, orc):
    import torch

    from rustscan_amd.evaluation import (SplatEvaluationConfig, SplatEvaluationRenderer, evaluate_frame,
                                         evaluate_splats, image_metrics)
    from rustscan_amd.rasterizer import DeviceSplats

    sc = scene.gaussian_scene(16384, seed=470, sh_degree=3)
    cams = scene.ring_cameras([(320, 240)] * 4, seed=471)
    targets = [torch.from_numpy(scene.Rng(480 + i).uniform(240 * 320 * 3).reshape(240, 320, 3).astype(np.float32))
               .cuda() for i in range(4)]
    ds = DeviceSplats(hip_ctx, sc)
    say("evaluate frame")
    m, c8 = evaluate_frame(hip_ctx, ds, cams[0], targets[0], 0.3, keep_color=True)
    rc8, _ = orc.gaussian_renderer(sc, cams[0], cov_blur=0.3, gpu_math=False)
    assert np.array_equal(c8.cpu().numpy(), rc8)
    rendered = SplatEvaluationRenderer(320, 240, 0.3).render(hip_ctx, ds, cams[0])
    exact = OE.rendered_from_u8(rc8)
    assert np.array_equal(rendered.cpu().numpy().view(np.uint32), exact.view(np.uint32))
    m2 = image_metrics(hip_ctx, rendered, targets[0])
    assert m == m2
    ref = OE.frame_metrics(OE.rendered_from_u8(rc8), targets[0].cpu().numpy(), "f64")
    for k, v in ref.items():
        np.testing.assert_allclose(m[k], float(v), rtol=1e-6, err_msg=k)
    say("evaluate splats (render scale 0.5, stride 2)")
    res = evaluate_splats(hip_ctx, ds, cams, targets, SplatEvaluationConfig(render_scale=0.5, frame_stride=2))
    ds.close()
    s = res.summary
    assert s["frame_count"] == 2 and (s["render_width"], s["render_height"]) == (160, 120)
    ps = [f["psnr_db"] for f in res.frame_metrics]
    assert s["psnr_min_db"] == min(ps) and s["psnr_max_db"] == max(ps)
    assert all(0.0 < p < 100.0 for p in ps)
    assert s["worst_frames"][0]["psnr_db"] == min(ps)
