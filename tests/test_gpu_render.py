"""HIP render path against the oracle through the C ABI.

Forward (config 1, and config 2 at reduced and full size): depth order, per-
Gaussian tile counts, sorted intersections and tile ranges bit-exact under the
canonical tie rule; projected splats and images within 1e-4 (the shared
deterministic math makes them equal in practice, the maximum is printed);
visibility flags equal.  Backward: per-group gradients within 1e-3 relative
L2 of the oracle's reference-semantics backward.
"""
import numpy as np
import pytest

from rustscan_amd import scene
from tests.gpu_util import rel_l2, say

pytestmark = pytest.mark.gpu


def render_both(hip_ctx, orc, sc, cam, blur=0.3, tile=16, batch=256):
    from rustscan_amd.rasterizer import DeviceSplats, render_forward

    hip_ctx.set_raster(tile, batch, True)
    ds = DeviceSplats(hip_ctx, sc)
    imgs = render_forward(hip_ctx, ds, [cam], cov_blur=blur)
    r = orc.Render(sc, cam, cov_blur=blur, tile_width=tile)
    return ds, imgs[0].cpu().numpy(), r


def check_forward(hip_ctx, orc, sc, cam, tile=16, batch=256):
    from rustscan_amd import _lib as L

    ds, img, r = render_both(hip_ctx, orc, sc, cam, tile=tile, batch=batch)
    ref = r.image()
    err = float(np.abs(img - ref).max())
    say(f"  image max abs err {err:.3g} (V={r.stats['num_visible']}, M={r.stats['num_intersections']})")
    assert err <= 1e-4
    order = hip_ctx.debug(L.RGS_DEBUG_DEPTH_ORDER)
    np.testing.assert_array_equal(order, r.depth_order())
    np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_TILE_COUNTS), r.counts())
    tiles, gids = r.isect()
    np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_ISECT_TILES), tiles)
    np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_ISECT_GIDS), gids)
    np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_TILE_RANGES), r.tile_offsets())
    proj = hip_ctx.debug(L.RGS_DEBUG_PROJECTED, dtype=np.float32).reshape(-1, 9)
    if r.depth_order().size:
        perr = float(np.abs(proj[r.depth_order()] - r.projected()).max())
        say(f"  projected max abs err {perr:.3g}")
        assert perr <= 1e-4 * max(1.0, float(np.abs(r.projected()).max()))
    np.testing.assert_array_equal(ds.visible(), r.visible())
    return ds, img, r


def test_forward_single_gaussian(hip_ctx, orc):
    # the reference's forward test scene (forward/tests.rs:8-39): one Gaussian at z=2, 64x64, f=500
    import math

    tf = np.array([[0, 0, 2, 1, 0, 0, 0, math.log(0.2), math.log(0.2), math.log(0.2)]], np.float32)
    sh = np.array([[[(c - 0.5) / 0.28209479177387814 for c in (1.0, 0.5, 0.25)]]], np.float32)
    sc = scene.Splats(tf, sh, np.zeros(1, np.float32), 0)
    say("forward single gaussian")
    check_forward(hip_ctx, orc, sc, scene.pinhole_identity(64, 64, 500.0))


def test_forward_config1(hip_ctx, orc):
    sc, cams = scene.config1()
    say("forward config1")
    check_forward(hip_ctx, orc, sc, cams[0])


def test_forward_config2_small(hip_ctx, orc):
    sc, cams, _ = scene.config2(n=16384, width=320, height=240)
    say("forward config2 small")
    check_forward(hip_ctx, orc, sc, cams[0])


def test_forward_config2_full(hip_ctx, orc):
    sc, cams, _ = scene.config2()
    say("forward config2 full")
    check_forward(hip_ctx, orc, sc, cams[0])


def grads_vs_oracle(hip_ctx, orc, sc, cam, target, tile=16, batch=256, tol=1e-3, reference="f32"):
    """Per-group gradients against the oracle's FP32 backward, or (reference
    "f64") against its FP64 replay of the same forward; the FP32 oracle's own
    distance to the FP64 replay is printed beside."""
    import torch

    from rustscan_amd.rasterizer import l1_loss_grad, render_backward

    ds, img, r = render_both(hip_ctx, orc, sc, cam, tile=tile, batch=batch)
    tgt = torch.from_numpy(target).cuda()
    loss, v_out = l1_loss_grad(hip_ctx, torch.from_numpy(img).cuda(), tgt)
    ref_loss, ref_vout = orc.l1_loss_grad(r.image(), target)
    assert abs(loss - ref_loss) <= 1e-5 * max(1.0, abs(ref_loss))
    np.testing.assert_array_equal(v_out.cpu().numpy(), ref_vout)
    vt, vs, vo, st = render_backward(hip_ctx, ds, [v_out])
    ref = r.backward(ref_vout, mode=0)
    if reference == "f64":
        r64 = r.backward_f64(ref_vout, mode=0)
        f32_vs_f64 = {
            "means": rel_l2(ref["v_transforms"][:, 0:3], r64["v_transforms"][:, 0:3]),
            "quats": rel_l2(ref["v_transforms"][:, 3:7], r64["v_transforms"][:, 3:7]),
            "log_scales": rel_l2(ref["v_transforms"][:, 7:10], r64["v_transforms"][:, 7:10]),
        }
        say("  FP32 oracle vs FP64 replay " + " ".join(f"{k}={v:.2e}" for k, v in f32_vs_f64.items()))
        ref = dict(ref, v_transforms=r64["v_transforms"], v_sh=r64["v_sh"], v_raw_opac=r64["v_raw_opac"])
    errs = {
        "means": rel_l2(vt[:, 0:3], ref["v_transforms"][:, 0:3]),
        "quats": rel_l2(vt[:, 3:7], ref["v_transforms"][:, 3:7]),
        "log_scales": rel_l2(vt[:, 7:10], ref["v_transforms"][:, 7:10]),
        "opacity": rel_l2(vo, ref["v_raw_opac"]),
        "sh": rel_l2(vs, ref["v_sh"]),
        "screen_stats": rel_l2(st, ref["screen_stats"]),
    }
    say("  grad rel-L2 " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()))
    for k, v in errs.items():
        assert v <= tol, f"{k}: rel L2 {v:.3e}"
    return errs


def test_backward_config1(hip_ctx, orc):
    sc, cams = scene.config1()
    target = scene.Rng(3).uniform(320 * 240 * 3).reshape(240, 320, 3).astype(np.float32)
    say("backward config1")
    grads_vs_oracle(hip_ctx, orc, sc, cams[0], target)


def test_backward_config2_small(hip_ctx, orc):
    sc, cams, tg = scene.config2(n=16384, width=320, height=240)
    say("backward config2 small")
    grads_vs_oracle(hip_ctx, orc, sc, cams[0], tg[0])


def test_backward_config2_full(hip_ctx, orc):
    sc, cams, tg = scene.config2()
    say("backward config2 full")
    grads_vs_oracle(hip_ctx, orc, sc, cams[0], tg[0])


# ---- rotated cameras (ring around the scene) and several views per launch


def ring_scene(n=65536, seed=31):
    sc = scene.gaussian_scene(n, seed=seed, sh_degree=3)
    cams = scene.ring_cameras([(640, 360), (320, 240), (480, 272)], seed=seed)
    return sc, cams


def test_forward_ring_cameras(hip_ctx, orc):
    sc, cams = ring_scene()
    for i, cam in enumerate(cams):
        say(f"forward ring camera {i} {cam.width}x{cam.height}")
        check_forward(hip_ctx, orc, sc, cam)


def test_backward_ring_camera(hip_ctx, orc):
    sc, cams = ring_scene()
    cam = cams[0]
    target = scene.Rng(41).uniform(cam.width * cam.height * 3).reshape(cam.height, cam.width, 3).astype(np.float32)
    say("backward ring camera")
    grads_vs_oracle(hip_ctx, orc, sc, cam, target)


def test_batched_forward_matches_oracle_per_view(hip_ctx, orc):
    """Several uneven views in one launch: every view equals the oracle's single-view render."""
    from rustscan_amd import _lib as L
    from rustscan_amd.rasterizer import DeviceSplats, render_forward

    sc, cams = ring_scene(n=32768, seed=32)
    hip_ctx.set_raster(16, 256, True)
    ds = DeviceSplats(hip_ctx, sc)
    imgs = render_forward(hip_ctx, ds, cams)
    vis_or = np.zeros(sc.n, np.float32)
    for v, cam in enumerate(cams):
        r = orc.Render(sc, cam)
        err = float(np.abs(imgs[v].cpu().numpy() - r.image()).max())
        say(f"  view {v}: image max abs err {err:.3g}")
        assert err <= 1e-4
        np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_DEPTH_ORDER, v), r.depth_order())
        tiles, gids = r.isect()
        np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_ISECT_TILES, v), tiles)
        np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_ISECT_GIDS, v), gids)
        np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_TILE_RANGES, v), r.tile_offsets())
        vis_or = np.maximum(vis_or, r.visible())
    np.testing.assert_array_equal(ds.visible(), vis_or)


# ---- needle-like splats: the compositing kernels skip a splat for an 8x8
# pixel block only when the conic's minimum over the block provably exceeds
# the alpha threshold; near-degenerate conics sit at the edge of that bound.
# The images (and binning) must still equal the oracle's.
#
# Their gradients are compared with the oracle's backward replayed in FP64
# over the same FP32 forward (same lists, splats and skip/stop decisions).
# The FP32 restatement of the reference's projection backward is not a usable
# reference here: on this scene it is 4.5e-2 (means) and 1.8e-1 (quaternions)
# rel-L2 away from the FP64 replay, while the FP32 compositing backward
# followed by an FP64 projection backward is within 1.5e-4 of it
# (tests/test_oracle_conditioning.py).  The kernel runs the projection
# backward's covariance chain in FP64, so it is held to 1e-3 of the FP64
# replay on every group.


def needle_scene():
    sc = scene.gaussian_scene(16384, seed=91, sh_degree=1, log_scale_mu=-3.2, log_scale_sigma=1.5)
    cam = scene.ring_cameras([(400, 300)], seed=92)[0]
    return sc, cam


@pytest.mark.parametrize("tile,batch", [(8, 128), (16, 256), (32, 128), (32, 256)])
def test_forward_needle_splats(hip_ctx, orc, tile, batch):
    sc, cam = needle_scene()
    say(f"forward needle splats, tile {tile} batch {batch}")
    check_forward(hip_ctx, orc, sc, cam, tile=tile, batch=batch)


@pytest.mark.parametrize("tile,batch", [(16, 256), (32, 256)])
def test_backward_needle_splats(hip_ctx, orc, tile, batch):
    sc, cam = needle_scene()
    target = scene.Rng(93).uniform(cam.width * cam.height * 3).reshape(cam.height, cam.width, 3).astype(np.float32)
    say(f"backward needle splats vs the FP64 replay, tile {tile} batch {batch}")
    grads_vs_oracle(hip_ctx, orc, sc, cam, target, tile=tile, batch=batch, reference="f64")


# ---- the fast paths against their plain forms: the packed tile sort and the
# count-carrying depth sort against plain (key, value) sorts, and the per-block
# culling against every wave walking every splat.  Results must be identical.


@pytest.mark.parametrize("tile,batch", [(16, 256), (32, 128)])
def test_packed_sorts_and_culling_match_plain(hip_ctx, tile, batch):
    import torch

    from rustscan_amd import _lib as L
    from rustscan_amd.rasterizer import DeviceSplats, l1_loss_grad, render_backward, render_forward

    sc, cams = ring_scene(n=32768, seed=33)
    target = torch.rand((cams[0].height, cams[0].width, 3), device="cuda",
                        generator=torch.Generator(device="cuda").manual_seed(34))
    hip_ctx.set_raster(tile, batch, True)
    what = [L.RGS_DEBUG_DEPTH_ORDER, L.RGS_DEBUG_TILE_COUNTS, L.RGS_DEBUG_ISECT_TILES, L.RGS_DEBUG_ISECT_GIDS,
            L.RGS_DEBUG_TILE_RANGES]

    def run(plain):
        hip_ctx.set_debug(plain_sorts=plain, no_cull=plain)
        try:
            ds = DeviceSplats(hip_ctx, sc)
            imgs = render_forward(hip_ctx, ds, cams)
            dbg = [[hip_ctx.debug(w, v) for w in what] for v in range(len(cams))]
            _, v_out = l1_loss_grad(hip_ctx, imgs[0], target)
            grads = render_backward(hip_ctx, ds, [v_out] + [torch.zeros_like(im) for im in imgs[1:]])
            out = ([im.cpu().numpy() for im in imgs], dbg, grads, ds.visible())
            ds.close()
            return out
        finally:
            hip_ctx.set_debug()

    say(f"packed sorts and culling vs plain, tile {tile} batch {batch}")
    fast, plain = run(False), run(True)
    for a, b in zip(fast[0], plain[0]):
        np.testing.assert_array_equal(a, b)
    for va, vb in zip(fast[1], plain[1]):
        for a, b in zip(va, vb):
            np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(fast[3], plain[3])
    errs = [rel_l2(a, b) for a, b in zip(fast[2], plain[2])]
    say("  grads rel-L2 " + " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) <= 1e-6
