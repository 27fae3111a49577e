"""The benchmark configurations at full size through the C ABI.

* 8-view batched step (1,048,576 Gaussians, SH3, 8 views of uneven size) at
  the configurations the bench's autotuner chooses between (32-pixel tiles,
  batch depth 128 and 256, persistent launch, HIP graph): images within 1e-4
  of the oracle, each gradient group within 1e-3 (rel-L2) of the oracle's
  backward summed over the views.
* The evaluation render (the CPU renderer's semantics) of the same scene's
  1280x720 view: colour and depth equal to the oracle's CPU renderer (with
  libm's expf, as the reference) bit for bit.
* Single-view training (1,048,576 Gaussians, 32 ring views at 1280x720, one
  view and one Adam update per step, graph replayed): the first step's
  gradients equal the oracle's, 40 replayed steps lower the L1 loss over the
  first 8 views, and no step is skipped.
* Large scene (4,194,304 Gaussians with a 1% x20 scale tail) on two
  1920x1080 views in one batched launch: no overflow, binning invariants
  (sorted keys, ranges covering every intersection, counts summing to M),
  and the first view's image equal to the oracle's.
* The same scene's full batch, 16 views at 1920x1080, in one persistent
  launch (graph-captured step) at the bench's configuration: loss and every
  gradient group equal to the per-view launches' (rel-L2 <= 1e-3), no step
  skipped; the first view's step alone equal to the oracle's gradients (the
  backward that recomputes tile ranges, at this scene's scale).
* The autotuner leaves parameters, Adam moments, step counters and the
  skipped-step count as it found them.
* Adam with weight decay and per-group step counters.
"""
import functools

import numpy as np
import pytest

from rustscan_amd import scene
from tests.gpu_util import rel_l2, say

pytestmark = pytest.mark.gpu

LRS = dict(lr_position=1.6e-4, lr_rotation=1e-3, lr_scale=5e-3, lr_opacity=5e-2, lr_sh_dc=2.5e-3, lr_sh_rest=1.25e-4)


@functools.lru_cache(maxsize=1)
def config5_scene():
    """The 4,194,304-Gaussian scene, generated once for the tests that use it."""
    return scene.config5_scene()


def group_errors(vt, vs, vo, rvt, rvs, rvo) -> dict:
    return dict(means=rel_l2(vt[:, :3], rvt[:, :3]), quats=rel_l2(vt[:, 3:7], rvt[:, 3:7]),
                log_scales=rel_l2(vt[:, 7:], rvt[:, 7:]), opacity=rel_l2(vo, rvo), sh=rel_l2(vs, rvs))


def test_config4_full_step_at_shipped_config(hip_ctx, orc):
    import torch

    from rustscan_amd.rasterizer import DeviceSplats, Trainer, render_forward, training_optimizer

    init, _ = scene.config3_scene()
    cams = scene.config4_cameras()
    tg = [scene.Rng(200 + i).uniform(c.width * c.height * 3).reshape(c.height, c.width, 3).astype(np.float32)
          for i, c in enumerate(cams)]
    tgt = [torch.from_numpy(t).cuda() for t in tg]
    w = 1.0 / len(cams)
    runs = {}
    try:
        for batch in (128, 256):
            hip_ctx.set_raster(32, batch, True)
            ds = DeviceSplats(hip_ctx, init)
            skipped0 = hip_ctx.skipped_steps()  # the context is shared with tests that skip on purpose
            say(f"config4: 1,048,576 Gaussians, 8 views, tile 32 batch {batch} persistent, graph")
            imgs = [im.cpu().numpy() for im in render_forward(hip_ctx, ds, cams)]
            tr = Trainer(hip_ctx, training_optimizer(3, **LRS), loss_weight=w)
            losses = tr.train_step(ds, cams, tgt, batched=True, graph=True, read_loss=True)
            runs[batch] = (imgs, losses, ds.grads())
            assert hip_ctx.skipped_steps() == skipped0
            ds.close()
    finally:
        hip_ctx.set_raster(16, 256, True)
    # the oracle's images and gradients depend on the tile width only
    rvt = rvs = rvo = None
    ref_imgs, ref_losses = [], []
    for cam, t in zip(cams, tg):
        r = orc.Render(init, cam, tile_width=32)
        img = r.image()
        loss, v = orc.l1_loss_grad(img, t, weight=w)
        g = r.backward(v, 0)
        rvt = g["v_transforms"] if rvt is None else rvt + g["v_transforms"]
        rvs = g["v_sh"] if rvs is None else rvs + g["v_sh"]
        rvo = g["v_raw_opac"] if rvo is None else rvo + g["v_raw_opac"]
        ref_imgs.append(img)
        ref_losses.append(loss)
        del r
    for batch, (imgs, losses, (vt, vs, vo, _)) in runs.items():
        worst_img = max(float(np.abs(a - b).max()) for a, b in zip(imgs, ref_imgs))
        np.testing.assert_allclose(losses, ref_losses, rtol=1e-5)
        errs = group_errors(vt, vs, vo, rvt, rvs, rvo)
        say(f"  batch {batch}: image max err {worst_img:.2e}; grads " +
            " ".join(f"{k}={e:.2e}" for k, e in errs.items()))
        assert worst_img <= 1e-4
        assert max(errs.values()) <= 1e-3


def test_config4_evaluation_render(hip_ctx, orc):
    from rustscan_amd.evaluation import TiledRenderer
    from rustscan_amd.rasterizer import DeviceSplats

    init, _ = scene.config3_scene()
    cam = scene.config4_cameras()[2]
    say("evaluation render: 1,048,576 Gaussians, 1280x720, the CPU renderer's semantics")
    ds = DeviceSplats(hip_ctx, init)
    gc, gd = TiledRenderer(cam.width, cam.height).with_raster_cov_blur(0.3).render_camera_splats(hip_ctx, ds, cam)
    gc, gd = gc.cpu().numpy(), gd.cpu().numpy()
    ds.close()
    # the reference's own arithmetic (libm expf)
    oc, od = orc.tiled_renderer(init, cam, cov_blur=0.3)
    say(f"  {int((gc != oc).sum())} colour values and {int((gd != od).sum())} depth pixels differ from the oracle")
    assert np.array_equal(gc, oc) and np.array_equal(gd, od)


def test_config3_graph_replayed_training(hip_ctx, orc):
    import torch

    from rustscan_amd.rasterizer import DeviceSplats, Trainer, render_forward, training_optimizer

    init, gt = scene.config3_scene()
    cams = scene.config3_cameras()
    g = DeviceSplats(hip_ctx, gt)
    targets = [render_forward(hip_ctx, g, [c])[0][..., :3].contiguous() for c in cams]
    g.close()
    ds = DeviceSplats(hip_ctx, init)
    tr = Trainer(hip_ctx, training_optimizer(3, **LRS), loss_weight=1.0)

    def eval_l1():
        out = []
        for i in range(8):
            img = render_forward(hip_ctx, ds, [cams[i]])[0]
            out.append(float((img[..., :3] - targets[i]).abs().mean()))
        return float(np.mean(out))

    say("config3: 1,048,576 Gaussians, 1280x720, graph-replayed single-view steps")
    before = eval_l1()
    skipped0 = hip_ctx.skipped_steps()
    loss0 = tr.train_step(ds, [cams[0]], [targets[0]], batched=True, graph=True, read_loss=True)
    vt, vs, vo, _ = ds.grads()
    r = orc.Render(init, cams[0])
    t0 = targets[0].cpu().numpy()
    ref_loss, v = orc.l1_loss_grad(r.image(), t0)
    ref = r.backward(v, 0)
    del r
    np.testing.assert_allclose(loss0[0], ref_loss, rtol=1e-5)
    errs = group_errors(vt, vs, vo, ref["v_transforms"], ref["v_sh"], ref["v_raw_opac"])
    say("  first step grads " + " ".join(f"{k}={e:.2e}" for k, e in errs.items()))
    assert max(errs.values()) <= 1e-3
    for it in range(1, 41):
        i = it % len(cams)
        tr.train_step(ds, [cams[i]], [targets[i]], batched=True, graph=True)
    hip_ctx.sync()
    after = eval_l1()
    skipped = hip_ctx.skipped_steps() - skipped0
    say(f"  L1 over the first 8 views {before:.5f} -> {after:.5f} after 41 steps; skipped {skipped}")
    assert skipped == 0
    assert after < before
    ds.close()


def test_config5_two_views(hip_ctx, orc):
    import torch

    from rustscan_amd._lib import RGS_DEBUG_ISECT_TILES, RGS_DEBUG_TILE_RANGES
    from rustscan_amd.rasterizer import DeviceSplats, render_forward

    sc = config5_scene()
    cams = scene.config5_cameras()[:2]
    hip_ctx.set_raster(32, 256, True)
    try:
        say("config5: 4,194,304 Gaussians (1% x20 scale tail), two 1920x1080 views in one launch")
        ds = DeviceSplats(hip_ctx, sc)
        imgs = render_forward(hip_ctx, ds, cams)
        counts = hip_ctx.counts(2)
        say(f"  (visible, intersections, tiles) per view {counts}")
        for v in range(2):
            keys = hip_ctx.debug(RGS_DEBUG_ISECT_TILES, v)
            ranges = hip_ctx.debug(RGS_DEBUG_TILE_RANGES, v).reshape(-1, 2).astype(np.int64)
            m = counts[v][1]
            assert len(keys) == m
            assert np.all(np.diff(keys.astype(np.int64)) >= 0)
            nonempty = ranges[ranges[:, 1] > ranges[:, 0]]
            assert int((ranges[:, 1] - ranges[:, 0]).sum()) == m
            assert nonempty[0, 0] == 0 and nonempty[-1, 1] == m
            assert np.all(nonempty[1:, 0] == nonempty[:-1, 1])
        r = orc.Render(sc, cams[0], tile_width=32)
        st = r.stats
        assert counts[0][0] == st["num_visible"] and counts[0][1] == st["num_intersections"]
        err = float(np.abs(imgs[0].cpu().numpy() - r.image()).max())
        say(f"  view 0 image max err {err:.2e} against the oracle ({st['num_intersections']} intersections)")
        assert err <= 1e-4
        del r
        ds.close()
    finally:
        hip_ctx.set_raster(16, 256, True)


def test_config5_full_batch_gradients(hip_ctx, orc):
    import torch

    from rustscan_amd.rasterizer import DeviceSplats, Trainer, training_optimizer

    sc = config5_scene()
    cams = scene.config5_cameras()
    gen = torch.Generator(device="cuda").manual_seed(10)
    tgt = [torch.rand((c.height, c.width, 3), device="cuda", generator=gen) for c in cams]
    w = 1.0 / len(cams)
    runs = {}
    hip_ctx.set_raster(32, 128, True)
    try:
        skipped0 = hip_ctx.skipped_steps()
        for batched in (True, False):
            say(f"config5: 4,194,304 Gaussians, 16 views 1920x1080, {'one batched launch, graph' if batched else 'per-view launches'}")
            ds = DeviceSplats(hip_ctx, sc)
            tr = Trainer(hip_ctx, training_optimizer(3, **LRS), loss_weight=w)
            losses = tr.train_step(ds, cams, tgt, batched=batched, graph=batched, read_loss=True)
            runs[batched] = (np.asarray(losses), ds.grads()[:3])
            ds.close()
        say("config5: view 0 alone")
        ds = DeviceSplats(hip_ctx, sc)
        tr = Trainer(hip_ctx, training_optimizer(3, **LRS), loss_weight=w)
        loss0 = tr.train_step(ds, cams[:1], tgt[:1], batched=True, graph=False, read_loss=True)[0]
        g0 = ds.grads()[:3]
        ds.close()
        assert hip_ctx.skipped_steps() == skipped0
    finally:
        hip_ctx.set_raster(16, 256, True)
    (lb, gb), (lp, gp) = runs[True], runs[False]
    np.testing.assert_allclose(lb, lp, rtol=1e-5)
    errs = group_errors(*gb, *gp)
    say("  batched vs per-view grads " + " ".join(f"{k}={e:.2e}" for k, e in errs.items()))
    assert max(errs.values()) <= 1e-3
    r = orc.Render(sc, cams[0], tile_width=32)
    loss, v = orc.l1_loss_grad(r.image(), tgt[0].cpu().numpy(), weight=w)
    g = r.backward(v, 0)
    del r
    assert abs(loss0 - loss) <= 1e-5 * abs(loss)
    errs = group_errors(*g0, g["v_transforms"], g["v_sh"], g["v_raw_opac"])
    say("  view 0 vs oracle grads " + " ".join(f"{k}={e:.2e}" for k, e in errs.items()))
    assert max(errs.values()) <= 1e-3


def test_autotune_restores_state(hip_ctx):
    import torch

    from rustscan_amd.rasterizer import AdamScaled, DeviceSplats, Trainer, training_optimizer

    sc = scene.gaussian_scene(32768, seed=81, sh_degree=3)
    cams = scene.ring_cameras([(480, 272), (320, 240)], seed=82)
    tgt = [torch.from_numpy(scene.Rng(83 + i).uniform(c.width * c.height * 3)
                            .reshape(c.height, c.width, 3).astype(np.float32)).cuda() for i, c in enumerate(cams)]
    hip_ctx.set_raster(16, 256, True)
    tr = Trainer(hip_ctx, training_optimizer(3, **LRS), loss_weight=0.5)
    a = DeviceSplats(hip_ctx, sc)
    for _ in range(2):  # moments and step counters away from zero
        tr.train_step(a, cams, tgt, batched=True, graph=False, read_loss=True)
    skipped = hip_ctx.skipped_steps()
    p0, (m0, v0), s0 = a.download(), a.adam_moments(), AdamScaled.steps(hip_ctx, a)
    say("autotune on a store with two steps of Adam state")
    chosen = tr.autotune(a, cams, tgt, batched=True, graph=True)
    say(f"  chose tile {chosen[0]}, batch {chosen[1]}")
    assert hip_ctx.skipped_steps() == skipped
    assert AdamScaled.steps(hip_ctx, a) == s0 == [2, 2, 2]
    p1, (m1, v1) = a.download(), a.adam_moments()
    for x, y in [(p0.transforms, p1.transforms), (p0.sh, p1.sh), (p0.raw_opac, p1.raw_opac), *zip(m0, m1),
                 *zip(v0, v1)]:
        np.testing.assert_array_equal(x, y)
    assert np.abs(m0[0]).max() > 0 and np.abs(v0[1]).max() > 0
    # the moments round-trip through import
    a.set_adam_moments(m0, v0)
    m2, v2 = a.adam_moments()
    for x, y in [*zip(m0, m2), *zip(v0, v2)]:
        np.testing.assert_array_equal(x, y)
    a.close()
    hip_ctx.set_raster(16, 256, True)


def test_adam_weight_decay_and_group_steps(hip_ctx, orc):
    """rgs_adam_step with weight decay equals the oracle's Adam on the same
    gradients, bit for bit; the three group step counters advance together,
    can be set (resume), and drive the bias corrections."""
    import torch

    from rustscan_amd.rasterizer import AdamScaled, DeviceSplats, render_backward, render_forward, \
        training_optimizer

    sc = scene.gaussian_scene(4096, seed=85, sh_degree=1)
    cam = scene.ring_cameras([(320, 240)], seed=86)[0]
    ds = DeviceSplats(hip_ctx, sc)
    img = render_forward(hip_ctx, ds, [cam])[0]
    gen = torch.Generator(device="cuda").manual_seed(3)
    vt, vs, vo, _ = render_backward(hip_ctx, ds, [torch.randn(img.shape, device="cuda", generator=gen)])
    opt = training_optimizer(1, **LRS)
    opt.config.weight_decay = 0.05
    AdamScaled.set_steps(hip_ctx, ds, [6, 6, 6])
    opt.step_device_splats(hip_ctx, ds)
    assert AdamScaled.steps(hip_ctx, ds) == [7, 7, 7]
    got = ds.download()
    flat = np.concatenate([sc.transforms, sc.raw_opac[:, None], sc.sh.reshape(sc.n, -1)], axis=1).copy()
    g = np.concatenate([vt, vo[:, None], vs.reshape(sc.n, -1)], axis=1).copy()
    lr = np.array([LRS["lr_position"]] * 3 + [LRS["lr_rotation"]] * 4 + [LRS["lr_scale"]] * 3 + [LRS["lr_opacity"]]
                  + [LRS["lr_sh_dc"]] * 3 + [LRS["lr_sh_rest"]] * 9, np.float32)
    m = np.zeros_like(flat)
    v = np.zeros_like(flat)
    orc.adam_group(flat, g, m, v, lr, 0.9, 0.999, 1e-8, 7, weight_decay=0.05)
    out = np.concatenate([got.transforms, got.raw_opac[:, None], got.sh.reshape(sc.n, -1)], axis=1)
    np.testing.assert_array_equal(out, flat)
    opt.reset(hip_ctx, ds)
    assert AdamScaled.steps(hip_ctx, ds) == [0, 0, 0]
    ds.close()
