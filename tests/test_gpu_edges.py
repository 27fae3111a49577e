"""Edge cases of the HIP path against the oracle, through the C ABI.

* Nothing in front of the camera: no visible instance and no intersection;
  the image is the background, every gradient is zero, and the Adam update
  leaves the parameters as they were.
* Images smaller than one tile and of ragged sizes, at every tile width.
* A context whose intersection capacity was sized for a sparse scene, asked
  to render (rgs_render_forward) or train on (rgs_train_step, synchronous) a
  scene with several times more intersections: the call detects the
  overflow, grows its buffers and redoes the work, and the result equals the
  oracle's.
"""
import numpy as np
import pytest

from rustscan_amd import scene
from tests.gpu_util import rel_l2, say
from tests.test_gpu_render import check_forward, grads_vs_oracle

pytestmark = pytest.mark.gpu


def test_nothing_visible(hip_ctx, orc):
    import torch

    from rustscan_amd.rasterizer import DeviceSplats, Trainer, training_optimizer

    sc = scene.gaussian_scene(1000, seed=41, sh_degree=1, center=(0.0, 0.0, -4.0))  # all behind the camera
    cam = scene.pinhole_identity(96, 64, 80.0)
    say("nothing visible: forward")
    ds, img, r = check_forward(hip_ctx, orc, sc, cam)
    assert r.stats["num_visible"] == 0 and r.stats["num_intersections"] == 0
    assert not np.any(img)
    target = np.full((64, 96, 3), 0.5, np.float32)
    say("nothing visible: backward")
    grads_vs_oracle(hip_ctx, orc, sc, cam, target)
    say("nothing visible: train step")
    tr = Trainer(hip_ctx, training_optimizer(1, 1.6e-4, 1e-3, 5e-3, 5e-2, 2.5e-3, 1.25e-4))
    ds = DeviceSplats(hip_ctx, sc)
    losses = tr.train_step(ds, [cam], [torch.from_numpy(target).cuda()], batched=True, graph=False, read_loss=True)
    np.testing.assert_allclose(losses, [0.5], rtol=1e-6)
    for g in ds.grads():
        assert not np.any(g)
    after = ds.download()
    np.testing.assert_array_equal(after.transforms, sc.transforms)
    np.testing.assert_array_equal(after.sh, sc.sh)
    np.testing.assert_array_equal(after.raw_opac, sc.raw_opac)


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_tiny_and_ragged_images(hip_ctx, orc, tile):
    sc = scene.gaussian_scene(4096, seed=72, sh_degree=1, center=(0.0, 0.0, 4.0))
    for w, h in [(1, 1), (5, 3), (17, 9), (33, 70), (129, 1)]:
        say(f"image {w}x{h}, tile {tile}")
        check_forward(hip_ctx, orc, sc, scene.pinhole_identity(w, h, 300.0), tile=tile, batch=128)
    target = scene.Rng(73).uniform(70 * 33 * 3).reshape(70, 33, 3).astype(np.float32)
    say(f"backward 33x70, tile {tile}")
    grads_vs_oracle(hip_ctx, orc, sc, scene.pinhole_identity(33, 70, 300.0), target, tile=tile, batch=128)


def sparse_and_dense():
    cam = scene.pinhole_identity(320, 240, 300.0)
    # 16,384 Gaussians each time: about 22 thousand intersections, then about 391 thousand
    sparse = scene.gaussian_scene(16384, seed=71, sh_degree=1, center=(0.0, 0.0, 4.0), log_scale_mu=-6.0)
    dense = scene.gaussian_scene(16384, seed=71, sh_degree=1, center=(0.0, 0.0, 4.0), log_scale_mu=-2.0)
    return cam, sparse, dense


def test_render_grows_its_capacity(orc):
    from rustscan_amd import _lib as L
    from rustscan_amd.rasterizer import Context, DeviceSplats, render_forward

    cam, sparse, dense = sparse_and_dense()
    ctx = Context(0)  # fresh: sized by the sparse scene only
    try:
        ctx.set_raster(16, 256, True)
        ds = DeviceSplats(ctx, sparse)
        render_forward(ctx, ds, [cam])
        m_sparse = ctx.counts(1)[0][1]
        ds.upload(dense)
        img = render_forward(ctx, ds, [cam])[0].cpu().numpy()
        m_dense = ctx.counts(1)[0][1]
        say(f"  intersections {m_sparse} then {m_dense}")
        r = orc.Render(dense, cam, tile_width=16)
        assert m_dense == r.stats["num_intersections"] > 4 * max(m_sparse, 1 << 16)
        assert float(np.abs(img - r.image()).max()) <= 1e-4
        tiles, gids = r.isect()
        np.testing.assert_array_equal(ctx.debug(L.RGS_DEBUG_ISECT_TILES), tiles)
        np.testing.assert_array_equal(ctx.debug(L.RGS_DEBUG_ISECT_GIDS), gids)
        np.testing.assert_array_equal(ctx.debug(L.RGS_DEBUG_TILE_RANGES), r.tile_offsets())
        ds.close()
    finally:
        ctx.close()


def test_train_step_grows_its_capacity(orc):
    import torch

    from rustscan_amd.rasterizer import Context, DeviceSplats, Trainer, training_optimizer

    cam, sparse, dense = sparse_and_dense()
    target = scene.Rng(74).uniform(240 * 320 * 3).reshape(240, 320, 3).astype(np.float32)
    tgt = [torch.from_numpy(target).cuda()]
    ctx = Context(0)
    try:
        ctx.set_raster(16, 256, True)
        tr = Trainer(ctx, training_optimizer(1, 1.6e-4, 1e-3, 5e-3, 5e-2, 2.5e-3, 1.25e-4))
        ds = DeviceSplats(ctx, sparse)
        tr.train_step(ds, [cam], tgt, batched=True, graph=False, read_loss=True)
        ds.upload(dense)
        say("train step on the dense scene after sizing for the sparse one")
        losses = tr.train_step(ds, [cam], tgt, batched=True, graph=False, read_loss=True)
        vt, vs, vo, _ = ds.grads()
        r = orc.Render(dense, cam)
        ref_loss, v = orc.l1_loss_grad(r.image(), target)
        g = r.backward(v, 0)
        np.testing.assert_allclose(losses, [ref_loss], rtol=1e-5)
        errs = [rel_l2(vt, g["v_transforms"]), rel_l2(vs, g["v_sh"][:, : vs.shape[1]]), rel_l2(vo, g["v_raw_opac"])]
        say("  grad rel-L2 " + " ".join(f"{e:.2e}" for e in errs))
        assert max(errs) <= 1e-3
        assert ctx.skipped_steps() == 0
        ds.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("graph", [True, False])
def test_async_steps_recover_from_overflow(orc, graph):
    """Asynchronous steps (no loss read back) on a context sized for the
    sparse scene, then given the dense one: the first dense step overflows
    and is skipped, the next call sees the overflow left by that step, grows
    and re-plans, and every later step trains; the following step's
    gradients equal the oracle's at the parameters it starts from."""
    import torch

    from rustscan_amd.rasterizer import Context, DeviceSplats, Trainer, training_optimizer

    cam, sparse, dense = sparse_and_dense()
    target = scene.Rng(75).uniform(240 * 320 * 3).reshape(240, 320, 3).astype(np.float32)
    tgt = [torch.from_numpy(target).cuda()]
    ctx = Context(0)
    try:
        ctx.set_raster(16, 256, True)
        tr = Trainer(ctx, training_optimizer(1, 1.6e-4, 1e-3, 5e-3, 5e-2, 2.5e-3, 1.25e-4))
        ds = DeviceSplats(ctx, sparse)
        tr.train_step(ds, [cam], tgt, batched=True, graph=graph, read_loss=True)  # plans for the sparse scene
        ds.upload(dense)
        before = ctx.skipped_steps()
        say(f"async dense steps after sizing for the sparse scene (graph={graph})")
        start = ds.download()
        for _ in range(5):
            tr.train_step(ds, [cam], tgt, batched=True, graph=graph)
        ctx.sync()
        skipped = ctx.skipped_steps() - before
        say(f"  skipped {skipped} of 5 asynchronous steps")
        assert skipped <= 1
        assert not np.array_equal(ds.download().transforms, start.transforms)
        cur = ds.download()
        losses = tr.train_step(ds, [cam], tgt, batched=True, graph=graph, read_loss=True)
        assert ctx.skipped_steps() - before == skipped
        vt, vs, vo, _ = ds.grads()
        r = orc.Render(cur, cam)
        ref_loss, v = orc.l1_loss_grad(r.image(), target)
        g = r.backward(v, 0)
        np.testing.assert_allclose(losses, [ref_loss], rtol=1e-5)
        errs = [rel_l2(vt, g["v_transforms"]), rel_l2(vs, g["v_sh"]), rel_l2(vo, g["v_raw_opac"])]
        say("  grad rel-L2 " + " ".join(f"{e:.2e}" for e in errs))
        assert max(errs) <= 1e-3
        ds.close()
    finally:
        ctx.close()
