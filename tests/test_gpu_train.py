"""Training step through the C ABI (rgs_train_step) against the oracle.

* Batched multi-view step (uneven views, one launch per stage): L1 losses and
  gradients summed over views equal the oracle's per-view backward (1e-3
  relative L2 per group); the Adam update equals the oracle's
  (engine/optimizer.rs:122-162) applied to the HIP gradients.
* Per-view launches and batched launches give the same gradients.
* A captured HIP graph replayed several times tracks eager steps.
* A non-finite loss skips the update (trainer.rs:297-302).
* Every tile width / LDS batch depth the autotuner chooses from renders the
  oracle's image at that tile width and matches its gradients.
"""
import numpy as np
import pytest

from rustscan_amd import scene
from tests.gpu_util import rel_l2, say

pytestmark = pytest.mark.gpu

LRS = dict(lr_position=1.6e-4, lr_rotation=1e-3, lr_scale=5e-3, lr_opacity=5e-2, lr_sh_dc=2.5e-3, lr_sh_rest=1.25e-4)


@pytest.fixture(scope="module")
def setup(hip_ctx):
    import torch

    sc = scene.gaussian_scene(32768, seed=51, sh_degree=3)
    cams = scene.ring_cameras([(480, 272), (320, 240), (640, 360)], seed=52)
    tg = [scene.Rng(60 + i).uniform(c.width * c.height * 3).reshape(c.height, c.width, 3).astype(np.float32)
          for i, c in enumerate(cams)]
    tgt = [torch.from_numpy(t).cuda() for t in tg]
    return sc, cams, tg, tgt


def oracle_step(orc, sc, cams, tg, weight):
    vt = np.zeros((sc.n, 10), np.float32)
    vs = np.zeros((sc.n, 16, 3), np.float32)
    vo = np.zeros(sc.n, np.float32)
    losses = []
    for cam, t in zip(cams, tg):
        r = orc.Render(sc, cam)
        loss, v = orc.l1_loss_grad(r.image(), t, weight=weight)
        g = r.backward(v, 0)
        vt += g["v_transforms"]
        vs += g["v_sh"]
        vo += g["v_raw_opac"]
        losses.append(loss)
    return losses, vt, vs, vo


def make_trainer(ctx, weight):
    from rustscan_amd.rasterizer import Trainer, training_optimizer

    return Trainer(ctx, training_optimizer(3, **LRS), loss_weight=weight)


@pytest.mark.parametrize("batched", [True, False])
def test_train_step_vs_oracle(hip_ctx, orc, setup, batched):
    from rustscan_amd.rasterizer import DeviceSplats

    sc, cams, tg, tgt = setup
    hip_ctx.set_raster(16, 256, True)
    w = 1.0 / len(cams)
    ds = DeviceSplats(hip_ctx, sc)
    tr = make_trainer(hip_ctx, w)
    say(f"train step batched={batched}")
    losses = tr.train_step(ds, cams, tgt, batched=batched, graph=False, read_loss=True)
    vt, vs, vo, _ = ds.grads()
    ref_losses, rvt, rvs, rvo = oracle_step(orc, sc, cams, tg, w)
    np.testing.assert_allclose(losses, ref_losses, rtol=1e-5)
    errs = dict(means=rel_l2(vt[:, :3], rvt[:, :3]), quats=rel_l2(vt[:, 3:7], rvt[:, 3:7]),
                scales=rel_l2(vt[:, 7:], rvt[:, 7:]), opacity=rel_l2(vo, rvo), sh=rel_l2(vs, rvs))
    say("  " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()))
    assert max(errs.values()) <= 1e-3
    # Adam applied to the HIP gradients, compared with the oracle's Adam on the same gradients
    after = ds.download()
    flat = np.concatenate([sc.transforms, sc.raw_opac[:, None], sc.sh.reshape(sc.n, -1)], axis=1).copy()
    g = np.concatenate([vt, vo[:, None], vs.reshape(sc.n, -1)], axis=1).copy()
    m = np.zeros_like(flat)
    v = np.zeros_like(flat)
    lr = np.array([LRS["lr_position"]] * 3 + [LRS["lr_rotation"]] * 4 + [LRS["lr_scale"]] * 3 + [LRS["lr_opacity"]] +
                  [LRS["lr_sh_dc"]] * 3 + [LRS["lr_sh_rest"]] * 45, np.float32)
    orc.adam_group(flat, g, m, v, lr, 0.9, 0.999, 1e-8, 1)
    got = np.concatenate([after.transforms, after.raw_opac[:, None], after.sh.reshape(sc.n, -1)], axis=1)
    np.testing.assert_array_equal(got, flat)


def test_per_view_equals_batched(hip_ctx, setup):
    from rustscan_amd.rasterizer import DeviceSplats

    sc, cams, _, tgt = setup
    tr = make_trainer(hip_ctx, 1.0 / len(cams))
    ds = DeviceSplats(hip_ctx, sc)
    tr.train_step(ds, cams, tgt, batched=True, graph=False)
    gb = ds.grads()
    ds.upload(sc)
    tr.train_step(ds, cams, tgt, batched=False, graph=False)
    gp = ds.grads()
    errs = [rel_l2(a, b) for a, b in zip(gb, gp)]
    say("  batched vs per-view " + " ".join(f"{e:.2e}" for e in errs))
    assert max(errs[:3]) <= 1e-5


def test_view_groups_equal_one_launch(hip_ctx, setup):
    """A launch limit below the step's intersection count splits the batched
    step into consecutive groups of views; the gradients and losses equal the
    single-launch step, eagerly and when captured as a graph."""
    from rustscan_amd.rasterizer import DeviceSplats

    sc, cams, _, tgt = setup
    tr = make_trainer(hip_ctx, 1.0 / len(cams))
    ds = DeviceSplats(hip_ctx, sc)
    l_one = tr.train_step(ds, cams, tgt, batched=True, graph=False, read_loss=True)
    g_one = ds.grads()
    counts = hip_ctx.counts(len(cams))
    m = [int(c[1]) for c in counts]
    say(f"  intersections per view {m}")
    try:
        hip_ctx.set_launch_limit(max(m) + 1)  # every view in its own group
        for graph in (False, True):
            ds.upload(sc)
            l_grp = tr.train_step(ds, cams, tgt, batched=True, graph=graph, read_loss=True)
            g_grp = ds.grads()
            errs = [rel_l2(a, b) for a, b in zip(g_grp, g_one)]
            say(f"  grouped (graph={graph}) vs one launch " + " ".join(f"{e:.2e}" for e in errs))
            assert max(errs[:3]) <= 1e-5
            np.testing.assert_allclose(l_grp, l_one, rtol=1e-5)
    finally:
        hip_ctx.set_launch_limit(1600000000)
    ds.close()


def test_graph_replay_tracks_eager(hip_ctx, setup):
    """Each replay of the captured step computes the gradients an eager step
    computes from the same parameters; after the first step (same Adam state)
    the parameters agree too.  Later steps re-seed the eager store from the
    graph's parameters: Adam's normalised updates would otherwise amplify the
    last-bit differences of atomic summation order step by step."""
    import torch

    from rustscan_amd.rasterizer import DeviceSplats

    sc, cams, _, tgt = setup
    tr = make_trainer(hip_ctx, 1.0 / len(cams))
    a = DeviceSplats(hip_ctx, sc)
    b = DeviceSplats(hip_ctx, sc)
    skipped = hip_ctx.skipped_steps()  # the context's running count (earlier tests may skip steps on purpose)
    for step in range(4):
        if step > 0:
            b.upload(a.download())
        tr.train_step(a, cams, tgt, batched=True, graph=True)
        hip_ctx.sync()
        torch.cuda.synchronize()
        tr.train_step(b, cams, tgt, batched=True, graph=False)
        ga, gb = a.grads(), b.grads()
        ge = max(rel_l2(ga[0], gb[0]), rel_l2(ga[1], gb[1]), rel_l2(ga[2], gb[2]))
        say(f"  replay {step}: grad rel {ge:.2e}")
        assert ge <= 1e-5
        if step == 0:
            pa, pb = a.download(), b.download()
            d = max(float(np.abs(pa.transforms - pb.transforms).max()), float(np.abs(pa.sh - pb.sh).max()),
                    float(np.abs(pa.raw_opac - pb.raw_opac).max()))
            say(f"  first step max param diff {d:.2e}")
            assert d <= 1e-5
    assert hip_ctx.skipped_steps() == skipped


def test_nonfinite_loss_skips_update(hip_ctx, setup):
    import torch

    from rustscan_amd.rasterizer import DeviceSplats

    sc, cams, _, tgt = setup
    bad = [t.clone() for t in tgt]
    bad[1][0, 0, 0] = float("nan")
    tr = make_trainer(hip_ctx, 1.0 / len(cams))
    ds = DeviceSplats(hip_ctx, sc)
    before = hip_ctx.skipped_steps()
    losses = tr.train_step(ds, cams, bad, batched=True, graph=False, read_loss=True)
    assert not np.isfinite(losses[1])
    after = ds.download()
    np.testing.assert_array_equal(after.transforms, sc.transforms)
    assert hip_ctx.skipped_steps() == before + 1
    tr.train_step(ds, cams, tgt, batched=True, graph=False, read_loss=True)
    assert not np.array_equal(ds.download().transforms, sc.transforms)


@pytest.mark.parametrize("tile,batch,persistent", [(8, 128, True), (8, 256, True), (8, 512, True), (16, 128, True),
                                                    (16, 512, True), (32, 128, True), (32, 256, True),
                                                    (32, 512, True), (16, 256, False), (32, 256, False)])
def test_raster_configs_vs_oracle(hip_ctx, orc, tile, batch, persistent):
    import torch

    from rustscan_amd.rasterizer import DeviceSplats, l1_loss_grad, render_backward, render_forward

    sc = scene.gaussian_scene(16384, seed=71, sh_degree=1)
    cam = scene.ring_cameras([(400, 300)], seed=72)[0]
    target = scene.Rng(73).uniform(400 * 300 * 3).reshape(300, 400, 3).astype(np.float32)
    hip_ctx.set_raster(tile, batch, persistent)
    ds = DeviceSplats(hip_ctx, sc)
    img = render_forward(hip_ctx, ds, [cam])[0]
    r = orc.Render(sc, cam, tile_width=tile)
    err = float(np.abs(img.cpu().numpy() - r.image()).max())
    _, v = l1_loss_grad(hip_ctx, img, torch.from_numpy(target).cuda())
    vt, vs, vo, _ = render_backward(hip_ctx, ds, [v])
    ref = r.backward(orc.l1_loss_grad(r.image(), target)[1], 0)
    rt = ref["v_transforms"]
    errs = dict(means=rel_l2(vt[:, :3], rt[:, :3]), quats=rel_l2(vt[:, 3:7], rt[:, 3:7]),
                log_scales=rel_l2(vt[:, 7:], rt[:, 7:]), opacity=rel_l2(vo, ref["v_raw_opac"]),
                sh=rel_l2(vs, ref["v_sh"]))
    say(f"  tile {tile} batch {batch} persistent {persistent}: image err {err:.2e}, grads "
        + " ".join(f"{k}={e:.1e}" for k, e in errs.items()))
    hip_ctx.set_raster(16, 256, True)
    assert err <= 1e-4 and max(errs.values()) <= 1e-3
