"""Seeded random scenes, cameras and raster configurations against the oracle:
for each case the training forward (depth order, tile counts, intersection
lists, tile ranges, projected splats, visibility and the image, exactly), the
backward (every gradient group within 1e-3 rel-L2 of the oracle's FP64 replay
of the same forward), and the evaluation render with the CPU renderer's
semantics (bit for bit against the oracle run with the library's
exponential).  Batched training steps over 1-8 random views in one
persistent launch, eager or graph-captured, against the oracle's images,
losses and FP64 gradient replay summed over the views.  Device topology
steps on random splats, statistics, state, schedules and thresholds, against
the topology oracle bit for bit.  The combined training loss on random
image sizes, loss configurations and weights, against the loss oracle.

The single-view cases cover 1 to 20,000 Gaussians, SH degrees 0-3, images
from 1 pixel to 700x500, log-scale spreads up to 1.5 (needle-shaped splats),
camera distances from inside to outside the scene, and every tile width with
every batch depth.
"""
import numpy as np
import pytest

from rustscan_amd import scene
from tests import test_gpu_eval as TE
from tests import test_gpu_render as TR
from tests.gpu_util import rel_l2, say

pytestmark = pytest.mark.gpu


def _cases(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        out.append(dict(
            n=int(rng.choice([1, 7, 100, 1000, 5000, 20000])),
            deg=int(rng.integers(0, 4)),
            wh=(int(rng.integers(1, 700)), int(rng.integers(1, 500))),
            mu=float(rng.uniform(-4.5, -1.5)),
            sigma=float(rng.uniform(0.1, 1.5)),
            tile=int(rng.choice([8, 16, 32])),
            batch=int(rng.choice([128, 256, 512])),
            radius=float(rng.uniform(2.5, 6.0)),
            seed=int(rng.integers(0, 1 << 30)),
        ))
    return out


@pytest.mark.parametrize("case", _cases(24, 20261203), ids=lambda c: f"n{c['n']}-sh{c['deg']}-{c['wh'][0]}x{c['wh'][1]}"
                         f"-t{c['tile']}b{c['batch']}")
def test_random_scene_vs_oracle(hip_ctx, orc, case):
    c = case
    say(f"random scene {c}")
    sc = scene.gaussian_scene(c["n"], seed=c["seed"], sh_degree=c["deg"], log_scale_mu=c["mu"],
                              log_scale_sigma=c["sigma"])
    cam = scene.ring_cameras([c["wh"]], radius=c["radius"], seed=c["seed"] + 1)[0]
    ds, _, _ = TR.check_forward(hip_ctx, orc, sc, cam, tile=c["tile"], batch=c["batch"])
    ds.close()
    w, h = c["wh"]
    target = scene.Rng(c["seed"] + 2).uniform(w * h * 3).reshape(h, w, 3).astype(np.float32)
    try:
        TR.grads_vs_oracle(hip_ctx, orc, sc, cam, target, tile=c["tile"], batch=c["batch"], reference="f64")
    finally:
        hip_ctx.set_raster(16, 256, True)
    col, dep = TE._tiled(hip_ctx, sc, cam)
    rc, rd = orc.tiled_renderer(sc, cam, cov_blur=0.3, gpu_math=False)
    TE._assert_same(col, dep, rc, rd, "evaluation render")


LRS = dict(lr_position=1.6e-4, lr_rotation=1e-3, lr_scale=5e-3, lr_opacity=5e-2, lr_sh_dc=2.5e-3, lr_sh_rest=1.25e-4)


def _step_cases(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        nv = int(rng.integers(1, 9))
        out.append(dict(
            n=int(rng.choice([50, 2000, 20000, 60000])),
            deg=int(rng.integers(0, 4)),
            sizes=[(int(rng.integers(8, 500)), int(rng.integers(8, 400))) for _ in range(nv)],
            tile=int(rng.choice([8, 16, 32])),
            batch=int(rng.choice([128, 256, 512])),
            graph=bool(rng.integers(0, 2)),
            mu=float(rng.uniform(-4.0, -2.0)),
            sigma=float(rng.uniform(0.1, 1.0)),
            radius=float(rng.uniform(3.0, 6.0)),
            seed=int(rng.integers(0, 1 << 30)),
        ))
    return out


@pytest.mark.parametrize("case", _step_cases(8, 20261204),
                         ids=lambda c: f"n{c['n']}-sh{c['deg']}-v{len(c['sizes'])}-t{c['tile']}b{c['batch']}"
                                       f"-{'graph' if c['graph'] else 'eager'}")
def test_random_batched_step_vs_oracle(hip_ctx, orc, case):
    """A batched training step over 1-8 random views in one persistent launch
    (eager or graph-captured): every view's image within 1e-4 of the oracle,
    every view's loss within 1e-5, and the gradients summed over the views
    within 1e-3 of the oracle's FP64 replay."""
    import torch

    from rustscan_amd.rasterizer import DeviceSplats, Trainer, render_forward, training_optimizer

    c = case
    say(f"random batched step {c}")
    sc = scene.gaussian_scene(c["n"], seed=c["seed"], sh_degree=c["deg"], log_scale_mu=c["mu"],
                              log_scale_sigma=c["sigma"])
    cams = scene.ring_cameras(c["sizes"], radius=c["radius"], seed=c["seed"] + 1)
    tg = [scene.Rng(c["seed"] + 10 + i).uniform(cam.width * cam.height * 3).reshape(cam.height, cam.width, 3)
          .astype(np.float32) for i, cam in enumerate(cams)]
    w = 1.0 / len(cams)
    hip_ctx.set_raster(c["tile"], c["batch"], True)
    try:
        ds = DeviceSplats(hip_ctx, sc)
        imgs = [im.cpu().numpy() for im in render_forward(hip_ctx, ds, cams)]
        tr = Trainer(hip_ctx, training_optimizer(c["deg"], **LRS), loss_weight=w)
        losses = tr.train_step(ds, cams, [torch.from_numpy(t).cuda() for t in tg], batched=True, graph=c["graph"],
                               read_loss=True)
        vt, vs, vo, _ = ds.grads()
        ds.close()
    finally:
        hip_ctx.set_raster(16, 256, True)
    rvt = rvs = rvo = None
    worst = 0.0
    for i, (cam, t) in enumerate(zip(cams, tg)):
        r = orc.Render(sc, cam, tile_width=c["tile"])
        img = r.image()
        worst = max(worst, float(np.abs(imgs[i] - img).max()))
        loss, v = orc.l1_loss_grad(img, t, weight=w)
        assert abs(losses[i] - loss) <= 1e-5 * max(1.0, abs(loss))
        g = r.backward_f64(v, mode=0)
        rvt = g["v_transforms"] if rvt is None else rvt + g["v_transforms"]
        rvs = g["v_sh"] if rvs is None else rvs + g["v_sh"]
        rvo = g["v_raw_opac"] if rvo is None else rvo + g["v_raw_opac"]
    errs = dict(means=rel_l2(vt[:, :3], rvt[:, :3]), quats=rel_l2(vt[:, 3:7], rvt[:, 3:7]),
                log_scales=rel_l2(vt[:, 7:], rvt[:, 7:]), opacity=rel_l2(vo, rvo), sh=rel_l2(vs, rvs))
    say(f"  image max err {worst:.2e}; grads " + " ".join(f"{k}={e:.2e}" for k, e in errs.items()))
    assert worst <= 1e-4
    assert max(errs.values()) <= 1e-3


TOPOLOGY_KEYS = ["before", "after", "added", "pruned", "clone_candidates", "split_candidates", "growth_candidates",
                 "prune_candidates", "replacement_count", "extra_growth_count", "selected_count", "mutated",
                 "adam_reset"]


def _topology_cases(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        cfg = dict(
            iterations=int(rng.choice([30, 1000, 30000])),
            max_initial_gaussians=int(rng.choice([1, 100, 100000])),
            refine_every=int(rng.choice([1, 2, 5, 160])),
            opacity_reset_interval=int(rng.choice([1, 2, 10])),
            opacity_reset_mode=int(rng.integers(0, 2)),
            growth_freeze_after_epoch=int(rng.choice([-1, 0, 3])),
            prune_until_epoch=int(rng.choice([-1, 2, 100])),
            split_score_mode=int(rng.integers(0, 4)),
            growth_grad_threshold=float(rng.choice([1e-5, 1.4e-4, 1e-3])),
            split_grad_threshold=float(rng.choice([1e-5, 1.4e-4, 1e-3])),
            growth_select_fraction=float(rng.choice([0.0, 0.25, 0.5, 1.0])),
            opacity_decay=float(rng.choice([0.0, 0.1])),
            scale_decay=float(rng.choice([0.0, 0.2])),
            prune_mode=int(rng.integers(0, 3)),
            prune_min_age=int(rng.choice([0, 5])),
            prune_invisible_epochs=int(rng.choice([1, 10])),
            prune_visibility_dry_run=int(rng.integers(0, 2)),
            prune_opacity_threshold=float(rng.choice([1.0 / 255.0, 0.05, 0.3])),
            prune_visibility_threshold=float(rng.choice([0.05, 0.3])),
        )
        out.append(dict(n=int(rng.choice([1, 2, 50, 1000, 20000, 70000])), deg=int(rng.integers(0, 4)),
                        iteration=int(rng.integers(1, 40)), frame_count=int(rng.integers(1, 6)),
                        mu=float(rng.uniform(-5.0, -1.0)), seed=int(rng.integers(0, 1 << 30)), cfg=cfg))
    return out


@pytest.mark.parametrize("case", _topology_cases(16, 20261205),
                         ids=lambda c: f"n{c['n']}-it{c['iteration']}-fc{c['frame_count']}-re{c['cfg']['refine_every']}"
                                       f"-pm{c['cfg']['prune_mode']}")
def test_random_topology_step_vs_oracle(hip_ctx, orc, case):
    """A device topology step on random splats, densification statistics,
    topology state, schedule and thresholds: counts, schedule, the rebuilt
    store and the remapped state equal the topology oracle's bit for bit."""
    from rustscan_amd.rasterizer import AdamScaled, DeviceSplats, TopologyConfig

    c = case
    n = c["n"]
    say(f"random topology step n={n} it={c['iteration']} fc={c['frame_count']} {c['cfg']}")
    rng = np.random.default_rng(c["seed"])
    sc = scene.gaussian_scene(n, seed=c["seed"], sh_degree=c["deg"], log_scale_mu=c["mu"])
    d = np.zeros((10, n), np.float32)
    d[0] = rng.exponential(3e-4, n)
    d[1] = rng.exponential(1e-3, n)
    d[2] = rng.exponential(2e-3, n)
    d[3] = rng.exponential(1e-2, n)
    d[4] = rng.uniform(0, 40, n)
    d[5] = d[4] * rng.uniform(1, 8, n)
    d[6] = rng.exponential(1e-4, n)
    d[7] = float(rng.integers(0, 6))
    d[8] = rng.integers(0, 5, n)
    d[9] = np.minimum(d[8], rng.integers(0, 5, n))
    birth = rng.integers(0, 19, n).astype(np.uint32)
    invis = rng.integers(0, 15, n).astype(np.uint32)
    ref = orc.topology_step(orc.topology_cfg(**c["cfg"]), sc, d, birth, invis, c["iteration"], c["frame_count"])
    ds = DeviceSplats(hip_ctx, sc)
    try:
        ds.set_densify_stats(d)
        ds.set_topology_state(birth, invis)
        AdamScaled.set_steps(hip_ctx, ds, [7, 7, 7])
        res = ds.topology_step(TopologyConfig(**c["cfg"]), c["iteration"], c["frame_count"])
        say("  " + " ".join(f"{k}={res[k]}" for k in TOPOLOGY_KEYS[:4]))
        assert {k: res[k] for k in TOPOLOGY_KEYS} == {k: int(ref[k]) for k in TOPOLOGY_KEYS}
        assert res["schedule"] == ref["schedule"]
        got = ds.download()
        np.testing.assert_array_equal(got.transforms, ref["transforms"])
        np.testing.assert_array_equal(got.sh, ref["sh"])
        np.testing.assert_array_equal(got.raw_opac, ref["raw_opac"])
        b, w = ds.topology_state()
        np.testing.assert_array_equal(b, ref["birth"])
        np.testing.assert_array_equal(w, ref["invisible_windows"])
    finally:
        ds.close()


LOSS_CONFIGS = [dict(), dict(l1_weight=1.0, ssim_weight=0.0), dict(l1_weight=0.0, ssim_weight=1.0),
                dict(l1_weight=0.3, ssim_weight=0.5, gradient_weight=0.2),
                dict(l1_weight=1.0, ssim_weight=0.2, robust_delta=0.2),
                dict(l1_weight=1.0, ssim_weight=0.2, outlier_threshold=0.3, outlier_weight=0.4)]


def _loss_cases(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        h = int(rng.integers(2, 80)) if rng.uniform() < 0.75 else int(rng.integers(80, 400))
        w = int(rng.integers(2, 80)) if rng.uniform() < 0.75 else int(rng.integers(80, 400))
        out.append(dict(hw=(h, w), cfg=int(rng.integers(0, len(LOSS_CONFIGS))), weight=float(rng.uniform(0.1, 2.0)),
                        seed=int(rng.integers(0, 1 << 30))))
    return out


@pytest.mark.parametrize("case", _loss_cases(12, 20261206),
                         ids=lambda c: f"{c['hw'][0]}x{c['hw'][1]}-cfg{c['cfg']}")
def test_random_image_loss_vs_oracle(hip_ctx, case):
    """The combined loss and its image gradient on a random image size, loss
    configuration and weight: loss within 1e-5 relative and gradient within
    1e-5 rel-L2 of the loss oracle."""
    import torch

    from oracle import loss as OL
    from rustscan_amd.rasterizer import LossConfig, image_loss_grad

    c = case
    h, w = c["hw"]
    kw = LOSS_CONFIGS[c["cfg"]]
    rng = np.random.default_rng(c["seed"])
    pred = rng.uniform(0, 1, (h, w, 4)).astype(np.float32)
    tgt = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    loss, v = image_loss_grad(hip_ctx, torch.from_numpy(pred).cuda(), torch.from_numpy(tgt).cuda(), c["weight"],
                              LossConfig(**kw))
    cfg_o = OL.LossConfig(**kw)
    ref_loss = c["weight"] * OL.forward(pred, tgt, cfg_o)
    ref_g = c["weight"] * OL.grad(pred, tgt, cfg_o)
    got = v.cpu().numpy()
    err = rel_l2(got[..., :3], ref_g)
    say(f"random loss {h}x{w} {kw} weight {c['weight']:.3f}: {loss:.6f} vs {ref_loss:.6f}, grad rel-L2 {err:.2e}")
    assert np.all(got[..., 3] == 0.0)
    np.testing.assert_allclose(loss, ref_loss, rtol=1e-5)
    assert err <= 1e-5
