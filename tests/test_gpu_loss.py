"""The combined training loss (combined_loss_with_kernel,
RustGS/src/training/engine/loss.rs:112-142: L1 variants + D-SSIM +
image-gradient difference) through the C ABI, against the loss oracle
(oracle/loss.py, pinned by tests/test_oracle_loss.py):

* rgs_image_loss_grad on ragged image sizes (1xW, Hx1, smaller than the
  11-tap window, not multiples of the 16-pixel block) and every loss
  variant: loss within 1e-5 relative, image gradient within 1e-5 rel-L2 of
  the oracle's float64 gradient;
* a batched 3-view training step with RustGS's default loss (L1 0.8, D-SSIM
  0.2): per-view losses and every gradient group within 1e-3 of the oracle
  (its compositing backward fed with the loss oracle's image gradient), and
  the same step captured as a HIP graph.
"""
import numpy as np
import pytest

from oracle import loss as OL
from rustscan_amd import scene
from tests.gpu_util import rel_l2, say

pytestmark = pytest.mark.gpu

CONFIGS = {
    "default": dict(),
    "l1_only": dict(l1_weight=1.0, ssim_weight=0.0),
    "ssim_only": dict(l1_weight=0.0, ssim_weight=1.0),
    "gradient": dict(l1_weight=0.3, ssim_weight=0.5, gradient_weight=0.2),
    "robust": dict(l1_weight=1.0, ssim_weight=0.2, robust_delta=0.2),
    "outlier": dict(l1_weight=1.0, ssim_weight=0.2, outlier_threshold=0.3, outlier_weight=0.4),
}


@pytest.mark.parametrize("name", list(CONFIGS))
@pytest.mark.parametrize("hw", [(37, 23), (1, 9), (7, 1), (5, 4), (150, 200)])
def test_image_loss_vs_oracle(hip_ctx, name, hw):
    import torch

    from rustscan_amd.rasterizer import LossConfig, image_loss_grad

    H, W = hw
    rng = np.random.default_rng(H * 1000 + W)
    pred = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    tgt = rng.uniform(0, 1, (H, W, 3)).astype(np.float32)
    kw = CONFIGS[name]
    cfg_o = OL.LossConfig(**kw)
    weight = 0.75
    loss, v = image_loss_grad(hip_ctx, torch.from_numpy(pred).cuda(), torch.from_numpy(tgt).cuda(), weight,
                              LossConfig(**kw))
    with np.errstate(invalid="ignore"):
        ref_loss = weight * OL.forward(pred, tgt, cfg_o)
    ref_g = weight * OL.grad(pred, tgt, cfg_o)
    got = v.cpu().numpy()
    assert np.all(got[..., 3] == 0.0)
    err = rel_l2(got[..., :3], ref_g)
    say(f"loss {name} {H}x{W}: {loss:.6f} vs {ref_loss:.6f}, grad rel-L2 {err:.2e}")
    if H == 1 or W == 1:  # the reference's loss is NaN there (an empty image-gradient mean)
        assert np.isnan(loss) and np.isnan(ref_loss)
    else:
        np.testing.assert_allclose(loss, ref_loss, rtol=1e-5)
    assert err <= 1e-5


@pytest.mark.parametrize("graph", [False, True])
def test_train_step_with_default_loss(hip_ctx, orc, graph):
    import torch

    from rustscan_amd.rasterizer import DeviceSplats, LossConfig, Trainer, training_optimizer

    sc = scene.gaussian_scene(32768, seed=171, sh_degree=3)
    cams = scene.ring_cameras([(480, 272), (320, 240), (333, 197)], seed=172)
    tg = [scene.Rng(180 + i).uniform(c.width * c.height * 3).reshape(c.height, c.width, 3).astype(np.float32)
          for i, c in enumerate(cams)]
    tgt = [torch.from_numpy(t).cuda() for t in tg]
    hip_ctx.set_raster(16, 256, True)
    w = 1.0 / len(cams)
    tr = Trainer(hip_ctx, training_optimizer(3, 1.6e-4, 1e-3, 5e-3, 5e-2, 2.5e-3, 1.25e-4), loss_weight=w,
                 loss=LossConfig())
    ds = DeviceSplats(hip_ctx, sc)
    say(f"train step with L1 0.8 + D-SSIM 0.2 (graph={graph})")
    if graph:  # a first replay, then check the second one from the restored parameters
        tr.train_step(ds, cams, tgt, batched=True, graph=True, read_loss=True)
        ds.upload(sc)
    losses = tr.train_step(ds, cams, tgt, batched=True, graph=graph, read_loss=True)
    vt, vs, vo, _ = ds.grads()
    rvt, rvs, rvo = np.zeros_like(vt), np.zeros_like(vs), np.zeros_like(vo)
    cfg = OL.LossConfig()
    for i, (cam, t) in enumerate(zip(cams, tg)):
        r = orc.Render(sc, cam)
        img = r.image()
        np.testing.assert_allclose(losses[i], w * OL.forward(img, t, cfg), rtol=1e-5)
        v = np.zeros((cam.height, cam.width, 4), np.float32)
        v[..., :3] = w * OL.grad(img, t, cfg)
        g = r.backward(v, 0)
        rvt += g["v_transforms"]
        rvs += g["v_sh"]
        rvo += g["v_raw_opac"]
    errs = dict(means=rel_l2(vt[:, :3], rvt[:, :3]), quats=rel_l2(vt[:, 3:7], rvt[:, 3:7]),
                log_scales=rel_l2(vt[:, 7:], rvt[:, 7:]), opacity=rel_l2(vo, rvo), sh=rel_l2(vs, rvs))
    say("  grads " + " ".join(f"{k}={e:.2e}" for k, e in errs.items()))
    assert max(errs.values()) <= 1e-3
    ds.close()
