"""Densification accumulators of the training step (accumulate_gradients,
RustGS/src/training/engine/trainer.rs:651-728) through the C ABI, against
the oracle's restatement fed with the oracle's per-view backward:

* a batched 3-view step and a per-view step accumulate once per view:
  observation and visibility counts exactly, the gradient statistics within
  1e-3 (rel-L2, the gradient tolerance), coverage exactly;
* a second step adds to the first; a step with a non-finite loss adds
  nothing (the reference returns before accumulating, trainer.rs:297-302);
* reset zeroes them (reset_accumulators, trainer.rs:861-871);
* with the flags off the accumulators are left alone.
"""
import numpy as np
import pytest

from rustscan_amd import scene
from tests.gpu_util import rel_l2, say

pytestmark = pytest.mark.gpu

LRS = dict(lr_position=1.6e-4, lr_rotation=1e-3, lr_scale=5e-3, lr_opacity=5e-2, lr_sh_dc=2.5e-3, lr_sh_rest=1.25e-4)
COUNTS = ("num_observations", "visible_observations", "actual_visible_observations")


@pytest.fixture(scope="module")
def setup():
    import torch

    sc = scene.gaussian_scene(32768, seed=151, sh_degree=3)
    cams = scene.ring_cameras([(480, 272), (320, 240), (640, 360)], seed=152)
    tg = [scene.Rng(160 + i).uniform(c.width * c.height * 3).reshape(c.height, c.width, 3).astype(np.float32)
          for i, c in enumerate(cams)]
    return sc, cams, tg, [torch.from_numpy(t).cuda() for t in tg]


def oracle_accumulate(orc, acc, sc, cams, tg, weight, flags):
    for cam, t in zip(cams, tg):
        r = orc.Render(sc, cam)
        _, v = orc.l1_loss_grad(r.image(), t, weight=weight)
        g = r.backward(v, 0)
        acc.accumulate(g["v_transforms"], g["screen_stats"], g["v_sh"], r.visible(),
                       use_actual_visibility=bool(flags & 2), collect_actual_visibility=bool(flags & 4))


def compare(got, ref, label):
    errs = {}
    for k, r in ref.items():
        if k in COUNTS:
            np.testing.assert_array_equal(got[k], r, err_msg=f"{label}: {k}")
        else:
            errs[k] = rel_l2(got[k], r)
    say(f"  {label}: " + " ".join(f"{k}={e:.1e}" for k, e in errs.items()))
    assert max(errs.values()) <= 1e-3
    assert errs["pixel_coverage"] <= 1e-6


@pytest.mark.parametrize("batched,flags", [(True, 7), (False, 7), (True, 1)])
def test_densify_stats_vs_oracle(hip_ctx, orc, setup, batched, flags):
    from oracle.oracle import DensifyAccumulators
    from rustscan_amd.rasterizer import DeviceSplats, Trainer, training_optimizer

    sc, cams, tg, tgt = setup
    hip_ctx.set_raster(16, 256, True)
    w = 1.0 / len(cams)
    tr = Trainer(hip_ctx, training_optimizer(3, **LRS), loss_weight=w, densify=flags)
    ds = DeviceSplats(hip_ctx, sc)
    say(f"densify stats, batched={batched}, flags={flags}")
    tr.train_step(ds, cams, tgt, batched=batched, graph=False, read_loss=True)
    acc = DensifyAccumulators(sc.n)
    oracle_accumulate(orc, acc, sc, cams, tg, w, flags)
    compare(ds.densify_stats(), acc.acc, "step 1")
    # a second step from the updated parameters adds to the first
    p1 = ds.download()
    tr.train_step(ds, cams, tgt, batched=batched, graph=True, read_loss=True)
    oracle_accumulate(orc, acc, p1, cams, tg, w, flags)
    compare(ds.densify_stats(), acc.acc, "step 2 (graph)")
    ds.reset_densify_stats()
    assert all(not v.any() for v in ds.densify_stats().values())
    ds.close()


def test_densify_skipped_and_disabled_steps_add_nothing(hip_ctx, setup):
    from rustscan_amd.rasterizer import DeviceSplats, Trainer, training_optimizer

    sc, cams, _, tgt = setup
    bad = [t.clone() for t in tgt]
    bad[0][3, 5, 1] = float("inf")
    ds = DeviceSplats(hip_ctx, sc)
    tr = Trainer(hip_ctx, training_optimizer(3, **LRS), loss_weight=1.0 / 3, densify=7)
    tr.train_step(ds, cams, tgt, batched=True, graph=False, read_loss=True)
    ref = ds.densify_stats()
    losses = tr.train_step(ds, cams, bad, batched=True, graph=False, read_loss=True)
    assert not np.isfinite(losses[0])
    after = ds.densify_stats()
    for k in ref:
        np.testing.assert_array_equal(after[k], ref[k], err_msg=k)
    off = Trainer(hip_ctx, training_optimizer(3, **LRS), loss_weight=1.0 / 3, densify=0)
    off.train_step(ds, cams, tgt, batched=True, graph=False, read_loss=True)
    for k, v in ds.densify_stats().items():
        np.testing.assert_array_equal(v, ref[k], err_msg=k)
    np.testing.assert_array_equal(ref["num_observations"], np.full(sc.n, 3.0, np.float32))
    ds.close()
