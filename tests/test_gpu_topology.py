"""The device topology step (rgs_topology_step: densify / prune / opacity
reset, RustGS/src/training/engine/trainer.rs:792-859 with topology/) against
the topology oracle (oracle/topology.cpp, pinned by
tests/test_oracle_topology.py), on the same splats, accumulators and state:

* counts, the selection, the rebuilt splats (bit for bit), the remapped
  birth and invisible-window state, cleared accumulators and the Adam reset,
  for several schedules and modes, and rows the reference prunes on sight
  (non-finite, out of bounds, scales below 1e-10);
* a short training run with the accumulators on, a topology step at its
  scheduled iteration on the accumulators the run gathered, and training
  continued on the new store (graph re-captured): gradients of the next step
  against the oracle, no skipped step.
"""
import numpy as np
import pytest

from oracle import oracle as O
from rustscan_amd import scene
from tests.gpu_util import rel_l2, say

pytestmark = pytest.mark.gpu

LRS = dict(lr_position=1.6e-4, lr_rotation=1e-3, lr_scale=5e-3, lr_opacity=5e-2, lr_sh_dc=2.5e-3, lr_sh_rest=1.25e-4)

CASES = {
    "densify": dict(refine_every=1),
    "threshold_prune": dict(refine_every=1, prune_mode=0, prune_opacity_threshold=0.05),
    "abs_pixel_depth_decay": dict(refine_every=1, split_score_mode=3, split_grad_threshold=5e-5, opacity_decay=0.1,
                                  scale_decay=0.2, max_initial_gaussians=1),
    "visibility_weight": dict(refine_every=1, prune_mode=2, prune_visibility_threshold=0.3,
                              growth_select_fraction=0.5),
    "reset_no_growth": dict(iterations=30, refine_every=2, opacity_reset_interval=2, opacity_reset_mode=1,
                            growth_freeze_after_epoch=0),
    "off_cadence": dict(refine_every=1000),
}


def random_state(n, seed):
    sc = scene.gaussian_scene(n, seed=seed, sh_degree=1)
    rng = np.random.default_rng(seed)
    # rows the reference prunes whatever the statistics say
    sc.transforms[5, 3] = np.nan  # a non-finite rotation: not retainable
    sc.transforms[6, 8] = 20.0    # a scale beyond 100 x the scene extent
    sc.transforms[7, 7] = -40.0   # a scale below 1e-10
    sc.sh[8, 2, 0] = np.inf
    d = np.zeros((10, n), np.float32)
    d[0] = rng.exponential(3e-4, n)
    d[1] = rng.exponential(1e-3, n)
    d[2] = rng.exponential(2e-3, n)
    d[3] = rng.exponential(1e-2, n)
    d[4] = rng.uniform(0, 40, n)
    d[5] = d[4] * rng.uniform(1, 8, n)
    d[6] = rng.exponential(1e-4, n)
    d[7] = 4.0
    d[8] = rng.integers(0, 5, n)
    d[9] = np.minimum(d[8], rng.integers(0, 5, n))
    d[0, 9] = np.nan
    d[8, 10] = np.inf
    birth = rng.integers(0, 19, n).astype(np.uint32)
    invis = rng.integers(0, 15, n).astype(np.uint32)
    return sc, d, birth, invis


@pytest.mark.parametrize("case", list(CASES))
def test_topology_step_vs_oracle(hip_ctx, case):
    from rustscan_amd.rasterizer import AdamScaled, DeviceSplats, TopologyConfig

    overrides = CASES[case]
    it, fc = (5, 2) if case == "reset_no_growth" else (20, 1)
    sc, d, birth, invis = random_state(20000, 31)
    ref = O.topology_step(O.topology_cfg(**overrides), sc, d, birth, invis, it, fc)
    ds = DeviceSplats(hip_ctx, sc)
    ds.set_densify_stats(d)
    ds.set_topology_state(birth, invis)
    AdamScaled.set_steps(hip_ctx, ds, [7, 7, 7])
    say(f"topology step {case}")
    res = ds.topology_step(TopologyConfig(**overrides), it, fc)
    keys = ["before", "after", "added", "pruned", "clone_candidates", "split_candidates", "growth_candidates",
            "prune_candidates", "replacement_count", "extra_growth_count", "selected_count", "mutated", "adam_reset"]
    say("  " + " ".join(f"{k}={res[k]}" for k in keys[:4]) + f" prune={res['prune_candidates']}")
    assert {k: res[k] for k in keys} == {k: int(ref[k]) for k in keys}
    assert res["schedule"] == ref["schedule"]
    got = ds.download()
    np.testing.assert_array_equal(got.transforms, ref["transforms"])
    np.testing.assert_array_equal(got.sh, ref["sh"])
    np.testing.assert_array_equal(got.raw_opac, ref["raw_opac"])
    b, w = ds.topology_state()
    np.testing.assert_array_equal(b, ref["birth"])
    np.testing.assert_array_equal(w, ref["invisible_windows"])
    assert all(not v.any() for v in ds.densify_stats().values())
    assert AdamScaled.steps(hip_ctx, ds) == ([0, 0, 0] if ref["adam_reset"] else [7, 7, 7])
    ds.close()


def test_train_topology_train(hip_ctx, orc):
    import torch

    from rustscan_amd._lib import RGS_DENSIFY_ACCUMULATE
    from rustscan_amd.rasterizer import DeviceSplats, TopologyConfig, Trainer, training_optimizer

    sc = scene.gaussian_scene(16384, seed=41, sh_degree=1)
    cams = scene.ring_cameras([(320, 240), (400, 300)], seed=42)
    tg = [scene.Rng(43 + i).uniform(c.width * c.height * 3).reshape(c.height, c.width, 3).astype(np.float32)
          for i, c in enumerate(cams)]
    tgt = [torch.from_numpy(t).cuda() for t in tg]
    hip_ctx.set_raster(16, 256, True)
    tr = Trainer(hip_ctx, training_optimizer(1, **LRS), loss_weight=0.5, densify=RGS_DENSIFY_ACCUMULATE)
    topo = TopologyConfig(refine_every=5, growth_grad_threshold=2e-6, growth_select_fraction=0.5,
                          max_initial_gaussians=1, prune_min_age=3)
    ds = DeviceSplats(hip_ctx, sc)
    skipped = hip_ctx.skipped_steps()
    frame_count = 8
    applied = None
    for it in range(1, 8):
        tr.train_step(ds, cams, tgt, batched=True, graph=True)
        if topo.should_apply(it, frame_count):
            hip_ctx.sync()
            p = ds.download()
            d = np.stack(list(ds.densify_stats().values()))
            b, w = ds.topology_state()
            ref = O.topology_step(O.topology_cfg(**{k: getattr(topo, k) for k in O.TOPOLOGY_DEFAULTS}), p, d, b, w,
                                  it, frame_count)
            res = ds.topology_step(topo, it, frame_count)
            say(f"  topology at iteration {it}: {res['before']} -> {res['after']} (+{res['added']} -{res['pruned']})")
            assert (res["after"], res["added"], res["pruned"]) == (ref["after"], ref["added"], ref["pruned"])
            np.testing.assert_array_equal(ds.download().transforms, ref["transforms"])
            applied = it
    assert applied == 6 and ds.n != sc.n
    # the next step on the rebuilt store, against the oracle at its parameters
    cur = ds.download()
    losses = tr.train_step(ds, cams, tgt, batched=True, graph=True, read_loss=True)
    vt, vs, vo, _ = ds.grads()
    rvt, rvo = np.zeros_like(vt), np.zeros_like(vo)
    for cam, t, loss in zip(cams, tg, losses):
        r = orc.Render(cur, cam)
        ref_loss, v = orc.l1_loss_grad(r.image(), t, weight=0.5)
        np.testing.assert_allclose(loss, ref_loss, rtol=1e-5)
        g = r.backward(v, 0)
        rvt += g["v_transforms"]
        rvo += g["v_raw_opac"]
    errs = [rel_l2(vt, rvt), rel_l2(vo, rvo)]
    say("  next step grads rel-L2 " + " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) <= 1e-3
    assert hip_ctx.skipped_steps() == skipped
    ds.close()
