"""Seeded random scenes, cameras and raster configurations against the oracle:
for each case the training forward (depth order, tile counts, intersection
lists, tile ranges, projected splats, visibility and the image, exactly), the
backward (every gradient group within 1e-3 rel-L2 of the oracle's FP64 replay
of the same forward), and the evaluation render with the CPU renderer's
semantics (bit for bit against the oracle run with the library's
exponential).

The cases cover 1 to 20,000 Gaussians, SH degrees 0-3, images from 1 pixel to
700x500, log-scale spreads up to 1.5 (needle-shaped splats), camera distances
from inside to outside the scene, and every tile width with every batch depth.
"""
import numpy as np
import pytest

from rustscan_amd import scene
from tests import test_gpu_eval as TE
from tests import test_gpu_render as TR
from tests.gpu_util import say

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
    rc, rd = orc.tiled_renderer(sc, cam, cov_blur=0.3, gpu_math=True)
    TE._assert_same(col, dep, rc, rd, "evaluation render")
