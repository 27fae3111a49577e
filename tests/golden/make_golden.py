"""Writes the regression fixtures of tests/golden/ from the oracle.

The reference publishes no golden images or gradients (its tests are
qualitative apart from the sort and scan vectors, which the tests hold
literally), so these fixtures freeze the oracle's outputs on two small seeded
cases once the oracle is pinned by the closed-form, brute-force and
finite-difference checks.  They catch any later drift of the oracle, and the
GPU tests compare the HIP path against the same files.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from rustscan_amd import scene  # noqa: E402


def cases():
    """(name, splats, camera, target) of each fixture."""
    sc1 = scene.gaussian_scene(512, seed=1, sh_degree=0, center=(0.0, 0.0, 4.0), log_scale_mu=-2.5)
    cam1 = scene.pinhole_identity(64, 48, 60.0)
    t1 = scene.Rng(31).uniform(48 * 64 * 3).reshape(48, 64, 3).astype(np.float32)
    sc2 = scene.gaussian_scene(768, seed=2, sh_degree=3, log_scale_mu=-2.8)
    cam2 = scene.ring_cameras([(80, 56)], radius=4.0, seed=7)[0]
    t2 = scene.Rng(32).uniform(56 * 80 * 3).reshape(56, 80, 3).astype(np.float32)
    return [("c1_tiny", sc1, cam1, t1), ("ring_sh3", sc2, cam2, t2)]


def compute(sc, cam, target):
    r = O.Render(sc, cam, cov_blur=0.3)
    img = r.image()
    loss, v = O.l1_loss_grad(img, target, weight=1.0)
    g = r.backward(v, 0)
    tiles, gids = r.isect()
    return dict(image=img, depth_order=r.depth_order(), isect_tiles=tiles, isect_gids=gids,
                tile_ranges=r.tile_offsets(), counts=r.counts(), visible=r.visible(), loss=np.float64(loss),
                v_transforms=g["v_transforms"], v_sh=g["v_sh"], v_raw_opac=g["v_raw_opac"])


def main():
    for name, sc, cam, target in cases():
        out = compute(sc, cam, target)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{path}: V={len(out['depth_order'])} M={len(out['isect_tiles'])} "
              f"size={os.path.getsize(path)} B")


if __name__ == "__main__":
    main()
