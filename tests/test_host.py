"""Host-side logic: random streams, synthetic scenes, cameras, learning-rate
schedule and optimizer configuration (CPU only)."""
import math

import numpy as np
import pytest

from rustscan_amd import scene
from rustscan_amd.rasterizer import AdamScaled, lr_at, training_optimizer


def test_splitmix64_known_answers():
    # SplitMix64 (Steele, Lea, Flood 2014) from state 0: the published first outputs
    r = scene.Rng(0)
    assert [int(x) for x in r.u64(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    # the stream is a counter: two draws of 2 equal one draw of 4
    a = scene.Rng(1234)
    b = scene.Rng(1234)
    assert np.array_equal(np.concatenate([a.u64(2), a.u64(2)]), b.u64(4))


def test_uniform_and_normal_statistics():
    r = scene.Rng(42)
    u = r.uniform(200000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5e-3
    z = scene.Rng(43).normal(200000)
    assert abs(z.mean()) < 1e-2 and abs(z.std() - 1.0) < 1e-2


def test_scene_layout_and_determinism():
    s1 = scene.gaussian_scene(1000, seed=5, sh_degree=3)
    s2 = scene.gaussian_scene(1000, seed=5, sh_degree=3)
    s3 = scene.gaussian_scene(1000, seed=6, sh_degree=3)
    assert s1.transforms.shape == (1000, 10) and s1.transforms.dtype == np.float32
    assert s1.sh.shape == (1000, 16, 3) and s1.raw_opac.shape == (1000,)
    assert np.array_equal(s1.transforms, s2.transforms) and np.array_equal(s1.sh, s2.sh)
    assert not np.array_equal(s1.transforms, s3.transforms)
    q = s1.transforms[:, 3:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)  # unit wxyz quaternions
    m = s1.transforms[:, :3]
    assert m.min() >= -1.0 and m.max() <= 1.0


def test_config1_geometry():
    sc, cams = scene.config1(n=4096)
    m = sc.transforms[:, :3]
    assert np.all(m[:, 2] >= 3.0) and np.all(m[:, 2] <= 5.0)  # U[-1,1]^3 + (0,0,4)
    assert sc.sh.shape[1] == 1
    c = cams[0]
    assert (c.width, c.height, float(c.fx), float(c.cx), float(c.cy)) == (320, 240, 300.0, 160.0, 120.0)
    assert np.array_equal(c.matrix(), np.eye(4, dtype=np.float32))


def test_heavy_tail_scene():
    base = scene.gaussian_scene(20000, seed=9, sh_degree=0)
    tail = scene.gaussian_scene(20000, seed=9, sh_degree=0, tail_fraction=0.01, tail_seed=8, tail_factor=20.0)
    d = tail.transforms[:, 7:10] - base.transforms[:, 7:10]
    moved = np.any(d != 0, axis=1)
    assert 150 <= moved.sum() <= 250  # about 1%
    assert np.allclose(d[moved], math.log(20.0), atol=1e-5)


def test_ring_cameras_look_at_origin():
    cams = scene.config4_cameras()
    assert [(c.width, c.height) for c in cams] == scene.C4_SIZES
    for c in cams:
        M = c.matrix().astype(np.float64)
        R, t = M[:3, :3], M[:3, 3]
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)
        assert abs(np.linalg.det(R) - 1.0) < 1e-6
        # camera centre = -R^T t, on the ring of radius 5 at height 1
        p = -R.T @ t
        assert np.allclose(p, c.cam_pos, atol=1e-5)
        assert abs(math.hypot(p[0], p[2]) - 5.0) < 1e-5 and abs(p[1] - 1.0) < 1e-6
        # the origin is in front, within the 10 degree yaw jitter of the optical axis
        o = R @ np.zeros(3) + t
        ang = math.degrees(math.acos(o[2] / np.linalg.norm(o)))
        assert o[2] > 0 and ang <= 10.0 + 1e-3
        assert float(c.fx) == pytest.approx(0.866 * c.width, rel=1e-6)


def test_lr_schedule_matches_reference_tests():
    # trainer_tests.rs:13-41 (position decay over 1000 iterations)
    at0 = lr_at(1.6e-4, 1.6e-6, 0, 1000)
    mid = lr_at(1.6e-4, 1.6e-6, 500, 1000)
    end = lr_at(1.6e-4, 1.6e-6, 1000, 1000)
    assert abs(at0 - 1.6e-4) < 1e-8
    assert abs(end - 1.6e-6) < 1.6e-6 * 0.01
    assert 1.6e-6 < mid < 1.6e-4
    assert mid == pytest.approx(math.sqrt(1.6e-4 * 1.6e-6), rel=1e-5)  # geometric midpoint
    # trainer_tests.rs:43-70 (horizon from lr_decay_iterations)
    assert abs(lr_at(5e-3, 5e-4, 10000, 30000, 10000) - 5e-4) < 1e-8
    assert abs(lr_at(1e-3, 1e-4, 10000, 30000, 10000) - 1e-4) < 1e-8
    assert abs(lr_at(5e-2, 5e-3, 10000, 30000, 10000) - 5e-3) < 1e-8
    assert abs(lr_at(2.5e-3, 2.5e-4, 30000, 30000, 10000) - 2.5e-4) < 1e-8
    # trainer.rs:179-181: constant when final <= 0, final >= initial, or no iterations
    assert lr_at(1e-3, 0.0, 50, 100) == pytest.approx(1e-3)
    assert lr_at(1e-3, 2e-3, 50, 100) == pytest.approx(1e-3)
    assert lr_at(1e-3, 1e-4, 50, 0) == pytest.approx(1e-3)


def test_optimizer_configuration_columns():
    opt = training_optimizer(3, lr_position=1.6e-4, lr_rotation=1e-3, lr_scale=5e-3, lr_opacity=5e-2,
                             lr_sh_dc=2.5e-3, lr_sh_rest=1.25e-4)
    c = opt.to_cfg()
    f = lambda x: float(np.float32(x))
    assert list(c.lr_transforms) == [f(1.6e-4)] * 3 + [f(1e-3)] * 4 + [f(5e-3)] * 3
    assert c.lr_opacity == f(5e-2)
    assert list(c.lr_sh)[:16] == [f(2.5e-3)] + [f(1.25e-4)] * 15
    assert (c.beta1, c.beta2, c.eps) == (f(0.9), f(0.999), f(1e-8))
    assert c.weight_decay == 0.0
    wd = AdamScaled()
    wd.config.weight_decay = 0.1
    assert wd.to_cfg().weight_decay == f(0.1)
