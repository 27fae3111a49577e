"""Regression fixtures in tests/golden/ (written by tests/golden/make_golden.py).

CPU: the oracle reproduces every fixture exactly (guards the pinned oracle
against drift).  GPU: the HIP path matches the fixtures under the parity bar:
depth order, intersections, tile ranges, counts and visibility bit-exact,
image within 1e-4, gradients within 1e-3 relative L2.
"""
import os

import numpy as np
import pytest

from tests.golden.make_golden import cases, compute

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = [c[0] for c in cases()]


def load(name):
    with np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def case(name):
    return next(c for c in cases() if c[0] == name)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(orc, name):
    _, sc, cam, target = case(name)
    got = compute(sc, cam, target)
    ref = load(name)
    assert set(got) == set(ref)
    for k in ref:
        np.testing.assert_array_equal(np.asarray(got[k]), ref[k], err_msg=k)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_matches_fixture(hip_ctx, name):
    import torch

    from rustscan_amd import _lib as L
    from rustscan_amd.rasterizer import DeviceSplats, l1_loss_grad, render_backward, render_forward
    from tests.gpu_util import rel_l2, say

    _, sc, cam, target = case(name)
    ref = load(name)
    say(f"golden {name}")
    hip_ctx.set_raster(16, 256, True)
    ds = DeviceSplats(hip_ctx, sc)
    img = render_forward(hip_ctx, ds, [cam], cov_blur=0.3)[0]
    assert float(np.abs(img.cpu().numpy() - ref["image"]).max()) <= 1e-4
    np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_DEPTH_ORDER), ref["depth_order"])
    np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_ISECT_TILES), ref["isect_tiles"])
    np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_ISECT_GIDS), ref["isect_gids"])
    np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_TILE_RANGES), ref["tile_ranges"])
    np.testing.assert_array_equal(hip_ctx.debug(L.RGS_DEBUG_TILE_COUNTS), ref["counts"])
    np.testing.assert_array_equal(ds.visible(), ref["visible"])
    loss, v_out = l1_loss_grad(hip_ctx, img, torch.from_numpy(target).cuda())
    assert abs(loss - float(ref["loss"])) <= 1e-5 * max(1.0, abs(float(ref["loss"])))
    vt, vs, vo, _ = render_backward(hip_ctx, ds, [v_out])
    for got, want, what in [(vt, ref["v_transforms"], "v_transforms"), (vs, ref["v_sh"], "v_sh"),
                            (vo, ref["v_raw_opac"], "v_raw_opac")]:
        e = rel_l2(got, want)
        say(f"  {what} rel-L2 {e:.2e}")
        assert e <= 1e-3, what
    ds.close()
