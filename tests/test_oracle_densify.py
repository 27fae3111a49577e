"""The oracle's restatement of the trainer's densification accumulators
(accumulate_gradients, RustGS/src/training/engine/trainer.rs:651-728),
pinned by hand-computed sums on constructed inputs: each statistic is fed
values whose norm or mean is known exactly, so a dropped term, a wrong
column, a missing square root or a wrong mean divisor fails."""
import numpy as np

from oracle.oracle import DensifyAccumulators


def one_view(n=3, k=4):
    vt = np.zeros((n, 10), np.float32)
    st = np.zeros((n, 7), np.float32)
    vs = np.zeros((n, k, 3), np.float32)
    # Gaussian 0: v_mean (3, -6, 0) -> mean |.| = 3; screen v_xy (3, 4) -> 5;
    # |v_xy| sums (6, 8) -> 10; abs-pixel 2, coverage 7, depth 21;
    # v_sh with |.| summing to 24 over 4 x 3 entries -> mean 2
    vt[0, 0:3] = (3.0, -6.0, 0.0)
    vt[0, 3:10] = 100.0  # quaternion / scale gradients do not enter
    st[0] = (3.0, 4.0, 6.0, 8.0, 2.0, 7.0, 21.0)
    vs[0] = 2.0
    vs[0, 0, 0] = -2.0
    # Gaussian 1: only the colour gradient, one nonzero entry 12 -> mean 1
    vs[1, 3, 2] = 12.0
    vis = np.array([1.0, 0.0, 1.0], np.float32)
    return vt, st, vs, vis


def test_accumulate_hand_sums():
    acc = DensifyAccumulators(3)
    vt, st, vs, vis = one_view()
    acc.accumulate(vt, st, vs, vis)
    a = acc.acc
    np.testing.assert_array_equal(a["grad_2d"], [3.0, 0.0, 0.0])
    np.testing.assert_array_equal(a["screen_grad_2d"], [5.0, 0.0, 0.0])
    np.testing.assert_array_equal(a["abs_grad_2d"], [10.0, 0.0, 0.0])
    np.testing.assert_array_equal(a["abs_pixel_grad_2d"], [2.0, 0.0, 0.0])
    np.testing.assert_array_equal(a["pixel_coverage"], [7.0, 0.0, 0.0])
    np.testing.assert_array_equal(a["camera_depth"], [21.0, 0.0, 0.0])
    np.testing.assert_array_equal(a["grad_color"], [2.0, 1.0, 0.0])
    # every Gaussian is observed once per step; visibility counts depend on the mode
    np.testing.assert_array_equal(a["num_observations"], [1.0, 1.0, 1.0])
    np.testing.assert_array_equal(a["visible_observations"], [1.0, 1.0, 1.0])
    np.testing.assert_array_equal(a["actual_visible_observations"], [0.0, 0.0, 0.0])


def test_accumulate_visibility_modes_and_sums_over_steps():
    acc = DensifyAccumulators(3)
    vt, st, vs, vis = one_view()
    acc.accumulate(vt, st, vs, vis, use_actual_visibility=True, collect_actual_visibility=True)
    acc.accumulate(vt, -st, vs, vis, use_actual_visibility=True, collect_actual_visibility=True)
    a = acc.acc
    # norms add (a sign flip of the stats does not cancel them); linear stats cancel
    np.testing.assert_array_equal(a["screen_grad_2d"], [10.0, 0.0, 0.0])
    np.testing.assert_array_equal(a["abs_grad_2d"], [20.0, 0.0, 0.0])
    np.testing.assert_array_equal(a["pixel_coverage"], [0.0, 0.0, 0.0])
    np.testing.assert_array_equal(a["grad_2d"], [6.0, 0.0, 0.0])
    np.testing.assert_array_equal(a["num_observations"], [2.0, 2.0, 2.0])
    np.testing.assert_array_equal(a["visible_observations"], [2.0, 0.0, 2.0])
    np.testing.assert_array_equal(a["actual_visible_observations"], [2.0, 0.0, 2.0])
    acc.reset()
    assert all(not v.any() for v in acc.acc.values())
