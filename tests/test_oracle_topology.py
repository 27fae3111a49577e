"""Pins of the topology oracle (oracle/topology.cpp, a restatement of
RustGS/src/training/topology/ and the trainer's topology state), mostly the
reference's own topology tests (topology/tests.rs, topology/apply.rs tests)
restated at the granularity of one whole topology step, plus:

* the Efraimidis-Spirakis keys log(u)/w pick index i first with probability
  w_i / sum(w) (a wrong key fails the frequency check);
* cos(2 pi u) of the refine offset against numpy; the Box-Muller scalars
  have mean 0 and variance 1;
* invariants of a full step on a random scene: row counts, survivors, the
  refined pairs placed symmetrically about their source, copied SH rows,
  state of new rows.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from rustscan_amd import scene

LN = lambda x: float(np.log(np.float32(x)))  # noqa: E731
IDENT = [1.0, 0.0, 0.0, 0.0]


def splats(positions, log_scales, rotations, logits, deg=0):
    n = len(logits)
    tf = np.zeros((n, 10), np.float32)
    tf[:, 0:3] = np.asarray(positions, np.float32).reshape(n, 3)
    tf[:, 3:7] = np.asarray(rotations, np.float32).reshape(n, 4)
    tf[:, 7:10] = np.asarray(log_scales, np.float32).reshape(n, 3)
    k = (deg + 1) ** 2
    sh = (np.arange(n * k * 3, dtype=np.float32).reshape(n, k, 3) * 0.01)
    return scene.Splats(tf, sh, np.asarray(logits, np.float32), deg)


def dens_for(n, grad_2d=None, visible=None, num_obs=None, actual=None, coverage=None, depth=None):
    d = np.zeros((10, n), np.float32)
    if grad_2d is not None:
        d[0] = grad_2d
    if coverage is not None:
        d[4] = coverage
    if depth is not None:
        d[5] = depth
    d[7] = 1.0 if num_obs is None else num_obs
    if visible is not None:
        d[8] = visible
    if actual is not None:
        d[9] = actual
    return d


# a schedule where the step densifies and prunes: refine every iteration, iteration 20, one frame
DENSIFY = dict(refine_every=1)
IT, FC = 20, 1


def step(sc, dens, cfg, age=10, invis=0, iteration=IT, frame_count=FC):
    n = sc.n
    assert iteration >= age
    birth = np.full(n, iteration - age, np.uint32)
    win = np.full(n, invis, np.uint32) if np.isscalar(invis) else np.asarray(invis, np.uint32)
    return O.topology_step(cfg, sc, dens, birth, win, iteration, frame_count)


# ---- schedule (topology/tests.rs:316-498)

def test_schedule_respects_refine_cadence_and_freeze():
    cfg = O.topology_cfg(iterations=30, refine_every=2, opacity_reset_interval=2, topology_freeze_after_epoch=4)
    e2 = O.topology_schedule(cfg, 7, 2)
    assert e2 == dict(completed_epoch=3, densify=1, prune=1, reset_opacity=0, allow_extra_growth=1)
    e4 = O.topology_schedule(cfg, 11, 2)
    assert e4["completed_epoch"] == 5 and not e4["densify"] and not e4["prune"] and not e4["reset_opacity"]


def test_reset_mode_schedules_opacity_reset_on_epoch_cadence():
    cfg = O.topology_cfg(iterations=30, refine_every=2, opacity_reset_interval=2, opacity_reset_mode=1)
    e2 = O.topology_schedule(cfg, 5, 2)
    assert e2["completed_epoch"] == 2 and e2["reset_opacity"]
    e3 = O.topology_schedule(cfg, 7, 2)
    assert e3["completed_epoch"] == 3 and not e3["reset_opacity"]


def test_growth_freeze_still_allows_pruning():
    cfg = O.topology_cfg(iterations=30000, growth_freeze_after_epoch=4, prune_until_epoch=30, refine_every=160)
    s = O.topology_schedule(cfg, 1601, 180)
    assert s["completed_epoch"] == 8 and not s["densify"] and s["prune"] and not s["allow_extra_growth"]


def test_short_run_schedule():
    cfg = O.topology_cfg(iterations=500, refine_every=200)
    early = O.topology_schedule(cfg, 100, 638)
    assert early["completed_epoch"] == 0 and not early["densify"] and not early["prune"]
    for it in (201, 401):
        s = O.topology_schedule(cfg, it, 638)
        assert s == dict(completed_epoch=0, densify=1, prune=1, reset_opacity=0, allow_extra_growth=1)


def test_should_apply_uses_refine_cadence():
    cfg = O.topology_cfg(iterations=500, refine_every=160)
    applies = lambda it: any(O.topology_schedule(cfg, it, 638)[k] for k in ("densify", "prune", "reset_opacity"))
    assert not applies(100) and applies(161) and not applies(201)
    assert O.topology_schedule(cfg, 0, 638)["completed_epoch"] == -1
    assert O.topology_schedule(cfg, 5, 0)["completed_epoch"] == -1


# ---- analysis and pruning (topology/tests.rs:578-827); the scene extent is 1
# (positions 2 apart around their centre), as the reference tests' policy

TWO = [0.0, 0.0, 1.0, 2.0, 0.0, 1.0]


def test_marks_growth_and_prune():
    cfg = O.topology_cfg(prune_min_age=1, prune_invisible_epochs=1, **DENSIFY)
    sc = splats(TWO, [LN(0.005)] * 3 + [LN(0.5)] * 3, IDENT * 2, [2.0, -10.0])
    r = step(sc, dens_for(2, grad_2d=[1.0, 0.0], visible=[1.0, 0.0]), cfg, age=1, invis=[0, 1])
    assert r["scene_extent"] == 1.0
    assert (r["clone_candidates"], r["prune_candidates"], r["growth_candidates"]) == (1, 1, 1)
    assert r["flags"][0] & 4 and r["flags"][1] & 8


@pytest.mark.parametrize("mode,densify,pruned", [(0, True, 1), (1, False, 0), (1, True, 1)])
def test_opacity_pruning_modes(mode, densify, pruned):
    """Configured opacity threshold (threshold mode); weight mode keeps a
    visible low-opacity splat when the step does not densify (growth frozen)
    and prunes it when it does (tests.rs:633-729)."""
    extra = DENSIFY if densify else dict(refine_every=1, growth_freeze_after_epoch=0)
    cfg = O.topology_cfg(prune_opacity_threshold=0.02, prune_mode=mode, prune_min_age=1, prune_invisible_epochs=100,
                         **extra)
    sc = splats([0.0, 0.0, 1.0], [LN(0.005)] * 3, IDENT, [math.log(0.01 / 0.99)])
    r = step(sc, dens_for(1, visible=[1.0]), cfg, age=10)
    assert r["schedule"]["densify"] == int(densify)
    assert r["prune_candidates"] == pruned


def test_invisible_pruning_requires_consecutive_windows():
    base = dict(prune_min_age=1, prune_invisible_epochs=3)
    sc = splats(TWO, [LN(0.005)] * 6, IDENT * 2, [1.0, 1.0])
    # windows 1 and 2 become 2 and 3 (invisible in this step)
    frozen = step(sc, dens_for(2, visible=[0.0, 0.0]), O.topology_cfg(refine_every=1, growth_freeze_after_epoch=0,
                                                                       **base), invis=[1, 2])
    assert list(frozen["flags"] & 8) == [0, 8]
    growing = step(sc, dens_for(2, visible=[0.0, 0.0]), O.topology_cfg(**base, **DENSIFY), invis=[1, 2])
    assert growing["prune_candidates"] == 0


def test_opacity_pruning_respects_min_age():
    cfg = O.topology_cfg(prune_min_age=5, prune_opacity_threshold=0.02, **DENSIFY)
    sc = splats(TWO, [LN(0.005)] * 6, IDENT * 2, [math.log(0.01 / 0.99)] * 2)
    birth = np.array([IT - 4, IT - 5], np.int64)
    r = O.topology_step(cfg, sc, dens_for(2, visible=[1.0, 1.0]), np.maximum(birth, 0).astype(np.uint32),
                        np.zeros(2, np.uint32), 10, 1)
    # ages at iteration 10 are 10 - birth: 12 and 13; use birth relative to 10 instead
    birth = np.array([10 - 4, 10 - 5], np.uint32)
    r = O.topology_step(cfg, sc, dens_for(2, visible=[1.0, 1.0]), birth, np.zeros(2, np.uint32), 10, 1)
    assert list(r["flags"] & 8) == [0, 8]


def test_out_of_bounds_tiny_scale_and_nonfinite_rows_are_pruned():
    pos = [0.0, 0.0, 1.0, 2.0, 0.0, 1.0, 1.0, 0.0, 1.0, 1.0, 1.0, 1.0]
    sc = splats(pos, [LN(0.05)] * 3 + [LN(0.05)] * 3 + [-40.0, LN(0.05), LN(0.05)] + [LN(0.05)] * 3, IDENT * 4,
                [2.0] * 4, deg=1)
    sc.sh[3, 2, 1] = np.nan
    r = step(sc, dens_for(4, visible=[1.0] * 4), O.topology_cfg(prune_min_age=100, **DENSIFY))
    assert list(r["flags"] & 8) == [0, 0, 8, 8]  # exp(-40) < 1e-10; a NaN SH coefficient is not retainable


# ---- selection (tests.rs:500-547) and the weighted sampler

def test_requested_additions_and_replacements_first():
    # one prune candidate (opacity logit -10, age 10) and three growth candidates
    cfg = O.topology_cfg(growth_select_fraction=1.0, max_initial_gaussians=1, **DENSIFY)
    pos = [0.0, 0.0, 1.0, 2.0, 0.0, 1.0, 1.0, 0.5, 1.0, 1.0, -0.5, 1.0]
    sc = splats(pos, [LN(0.005)] * 12, IDENT * 4, [-10.0, 2.0, 1.5, 1.0])
    r = step(sc, dens_for(4, grad_2d=[0.0, 2.0, 3.0, 1.0], visible=[1.0] * 4), cfg)
    assert (r["prune_candidates"], r["growth_candidates"]) == (1, 3)
    assert (r["replacement_count"], r["extra_growth_count"], r["selected_count"]) == (1, 2, 3)
    assert all(i > 0 for i in r["selected"])
    assert len(set(r["selected"].tolist())) == 3
    assert (r["before"], r["pruned"], r["added"], r["after"]) == (4, 1, 3, 6)
    again = step(sc, dens_for(4, grad_2d=[0.0, 2.0, 3.0, 1.0], visible=[1.0] * 4), cfg)
    np.testing.assert_array_equal(again["selected"], r["selected"])
    # no extra growth when the schedule does not allow it: replacements only
    cfg2 = O.topology_cfg(growth_select_fraction=1.0, max_initial_gaussians=1, growth_stop_iter=0, **DENSIFY)
    r2 = step(sc, dens_for(4, grad_2d=[0.0, 2.0, 3.0, 1.0], visible=[1.0] * 4), cfg2)
    assert (r2["replacement_count"], r2["extra_growth_count"]) == (1, 0)


def test_weighted_keys_pick_in_proportion_to_weight():
    w = np.array([1.0, 2.0, 3.0, 4.0], np.float32)
    L = O._topology_lib()
    first = np.zeros(4)
    trials = 20000
    for seed in range(trials):
        keys = [L.orc_topology_key(seed * 7919 + 13, 1, i, float(w[i])) for i in range(4)]
        first[int(np.argmax(keys))] += 1
    np.testing.assert_allclose(first / trials, w / w.sum(), atol=0.012)
    assert L.orc_topology_key(1, 1, 0, 0.0) == -np.inf


def test_refine_scalar_is_standard_normal_and_cos_is_accurate():
    u = np.linspace(-1.5, 1.5, 20001).astype(np.float32)
    L = O._topology_lib()
    got = np.array([L.orc_cos_2pi(float(x)) for x in u])
    assert np.abs(got - np.cos(2 * np.pi * u.astype(np.float64))).max() < 3e-7
    s = np.array([L.orc_topology_refine_scalar(12345, j) for j in range(20000)])
    assert abs(s.mean()) < 0.03 and abs(s.var() - 1.0) < 0.04


def test_topology_seed_formula():
    # base * 0x9e3779b97f4a7c15 + iteration, rotated left by 17 (mod.rs:1031-1036)
    base, it = 0x1234_5678_9ABC_DEF0, 321
    x = (base * 0x9E3779B97F4A7C15 + it) % (1 << 64)
    assert O._topology_lib().orc_topology_seed(base, it) == ((x << 17) | (x >> 47)) % (1 << 64)


# ---- mutations (apply.rs, tests.rs:976-1243)

def test_refine_offset_examples():
    # apply.rs tests: a doubled 90-degree z rotation is normalised; zero rotation is the identity
    h = math.sqrt(0.5)
    off = O.brush_refine_offset([2 * h, 0.0, 0.0, 2 * h], [2.0, 0.0, 0.0], 1.0)
    np.testing.assert_allclose(off, [0.0, 2.0, 0.0], atol=1e-5)
    np.testing.assert_array_equal(O.brush_refine_offset([0.0] * 4, [1.0, 2.0, 3.0], 0.5), [0.5, 1.0, 1.5])


def test_prune_and_densify_rebuild():
    cfg = O.topology_cfg(growth_select_fraction=1.0, max_initial_gaussians=1, **DENSIFY)
    sc = splats(TWO, [LN(0.05), LN(0.02), LN(0.01)] * 2, IDENT * 2, [2.0, -10.0], deg=1)
    r = step(sc, dens_for(2, grad_2d=[1.0, 0.0], visible=[1.0, 1.0]), cfg)
    assert (r["added"], r["pruned"], r["after"], r["mutated"], r["adam_reset"]) == (1, 1, 2, 1, 1)
    assert list(r["row_src"]) == [0, 0] and list(r["row_kind"]) == [1, 2]
    t0 = sc.transforms[0]
    a, b = r["transforms"]
    # the pair sits symmetrically about the source along the rotated scale
    np.testing.assert_allclose((a[0:3] + b[0:3]) / 2, t0[0:3], atol=1e-6)
    assert np.abs(a[0:3] - t0[0:3]).max() > 0
    # largest axis halved, others kept
    np.testing.assert_allclose(np.exp(a[7:10]), [0.025, 0.02, 0.01], rtol=1e-5)
    np.testing.assert_array_equal(a[7:10], b[7:10])
    # opacity 1 - sqrt(1 - o)
    o = 1 / (1 + np.exp(-2.0))
    np.testing.assert_allclose(1 / (1 + np.exp(-r["raw_opac"])), 1 - np.sqrt(1 - o), rtol=1e-5)
    np.testing.assert_array_equal(r["sh"][0], sc.sh[0])
    np.testing.assert_array_equal(r["sh"][1], sc.sh[0])
    assert list(r["birth"]) == [IT - 10, IT] and list(r["invisible_windows"]) == [0, 0]


def test_opacity_reset_caps_high_opacity_only():
    cfg = O.topology_cfg(iterations=30, refine_every=2, opacity_reset_interval=2, opacity_reset_mode=1,
                         growth_freeze_after_epoch=0)
    sc = splats(TWO, [LN(0.05)] * 6, IDENT * 2, [math.log(0.8 / 0.2), math.log(0.005 / 0.995)])
    r = step(sc, dens_for(2, visible=[1.0, 1.0]), cfg, age=2, iteration=5, frame_count=2)
    assert r["schedule"]["reset_opacity"] and (r["added"], r["pruned"]) == (0, 0)
    assert r["mutated"] and r["adam_reset"]
    op = 1 / (1 + np.exp(-r["raw_opac"].astype(np.float64)))
    np.testing.assert_allclose(op, [0.01, 0.005], atol=1e-6)


def test_no_op_step_keeps_splats_and_adam():
    cfg = O.topology_cfg(refine_every=1000)  # not on the refine cadence
    sc = splats(TWO, [LN(0.05)] * 6, IDENT * 2, [2.0, 2.0])
    r = step(sc, dens_for(2, visible=[0.0, 1.0]), cfg, invis=[4, 4])
    assert (r["added"], r["pruned"], r["mutated"], r["adam_reset"]) == (0, 0, 0, 0)
    np.testing.assert_array_equal(r["transforms"], sc.transforms)
    np.testing.assert_array_equal(r["raw_opac"], sc.raw_opac)
    assert list(r["invisible_windows"]) == [5, 0]  # the visibility windows still advance


def test_refine_decay_shrinks_without_structural_change():
    # iteration 15 of 30 -> train_t 0.5; scale decay 0.2, opacity decay 0.1 (tests.rs:1203-1243)
    cfg = O.topology_cfg(iterations=30, refine_every=2, opacity_decay=0.1, scale_decay=0.2, growth_freeze_after_epoch=0)
    sc = splats([0.0, 0.0, 1.0], [LN(0.1)] * 3, IDENT, [0.0])
    r = step(sc, dens_for(1, visible=[1.0]), cfg, age=2, iteration=15, frame_count=1)
    assert (r["added"], r["pruned"], r["mutated"]) == (0, 0, 1)
    np.testing.assert_allclose(np.exp(r["transforms"][0, 7:10]), [0.09] * 3, rtol=1e-5)
    np.testing.assert_allclose(1 / (1 + np.exp(-float(r["raw_opac"][0]))), 0.45, atol=1e-6)


def test_full_step_invariants_random_scene():
    sc = scene.gaussian_scene(4000, seed=31, sh_degree=2)
    rng = np.random.default_rng(5)
    n = sc.n
    dens = dens_for(n, grad_2d=rng.exponential(2e-4, n) * 3, visible=rng.integers(0, 3, n).astype(np.float32),
                    num_obs=3.0, coverage=rng.uniform(0, 50, n), depth=rng.uniform(0, 200, n))
    cfg = O.topology_cfg(**DENSIFY)
    r = step(sc, dens, cfg, age=20, invis=rng.integers(0, 12, n))
    keep = (r["flags"] & 8) == 0
    assert r["pruned"] == n - keep.sum() and r["after"] == n - r["pruned"] + r["added"]
    assert r["added"] == r["selected_count"] > 0
    S = int(keep.sum())
    np.testing.assert_array_equal(r["row_src"][:S], np.nonzero(keep)[0])
    np.testing.assert_array_equal(r["row_src"][S:], r["selected"][:r["added"]])
    assert set(r["row_kind"][:S]) <= {0, 1} and set(r["row_kind"][S:]) == {2}
    plain = r["row_kind"][:S] == 0
    np.testing.assert_array_equal(r["transforms"][:S][plain], sc.transforms[r["row_src"][:S][plain]])
    np.testing.assert_array_equal(r["sh"], sc.sh[r["row_src"]])
    assert (r["birth"][S:] == IT).all() and (r["invisible_windows"][S:] == 0).all()
