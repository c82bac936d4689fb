"""Pins the oracle's temporal history merge (``oracle.temporal_merge``) against the rules of
update_temporal_state (full_pyramid.wgsl:519-565) and reset_temporal_state
(full_pyramid_backend.py:660-677), case by case."""

from __future__ import annotations

import numpy as np
import pytest

import oracle


def _state(statuses, value):
    s = np.zeros((len(statuses), 8), np.float32)
    s[:, 0] = value
    s[:, 6] = statuses
    return s


def test_merge_rules_filtered_and_raw():
    hist = _state([-1, -1, -1, -1, -1], 0.0)
    ages = np.full(5, oracle.NEVER, np.uint64)
    raw = _state([1, 2, 0, 1, -1], 5.0)
    out = _state([1, 2, 2, 0, 1], 7.0)
    # Filtered output (replace policy): final accepted is enough.
    h, a = hist.copy(), ages.copy()
    n = oracle.temporal_merge(h, a, raw, out, True, 0xFFFFFFFE)
    assert n == 4
    np.testing.assert_array_equal(h[:, 0], [7, 7, 7, 0, 7])
    np.testing.assert_array_equal(a, [0, 0, 0, oracle.NEVER, 0])
    # Raw state: both raw and final must be accepted, and the raw values are stored.
    h, a = hist.copy(), ages.copy()
    n = oracle.temporal_merge(h, a, raw, out, False, 0xFFFFFFFE)
    assert n == 2
    np.testing.assert_array_equal(h[:, 0], [5, 5, 0, 0, 0])


def test_ages_saturate_and_expire():
    hist = _state([1, 1, 1], 3.0)
    ages = np.array([0, 0xFFFFFFFD, oracle.NEVER], np.uint64)
    rejected = _state([0, 0, 0], 9.0)
    n = oracle.temporal_merge(hist, ages, rejected, rejected, True, 0)
    assert n == 0
    # Age 0 -> 1 > max_age 0: expired (status -1, values kept); the saturated age stays below the
    # sentinel; a never-set entry stays never-set.
    np.testing.assert_array_equal(ages, [1, 0xFFFFFFFE, oracle.NEVER])
    np.testing.assert_array_equal(hist[:, 6], [-1, -1, 1])
    np.testing.assert_array_equal(hist[:, 0], [3, 3, 3])
    hist = _state([1], 3.0)
    ages = np.array([0], np.uint64)
    oracle.temporal_merge(hist, ages, _state([0], 0.0), _state([0], 0.0), True, 1)
    assert hist[0, 6] == 1 and ages[0] == 1        # within max_age 1


@pytest.mark.parametrize("case,levels", [("translation-l2-none-temporal", 2), ("translation-l3-none-temporal", 3)])
def test_temporal_translation_goldens(case, levels):
    """WGSL temporal goldens (before.npz, qualify_binding_budget.py:36-90): frames np.roll(reference,
    (-i, 2i)), i = 0, 1, 2, through one temporal engine; frame 2 is seeded from frame 1's history."""
    from helpers import GOLDEN, compare_fields, load_gray

    ref = load_gray("uhd_reference.png")
    goldens = np.load(GOLDEN / "uhd_goldens.npz", allow_pickle=False)
    pipe = oracle.Pipeline(pyramid_levels=levels, motion_model="translation", outlier_policy="median",
                           median_filter_scope="all", initialization="temporal")
    pipe.set_reference(ref)
    for frame in range(3):
        out = pipe.register(np.roll(ref, (-frame, 2 * frame), (0, 1)))
        if frame == 0:
            continue
        gold = np.concatenate([goldens[f"{case}-f{frame}-parameters"], goldens[f"{case}-f{frame}-status"][..., None],
                               goldens[f"{case}-f{frame}-iterations"][..., None]], axis=-1)
        stats = compare_fields(out, gold, 6)
        print(case, frame, stats, pipe.last_counts)
        assert pipe.last_counts[0] > 0
        assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999


def test_temporal_affine_golden():
    """affine-l2-none-temporal frame 1 (the tensile pair seeded by the identity frame): statuses and
    parameters.  The capture's "median" policy filtered every subset (median_filter_scope="all")."""
    from helpers import GOLDEN, compare_fields, load_gray

    ref = load_gray("uhd_reference.png")
    goldens = np.load(GOLDEN / "uhd_goldens.npz", allow_pickle=False)
    pipe = oracle.Pipeline(pyramid_levels=2, outlier_policy="median", median_filter_scope="all",
                           initialization="temporal")
    pipe.set_reference(ref)
    pipe.register(ref)
    out = pipe.register(load_gray("uhd_frame_040.png"))
    key = "affine-l2-none-temporal-f1"
    gold = np.concatenate([goldens[f"{key}-parameters"], goldens[f"{key}-status"][..., None],
                           goldens[f"{key}-iterations"][..., None]], axis=-1)
    stats = compare_fields(out, gold, 6)
    print(stats)
    assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999


# -- latency profile: velocity prediction, fine correction, divergence patience ------------------
def _hist(n, P=6):
    h = np.zeros((n, P + 2), np.float32)
    h[:, P] = -1.0
    return h


def test_velocity_prediction_rule():
    """predict_temporal_state (full_pyramid.wgsl:83-113): only fresh accepted pairs extrapolate."""
    h, p = _hist(4), _hist(4)
    h[:, :6] = np.arange(24, dtype=np.float32).reshape(4, 6)
    p[:, :6] = 1.0
    h[:, 6] = [1, 1, 1, -1]
    p[:, 6] = [1, 2, 1, 1]
    h[:, 7] = 5.0
    ages = np.array([0, 0, 1, 0], np.uint64)
    pages = np.array([0, 0, 0, 0], np.uint64)
    out, count = oracle.temporal_predict(h, p, ages, pages)
    assert count == 2
    np.testing.assert_array_equal(out[:2, :6], 2 * h[:2, :6] - 1.0)
    np.testing.assert_array_equal(out[2:], h[2:])
    np.testing.assert_array_equal(out[:, 6:], h[:, 6:])


def test_save_previous_rule():
    """save_previous_temporal_state (full_pyramid.wgsl:115-137)."""
    h = _hist(4)
    h[:, 0] = [1, 2, 3, 4]
    prev = _hist(4)
    ages = np.array([0, 3, 0, 0], np.uint64)
    pages = np.array([7, oracle.NEVER, 0xFFFFFFFE, 2], np.uint64)
    raw = _hist(4)
    out = _hist(4)
    raw[:, 6] = [1, 2, 1, 0]
    out[:, 6] = [2, 1, 0, 1]
    oracle.temporal_save_previous(h, ages, raw, out, prev, pages)
    np.testing.assert_array_equal(prev[:2, 0], [1, 2])
    assert list(pages) == [0, 3, 0xFFFFFFFE, 3]


def test_fine_correction_adds_the_coarse_increment():
    """refine_temporal_seed (full_pyramid.wgsl:451-515): with a uniform coarse history of the fine
    subset's age, the seed becomes fine history + (absolute seed - coarse history interpolated);
    an age mismatch, an inactive or unavailable seed, or an expired history keeps the absolute seed."""
    cgw, cgh, fgw, fgh = 6, 5, 12, 10
    cfg = oracle.SeedCfg(cgw, cgh, fgw, fgw * fgh, 2.0, 1.0, 16.0, 16.0, 0, 1)
    coarse = _hist(cgw * cgh)
    coarse[:, 0] = 0.5                    # coarse-grid translation; x2 at the fine scale
    coarse[:, 6] = 1.0
    cages = np.zeros(cgw * cgh, np.uint64)
    fine = _hist(fgw * fgh)
    fine[:, 0] = 3.0
    fine[:, 3] = -2.0
    fine[:, 6] = 1.0
    fages = np.zeros(fgw * fgh, np.uint64)
    seed = np.zeros((fgw * fgh, 8), np.float32)
    seed[:, 0] = 1.25
    seed[:, 3] = 0.5
    active = np.ones(fgw * fgh, np.uint32)
    active[5] = 0
    seed[6, 6] = -1.0
    fages[7] = 2
    fages[8] = oracle.NEVER
    cages[0] = 4                          # the corner shared by fine subsets near the origin
    before = seed.copy()
    n = oracle.temporal_refine(cfg, coarse, cages, fine, fages, 1, None, seed, active)
    changed = np.flatnonzero(np.any(seed != before, axis=1))
    unchanged = {5, 6, 7, 8, 0, 1, 12, 13}
    assert n == len(changed) and not unchanged & set(changed.tolist())
    k = changed[-1]
    assert seed[k, 0] == pytest.approx(3.0 + (1.25 - 1.0)) and seed[k, 3] == pytest.approx(-2.0 + 0.5)


@pytest.mark.parametrize("case,frames,roi", [("translation-l2-none-latency", (1, 2), "none"),
                                             ("translation-l2-full-latency", (2,), "full"),
                                             ("translation-l2-partial-latency", (2,), "partial")])
def test_latency_translation_goldens(case, frames, roi):
    """WGSL latency-profile goldens (qualify_binding_budget.py:54-60: velocity prediction, fine
    correction, divergence patience 3 from iteration 4, median), frames np.roll(reference, (-i, 2i))."""
    from helpers import GOLDEN, compare_fields, load_gray, oracle_rois
    from oracle.roi_plan import seed_permissions

    ref = load_gray("uhd_reference.png")
    goldens = np.load(GOLDEN / "uhd_goldens.npz", allow_pickle=False)
    rois = links = None
    if roi != "none":
        mask = np.zeros(ref.shape, bool)
        mask[160:2000, 160:3680] = True
        mask[880:1200, 1800:2040] = False
        rois = oracle_rois(16, 16, mask, 2, policy=roi)
        links = seed_permissions(rois[0]["plan"], rois[1]["plan"], source_scale=1, destination_scale=2, size=16,
                                 stride=16, reference_mask=mask)
    pipe = oracle.Pipeline(pyramid_levels=2, motion_model="translation", outlier_policy="median",
                           median_filter_scope="all", initialization="temporal", temporal_prediction="velocity", temporal_fine_correction=True,
                           divergence_patience=3, divergence_min_iterations=4)
    pipe.set_reference(ref, rois=rois, temporal_links=links)
    for frame in range(3):
        out = pipe.register(np.roll(ref, (-frame, 2 * frame), (0, 1)))
        if frame not in frames:
            continue
        key = f"{case}-f{frame}"
        gold = np.concatenate([goldens[f"{key}-parameters"], goldens[f"{key}-status"][..., None],
                               goldens[f"{key}-iterations"][..., None]], axis=-1)
        stats = compare_fields(out, gold, 6)
        print(key, stats, pipe.last_counts, pipe.last_predicted)
        assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999
        assert np.mean(out[..., 7] == gold[..., 7]) >= 0.98
        if frame == 2:
            assert pipe.last_predicted > 0


def test_adaptive_pyramid_rule_on_a_sequence():
    """The adaptive feedback (full_pyramid_backend.py:1775-1791) on the oracle: a good frame
    switches the next one to the finest level alone, a corrupted frame (whose expired
    history is counted) switches back, and a reset restarts with both levels."""
    from helpers import load_gray
    from test_oracle_partial import fractional_shift

    src = load_gray("frame_000.png")[:256, :256]
    targets = [fractional_shift(src, 0.2 * i, -0.1 * i) for i in range(4)]
    sequence = [*targets, np.zeros_like(src), targets[-1], targets[-1]]
    pipe = oracle.Pipeline(pyramid_levels=2, initialization="temporal", temporal_prediction="velocity",
                           pyramid_policy="adaptive", adaptive_acceptance=0.7, temporal_max_age=0,
                           outlier_policy="median", temporal_fine_correction=True)
    pipe.set_reference(src)
    levels, expired = [], []
    for t in sequence:
        pipe.register(t)
        levels.append(pipe.last_levels)
        expired.append(pipe.last_expired)
    print(levels, expired)
    assert levels[:2] == [2, 1] and levels[5] == 2
    assert expired[4] > 0
    pipe.reset_temporal_state()
    pipe.register(targets[0])
    assert pipe.last_levels == 2
