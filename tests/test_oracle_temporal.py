"""Pins the oracle's temporal history merge (``oracle.temporal_merge``) against the rules of
update_temporal_state (full_pyramid.wgsl:519-565) and reset_temporal_state
(full_pyramid_backend.py:660-677), case by case."""

from __future__ import annotations

import numpy as np

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
