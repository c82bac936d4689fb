"""The oracle's ROI planner against brute force: selection, reduced masks, neighbour and seed
links (roi.py semantics).  The library's device planner is compared with it on the GPU."""

from __future__ import annotations

import numpy as np
import pytest

from live_dic.roi import validate_roi
from oracle.roi_plan import integral, prepare_roi, reduce_mask, seed_permissions, select_windows


def test_full_window_selection_matches_brute_force():
    rng = np.random.default_rng(0)
    mask = rng.random((60, 70)) > 0.02
    sel = select_windows(mask, 8, 4)
    gh, gw = sel.shape
    for y in range(gh):
        for x in range(gw):
            assert sel[y, x] == mask[y * 4 : y * 4 + 8, x * 4 : x * 4 + 8].all()


def test_reduce_mask_requires_all_25_reflected_pixels():
    mask = np.ones((20, 20), bool)
    mask[10, 10] = False
    red = reduce_mask(mask)
    assert red.shape == (10, 10)
    # (2x, 2y) windows of radius 2 that touch (10, 10): x, y in {4, 5, 6}.
    assert not red[4:7, 4:7].any()
    assert red.sum() == 100 - 9


def test_neighbour_bits_cut_across_holes():
    mask = np.ones((64, 64), bool)
    mask[:, 30:34] = False  # a vertical slit
    plan = prepare_roi(mask, 8, 8)
    assert plan.effective
    bits = plan.neighbour_bits
    # Centre bit (4) is set for selected subsets only.
    assert np.array_equal((bits >> 4) & 1 == 1, plan.selected)
    left, right = 2, 4  # grid columns flanking the slit (x 16..23, 32..39 -> 32 is masked)
    assert plan.selected[3, left] and not plan.selected[3, 3]
    assert ((bits[3, left] >> 5) & 1) == 0  # no right link across the slit
    assert plan.selected_count == int(plan.selected.sum())


def test_all_true_mask_is_not_effective():
    plan = prepare_roi(np.ones((40, 40), bool), 8, 8)
    assert not plan.effective and plan.selected.all()


def test_validation():
    with pytest.raises(ValueError):
        validate_roi(np.ones((4, 4), bool), "window", (4, 4), "affine")
    with pytest.raises(TypeError):
        validate_roi(np.ones((4, 4), np.uint8), "full", (4, 4), "affine")
    with pytest.raises(ValueError):
        validate_roi(np.ones((4, 5), bool), "full", (4, 4), "affine")
    assert validate_roi(None, "full", (4, 4), "affine") is None


def test_seed_permissions_select_corners_and_fallback():
    mask = np.ones((128, 128), bool)
    mask[60:68, :] = False
    fine = prepare_roi(mask, 8, 8)
    coarse = prepare_roi(reduce_mask(mask), 8, 8)
    links = seed_permissions(coarse, fine, source_scale=2, destination_scale=1, size=8, stride=8,
                             reference_mask=mask, reference_prefix=integral(mask))
    assert links.shape == fine.selected.shape
    assert np.array_equal((links >> 4) & 1 == 1, fine.selected)
    # Unselected destinations carry no corner links.
    assert not ((links & 15)[~fine.selected]).any()
