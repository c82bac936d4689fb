"""GPU: the device ROI planner and partial-support statistics against the oracle's restatements.

The planner's tables are integer work and must equal the oracle planner's bit for bit: selection,
3x3 link bits, median link words, partial supports (indices, bits, boxes, coverage), transfer links
between pyramid levels and onto the coarsest level, reduced masks and their hashes.  The float64
support statistics must give the same admissions and ranks as the oracle's and the same metadata
rows and inverses to float32 rounding.
"""

from __future__ import annotations

import time

import numpy as np
import pytest

from helpers import load_gray
from oracle import partial as opartial
from oracle import roi_plan as oplan

pytestmark = pytest.mark.gpu


def _say(*args):
    print("[case]", *args, flush=True)


def _masks(h=200, w=260):
    rng = np.random.default_rng(77)
    yy, xx = np.mgrid[:h, :w]
    holes = np.ones((h, w), bool)
    for cx, cy, r in ((70, 60, 25), (180, 130, 33), (40, 160, 12)):
        holes &= (xx - cx) ** 2 + (yy - cy) ** 2 > r * r
    slit = np.ones((h, w), bool)
    slit[30:170, 128:130] = False
    slit[100, :90] = False
    noisy = rng.random((h, w)) > 0.004
    ellipse = ((xx - w / 2) / (0.45 * w)) ** 2 + ((yy - h / 2) / (0.42 * h)) ** 2 <= 1
    return {"holes": holes, "slit": slit, "noisy": noisy, "ellipse": ellipse & slit, "all": np.ones((h, w), bool)}


CASES = [(16, 16, "full", "none", 3), (17, 8, "partial", "none", 3), (16, 8, "partial", "visibility", 5),
         (25, 4, "full", "none", 9), (21, 7, "partial", "visibility", 3), (12, 12, "partial", "none", 7)]


@pytest.mark.parametrize("size,stride,policy,truncation,median", CASES)
def test_planner_tables_equal_oracle(size, stride, policy, truncation, median):
    from live_dic.roi import mask_digest, plan_pyramid

    levels = 2
    for name, mask in _masks().items():
        _say("roi plan", name, size, stride, policy, truncation, median)
        got = plan_pyramid(mask, levels, size=size, stride=stride, policy=policy, truncation=truncation,
                           median_filter_size=median, temporal=True)
        masks = [mask, oplan.reduce_mask(mask)]
        want = [oplan.prepare_roi(m, size, stride, policy=policy, truncation=truncation, median_filter_size=median)
                for m in masks]
        for k in range(levels):
            g, w = got.plans[k], want[k]
            np.testing.assert_array_equal(g.mask, masks[k], err_msg=f"{name} mask {k}")
            assert g.mask_hash == mask_digest(masks[k]) == w.mask_hash
            assert g.effective == w.effective and g.selected_count == w.selected_count
            np.testing.assert_array_equal(g.selected, w.selected, err_msg=f"{name} selected {k}")
            np.testing.assert_array_equal(g.neighbour_bits, w.neighbour_bits, err_msg=f"{name} bits {k}")
            if median > 3:
                np.testing.assert_array_equal(g.median_neighbour_words, w.median_neighbour_words)
            if policy == "partial":
                np.testing.assert_array_equal(g.partial_indices, w.partial_indices, err_msg=f"{name} partial {k}")
                np.testing.assert_array_equal(g.support_words, w.support_words)
                np.testing.assert_array_equal(g.support_boxes, w.support_boxes)
                np.testing.assert_array_equal(g.coverage, w.coverage)
        if want[0].effective:
            links = oplan.seed_permissions(want[1], want[0], source_scale=2, destination_scale=1, size=size,
                                           stride=stride, reference_mask=mask)
            np.testing.assert_array_equal(got.seed_links[0], links, err_msg=f"{name} seed links")
            proj = oplan.seed_permissions(want[0], want[1], source_scale=1, destination_scale=2, size=size,
                                          stride=stride, reference_mask=mask)
            np.testing.assert_array_equal(got.temporal_links, proj, err_msg=f"{name} temporal links")
        else:
            assert got.seed_links[0] is None and got.temporal_links is None


def test_partial_supports_exist_for_the_cases_above():
    """The planner test above is only meaningful with partial supports present."""
    from live_dic.roi import plan_pyramid

    plan = plan_pyramid(_masks()["ellipse"], 1, size=17, stride=8, policy="partial").plans[0]
    assert plan.partial_indices.size > 20
    vis = plan_pyramid(_masks()["slit"], 1, size=16, stride=8, policy="partial", truncation="visibility").plans[0]
    full = plan_pyramid(_masks()["slit"], 1, size=16, stride=8, policy="partial").plans[0]
    assert not np.array_equal(vis.support_words, full.support_words)


@pytest.mark.parametrize("model,weighting,size,stride", [("translation", "uniform", 17, 8),
                                                         ("affine", "gaussian", 17, 8),
                                                         ("affine", "uniform", 16, 16),
                                                         ("translation", "gaussian", 21, 10)])
def test_partial_statistics_equal_oracle(model, weighting, size, stride):
    """Metadata rows, inverses, admissions and ranks of the device's float64 support statistics
    against the oracle's, both computed from the engine's prepared plane."""
    from live_dic import ICGNWGPU

    img = load_gray("frame_000.png")[:256, :256]
    yy, xx = np.mgrid[:256, :256]
    mask = ((xx - 128) ** 2 / 110.0 ** 2 + (yy - 128) ** 2 / 95.0 ** 2 <= 1) & ~((xx > 120) & (xx < 123) & (yy < 140))
    _say("partial statistics", model, weighting, size, stride)
    with ICGNWGPU(motion_model=model, subset_size=size, subset_stride=stride, spatial_weighting=weighting,
                  interpolation="linear") as eng:
        info = eng.set_reference(img, roi_mask=mask, roi_policy="partial")
        plane, meta, inv = eng.export_reference(0)
        plan = info.roi
    idx = np.asarray(plan.partial_indices, np.int64)
    assert idx.size > 10
    want_meta = meta.copy()
    want_inv = inv.copy()
    diag = opartial.partial_reference(plane, want_meta, want_inv, plan.partial_indices, plan.support_words,
                                      grid_width=plan.selected.shape[1], size=size, stride=stride, motion_model=model,
                                      weighting=weighting)
    np.testing.assert_array_equal(meta[idx, 1], want_meta[idx, 1])
    admitted = meta[idx, 1] == -4.0
    assert admitted.sum() > 5
    np.testing.assert_array_equal(plan.support_rank.ravel()[idx], [d[0] for d in diag])
    np.testing.assert_allclose(meta[idx], want_meta[idx], rtol=2e-6, atol=1e-4)
    # Inverses: Gauss-Jordan on the device, LAPACK in the oracle; both float64 rounded to float32,
    # so each row agrees to a small multiple of float32 rounding of its largest entry.
    scale = np.abs(want_inv[idx]).max(axis=1, keepdims=True) + 1e-30
    assert np.max(np.abs(inv[idx] - want_inv[idx]) / scale) < 1e-5
    contrast = plan.support_contrast.ravel()[idx]
    np.testing.assert_allclose(contrast, [d[2] for d in diag], rtol=1e-6)
    condition = plan.support_condition.ravel()[idx]
    np.testing.assert_allclose(condition[admitted], np.array([d[1] for d in diag])[admitted], rtol=1e-4)


def test_4k_planning_time():
    """Plans of the 4K plate with three holes (subset 25, step 4, two levels, median 9) and of a
    partial-policy 4K ellipse with a slit (16/8, visibility), timed."""
    from live_dic import synthetic
    from live_dic.roi import plan_pyramid

    mask = synthetic.plate_with_holes()
    plan_pyramid(mask, 2, size=25, stride=4, median_filter_size=9)  # first use: code objects load
    t0 = time.perf_counter()
    plan = plan_pyramid(mask, 2, size=25, stride=4, median_filter_size=9, temporal=True)
    full_ms = (time.perf_counter() - t0) * 1e3
    yy, xx = np.mgrid[:2160, :3840]
    ellipse = ((xx - 1920) / 1600.0) ** 2 + ((yy - 1080) / 900.0) ** 2 <= 1
    ellipse[400:1205, 1930:1932] = False
    t0 = time.perf_counter()
    part = plan_pyramid(ellipse, 2, size=16, stride=8, policy="partial", truncation="visibility")
    partial_ms = (time.perf_counter() - t0) * 1e3
    print(f"4K planning: plate S25/D4 2 levels {full_ms:.1f} ms ({plan.plans[0].selected_count} selected); "
          f"ellipse partial 16/8 visibility {partial_ms:.1f} ms ({part.plans[0].partial_indices.size} partial)")
    assert plan.plans[0].selected_count > 390_000 and part.plans[0].partial_indices.size > 100
    assert full_ms < 2000 and partial_ms < 2000
