"""Partial-support ROI on the CPU: planner geometry, FP64 statistics and the oracle's partial pass.

Mirrors the reference's geometry tests (tests/test_partial_roi.py:89-197) on the oracle's
planner, pins the oracle's FP64 support statistics (a full support must give back the full-window
preparation; the thin-band admission case), and pins the oracle's partial solve against an
independent float64 matrix IC-GN.  The library computes both on the device; the GPU tests compare
them with these.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import load_gray, oracle_rois
from oracle import partial as opartial
from oracle import reference_prep as prep


def fractional_shift(source: np.ndarray, dx: float, dy: float) -> np.ndarray:
    """Bilinear sampling of source(x - dx, y - dy) with mirrored borders, rounded to uint8: a target
    whose true displacement is (dx, dy) (the rendering of tests/uhd_partial_cases.py:26-54)."""
    h, w = source.shape

    def mirror(i, n):
        i = np.abs(i) % (2 * (n - 1))
        return np.minimum(i, 2 * (n - 1) - i)

    x = np.arange(w, dtype=np.float64) - dx
    y = np.arange(h, dtype=np.float64) - dy
    ix, iy = np.floor(x).astype(int), np.floor(y).astype(int)
    fx, fy = x - ix, (y - iy)[:, None]
    x0, x1, y0, y1 = mirror(ix, w), mirror(ix + 1, w), mirror(iy, h)[:, None], mirror(iy + 1, h)[:, None]
    s = source.astype(np.float64)
    v = (1 - fy) * ((1 - fx) * s[y0, x0] + fx * s[y0, x1]) + fy * ((1 - fx) * s[y1, x0] + fx * s[y1, x1])
    return np.clip(np.rint(v), 0, 255).astype(np.uint8)


@pytest.fixture(scope="module")
def speckle():
    return load_gray("frame_000.png")


# -- geometry (test_partial_roi.py:89-156) -------------------------------------------------------
def test_support_bit_layout():
    from oracle.roi_plan import pack_support, unpack_support

    mask = np.zeros((33, 33), bool)
    mask.ravel()[[0, 31, 32, 1088]] = True
    words = pack_support(mask)
    assert words[0] == 0x80000001 and words[1] == 1 and words[-1] == 1
    np.testing.assert_array_equal(unpack_support(words, 33), mask)
    np.testing.assert_array_equal(opartial.unpack(words, 33), mask)


@pytest.mark.parametrize("size", [16, 17])
def test_central_component_and_even_centre(size):
    from oracle.roi_plan import prepare_roi, unpack_support

    mask = np.ones((size, size), bool)
    mask[:, 7] = False
    plan = prepare_roi(mask, size, size, policy="partial")
    if size == 16:
        assert not plan.selected.any()
    else:
        assert plan.selected.all()
        support = unpack_support(plan.support_words[0], size)
        assert not support[:, :8].any() and support[:, 8:].all()
        np.testing.assert_array_equal(plan.support_boxes[0], [8, 0, 16, 16])
        assert plan.coverage[0, 0] == pytest.approx(9 * 17 / 289)


def test_visibility_cuts_the_wrap_around_a_crack_tip():
    from oracle.roi_plan import prepare_roi, unpack_support

    mask = np.ones((17, 17), bool)
    mask[:14, 7] = False
    before = unpack_support(prepare_roi(mask, 17, 17, policy="partial").support_words[0], 17)
    after = unpack_support(prepare_roi(mask, 17, 17, policy="partial", truncation="visibility").support_words[0], 17)
    assert before[8, 3] and not after[8, 3] and after[8, 8]
    assert np.all(~after | before)


@pytest.mark.parametrize("size", [5, 6])
def test_visibility_paths_against_crossing_events(size):
    """Each path must be the set of closed cells the centre-to-pixel segment touches, found here by
    probing the segment at every cell-boundary crossing and between consecutive crossings."""
    from oracle.roi_plan import unpack_support, visibility_paths

    centre = (size - 1) / 2
    paths = visibility_paths(size)
    yy, xx = np.indices((size, size))
    for y in range(size):
        for x in range(size):
            delta = np.array([x - centre, y - centre])
            times = {0.0, 1.0}
            for d in delta:
                if d:
                    times.update(float(t) for t in (np.arange(size + 1) - 0.5 - centre) / d if 0 <= t <= 1)
            times = sorted(times)
            probes = times + [(a + b) / 2 for a, b in zip(times[:-1], times[1:])]
            expected = np.zeros((size, size), bool)
            for t in probes:
                q = centre + t * delta
                expected |= (np.abs(xx - q[0]) <= 0.5 + 1e-12) & (np.abs(yy - q[1]) <= 0.5 + 1e-12)
            np.testing.assert_array_equal(unpack_support(paths[y * size + x], size), expected)


def test_truncation_validation():
    from live_dic.roi import validate_roi

    mask = np.ones((32, 32), bool)
    with pytest.raises(ValueError):
        validate_roi(mask, "full", mask.shape, "affine", "visibility")
    with pytest.raises(ValueError):
        validate_roi(mask, "partial", mask.shape, "affine", "unknown")
    with pytest.raises(ValueError, match="translation and affine"):
        validate_roi(mask, "partial", mask.shape, "quadratic")


# -- FP64 statistics ---------------------------------------------------------------------------
@pytest.mark.parametrize("model", ["translation", "affine"])
@pytest.mark.parametrize("size,stride,weighting", [(16, 16, "uniform"), (17, 8, "uniform"), (17, 8, "gaussian")])
def test_full_support_reproduces_full_window_preparation(speckle, model, size, stride, weighting):
    """A support covering the whole window is the full window: energy, J^T f, sum J, weight sum
    and inverse Hessian must equal the full-window preparation (which the goldens pin)."""
    img = speckle[:96, :96]
    ref, inv, meta = prep.reference_arrays(img, interpolation="linear", motion_model=model, subset_size=size,
                                           subset_stride=stride, spatial_weighting=weighting)
    gw, gh = prep.grid_shape(96, 96, size, stride)
    from oracle.roi_plan import pack_support

    idx = np.arange(gw * gh, dtype=np.uint32)
    words = np.tile(pack_support(np.ones((size, size), bool)), (idx.size, 1))
    m2, i2 = meta.copy(), inv.copy()
    opartial.partial_reference(ref, m2, i2, idx, words, grid_width=gw, size=size, stride=stride, motion_model=model,
                               weighting=weighting)
    ok = meta[:, 1] > 0.5
    assert ok.all() and np.all(m2[:, 1] == -4.0)
    cols = [0, 2, 5, 8, 11, 15] if model == "translation" else [0] + list(range(2, 14)) + [15]
    np.testing.assert_allclose(m2[:, cols], meta[:, cols], rtol=2e-5, atol=1e-3)
    if stride != size:
        np.testing.assert_allclose(m2[:, 14], meta[:, 14], rtol=1e-6, atol=1e-4)
    else:
        assert np.abs(m2[:, 14]).max() < 1e-3          # the stored plane is already centred
    scale = np.abs(inv).max(axis=1, keepdims=True)
    assert np.abs(i2 - inv).max() <= 1e-4 * scale.max()


def test_thin_band_is_informative_for_translation_not_affine(speckle):
    """test_partial_roi.py:158-188: a one-pixel-high band supports translation only."""
    img = speckle[100:151, 100:151]
    mask = np.zeros(img.shape, bool)
    mask[25] = True
    from oracle.roi_plan import prepare_roi

    plan = prepare_roi(mask, 17, 17, policy="partial")
    assert plan.partial_indices.size == 3
    for model in ("translation", "affine"):
        ref, inv, meta = prep.reference_arrays(img, interpolation="linear", motion_model=model, subset_size=17)
        diag = opartial.partial_reference(ref, meta, inv, plan.partial_indices, plan.support_words, grid_width=3,
                                          size=17, stride=17, motion_model=model)
        rows = meta[plan.partial_indices, 1]
        if model == "translation":
            assert np.all(rows == -4)
        else:
            assert np.all(rows == 0) and all(r < 6 for r, _, _ in diag)
        assert np.isfinite(inv).all()


# -- the partial solve ---------------------------------------------------------------------------
@pytest.mark.parametrize("size,stride", [(16, 16), (17, 8)])
@pytest.mark.parametrize("model", ["translation", "affine"])
@pytest.mark.parametrize("weighting", ["uniform", "gaussian"])
def test_oracle_partial_solve_matches_double_precision_icgn(speckle, size, stride, model, weighting):
    """test_partial_roi.py:417-497: a partially covered admitted subset solved by the oracle's partial
    pass agrees with the float64 matrix IC-GN to 2e-3 in every parameter."""
    src = speckle[:128, :128]
    tgt = fractional_shift(src, 0.35, -0.2)
    y, x = np.indices(src.shape)
    mask = (x + y >= 72) & (x < 112) & (y < 112)
    rois = oracle_rois(size, stride, mask, 1, policy="partial")
    pipe = oracle.Pipeline(pyramid_levels=1, iterations_per_level=(30,), tolerance_per_level=(1e-4,),
                           motion_model=model, interpolation="linear", subset_size=size, subset_stride=stride,
                           spatial_weighting=weighting, outlier_policy="disabled")
    pipe.set_reference(src, rois=rois)
    out = pipe.register(tgt).reshape(-1, 8)
    lv = pipe.levels[0]
    plan = rois[0]["plan"]
    gw = lv.grid_width
    cover = plan.coverage.ravel()
    picks = [(k, int(i)) for k, i in enumerate(plan.partial_indices)
             if 0.5 < cover[i] < 0.95 and lv.metadata[i, 1] == -4 and (i // gw) * stride >= 2 and (i % gw) * stride >= 2]
    assert picks
    assert not np.any(out[:, 6] == -4)
    assert np.all(out[~plan.selected.ravel(), 6] == -3)
    plane = lv.reference.reshape(128, 128, 4)
    yy, xx = np.indices((size, size), dtype=np.float64)
    h = (size - 1) / 2
    weights = (np.exp(-0.5 * ((xx - h) ** 2 + (yy - h) ** 2) / (size / 4) ** 2) if weighting == "gaussian"
               else np.ones((size, size)))
    checked = 0
    for k, i in picks[:6]:
        gy, gx = divmod(i, gw)
        tile = plane[gy * stride:gy * stride + size, gx * stride:gx * stride + size].astype(np.float64)
        expected = opartial.double_icgn(tile, tgt, opartial.unpack(plan.support_words[k], size),
                                        origin=(gx * stride, gy * stride), model=model, weights=weights)
        assert out[i, 6] == 1.0
        np.testing.assert_allclose(out[i, :6], expected, atol=2e-3, rtol=0)
        checked += 1
    assert checked >= 1


def test_all_true_partial_equals_no_roi(speckle):
    """test_partial_roi.py:355-385 on the oracle: an all-true mask has no partial subsets and gives
    the field without ROI."""
    src = speckle[:128, :128]
    tgt = fractional_shift(src, 0.35, -0.2)
    rois = oracle_rois(16, 16, np.ones(src.shape, bool), 1, policy="partial")
    assert rois[0]["partial_indices"].size == 0
    a = oracle.Pipeline(pyramid_levels=1, motion_model="translation", outlier_policy="disabled")
    b = oracle.Pipeline(pyramid_levels=1, motion_model="translation", outlier_policy="disabled")
    np.testing.assert_array_equal(a.set_reference(src).register(tgt), b.set_reference(src, rois=rois).register(tgt))
