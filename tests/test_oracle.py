"""Pins the CPU oracle against the reference's own fixtures, goldens and closed forms."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import compare_fields, load_gray
from oracle import reference_prep as prep


def test_fir_taps_sum_to_one_and_drop_negligible_tails():
    """FIR of Gaussian(sigma=1) * sqrt(3) z^|k|: unit DC gain; taps beyond 12 below 7e-8
    (test_wgpu_backend.py:800-808)."""
    w = prep.fir_weights().astype(np.float64)
    assert w.shape == (31,)
    assert abs(w.sum() - 1.0) < 1e-6
    assert np.all(np.abs(w[:3]) < 7e-8) and np.all(np.abs(w[-3:]) < 7e-8)
    np.testing.assert_allclose(w, w[::-1], rtol=0, atol=0)


def test_inverse_spline_identity():
    """B-spline coefficients interpolate: (c[i-1] + 4 c[i] + c[i+1]) / 6 = input."""
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 255, size=(64, 80))
    c = prep.bspline_coefficients(img)
    rec = (c[:, :-2] + 4 * c[:, 1:-1] + c[:, 2:]) / 6
    rec = (rec[:-2] + 4 * rec[1:-1] + rec[2:]) / 6
    np.testing.assert_allclose(rec, img[1:-1, 1:-1], atol=1e-9)


def test_prefilter_reproduces_gaussian_away_from_edges():
    """FIR coefficients (target side) evaluated at pixel centres give the Gaussian image."""
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, size=(96, 96), dtype=np.uint8)
    c = oracle.prefilter(img).astype(np.float64)
    rec = (c[:, :-2] + 4 * c[:, 1:-1] + c[:, 2:]) / 6
    rec = (rec[:-2] + 4 * rec[1:-1] + rec[2:]) / 6
    g = prep.gaussian_sigma_one(img)[1:-1, 1:-1]
    inner = (slice(20, -20), slice(20, -20))
    np.testing.assert_allclose(rec[inner], g[inner], atol=2e-3)


def test_integer_reduction_equals_float_two_pass():
    """reduce_fused (integer) equals the reference's float32 _reduce_half (full_pyramid.wgsl:219)."""
    rng = np.random.default_rng(5)
    for shape in ((37, 29), (128, 200), (255, 17)):
        img = rng.integers(0, 256, size=shape, dtype=np.uint8)
        np.testing.assert_array_equal(oracle.reduce(img), prep.reduce_half_float(img))
    img = np.full((40, 40), 255, np.uint8)
    assert np.all(oracle.reduce(img) == 255)


def test_flat_image_gives_invalid_subsets():
    """test_wgpu_backend.py:906-912: no texture -> invalid, finite zero inverse."""
    img = np.full((64, 64), 128, np.uint8)
    _, inv, meta = prep.reference_arrays(img, subset_size=16)
    assert np.all(meta[:, 1] == 0)
    assert np.all(inv == 0)


@pytest.mark.parametrize("model", ["affine", "translation", "quadratic"])
def test_hessian_inverse_identity(model):
    rng = np.random.default_rng(6)
    img = rng.integers(0, 256, size=(96, 96), dtype=np.uint8)
    size = 25 if model == "quadratic" else 16
    ref, inv, meta = prep.reference_arrays(img, motion_model=model, subset_size=size, subset_stride=8)
    P = 12 if model == "quadratic" else 6
    # Rebuild H for subset 0 in float64 from the reference plane and check H H^-1 = I.
    half = 0.5 * (size - 1)
    ly, lx = np.mgrid[:size, :size].astype(np.float64) - half
    gx = ref[:size, :size, 1].astype(np.float64)
    gy = ref[:size, :size, 2].astype(np.float64)
    comps = [gx, gx * lx, gx * ly, gy, gy * lx, gy * ly]
    if model == "quadratic":
        comps += [gx * 0.5 * lx * lx, gx * lx * ly, gx * 0.5 * ly * ly, gy * 0.5 * lx * lx, gy * lx * ly,
                  gy * 0.5 * ly * ly]
    J = np.stack([c.ravel() for c in comps], axis=1)
    if model == "translation":
        J = J[:, [0, 3]]
        Hinv = inv[0].reshape(6, 6)[np.ix_([0, 3], [0, 3])]
    else:
        Hinv = inv[0].reshape(P, P)
    prod = (J.T @ J) @ Hinv.astype(np.float64)
    np.testing.assert_allclose(prod, np.eye(prod.shape[0]), atol=2e-3)


def test_identity_registration_is_zero():
    rng = np.random.default_rng(7)
    y, x = np.mgrid[:128, :128]
    img = np.clip(127 + 40 * np.sin(0.3 * x + 0.1 * y) + 30 * np.cos(0.17 * x - 0.23 * y)
                  + rng.normal(0, 3, x.shape), 0, 255).astype(np.uint8)
    out = oracle.Pipeline(pyramid_levels=1, iterations_per_level=(30,), tolerance_per_level=(1e-4,)).set_reference(
        img).register(img)
    # Edge subsets touch the 2-pixel sampling margin and are rejected (status -1).
    interior = out[1:-1, 1:-1]
    ok = interior[..., 6] == 1
    assert ok.all()
    assert np.all(out[0, :, 6] == -1)
    assert np.abs(interior[..., :6][ok]).max() < 1e-4


def _fractional_pair(width, height, dx, dy):
    """Sum-of-sinusoids texture before and after an exact translation
    (test_gpu_regressions.py:67-87)."""
    y, x = np.mgrid[:height, :width].astype(np.float64)

    def texture(sx, sy):
        v = (127.0 + 34.0 * np.sin(0.173 * sx + 0.071 * sy) + 29.0 * np.cos(0.043 * sx - 0.191 * sy)
             + 23.0 * np.sin(0.311 * sx - 0.137 * sy + 0.4) + 17.0 * np.cos(0.229 * sx + 0.257 * sy - 0.8))
        return np.clip(np.rint(v), 0, 255).astype(np.uint8)

    return texture(x, y), texture(x - dx, y - dy)


@pytest.mark.parametrize("interp,limit", [("bspline", 0.03), ("cubic", 0.03), ("linear", 0.05)])
def test_fractional_translation_accuracy(interp, limit):
    """Accuracy gates of test_gpu_regressions.py:411-424 (128x128, truth (0.5, -0.375))."""
    ref, tgt = _fractional_pair(128, 128, 0.5, -0.375)
    out = oracle.Pipeline(pyramid_levels=1, iterations_per_level=(30,), tolerance_per_level=(1e-3,),
                          interpolation=interp).set_reference(ref).register(tgt)
    ok = out[..., 6] > 0.5
    err = np.hypot(out[..., 0][ok] - 0.5, out[..., 3][ok] + 0.375)
    assert ok.sum() > 0
    assert np.sqrt(np.mean(err**2)) < limit


def test_small_fixture_known_answers():
    """frame_000 -> frame_010, affine 16/16, B-spline, 30 iterations, tol 1e-3, replace:
    900 accepted, mean translation (1.99999318, -0.99999635), mean iterations 5.9567
    (benchmarks/results/2026-10-30-intel-arc-130v-portable.json)."""
    a = load_gray("frame_000.png")
    b = load_gray("frame_010.png")
    out = oracle.Pipeline(pyramid_levels=1, iterations_per_level=(30,), tolerance_per_level=(1e-3,)).set_reference(
        a).register(b)
    ok = out[..., 6] > 0.5
    assert ok.sum() == 900
    assert abs(out[..., 0][ok].astype(np.float64).mean() - 1.99999318) < 2e-6
    assert abs(out[..., 3][ok].astype(np.float64).mean() + 0.99999635) < 2e-6
    assert abs(out[..., 7][ok].astype(np.float64).mean() - 5.9567) < 1e-3


@pytest.fixture(scope="module")
def uhd():
    return load_gray("uhd_reference.png"), load_gray("uhd_frame_040.png"), np.load(
        oracle_goldens(), allow_pickle=False)


def oracle_goldens():
    from helpers import GOLDEN

    return GOLDEN / "uhd_goldens.npz"


def _golden(goldens, key):
    return np.concatenate([goldens[key + "-parameters"], goldens[key + "-status"][..., None],
                           goldens[key + "-iterations"][..., None]], axis=-1)


# The capture (before.npz, qualify_binding_budget.py:46-50) ran outlier_policy="median" before the
# reference added median_filter_scope; its "median" then filtered every subset, which is the
# present median_filter_scope="all".  With that scope the oracle reproduces the capture's
# parameters (below); with the present default ("rejected") it differs by the 3x3 median's
# smoothing, about 0.02 px RMS at 16/16, while statuses and iteration counts stay identical.
GOLDEN_OPTIONS = dict(outlier_policy="median", median_filter_scope="all")


def test_uhd_translation_golden(uhd):
    """WGSL golden (before.npz): translation l1, np.roll(reference, (-1, 2)) -> truth (2, -1)."""
    ref, _, goldens = uhd
    out = oracle.Pipeline(pyramid_levels=1, motion_model="translation", **GOLDEN_OPTIONS).set_reference(
        ref).register(np.roll(ref, (-1, 2), (0, 1)))
    stats = compare_fields(out, _golden(goldens, "translation-l1-none-zero-f1"), 6)
    print(stats)
    assert stats["status_equal"] == 1.0
    assert stats["disp_within_1e3"] == 1.0 and stats["param_rmse"] < 1e-5


def test_uhd_affine_golden(uhd):
    """WGSL golden affine l1 on the tensile pair: identical statuses, |du| <= 1e-3 px on >= 99.9 %
    of the accepted subsets."""
    ref, tgt, goldens = uhd
    out = oracle.Pipeline(pyramid_levels=1, **GOLDEN_OPTIONS).set_reference(ref).register(tgt)
    stats = compare_fields(out, _golden(goldens, "affine-l1-none-zero-f1"), 6)
    print(stats)
    assert stats["status_equal"] == 1.0
    assert stats["disp_within_1e3"] >= 0.999 and stats["param_rmse"] < 1e-4


def test_uhd_affine_golden_needs_the_capture_scope(uhd):
    """The present default scope ("rejected") leaves accepted subsets unfiltered: same statuses and
    iterations as the capture, parameters 0.01-0.02 px away.  Pins that the scope, not the
    solver arithmetic, separates the two."""
    ref, tgt, goldens = uhd
    out = oracle.Pipeline(pyramid_levels=1, outlier_policy="median").set_reference(ref).register(tgt)
    gold = _golden(goldens, "affine-l1-none-zero-f1")
    stats = compare_fields(out, gold, 6)
    assert stats["status_equal"] == 1.0 and np.mean(out[..., 7] == gold[..., 7]) > 0.998
    assert stats["disp_within_1e3"] < 0.2


@pytest.mark.slow
def test_uhd_quadratic_and_pyramid_goldens(uhd):
    """Affine l2/l3 at the 99.9 % bar.  Quadratic l1: statuses >= 99.9 %, but |du| <= 1e-3 on only
    98.6 %: 1.4 % of its subsets stop one iteration apart at the 5e-3 tolerance (summation order
    of the 12x12 update), and the all-subsets median spreads each to its 3x3 neighbourhood.
    Where the neighbourhood's iteration counts are identical, 99.89 % agree."""
    ref, tgt, goldens = uhd
    for key, kw, P, bar in (("affine-l2-none-zero-f1", dict(pyramid_levels=2), 6, 0.999),
                            ("affine-l3-none-zero-f1", dict(pyramid_levels=3), 6, 0.998),
                            ("quadratic-l1-none-zero-f1", dict(pyramid_levels=1, motion_model="quadratic"), 12,
                             0.985)):
        out = oracle.Pipeline(**GOLDEN_OPTIONS, **kw).set_reference(ref).register(tgt)
        gold = _golden(goldens, key)
        stats = compare_fields(out, gold, P)
        print(key, stats)
        assert stats["status_equal"] >= 0.999
        assert stats["disp_within_1e3"] >= bar
        if P == 12:
            same = out[..., 13] == gold[..., 13]
            pad = np.pad(same, 1, constant_values=True)
            calm = np.ones_like(same)
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    calm &= pad[1 + dy:1 + dy + same.shape[0], 1 + dx:1 + dx + same.shape[1]]
            both = calm & (out[..., 12] > 0.5) & (gold[..., 12] > 0.5)
            du = np.maximum(np.abs(out[..., 0] - gold[..., 0]), np.abs(out[..., 3] - gold[..., 3]))[both]
            assert np.mean(du <= 1e-3) >= 0.998


def _uhd_mask():
    """ROI of the capture (benchmarks/qualify_binding_budget.py:30-32)."""
    mask = np.zeros((2160, 3840), dtype=bool)
    mask[160:2000, 160:3680] = True
    mask[880:1200, 1800:2040] = False
    return mask


@pytest.mark.slow
@pytest.mark.parametrize("key,model,levels,bar", [("affine-l1-full-zero", "affine", 1, 0.999),
                                                  ("translation-l2-full-zero", "translation", 2, 0.999),
                                                  ("translation-l1-partial-zero", "translation", 1, 0.999),
                                                  ("affine-l2-partial-zero", "affine", 2, 0.998)])
def test_uhd_roi_goldens(uhd, key, model, levels, bar):
    """ROI goldens of the capture with the host planner's tables: the -3 pattern, statuses and
    parameters."""
    from helpers import oracle_rois

    ref, tgt, goldens = uhd
    rois = oracle_rois(16, 16, _uhd_mask(), levels, policy=key.split("-")[2])
    target = tgt if model != "translation" else np.roll(ref, (-1, 2), (0, 1))
    out = oracle.Pipeline(pyramid_levels=levels, motion_model=model, **GOLDEN_OPTIONS).set_reference(
        ref, rois=rois).register(target)
    gold = _golden(goldens, f"{key}-f1")
    stats = compare_fields(out, gold, 6)
    print(key, stats)
    assert np.array_equal(out[..., 6] == -3, gold[..., 6] == -3)
    assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= bar


@pytest.mark.slow
def test_quadratic_uhd_summation_order_sensitivity(uhd):
    """The quadratic tensile solve at the reference's 5e-3 tolerance cannot meet |du| <= 1e-3 on
    99.9 % between two legitimate summation orders: the oracle with the reference's 16 lanes against
    the same oracle with 64 lanes (a wave's strided order) stays below the bar, which is why the
    GPU-vs-oracle quadratic comparison at this tolerance is held to 99.8 %."""
    ref, tgt, _ = uhd
    runs = {lanes: oracle.Pipeline(pyramid_levels=1, motion_model="quadratic", outlier_policy="disabled",
                                   lanes=lanes).set_reference(ref).register(tgt) for lanes in (16, 64)}
    stats = compare_fields(runs[64], runs[16], 12)
    print(stats)
    assert stats["status_equal"] >= 0.999
    assert 0.997 <= stats["disp_within_1e3"] < 0.999
