"""Pins the oracle's parameter values against analytic truths the reference and the mathematics fix.

- The UHD tensile pair (assets/uhd_tensile, forward map about (1920, 1080), machine_truth.csv frame
  40) under the manifest's regression_limits, and the reference's published accuracy of the same
  field (benchmarks/results/browser-gpu-reference-20261101/affine-uhd.json), which the oracle
  reproduces to four or five significant digits.
- The 4-level capture of a (16.35, -8.2) px bilinear shift (python/tests/test_uhd_fixtures.py:86-108).
- An exact model: the target is a cubic B-spline surface T of given coefficients and the reference
  is T(X + u(X)) under an affine or quadratic displacement field u, evaluated in float64, so IC-GN
  must give back u's local expansion to float32 rounding.  A dropped term, a wrong sign or a
  transposed operand in the increment, the inverse composition or the quadratic lifted operator
  moves it by far more.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import bilinear_shift, interior_centres, load_gray, tensile_truth
from oracle import reference_prep as prep


@pytest.fixture(scope="module")
def tensile():
    truth = tensile_truth()
    return load_gray("uhd_reference.png"), load_gray("uhd_frame_040.png"), truth


def _frame_040_errors(out, truth, margin):
    ex = truth["frame_040"]["stretch_x"] - 1.0
    ey = truth["frame_040"]["stretch_y"] - 1.0
    inside, x, y = interior_centres(out.shape[:2], 16, 16, 3840, 2160, margin)
    acc = inside & (out[..., 6] > 0.5)
    dist = np.hypot(out[..., 0] - ex * (x - 1920), out[..., 3] - ey * (y - 1080))[acc].astype(np.float64)
    grad = np.stack([out[..., 1] - ex, out[..., 2], out[..., 4], out[..., 5] - ey], -1)[acc].astype(np.float64)
    return inside, acc, dist, grad


@pytest.mark.slow
def test_tensile_frame_040_regression_limits(tensile):
    """test_uhd_fixtures.py:357-401: raw and median outputs meet the manifest's limits, and the
    median pass (scope "rejected") leaves statuses and accepted parameters untouched."""
    ref, tgt, truth = tensile
    lim = truth["regression_limits"]
    fields = {}
    for policy in ("disabled", "median"):
        out = oracle.Pipeline(pyramid_levels=2, outlier_policy=policy).set_reference(ref).register(tgt)
        inside, acc, dist, _ = _frame_040_errors(out, truth, lim["interior_margin_px"])
        assert acc.sum() / inside.sum() >= lim["affine_frame_040_minimum_accepted_fraction"]
        assert np.sqrt(np.mean(dist**2)) < lim["affine_frame_040_displacement_rmse_px"]
        assert np.percentile(dist, 95) < lim["affine_frame_040_displacement_p95_px"]
        fields[policy] = out
    raw, med = fields["disabled"], fields["median"]
    np.testing.assert_array_equal(med[..., 6], raw[..., 6])
    ok = med[..., 6] > 0.5
    np.testing.assert_array_equal(med[ok][:, :6], raw[ok][:, :6])


@pytest.mark.slow
def test_tensile_frame_040_matches_published_accuracy(tensile):
    """The reference's published frame-40 field (2 levels, 16/16, B-spline, median over every
    subset as its median policy then was): 31,648 accepted, 30,910 of 30,916 interior subsets,
    displacement RMSE 0.0242014 px, p95 0.04554 px, gradient RMSE per component 0.0010642, and
    the manifest's median-output gradient bound 0.002."""
    ref, tgt, truth = tensile
    pub = truth["published_frame_040_affine"]
    out = oracle.Pipeline(pyramid_levels=2, outlier_policy="median", median_filter_scope="all").set_reference(
        ref).register(tgt)
    inside, acc, dist, grad = _frame_040_errors(out, truth, 32)
    cpu = pub["analytical"]["cpu"]
    assert inside.sum() == cpu["population"]
    assert abs(int(np.count_nonzero(out[..., 6] > 0.5)) - pub["accepted"]) <= 2
    assert abs(int(acc.sum()) - cpu["accepted"]) <= 2
    assert abs(np.sqrt(np.mean(dist**2)) - cpu["displacementRmse"]) < 1e-5
    assert abs(np.percentile(dist, 95) - cpu["displacementP95"]) < 2e-5
    assert abs(np.sqrt(np.mean(grad**2)) - cpu["gradientRmse"]) < 1e-6
    per_component = np.sqrt(np.mean(grad**2, axis=0))
    assert np.all(per_component < truth["regression_limits"]["affine_frame_040_median_gradient_rmse_per_component"])


@pytest.mark.slow
@pytest.mark.parametrize("model", ["translation", "affine"])
def test_four_level_pyramid_captures_large_fractional_shift(tensile, model):
    """test_uhd_fixtures.py:86-108: (16.35, -8.2) px, 4 levels, outlier pass disabled; 192 px
    interior: >= 90 % accepted, RMSE < 0.05 px, p95 < 0.10 px."""
    ref, _, _ = tensile
    tgt = bilinear_shift(ref, 16.35, -8.2)
    out = oracle.Pipeline(pyramid_levels=4, motion_model=model, outlier_policy="disabled").set_reference(
        ref).register(tgt)
    inside, _, _ = interior_centres(out.shape[:2], 16, 16, 3840, 2160, 192)
    acc = inside & (out[..., 6] > 0.5)
    dist = np.hypot(out[..., 0] - 16.35, out[..., 3] + 8.2)[acc]
    assert acc.sum() / inside.sum() >= 0.90
    assert np.sqrt(np.mean(dist**2)) < 0.05
    assert np.percentile(dist, 95) < 0.10


def test_bilinear_shift_is_exact_for_integer_shifts_and_linear_ramps():
    """The fixture renderer: an integer shift is a mirrored roll; a linear ramp moves by dx."""
    rng = np.random.default_rng(11)
    img = rng.integers(0, 256, (40, 50), dtype=np.uint8)
    np.testing.assert_array_equal(bilinear_shift(img, 3, -2)[5:-5, 5:-5], np.roll(img, (-2, 3), (0, 1))[5:-5, 5:-5])
    ramp = np.tile(np.arange(50, dtype=np.float64) * 4.0, (40, 1)).astype(np.uint8)
    out = bilinear_shift(ramp, 0.25, 0.0).astype(np.float64)
    np.testing.assert_array_equal(out[:, 2:-2], np.rint(ramp[:, 2:-2].astype(np.float64) - 1.0))


# -- exact model ----------------------------------------------------------------------------------
_W = _H = 112
_C = 55.5


def _bspline_weights(t):
    """Cubic B-spline weights of taps floor-1 .. floor+2 at fraction t, and their derivatives."""
    w = np.stack([(1 - t) ** 3 / 6, (3 * t**3 - 6 * t**2 + 4) / 6, (-3 * t**3 + 3 * t**2 + 3 * t + 1) / 6, t**3 / 6])
    d = np.stack([-0.5 * (1 - t) ** 2, 1.5 * t**2 - 2 * t, -1.5 * t**2 + t + 0.5, 0.5 * t**2])
    return w, d


def _spline_surface(c, px, py):
    """The cubic B-spline surface of coefficients ``c`` (pixel centres at integers) and its
    gradient at (px, py), in float64."""
    fx, fy = np.floor(px), np.floor(py)
    wx, dx = _bspline_weights(px - fx)
    wy, dy = _bspline_weights(py - fy)
    ix, iy = fx.astype(np.int64) - 1, fy.astype(np.int64) - 1
    val = np.zeros(px.shape)
    gx = np.zeros(px.shape)
    gy = np.zeros(px.shape)
    for j in range(4):
        for i in range(4):
            cij = c[np.clip(iy + j, 0, c.shape[0] - 1), np.clip(ix + i, 0, c.shape[1] - 1)]
            val += cij * wx[i] * wy[j]
            gx += cij * dx[i] * wy[j]
            gy += cij * wx[i] * dy[j]
    return val, gx, gy


def _field(x, y, p):
    """Global displacement u(X) = p0 + p1 X + p2 Y + p6 X^2/2 + p7 X Y + p8 Y^2/2 (and v likewise,
    p3..p5, p9..p11) about the image centre; returns u, v and the Jacobian."""
    X, Y = x - _C, y - _C
    u = p[0] + p[1] * X + p[2] * Y + 0.5 * p[6] * X * X + p[7] * X * Y + 0.5 * p[8] * Y * Y
    v = p[3] + p[4] * X + p[5] * Y + 0.5 * p[9] * X * X + p[10] * X * Y + 0.5 * p[11] * Y * Y
    ux, uy = p[1] + p[6] * X + p[7] * Y, p[2] + p[7] * X + p[8] * Y
    vx, vy = p[4] + p[9] * X + p[10] * Y, p[5] + p[10] * X + p[11] * Y
    return u, v, ux, uy, vx, vy


_QUAD = (0.71, 0.01, -0.006, -0.55, 0.004, 0.012, 2e-4, -1.5e-4, 1e-4, -1e-4, 1.8e-4, -2.2e-4)


@pytest.mark.parametrize("model,size,stride,params,weighting", [
    ("translation", 21, 15, (1.37, 0, 0, -0.81, 0, 0, 0, 0, 0, 0, 0, 0), "uniform"),
    ("affine", 21, 15, (0.93, 0.012, -0.008, -1.21, 0.006, -0.015, 0, 0, 0, 0, 0, 0), "uniform"),
    ("affine", 31, 20, (-0.47, -0.02, 0.011, 0.66, 0.009, 0.018, 0, 0, 0, 0, 0, 0), "uniform"),
    ("quadratic", 25, 15, _QUAD, "uniform"),
    # Gaussian weights change the sums, not the solution of an exact model: this pins the
    # weighted path (weights in the preparation and in every iteration's sums).
    ("affine", 21, 15, (0.93, 0.012, -0.008, -1.21, 0.006, -0.015, 0, 0, 0, 0, 0, 0), "gaussian"),
    ("quadratic", 25, 15, _QUAD, "gaussian"),
])
def test_exact_spline_warp_is_recovered(model, size, stride, params, weighting):
    """The target is the cubic B-spline surface T of random coefficients c (so the solver's target
    plane is c itself, with no prefilter error); the reference is R(X) = T(X + u(X)) with its
    chain-rule gradient, both evaluated in float64.  IC-GN from zero must return u's local expansion
    at every interior subset centre to float32 rounding."""
    p = np.asarray(params, np.float64)
    rng = np.random.default_rng(20261204)
    coeff = rng.uniform(40.0, 215.0, size=(_H, _W)).astype(np.float32)
    y, x = np.mgrid[:_H, :_W].astype(np.float64)
    u, v, ux, uy, vx, vy = _field(x, y, p)
    r, tx, ty = _spline_surface(coeff.astype(np.float64), x + u, y + v)
    gx = tx * (1 + ux) + ty * vx
    gy = tx * uy + ty * (1 + vy)
    reference, inverse, metadata = prep.subset_arrays(r, gx, gy, motion_model=model, subset_size=size,
                                                      subset_stride=stride, spatial_weighting=weighting,
                                                      gaussian_sigma=0.25 * size)
    gw, gh = prep.grid_shape(_W, _H, size, stride)
    cfg = oracle.SolveCfg(_W, _H, gw, gh, size, stride, oracle.MODELS[model], 0, 1 if weighting == "gaussian" else 0, 0,
                          0.25 * size, 30, 2e-5, 0.8, 16, 0, 0, 4)
    state = oracle.solve(cfg, reference.reshape(-1, 4), metadata, inverse, coeff, None).reshape(gh, gw, -1)
    P = 12 if model == "quadratic" else 6
    ok = state[..., P] == 1
    assert ok[1:-1, 1:-1].all()
    cx = np.arange(gw) * stride + 0.5 * (size - 1)
    cy = np.arange(gh) * stride + 0.5 * (size - 1)
    CX, CY = np.meshgrid(cx, cy)
    u, v, ux, uy, vx, vy = _field(CX, CY, p)
    want = np.stack([u, ux, uy, v, vx, vy], -1)
    got = state[..., :6].astype(np.float64)
    print(model, "max |du|", np.abs(got[..., [0, 3]] - want[..., [0, 3]])[ok].max(),
          "max |dgrad|", np.abs(got[..., [1, 2, 4, 5]] - want[..., [1, 2, 4, 5]])[ok].max())
    if model == "affine" and np.any(p[6:]):
        pytest.fail("the affine case must have no curvature")
    assert np.abs(got[..., [0, 3]] - want[..., [0, 3]])[ok].max() < 2e-5
    assert np.abs(got[..., [1, 2, 4, 5]] - want[..., [1, 2, 4, 5]])[ok].max() < 1e-6
    if model == "quadratic":
        print("max |dcurv|", np.abs(state[..., 6:12][ok] - p[6:]).max())
        np.testing.assert_allclose(state[..., 6:12][ok], np.broadcast_to(p[6:], (ok.sum(), 6)), rtol=0, atol=1e-7)
    if model == "translation":
        assert np.all(state[..., [1, 2, 4, 5]] == 0)
