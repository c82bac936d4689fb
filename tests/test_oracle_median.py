"""Pins the oracle's outlier pass (``oracle.median``: UOD and the recentred neighbour median,
icgn.wgsl:1087-1605) against independent restatements and the reference's behavioural tests."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

POLICY = {"flag": 0, "replace": 1, "disabled": 2, "median": 3}


def _median_cfg(side, stride, model, policy, size, scope="rejected", roi=False, seed_available=False):
    return oracle.MedianCfg(side, side * side, stride, oracle.MODELS[model], POLICY[policy], size,
                            1 if scope == "all" else 0, 1 if roi else 0, 1 if seed_available else 0)


def _valid_metadata(count, model):
    meta = np.zeros((count, 28 if model == "quadratic" else 16), np.float32)
    meta[:, 1] = 1.0
    return meta


def _recentred_samples(state, side, target, P, model, stride):
    """Every neighbour's warp moved to the target centre (the polynomial re-expansion of its
    displacement), excluding the target itself."""
    ty, tx = divmod(target, side)
    out = []
    for sy in range(side):
        for sx in range(side):
            if (sx, sy) == (tx, ty):
                continue
            q = state[sy * side + sx, :P].astype(np.float64).copy()
            ox, oy = -float(stride) * (sx - tx), -float(stride) * (sy - ty)
            if model == "quadratic":
                u = q[0] + q[1] * ox + q[2] * oy + 0.5 * q[6] * ox * ox + q[7] * ox * oy + 0.5 * q[8] * oy * oy
                ux, uy = q[1] + q[6] * ox + q[7] * oy, q[2] + q[7] * ox + q[8] * oy
                v = q[3] + q[4] * ox + q[5] * oy + 0.5 * q[9] * ox * ox + q[10] * ox * oy + 0.5 * q[11] * oy * oy
                vx, vy = q[4] + q[9] * ox + q[10] * oy, q[5] + q[10] * ox + q[11] * oy
                q[:6] = (u, ux, uy, v, vx, vy)
            elif model == "affine":
                q[0] += q[1] * ox + q[2] * oy
                q[3] += q[4] * ox + q[5] * oy
            out.append(q)
    return np.asarray(out)


@pytest.mark.parametrize("size,model", [(3, "affine"), (5, "translation"), (7, "affine"), (9, "quadratic"),
                                        (5, "quadratic")])
def test_rejected_subset_takes_recentred_neighbour_median(size, model):
    """test_gpu_regressions.py:1134-1208: a rejected centre of a size x size accepted field takes,
    per active parameter, the median of its neighbours' warps re-expanded at its own centre;
    statuses, iteration counts and every accepted subset stay untouched."""
    P = 12 if model == "quadratic" else 6
    stride_floats = 16 if model == "quadratic" else 8
    side = size
    rng = np.random.default_rng(20261102 + size)
    scale = np.full(P, 1e-3, np.float32)
    scale[[1, 2, 4, 5]] = 0.0 if model == "translation" else 1e-5
    if model == "quadratic":
        scale[6:] = 1e-7
    state = np.zeros((side * side, stride_floats), np.float32)
    state[:, :P] = rng.normal(size=(side * side, P)).astype(np.float32) * scale
    state[:, P] = 1.0
    state[:, P + 1] = 4.0
    target = side * (side // 2) + side // 2
    state[target, P] = -2.0
    active = [0, 3] if model == "translation" else list(range(P))
    want = np.median(_recentred_samples(state, side, target, P, model, 16), axis=0)
    out = oracle.median(_median_cfg(side, 16, model, "median", size), state, _valid_metadata(side * side, model))
    np.testing.assert_allclose(out[target, active], want[active], rtol=0, atol=1e-6)
    np.testing.assert_array_equal(out[:, P:], state[:, P:])
    keep = np.arange(side * side) != target
    np.testing.assert_array_equal(out[keep], state[keep])


def test_scope_all_filters_accepted_subsets_too():
    """median_filter_scope="all": every accepted subset is replaced by its neighbourhood median
    (size 3 on an affine field whose warps agree, so the median reproduces the field)."""
    side = 5
    state = np.zeros((side * side, 8), np.float32)
    gy, gx = np.divmod(np.arange(side * side), side)
    state[:, 0] = 0.5 + 1e-3 * 16 * gx
    state[:, 1] = 1e-3
    state[:, 3] = -0.25 + 2e-3 * 16 * gy
    state[:, 5] = 2e-3
    state[:, 6] = 1.0
    state[:, 7] = 5.0
    noisy = state.copy()
    noisy[12, 0] += 0.4  # one accepted outlier
    cfg = _median_cfg(side, 16, "affine", "median", 3, scope="all")
    out = oracle.median(cfg, noisy, _valid_metadata(side * side, "affine"))
    interior = [6, 7, 8, 11, 12, 13, 16, 17, 18]
    np.testing.assert_allclose(out[interior, 0], state[interior, 0], rtol=0, atol=2e-6)
    rejected_only = oracle.median(_median_cfg(side, 16, "affine", "median", 3), noisy,
                                  _valid_metadata(side * side, "affine"))
    np.testing.assert_array_equal(rejected_only, noisy)


def _fractional_texture(x, y):
    """Sum of sinusoids (test_gpu_regressions.py:67-87)."""
    v = (127.0 + 34.0 * np.sin(0.173 * x + 0.071 * y) + 29.0 * np.cos(0.043 * x - 0.191 * y)
         + 23.0 * np.sin(0.311 * x - 0.137 * y + 0.4) + 17.0 * np.cos(0.229 * x + 0.257 * y - 0.8))
    return np.clip(np.rint(v), 0, 255).astype(np.uint8)


def noisy_fractional_pair(size=256, shift=(0.3, -0.2), seed=20261102):
    y, x = np.mgrid[:size, :size].astype(np.float64)
    ref = _fractional_texture(x, y)
    tgt = _fractional_texture(x - shift[0], y - shift[1])
    rng = np.random.default_rng(seed)
    tgt = np.clip(np.rint(tgt.astype(np.float32) + rng.normal(0.0, 1.0, tgt.shape)), 0, 255).astype(np.uint8)
    return ref, tgt


def affine_band_pair(width=512, height=512, translation=0.3, slope=0.004, band=16.0):
    """A horizontal displacement with gradient `slope` inside one 16-px band centred on a subset
    column and zero outside (test_gpu_regressions.py:90-130); the inverse map by fixed point."""
    ty, tx = np.mgrid[:height, :width].astype(np.float64)
    origin = 0.5 * (16 - 1)
    centre = origin + 16.0 * round((0.5 * width - origin) / 16.0)
    left = centre - 0.5 * band
    rx = tx - translation
    for _ in range(12):
        rx = tx - translation - slope * np.clip(rx - left, 0.0, band)
    return _fractional_texture(tx, ty), _fractional_texture(rx, ty), (left, left + band)


def single_level(policy):
    """ICGNWGPU's defaults: one level, 16/16, 30 iterations, tolerance 1e-3, zncc_min 0.8."""
    return dict(pyramid_levels=1, iterations_per_level=(30,), tolerance_per_level=(1e-3,), outlier_policy=policy)


@pytest.mark.parametrize("policy", ["replace", "flag"])
def test_uod_false_positives_on_a_healthy_noisy_field(policy):
    """test_gpu_regressions.py:1210-1233: at most 1 % of the accepted subsets of a smooth
    fractional translation with 1-grey noise are flagged as outliers."""
    ref, tgt = noisy_fractional_pair()
    out = oracle.Pipeline(**single_level(policy)).set_reference(ref).register(tgt)
    status = out[..., 6]
    accepted = status >= 1.0
    assert accepted.sum() > 100
    assert np.mean(status[accepted] == 2.0) <= 0.01


def test_uod_preserves_a_localized_affine_band():
    """test_gpu_regressions.py:1235-1262: the band's gradient survives the outlier pass untouched."""
    ref, tgt, (left, right) = affine_band_pair()
    raw = oracle.Pipeline(**single_level("disabled")).set_reference(ref).register(tgt)
    filt = oracle.Pipeline(**single_level("replace")).set_reference(ref).register(tgt)
    cx = 0.5 * (16 - 1) + 16.0 * np.arange(raw.shape[1])
    band = (cx >= left) & (cx < right)
    assert band.sum() == 1
    usable = (raw[:, band, 6] >= 1.0) & (filt[:, band, 6] >= 1.0)
    assert usable.sum() > 20
    assert np.median(raw[:, band, 1][usable]) > 0.003
    np.testing.assert_allclose(filt[:, band, 1][usable], raw[:, band, 1][usable], rtol=0, atol=1e-7)
    assert not np.any(filt[:, band, 6][usable] == 2.0)


def hole_edge_field(model="quadratic"):
    """A 5x5 grid with a hole to the right of and below the centre subset: the five neighbours
    across it (column > 2 or row > 2) moved 3 px more.  The centre agrees with its own side, and
    its link word allows only the neighbours up and to the left."""
    P = 12 if model == "quadratic" else 6
    side = 5
    st = np.zeros((side * side, 16 if model == "quadratic" else 8), np.float32)
    gy, gx = np.divmod(np.arange(side * side), side)
    st[:, 0] = 0.4 + np.where((gx > 2) | (gy > 2), 3.0, 0.0)
    st[:, 3] = -0.3
    st[:, P] = 1.0
    st[:, P + 1] = 5.0
    words = np.zeros((side * side, 4), np.float32)
    words[:, 0] = sum(1 << ((dy + 1) * 3 + dx + 1) for dy in (-1, 0) for dx in (-1, 0) if dx or dy)
    return st, words, side * 2 + 2, P


@pytest.mark.parametrize("model", ["quadratic", "affine"])
@pytest.mark.parametrize("policy", ["flag", "replace"])
def test_uod_beside_a_hole_uses_only_linked_neighbours(model, policy):
    """The UOD of a subset beside a hole: with the ROI links its neighbours across the hole are
    ignored (three linked neighbours are too few to judge it) and it stays accepted and unchanged;
    without them the five across-hole neighbours outvote it.  Quadratic follows the affine rule
    (the reference defines no quadratic ROI)."""
    st, words, centre, P = hole_edge_field(model)
    meta = _valid_metadata(len(st), model)
    linked = oracle.median(_median_cfg(5, 16, model, policy, 3, roi=True), st, meta, words)
    np.testing.assert_array_equal(linked[centre], st[centre])
    unlinked = oracle.median(_median_cfg(5, 16, model, policy, 3), st, meta)
    assert unlinked[centre, P] == 2.0
