"""Pins the oracle's seed transfer and final ZNCC against the reference's known answers and an
independent NumPy computation."""

from __future__ import annotations

import numpy as np

import oracle
from oracle import reference_prep as prep


def test_quadratic_transfer_recentres_and_rescales_curvature():
    """test_wgpu_backend.py:659-687: a 2x2 coarse grid (scale 2, 16/16) with d2u/dx2 = 0.02 seeds
    the one fine subset (scale 1) with tx 0.28125, du/dx -0.075 and d2u/dx2 0.01."""
    src = np.zeros((4, 16), np.float32)
    src[:, 6] = 0.02
    src[:, 12] = 1.0
    src[:, 13] = 1.0
    cfg = oracle.SeedCfg(2, 2, 1, 1, 2.0, 1.0, 16.0, 16.0, oracle.MODELS["quadratic"], 0)
    seed, active = oracle.transfer_seed(cfg, src)
    assert active[0] == 1
    np.testing.assert_allclose(seed[0, 0], 0.28125, atol=1e-6)
    np.testing.assert_allclose(seed[0, 1], -0.075, atol=1e-6)
    np.testing.assert_allclose(seed[0, 6], 0.01, atol=1e-6)
    np.testing.assert_array_equal(seed[0, [2, 3, 4, 5, 7, 8, 9, 10, 11]], 0.0)


def test_transfer_without_accepted_sources_gives_a_zero_seed():
    """test_wgpu_backend.py:630-657: no usable coarse source -> zero seed, still active."""
    src = np.zeros((4, 8), np.float32)
    src[:, 0] = 3.0
    src[:, 6] = -1.0
    cfg = oracle.SeedCfg(2, 2, 2, 4, 2.0, 1.0, 16.0, 16.0, oracle.MODELS["affine"], 0)
    seed, active = oracle.transfer_seed(cfg, src)
    assert np.all(active == 1)
    np.testing.assert_array_equal(seed[:, :6], 0.0)


def _seeded_pair():
    rng = np.random.default_rng(132)
    reference = rng.integers(20, 220, (128, 128), dtype=np.uint8)
    target = np.clip(np.roll(reference.astype(np.int16), (1, -2), axis=(0, 1)) + rng.integers(-8, 9, reference.shape),
                     0, 255).astype(np.uint8)
    return reference, target


def test_final_zncc_matches_independent_numpy_bilinear():
    """test_public_seeded.py:75-104: bilinear interpolation, one subset seeded at (-2, 1): the
    final ZNCC equals NumPy's Pearson coefficient of the reference window and the bilinearly
    sampled target window at the solved warp, to 2e-5."""
    reference, target = _seeded_pair()
    S = 16
    ref_plane, inverse, metadata = prep.reference_arrays(reference, interpolation="linear", subset_size=S)
    gw = gh = 8
    cfg = oracle.SolveCfg(128, 128, gw, gh, S, S, 0, oracle.INTERP["linear"], 0, 0, 4.0, 30, 1e-3, 0.65, 16, 0, 0,
                          4)
    seed = np.zeros((gw * gh, 8), np.float32)
    seed[:, 0], seed[:, 3] = -2, 1
    active = np.zeros(gw * gh, np.uint32)
    active[3 * gw + 3] = 1
    state = oracle.solve(cfg, ref_plane.reshape(-1, 4), metadata, inverse, None, target, seed, active)
    assert state[3 * gw + 3, 6] == 1
    zncc = oracle.affine_zncc(128, 128, gw, gh, S, S, oracle.INTERP["linear"], ref_plane, state, None, target)
    p = state[3 * gw + 3, :6].astype(np.float64)
    y, x = np.mgrid[48:64, 48:64].astype(np.float64)
    dx, dy = x - 55.5, y - 55.5
    xx = x + p[0] + p[1] * dx + p[2] * dy
    yy = y + p[3] + p[4] * dx + p[5] * dy
    ix, iy = np.floor(xx).astype(int), np.floor(yy).astype(int)
    fx, fy = xx - ix, yy - iy
    t = target.astype(np.float64)
    sample = ((1 - fx) * (1 - fy) * t[iy, ix] + fx * (1 - fy) * t[iy, ix + 1] + (1 - fx) * fy * t[iy + 1, ix]
              + fx * fy * t[iy + 1, ix + 1])
    want = np.corrcoef(reference[48:64, 48:64].ravel(), sample.ravel())[0, 1]
    np.testing.assert_allclose(zncc[3 * gw + 3], want, rtol=0, atol=2e-5)
    # Only the solved, accepted subset has a correlation.
    others = np.arange(gw * gh) != 3 * gw + 3
    assert np.all(np.isnan(zncc[others]))


def test_final_zncc_is_one_on_identity_for_every_interpolant():
    """test_public_seeded.py:49-73: registering an image against itself gives ZNCC 1 +- 1e-5 on
    every accepted subset (Keys cubic there; B-spline and bilinear as well)."""
    image = np.random.default_rng(137).integers(0, 256, (128, 128), dtype=np.uint8)
    for interp in ("cubic", "bspline", "linear"):
        out = oracle.Pipeline(pyramid_levels=1, interpolation=interp, outlier_policy="disabled").set_reference(
            image).register(image)
        lv = oracle.Pipeline(pyramid_levels=1, interpolation=interp).set_reference(image).levels[0]
        coeff = oracle.prefilter(image) if interp == "bspline" else None
        zncc = oracle.affine_zncc(128, 128, lv.grid_width, lv.grid_height, 16, 16, oracle.INTERP[interp],
                                  lv.reference, out.reshape(-1, 8), coeff, image)
        ok = out.reshape(-1, 8)[:, 6] == 1
        assert ok.sum() > 0
        np.testing.assert_allclose(zncc[ok], 1.0, rtol=0, atol=1e-5)
