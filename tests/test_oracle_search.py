"""Pins the oracle's exhaustive integer ZNCC search (``ora_zncc_search``).

The search has no reference implementation; its nearest formulation is the
candidate ZNCC of shaders/spatial_seed.wgsl:65-190.  The checks below fix it by
an independent float64 correlation coefficient (numpy ``corrcoef``), exact
recoveries of integer shifts, the affine intensity invariance of ZNCC and the
documented edge rules.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import reference_prep as prep


def _pair(seed=0, h=72, w=90):
    rng = np.random.default_rng(seed)
    # Smooth-ish texture so that neighbouring shifts correlate (parabola defined).
    base = rng.uniform(0, 255, size=(h + 4, w + 4))
    k = np.array([1.0, 4.0, 6.0, 4.0, 1.0]) / 16.0
    base = np.apply_along_axis(lambda r: np.convolve(r, k, "valid"), 1, base)
    base = np.apply_along_axis(lambda c: np.convolve(c, k, "valid"), 0, base)
    base = (base - base.min()) / (base.max() - base.min()) * 255.0
    return np.clip(np.rint(base), 0, 255).astype(np.uint8), rng


def _brute_scores(ref, tgt, ox, oy, S, R, px=0, py=0):
    """float64 Pearson correlation of every shift, -2 where the window leaves the image."""
    h, w = tgt.shape
    f = ref[oy:oy + S, ox:ox + S].astype(np.float64).ravel()
    K = 2 * R + 1
    out = np.full((K, K), -2.0)
    for dy in range(K):
        for dx in range(K):
            x0, y0 = ox + px - R + dx, oy + py - R + dy
            if x0 < 0 or y0 < 0 or x0 + S > w or y0 + S > h:
                continue
            g = tgt[y0:y0 + S, x0:x0 + S].astype(np.float64).ravel()
            if g.std() == 0 or f.std() == 0:
                continue
            out[dy, dx] = np.corrcoef(f, g)[0, 1]
    return out


def test_scores_equal_an_independent_correlation_coefficient():
    ref, rng = _pair(1)
    tgt = np.roll(ref, (2, -3), (0, 1))
    tgt = np.clip(tgt.astype(int) + rng.integers(-6, 7, size=tgt.shape), 0, 255).astype(np.uint8)
    S, D, R = 11, 9, 4
    shift, score = oracle.zncc_search(ref, tgt, S, D, R)
    gw, gh = prep.grid_shape(ref.shape[1], ref.shape[0], S, D)
    for gy in range(gh):
        for gx in range(gw):
            brute = _brute_scores(ref, tgt, gx * D, gy * D, S, R)
            k = int(np.argmax(brute))
            # The oracle's peak value is the brute-force maximum (float32 rounding).
            assert abs(score[gy, gx] - brute.flat[k]) < 2e-6
            if np.sort(brute.ravel())[-2] < brute.flat[k] - 1e-5:
                assert np.floor(shift[gy, gx] + 0.5).tolist() == [k % (2 * R + 1) - R, k // (2 * R + 1) - R] \
                    or np.all(np.abs(shift[gy, gx] - [k % (2 * R + 1) - R, k // (2 * R + 1) - R]) <= 0.5)


def test_integer_shift_is_recovered_exactly_with_unit_score():
    ref, _ = _pair(2)
    tgt = np.roll(ref, (-5, 3), (0, 1))  # content moves by (+3, -5)
    S, D, R = 15, 8, 7
    shift, score = oracle.zncc_search(ref, tgt, S, D, R)
    gw, gh = prep.grid_shape(ref.shape[1], ref.shape[0], S, D)
    interior = 0
    for gy in range(gh):
        for gx in range(gw):
            ox, oy = gx * D, gy * D
            if ox + 3 + S <= ref.shape[1] and oy - 5 >= 0:
                interior += 1
                assert score[gy, gx] == 1.0
                # The peak is exact; the parabola of a symmetric peak adds at most a small offset.
                assert np.all(np.abs(shift[gy, gx] - [3.0, -5.0]) < 0.5)
    assert interior > 10


def test_affine_intensity_change_leaves_scores_unchanged():
    ref, rng = _pair(3)
    ref = (ref // 2).astype(np.uint8)
    tgt = np.roll(ref, (1, 2), (0, 1))
    tgt2 = 2 * tgt.astype(int) + 1
    assert tgt2.max() <= 255
    tgt2 = tgt2.astype(np.uint8)
    a = oracle.zncc_search(ref, tgt, 13, 10, 3)
    b = oracle.zncc_search(ref, tgt2, 13, 10, 3)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_prior_recentres_the_window_and_bad_priors_count_as_zero():
    ref, _ = _pair(4, 80, 96)
    tgt = np.roll(ref, (6, 9), (0, 1))
    S, D = 15, 16
    gw, gh = prep.grid_shape(96, 80, S, D)
    prior = np.zeros((gh * gw, 8), np.float32)
    prior[:, 0], prior[:, 3] = 9.2, 5.7   # rounds to (9, 6)
    shift, score = oracle.zncc_search(ref, tgt, S, D, 0, prior=prior)
    ok = score > -1.5
    assert ok.sum() > 3
    np.testing.assert_array_equal(shift[ok], np.tile([9.0, 6.0], (ok.sum(), 1)))
    prior[:, 0] = np.nan
    prior[:, 3] = 1e9
    shift, _ = oracle.zncc_search(ref, tgt, S, D, 0, prior=prior)
    assert np.all(shift == 0.0)


def test_edges_masks_and_flat_references():
    ref, _ = _pair(5, 48, 48)
    ref[:16, :16] = 100      # flat first subset
    tgt = ref.copy()
    mask = np.ones(prep.grid_shape(48, 48, 16, 16)[::-1], np.uint8)
    mask[1, 1] = 0
    shift, score = oracle.zncc_search(ref, tgt, 16, 16, 2, mask=mask)
    assert score[0, 0] == -2.0            # vf == 0
    assert np.isnan(score[1, 1]) and np.all(shift[1, 1] == 0)
    assert score[1, 0] == 1.0 and np.all(np.abs(shift[1, 0]) < 0.5)
    # Every window of the last subset leaves the image when searched 20 px to the right.
    prior = np.zeros((9, 8), np.float32)
    prior[:, 0] = 20.0
    _, score = oracle.zncc_search(ref, tgt, 16, 16, 2, prior=prior)
    assert score[0, 2] == -2.0


@pytest.mark.parametrize("true_shift", [(0.3, -0.4), (-0.25, 0.45)])
def test_parabola_refines_a_smooth_peak(true_shift):
    """Sub-pixel shifts of a smooth analytic texture are recovered to a fraction of a pixel."""
    h = w = 96
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)

    def tex(xx, yy):
        return 128 + 50 * np.sin(xx / 3.1) * np.cos(yy / 2.7) + 40 * np.sin((xx + yy) / 4.3)

    ref = np.rint(tex(x, y)).astype(np.uint8)
    tgt = np.rint(tex(x - true_shift[0], y - true_shift[1])).astype(np.uint8)
    shift, score = oracle.zncc_search(ref, tgt, 21, 12, 3)
    # Interior subsets: every neighbour of the peak is a valid window.
    shift, score = shift[1:-1, 1:-1], score[1:-1, 1:-1]
    good = score > 0.9
    err = np.abs(shift[good] - np.array(true_shift))
    integer_err = np.abs(np.rint(shift[good]) - np.array(true_shift))
    # The parabola's known bias on broad peaks stays well inside the rounding error it removes.
    assert good.sum() > 10 and float(err.max()) < 0.33
    assert float(np.median(err)) < 0.75 * float(np.median(integer_err))


# -- neighbour re-seeding ----------------------------------------------------------------------
@pytest.mark.parametrize("gw,gh,density,seed", [(40, 30, 0.02, 1), (97, 53, 0.005, 2), (64, 64, 0.3, 3), (7, 200, 0.01, 4)])
def test_jump_flood_finds_nearest_sources(gw, gh, density, seed):
    """Jump flooding against brute force: every point gets a source, and the found distance is the
    exact nearest distance for (almost) every point."""
    rng = np.random.default_rng(seed)
    src = rng.random(gw * gh) < density
    src[rng.integers(gw * gh)] = True
    near = oracle.jump_flood_nearest(src, gw, gh)
    assert np.all(near >= 0) and np.all(src[near])
    ys, xs = np.divmod(np.arange(gw * gh), gw)
    sy, sx = np.divmod(np.flatnonzero(src), gw)
    exact = ((xs[:, None] - sx[None]) ** 2 + (ys[:, None] - sy[None]) ** 2).min(1)
    found = (xs - near % gw) ** 2 + (ys - near // gw) ** 2
    assert np.all(found >= exact)
    assert np.mean(found == exact) >= 0.99
    np.testing.assert_array_equal(near[src], np.flatnonzero(src))


def test_reseed_extrapolates_accepted_neighbours_and_ignores_lone_acceptances():
    gw, gh, D = 12, 9, 8.0
    n = gw * gh
    yy, xx = np.divmod(np.arange(n), gw)
    cx, cy = 15.5 + D * xx, 15.5 + D * yy
    # An exact affine field: u = 3 + 0.01 x - 0.002 y, v = -1 + 0.003 x + 0.004 y.
    state = np.zeros((n, 8), np.float32)
    state[:, 0] = 3 + 0.01 * cx - 0.002 * cy
    state[:, 1], state[:, 2] = 0.01, -0.002
    state[:, 3] = -1 + 0.003 * cx + 0.004 * cy
    state[:, 4], state[:, 5] = 0.003, 0.004
    state[:, 6] = 1.0
    failed = (xx >= 4) & (xx < 9) & (yy >= 2) & (yy < 7)
    state[failed, 6] = 0.0
    lone = 6 * gw + 10
    state[:, 6][(yy >= 5) & (xx >= 9)] = -1.0          # a rejected corner around one lone acceptance
    state[lone, 6] = 1.0
    state[lone, 0] = 99.0
    meta = np.zeros((n, 16), np.float32)
    meta[:, 1] = 1.0
    meta[3, 1] = 0.0                                  # an invalid reference subset is never re-seeded
    state[3, 6] = 0.0
    seed, active = oracle.reseed(state, meta, gw, gh, D)
    assert active[3] == 0 and active[lone] == 0
    on = active.astype(bool)
    assert np.all(on == ((state[:, 6] < 0.5) & (np.arange(n) != 3)))
    # Every seed reproduces the exact field at its own centre, and none comes from the lone subset.
    np.testing.assert_allclose(seed[on, 0], (3 + 0.01 * cx - 0.002 * cy)[on], atol=2e-5)
    np.testing.assert_allclose(seed[on, 3], (-1 + 0.003 * cx + 0.004 * cy)[on], atol=2e-5)
    np.testing.assert_array_equal(seed[on, 1], np.float32(0.01))
