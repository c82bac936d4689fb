"""GPU: the exhaustive integer ZNCC search against the oracle, as the coarsest-level
initialiser of the pyramid engine and as the fallback of seeded registration.

The search has no reference implementation (nearest: shaders/spatial_seed.wgsl:
65-190); the oracle's ``ora_zncc_search`` is pinned on the CPU by
tests/test_oracle_search.py.  Integer moments make the scores bit-exact.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from oracle import reference_prep as prep

pytestmark = pytest.mark.gpu


def _say(*args):
    print("[case]", *args, flush=True)


def _op_search(ref, tgt, S, D, R, prior=None, mask=None):
    from live_dic import _native

    h, w = ref.shape
    gw, gh = prep.grid_shape(w, h, S, D)
    shift = np.empty((gh, gw, 2), np.float32)
    score = np.empty((gh, gw), np.float32)
    pr = None if prior is None else np.ascontiguousarray(prior, np.float32).reshape(gh * gw, -1)
    mk = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    _native.check(_native.library().ldic_op_zncc_search(
        np.ascontiguousarray(ref).ctypes.data, np.ascontiguousarray(tgt).ctypes.data, w, h, S, D, R,
        None if pr is None else pr.ctypes.data, 0 if pr is None else pr.shape[1],
        None if mk is None else mk.ctypes.data, shift.ctypes.data, score.ctypes.data))
    return shift, score


def _speckle_pair(w, h, fld, seed, noise=1.0):
    from live_dic import synthetic

    count = int(round(synthetic.BLOB_DENSITY * w * h))
    tex = synthetic.SpeckleTexture(w, h, synthetic.blobs(w, h, count, seed))
    ref = tex.render(None, noise_sigma=noise, noise_seed=seed * 2 + 1)
    tgt = tex.render(fld, noise_sigma=noise, noise_seed=seed * 2 + 2)
    return ref, tgt


@pytest.mark.parametrize("S,D,R,with_prior", [(21, 8, 6, False), (31, 8, 16, True), (15, 4, 0, True),
                                               (31, 16, 40, False), (41, 16, 9, True), (16, 7, 3, False)])
def test_op_search_is_bit_exact(S, D, R, with_prior):
    from live_dic import synthetic

    _say("zncc search op", S, D, R, with_prior)
    fld = synthetic.affine_field(240, 200, 4.6, -3.2, math.radians(2.0), 0.01, -0.01)
    ref, tgt = _speckle_pair(240, 200, fld, 7)
    gw, gh = prep.grid_shape(240, 200, S, D)
    prior = mask = None
    if with_prior:
        rng = np.random.default_rng(S + R)
        prior = np.zeros((gh * gw, 8), np.float32)
        prior[:, 0] = rng.uniform(-12, 12, gh * gw)
        prior[:, 3] = rng.uniform(-12, 12, gh * gw)
        prior[::7, 0] = np.nan
        prior[::11, 3] = 1e12
        mask = (rng.uniform(size=(gh, gw)) > 0.2).astype(np.uint8)
    g_shift, g_score = _op_search(ref, tgt, S, D, R, prior, mask)
    o_shift, o_score = oracle.zncc_search(ref, tgt, S, D, R, prior=prior, mask=mask)
    np.testing.assert_array_equal(g_score, o_score)
    np.testing.assert_array_equal(g_shift, o_shift)
    if not with_prior:
        assert np.mean(o_score > 0.5) > 0.5


def test_op_search_flat_and_border_cases():
    _say("zncc search op: flat reference, windows outside the image")
    rng = np.random.default_rng(3)
    ref = rng.integers(0, 256, size=(64, 64), dtype=np.uint8)
    ref[:16, :16] = 77
    tgt = np.roll(ref, (2, -1), (0, 1))
    prior = np.zeros((16, 8), np.float32)
    prior[:, 0] = 50.0   # pushes every window of the right-hand subsets out of the image
    for pr in (None, prior):
        g = _op_search(ref, tgt, 16, 16, 5, pr)
        o = oracle.zncc_search(ref, tgt, 16, 16, 5, prior=pr)
        np.testing.assert_array_equal(g[1], o[1])
        np.testing.assert_array_equal(g[0], o[0])
        assert g[1][0, 0] == -2.0


def test_zncc_initialisation_recovers_large_displacements():
    """A 3-level pyramid seeded by the coarse ZNCC search registers a pair whose
    displacement (up to ~80 px with 3 degrees of rotation) defeats the zero start."""
    from live_dic import ICGNWGPUFullPyramid, synthetic

    _say("pyramid with ZNCC initialisation, 1024^2, |u| <= 80 px")
    w = h = 1024
    fld = synthetic.affine_field(w, h, 37.6, -21.3, math.radians(3.0), 0.01, -0.005)
    ref, tgt = _speckle_pair(w, h, fld, 11)
    S, D = 21, 8
    X, Y = synthetic.subset_centres(w, h, S, D)
    u, v = fld.displacement(X, Y)
    # Subsets whose deformed window stays inside the rendered texture.
    inside = (X + u > 1.5 * S) & (X + u < w - 1.5 * S) & (Y + v > 1.5 * S) & (Y + v < h - 1.5 * S)

    results = {}
    for init in ("zero", "zncc"):
        with ICGNWGPUFullPyramid(pyramid_levels=3, subset_size=S, subset_stride=D, initialization=init,
                                 seed_search_radius=96) as eng:
            eng.set_reference(ref)
            res = eng.register(tgt)
            p = res.parameters
            ok = (res.status > 0.5) & inside
            err = np.hypot(p[..., 0] - u, p[..., 3] - v)[ok]
            results[init] = (ok.sum() / inside.sum(), float(np.median(err)) if err.size else np.inf,
                             float(np.mean(err < 0.05)) if err.size else 0.0)
            if init == "zncc":
                shift, score = eng.export_search(2)
                r2 = oracle.reduce(oracle.reduce(ref))
                t2 = oracle.reduce(oracle.reduce(tgt))
                o_shift, o_score = oracle.zncc_search(r2, t2, S, D, eng.coarse_search_radius)
                np.testing.assert_array_equal(score, o_score)
                np.testing.assert_array_equal(shift, o_shift)
    _say("accepted fraction / median error / fraction within 0.05 px", results)
    # Coarse subsets whose true window leaves the image find wrong peaks; seeds
    # transferred from their unconverged solves (status 0 sources take part in
    # the transfer, full_pyramid.wgsl:251-425) fail at the fine level, and a
    # handful converge to a distorted warp: hence a robust accuracy bar.
    assert results["zncc"][0] > 0.93 and results["zncc"][1] < 0.01 and results["zncc"][2] > 0.999
    assert results["zero"][0] < 0.5 * results["zncc"][0]


def test_seeded_fallback_recovers_bad_seeds():
    """Seeds 9 px off on a block of subsets fail; the ZNCC fallback re-solves them."""
    from live_dic import ICGNWGPU, synthetic

    _say("register_seeded with ZNCC fallback, 512^2")
    w = h = 512
    fld = synthetic.affine_field(w, h, 2.3, -1.7, math.radians(0.5), 0.004, 0.0)
    ref, tgt = _speckle_pair(w, h, fld, 13)
    S, D = 21, 8
    X, Y = synthetic.subset_centres(w, h, S, D)
    u, v = fld.displacement(X, Y)
    grad = fld.gradient(X, Y)
    params = np.zeros((*X.shape, 6), np.float32)
    params[..., 0], params[..., 3] = u, v
    params[..., 1], params[..., 2], params[..., 4], params[..., 5] = (grad[..., k] for k in range(4))
    bad = np.zeros(X.shape, bool)
    bad[10:30, 10:40] = True
    params[bad, 0] += 9.0
    params[bad, 3] -= 6.0
    active = np.ones(X.shape, bool)
    interior = np.zeros(X.shape, bool)
    interior[2:-2, 2:-2] = True
    with ICGNWGPU(subset_size=S, subset_stride=D, outlier_policy="disabled") as eng:
        eng.set_reference(ref, iterations=20, tolerance=1e-3)
        plain = eng.register_seeded(tgt, params, active, include_zncc=True)
        fixed = eng.register_seeded(tgt, params, active, fallback_zncc=0.9, fallback_radius=12,
                                    include_zncc=True)
        again = eng.register_seeded(tgt, params, active, include_zncc=True)   # fallback is per call
    err = lambda r: np.hypot(r.parameters[..., 0] - u, r.parameters[..., 3] - v)
    good_plain = (plain.status > 0.5) & (err(plain) < 0.05)
    good_fixed = (fixed.status > 0.5) & (err(fixed) < 0.05)
    _say("bad block recovered:", good_plain[bad].mean(), "->", good_fixed[bad].mean())
    assert good_plain[bad].mean() < 0.2
    assert good_fixed[bad & interior].mean() > 0.98
    # Subsets that converged from their own seed are untouched by the retry.
    keep = (plain.status == 1.0) & (np.nan_to_num(plain.zncc, nan=0.0) >= 0.9)
    np.testing.assert_array_equal(fixed.parameters[keep], plain.parameters[keep])
    np.testing.assert_array_equal(again.parameters, plain.parameters)


def test_neighbour_reseeding_recovers_a_failed_region_and_matches_oracle():
    """Seeds 14 px off on a block fail; two re-seeding passes restart them from their nearest
    accepted neighbours' warps.  The whole sequence (seeded solve, re-seed, preserving retry,
    twice) matches the oracle's, and subsets accepted by the first solve are untouched."""
    import oracle
    from helpers import compare_fields
    from live_dic import ICGNWGPU, synthetic

    _say("register_seeded with neighbour re-seeding, 512^2")
    w = h = 512
    fld = synthetic.affine_field(w, h, 14.0, -9.0, math.radians(1.0), 0.004, 0.0)
    ref, tgt = _speckle_pair(w, h, fld, 17)
    S, D = 21, 8
    X, Y = synthetic.subset_centres(w, h, S, D)
    u, v = fld.displacement(X, Y)
    grad = fld.gradient(X, Y)
    params = np.zeros((*X.shape, 6), np.float32)
    params[..., 0], params[..., 3] = u, v
    params[..., 1], params[..., 2], params[..., 4], params[..., 5] = (grad[..., k] for k in range(4))
    bad = np.zeros(X.shape, bool)
    bad[10:30, 10:40] = True
    params[bad] = 0.0
    active = np.ones(X.shape, bool)
    interior = np.zeros(X.shape, bool)
    interior[2:-2, 2:-2] = True
    with ICGNWGPU(subset_size=S, subset_stride=D, outlier_policy="disabled") as eng:
        eng.set_reference(ref, iterations=20, tolerance=1e-3)
        plain = eng.register_seeded(tgt, params, active)
        fixed = eng.register_seeded(tgt, params, active, reseed_passes=2)
        plane, meta, inv = eng.export_reference(0)
        coeff = eng.export_coefficients(0)
        gw, gh = eng.grid_width, eng.grid_height
    err = lambda r: np.hypot(r.parameters[..., 0] - u, r.parameters[..., 3] - v)
    good_plain = (plain.status > 0.5) & (err(plain) < 0.05)
    good_fixed = (fixed.status > 0.5) & (err(fixed) < 0.05)
    _say("bad block recovered:", good_plain[bad].mean(), "->", good_fixed[bad].mean())
    assert good_plain[bad].mean() < 0.2
    assert good_fixed[bad & interior].mean() > 0.98
    keep = plain.status == 1.0
    np.testing.assert_array_equal(fixed.parameters[keep], plain.parameters[keep])
    cfg = oracle.SolveCfg(w, h, gw, gh, S, D, 0, 0, 0, 0, S / 4, 20, 1e-3, 0.8, 16, 0)
    seeds = np.zeros((gw * gh, 8), np.float32)
    seeds[:, :6] = params.reshape(-1, 6)
    state = oracle.solve(cfg, plane, meta, inv, coeff, tgt, seeds, np.ones(gw * gh, np.uint32))
    cfg.preserve_disabled = 1
    for _ in range(2):
        seed, act = oracle.reseed(state, meta, gw, gh, D)
        oracle.solve(cfg, plane, meta, inv, coeff, tgt, seed, act, state=state)
    got = np.concatenate([fixed.parameters, fixed.status[..., None], fixed.iterations[..., None]], -1)
    stats = compare_fields(got, state.reshape(gh, gw, 8), 6)
    _say("vs oracle", stats)
    assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999
