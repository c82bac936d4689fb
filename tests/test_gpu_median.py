"""GPU outlier pass (UOD and neighbour medians, icgn.wgsl:1325-1605) against the oracle.

Injected packed states go through ``ldic_op_median`` (the pass alone, as the reference's own tests
drive its median pipeline) and through ``oracle.median``; whole registrations with the policies
flag / replace / median, median sizes 5-9 and scope "all" are compared with the oracle's pipeline.
"""

from __future__ import annotations

import ctypes
import zlib

import numpy as np
import pytest

import oracle
from helpers import compare_fields
from test_oracle_median import (_recentred_samples, _valid_metadata, affine_band_pair, noisy_fractional_pair,
                                single_level)

pytestmark = pytest.mark.gpu

POLICY = {"flag": 0, "replace": 1, "disabled": 2, "median": 3}


def _say(*args):
    print("[case]", *args, flush=True)


def _gpu_median(state, metadata, side_w, side_h, stride, model, policy, size, scope="rejected", words=None,
                seed_available=False):
    from live_dic import _native

    cfg = _native.MedianOp(side_w, side_h, stride, oracle.MODELS[model], POLICY[policy], size,
                           1 if scope == "all" else 0, 1 if seed_available else 0)
    st = np.ascontiguousarray(state, np.float32)
    meta = np.ascontiguousarray(metadata, np.float32)
    out = np.empty_like(st)
    w = None if words is None else np.ascontiguousarray(words, np.float32)
    _native.check(_native.library().ldic_op_median(ctypes.byref(cfg), st.ctypes.data, meta.ctypes.data,
                                                   None if w is None else w.ctypes.data, out.ctypes.data))
    return out


def _random_field(rng, gw, gh, model):
    P = 12 if model == "quadratic" else 6
    stride = 16 if model == "quadratic" else 8
    st = np.zeros((gw * gh, stride), np.float32)
    gy, gx = np.divmod(np.arange(gw * gh), gw)
    st[:, 0] = 0.3 + 1e-3 * 16 * gx + rng.normal(0, 0.01, gw * gh)
    st[:, 3] = -0.2 + 2e-3 * 16 * gy + rng.normal(0, 0.01, gw * gh)
    if model != "translation":
        st[:, 1], st[:, 5] = 1e-3, 2e-3
        st[:, [2, 4]] = rng.normal(0, 1e-4, (gw * gh, 2))
    if model == "quadratic":
        st[:, 6:12] = rng.normal(0, 1e-6, (gw * gh, 6))
    # Outliers and rejections.
    bad = rng.random(gw * gh) < 0.06
    st[bad, 0] += rng.choice([-1, 1], bad.sum()) * rng.uniform(0.5, 3.0, bad.sum())
    st[:, P] = rng.choice([1.0, 1.0, 1.0, 1.0, 0.0, -1.0, -2.0], gw * gh)
    st[:, P + 1] = rng.integers(1, 20, gw * gh)
    meta = np.zeros((gw * gh, 28 if model == "quadratic" else 16), np.float32)
    meta[:, 1] = rng.choice([1.0, 1.0, 1.0, 0.0, -3.0], gw * gh)
    return st, meta


CASES = [(model, policy, size, scope)
         for model in ("translation", "affine", "quadratic")
         for policy, size, scope in (("flag", 3, "rejected"), ("replace", 3, "rejected"), ("median", 3, "rejected"),
                                     ("median", 5, "rejected"), ("median", 7, "rejected"), ("median", 9, "rejected"),
                                     ("median", 3, "all"), ("median", 5, "all"), ("median", 9, "all"))]


@pytest.mark.parametrize("model,policy,size,scope", CASES)
def test_injected_field_matches_oracle(model, policy, size, scope):
    _say("median op", model, policy, size, scope)
    rng = np.random.default_rng(zlib.crc32(f"{model}-{policy}-{size}-{scope}".encode()))
    gw, gh = 23, 17
    st, meta = _random_field(rng, gw, gh, model)
    P = 12 if model == "quadratic" else 6
    for words, seed_available in ((None, False), (None, True), ("random", False)):
        w = None
        if words is not None:
            w = np.zeros((gw * gh, 4), np.float32)
            w[:, 0] = rng.integers(0, 1 << 9, gw * gh)  # 3x3 neighbour permission bits
            if size > 3:
                w[:, 1:] = rng.integers(0, 1 << 24, (gw * gh, 3))
        cfg = oracle.MedianCfg(gw, gw * gh, 16, oracle.MODELS[model], POLICY[policy], size,
                               1 if scope == "all" else 0, 0 if w is None else 1, 1 if seed_available else 0)
        want = oracle.median(cfg, st, meta, w)
        got = _gpu_median(st, meta, gw, gh, 16, model, policy, size, scope, w, seed_available)
        np.testing.assert_array_equal(got[:, P], want[:, P])
        np.testing.assert_array_equal(got[:, P + 1], want[:, P + 1])
        np.testing.assert_allclose(got[:, :P], want[:, :P], rtol=2e-6, atol=1e-7)


@pytest.mark.parametrize("size,model", [(5, "translation"), (7, "affine"), (9, "quadratic")])
def test_rejected_subset_takes_recentred_neighbour_median(size, model):
    """test_gpu_regressions.py:1134-1208 on the device."""
    P = 12 if model == "quadratic" else 6
    side = size
    rng = np.random.default_rng(20261102)
    scale = np.full(P, 1e-3, np.float32)
    scale[[1, 2, 4, 5]] = 0.0 if model == "translation" else 1e-5
    if model == "quadratic":
        scale[6:] = 1e-7
    state = np.zeros((side * side, 16 if model == "quadratic" else 8), np.float32)
    state[:, :P] = rng.normal(size=(side * side, P)).astype(np.float32) * scale
    state[:, P], state[:, P + 1] = 1.0, 4.0
    target = side * (side // 2) + side // 2
    state[target, P] = -2.0
    active = [0, 3] if model == "translation" else list(range(P))
    want = np.median(_recentred_samples(state, side, target, P, model, 16), axis=0)
    got = _gpu_median(state, _valid_metadata(side * side, model), side, side, 16, model, "median", size)
    np.testing.assert_allclose(got[target, active], want[active], rtol=0, atol=1e-6)
    keep = np.arange(side * side) != target
    np.testing.assert_array_equal(got[keep], state[keep])
    np.testing.assert_array_equal(got[:, P:P + 2], state[:, P:P + 2])


@pytest.mark.parametrize("policy", ["replace", "flag"])
def test_uod_false_positives_on_a_healthy_noisy_field(policy):
    from live_dic import ICGNWGPU

    _say("uod false positives", policy)
    ref, tgt = noisy_fractional_pair()
    with ICGNWGPU(outlier_policy=policy) as eng:
        eng.set_reference(ref)
        res = eng.register(tgt)
    accepted = res.status >= 1.0
    assert accepted.sum() > 100
    assert np.mean(res.status[accepted] == 2.0) <= 0.01
    want = oracle.Pipeline(**single_level(policy)).set_reference(ref).register(tgt)
    assert np.mean(res.status == want[..., 6]) >= 0.99


def test_uod_preserves_a_localized_affine_band():
    from live_dic import ICGNWGPU

    _say("uod band")
    ref, tgt, (left, right) = affine_band_pair()
    out = {}
    for policy in ("disabled", "replace"):
        with ICGNWGPU(outlier_policy=policy) as eng:
            eng.set_reference(ref)
            out[policy] = eng.register(tgt)
    raw, filt = out["disabled"], out["replace"]
    cx = 0.5 * (16 - 1) + 16.0 * np.arange(raw.parameters.shape[1])
    band = (cx >= left) & (cx < right)
    usable = (raw.status[:, band] >= 1.0) & (filt.status[:, band] >= 1.0)
    assert usable.sum() > 20
    assert np.median(raw.parameters[:, band, 1][usable]) > 0.003
    np.testing.assert_allclose(filt.parameters[:, band, 1][usable], raw.parameters[:, band, 1][usable], atol=1e-7)
    assert not np.any(filt.status[:, band][usable] == 2.0)


@pytest.mark.parametrize("policy,size,scope", [("flag", 3, "rejected"), ("median", 5, "rejected"),
                                               ("median", 7, "all"), ("median", 9, "rejected"),
                                               ("median", 3, "all")])
def test_registration_with_outlier_policies_matches_oracle(policy, size, scope):
    """Two-level registrations of the 512 px pair (subset 21, step 8) with each policy."""
    from live_dic import ICGNWGPUFullPyramid, synthetic

    _say("policy", policy, size, scope)
    ref, tgt, _ = synthetic.small_affine_pair()
    kw = dict(pyramid_levels=2, subset_size=21, subset_stride=8, outlier_policy=policy, median_filter_size=size,
              median_filter_scope=scope)
    with ICGNWGPUFullPyramid(**kw) as eng:
        eng.set_reference(ref)
        res = eng.register(tgt)
    got = np.concatenate([res.parameters, res.status[..., None], res.iterations[..., None]], -1)
    want = oracle.Pipeline(**kw).set_reference(ref).register(tgt)
    stats = compare_fields(got, want, 6)
    print(stats)
    assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999


@pytest.mark.parametrize("model", ["quadratic", "affine"])
@pytest.mark.parametrize("policy", ["flag", "replace"])
def test_uod_beside_a_hole_uses_only_linked_neighbours(model, policy):
    """A subset beside a hole keeps its state when the ROI links exclude the neighbours across
    the hole, for the quadratic model as for affine; without links it is outvoted."""
    from test_oracle_median import hole_edge_field

    st, words, centre, P = hole_edge_field(model)
    meta = _valid_metadata(len(st), model)
    linked = _gpu_median(st, meta, 5, 5, 16, model, policy, 3, words=words)
    np.testing.assert_array_equal(linked[centre], st[centre])
    unlinked = _gpu_median(st, meta, 5, 5, 16, model, policy, 3)
    assert unlinked[centre, P] == 2.0
