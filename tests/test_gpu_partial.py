"""GPU: partial-support ROI (roi_policy="partial", roi.py:266-355, partial.py, icgn.wgsl:1040-1085).

The partial pass must match the oracle's partial pass (statuses identical, |du| <= 1e-3 px) with
and without overlap, uniform and Gaussian weights, in one level and a pyramid; and the reference's
contracts hold (tests/test_partial_roi.py:196-548, tests/test_uhd_fixtures.py:131-199): no -4
status leaks, an all-true mask equals no ROI, visibility truncation stops mixing across a
displacement jump, streams equal synchronous registration, exports carry the support tables, and
boundary subsets of a UHD ellipse keep metrology accuracy.
"""

from __future__ import annotations

import json

import numpy as np
import pytest

import oracle
from helpers import compare_fields, load_gray, oracle_rois
from test_oracle_partial import fractional_shift

pytestmark = pytest.mark.gpu

TRUTH = (0.35, -0.2)


def _say(*args):
    print("[case]", *args, flush=True)


@pytest.fixture(scope="module")
def speckle():
    return load_gray("frame_000.png")


def _packed(res):
    return np.concatenate([res.parameters, res.status[..., None], res.iterations[..., None]], axis=-1)


def _mask256():
    mask = np.zeros((256, 256), bool)
    mask[21:237, 19:235] = True
    mask[80:101, 112:115] = False
    return mask


@pytest.mark.parametrize("size,stride,model,interp,weighting", [
    (16, 16, "translation", "bspline", "uniform"),
    (15, 7, "translation", "cubic", "uniform"),
    (17, 8, "affine", "linear", "uniform"),
    (31, 16, "affine", "cubic", "uniform"),
    (33, 16, "affine", "bspline", "gaussian"),
])
def test_partial_pass_matches_oracle(speckle, size, stride, model, interp, weighting):
    from live_dic import ICGNWGPU

    _say("partial vs oracle", size, stride, model, interp, weighting)
    src = speckle[:256, :256].copy()
    tgt = fractional_shift(src, *TRUTH)
    mask = _mask256()
    with ICGNWGPU(subset_size=size, subset_stride=stride, motion_model=model, interpolation=interp,
                  spatial_weighting=weighting, outlier_policy="disabled") as eng:
        info = eng.set_reference(src, roi_mask=mask, roi_policy="partial", tolerance=0.005)
        res = eng.register(tgt)
    partial = np.zeros(res.status.shape, bool)
    partial.ravel()[info.roi.partial_indices] = True
    assert partial.any()
    assert not np.any(res.status == -4)
    assert np.all(res.status[~info.roi_selected] == -3)
    admitted = np.zeros(res.status.shape, bool)
    rois = oracle_rois(size, stride, mask, 1, policy="partial")
    pipe = oracle.Pipeline(pyramid_levels=1, iterations_per_level=(30,), tolerance_per_level=(0.005,),
                           motion_model=model, interpolation=interp, subset_size=size, subset_stride=stride,
                           spatial_weighting=weighting, outlier_policy="disabled")
    pipe.set_reference(src, rois=rois)
    admitted.ravel()[:] = pipe.levels[0].metadata[:, 1] == -4
    np.testing.assert_array_equal(info.reference_valid & partial, admitted)
    want = pipe.register(tgt)
    got = _packed(res)
    stats = compare_fields(got, want, 6)
    part = compare_fields(got[partial], want[partial], 6)
    _say("all", stats, "partial", part, "partial subsets", int(partial.sum()), "admitted", int(admitted.sum()))
    assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999
    assert part["status_equal"] >= 0.99 and part["disp_within_1e3"] >= 0.99
    acc = res.accepted_mask & partial
    assert acc.sum() > 0
    err = np.linalg.norm(res.displacement[acc] - TRUTH, axis=-1)
    _say("partial accepted", int(acc.sum()), "rms", float(np.sqrt(np.mean(err ** 2))))
    if size >= 15:
        assert np.sqrt(np.mean(err ** 2)) < 0.05


@pytest.mark.parametrize("model", ["translation", "affine"])
@pytest.mark.parametrize("truncation", ["none", "visibility"])
def test_pyramid_partial_matches_oracle_and_streams(speckle, model, truncation):
    """test_partial_roi.py:520-548, and the two-level pipeline against the oracle's."""
    from live_dic import ICGNWGPUFullPyramid

    _say("pyramid partial", model, truncation)
    src = speckle[:256, :256].copy()
    tgt = fractional_shift(src, *TRUTH)
    mask = np.zeros(src.shape, bool)
    mask[19:235, 21:231] = True
    mask[90:150, 111:114] = False
    with ICGNWGPUFullPyramid(motion_model=model, outlier_policy="median") as eng:
        info = eng.set_reference(src, roi_mask=mask, roi_policy="partial", roi_truncation=truncation)
        assert info.roi.partial_indices.size > 0
        sync = eng.register(tgt)
        identity = eng.register(src)
        assert sync.accepted / info.roi.selected_count > 0.9
        assert sync.rmse(TRUTH) < 0.03
        for result, expected in zip(eng.register_stream([tgt, src, tgt], pipeline_depth=3),
                                    (sync, identity, sync), strict=True):
            np.testing.assert_array_equal(result.parameters, expected.parameters)
            np.testing.assert_array_equal(result.status, expected.status)
            assert np.all(result.status[~info.roi_selected] == -3)
    rois = oracle_rois(16, 16, mask, 2, policy="partial", truncation=truncation)
    want = oracle.Pipeline(pyramid_levels=2, motion_model=model, outlier_policy="median").set_reference(
        src, rois=rois).register(tgt)
    stats = compare_fields(_packed(sync), want, 6)
    _say("vs oracle", stats)
    assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999


def test_all_true_and_empty_partial_keep_the_full_contract(speckle):
    """test_partial_roi.py:355-385."""
    from live_dic import ICGNWGPU, ICGNWGPUFullPyramid

    src = speckle[:128, :128].copy()
    tgt = fractional_shift(src, *TRUTH)
    for factory in (ICGNWGPU, ICGNWGPUFullPyramid):
        _say("all-true / empty partial", factory.__name__)
        with factory(motion_model="translation", outlier_policy="disabled") as eng:
            eng.set_reference(src)
            baseline = eng.register(tgt)
            info = eng.set_reference(src, roi_mask=np.ones(src.shape, bool), roi_policy="partial")
            res = eng.register(tgt)
            np.testing.assert_array_equal(res.parameters, baseline.parameters)
            np.testing.assert_array_equal(res.status, baseline.status)
            assert info.roi.partial_indices.size == 0
            eng.set_reference(src, roi_mask=np.zeros(src.shape, bool), roi_policy="partial")
            empty = eng.register(tgt)
            assert empty.accepted == 0


def test_visibility_limits_mixing_across_a_displacement_jump(speckle):
    """test_partial_roi.py:227-258: a one-pixel slit ending inside subset (3, 3) separates opposite
    translations; the connected support mixes them, the visible support does not."""
    from live_dic import ICGNWGPU

    truth = np.array(TRUTH)
    src = speckle[100:236, 100:236].copy()
    right = fractional_shift(src, *truth)
    left = fractional_shift(src, *(-truth))
    tgt = right.copy()
    tgt[:, :56] = left[:, :56]
    mask = np.ones(src.shape, bool)
    mask[:65, 56] = False
    for model in ("translation", "affine"):
        errors = {}
        for truncation in ("none", "visibility"):
            _say("displacement jump", model, truncation)
            with ICGNWGPU(motion_model=model, subset_size=17, interpolation="linear",
                          outlier_policy="disabled") as eng:
                eng.set_reference(src, roi_mask=mask, roi_policy="partial", roi_truncation=truncation,
                                  tolerance=0.005)
                res = eng.register(tgt)
            assert res.status[3, 3] == 1
            errors[truncation] = float(np.linalg.norm(res.displacement[3, 3] - truth))
        _say(model, errors)
        assert errors["none"] > 0.1 and errors["visibility"] < 0.05


def test_seeded_partial_solve_equals_register(speckle):
    """The partial pass runs in the seeded solve too: zero seeds everywhere reproduce register()
    bit for bit, and inactive subsets keep the status of an inactive subset."""
    from live_dic import ICGNWGPU

    _say("seeded partial")
    src = speckle[:256, :256].copy()
    tgt = fractional_shift(src, *TRUTH)
    with ICGNWGPU(outlier_policy="disabled") as eng:
        info = eng.set_reference(src, roi_mask=_mask256(), roi_policy="partial")
        res = eng.register(tgt)
        shape = res.status.shape
        seeded = eng.register_seeded(tgt, np.zeros((*shape, 6), np.float32), np.ones(shape, bool))
        np.testing.assert_array_equal(seeded.parameters, res.parameters)
        np.testing.assert_array_equal(seeded.status, res.status)
        active = np.zeros(shape, bool)
        active.ravel()[info.roi.partial_indices] = True
        only = eng.register_seeded(tgt, res.parameters, active)
        acc = only.accepted_mask & active
        assert acc.sum() > 0.9 * (info.reference_valid & active).sum()
        np.testing.assert_allclose(only.parameters[acc], res.parameters[acc], atol=2e-3, rtol=0)
        assert np.all(only.status[~active & info.roi_selected] == -1)
        with pytest.raises(ValueError, match="include_zncc"):
            eng.register(tgt, include_zncc=True)


def test_partial_exports_carry_supports_and_diagnostics(speckle, tmp_path):
    """test_partial_roi.py:395-415."""
    from live_dic import ICGNWGPUFullPyramid

    _say("exports")
    src = speckle[:128, :128].copy()
    mask = np.ones(src.shape, bool)
    mask[40:80, 45:49] = False
    with ICGNWGPUFullPyramid(motion_model="affine") as eng:
        info = eng.set_reference(src, roi_mask=mask, roi_policy="partial", roi_truncation="visibility")
        res = eng.register(src)
    assert info.roi.support_condition is not None and info.roi.support_rank is not None
    path = tmp_path / "partial.npz"
    res.save_npz(path)
    with np.load(path, allow_pickle=False) as archive:
        for key, expected in info.roi.archive_fields().items():
            np.testing.assert_array_equal(archive[key], expected)
        config = json.loads(str(archive["config_json"]))
    assert config["roi"]["policy"] == "partial" and config["roi"]["truncation"] == "visibility"
    assert config["roi"]["supportConditionLimit"] == 1e8


def _boundary_mask():
    """tests/uhd_partial_cases.py:17-21: an ellipse with a two-pixel slit."""
    y, x = np.indices((2160, 3840))
    mask = ((x - 1920) / 1600) ** 2 + ((y - 1080) / 900) ** 2 <= 1
    mask[400:1205, 1930:1932] = False
    return mask


def test_uhd_boundary_partial_supports_keep_metrology_accuracy():
    """tests/test_uhd_fixtures.py:166-199 on the archived speckle: a fractional (0.35, -0.2)
    translation through a two-level pyramid; both the partial subsets and the complete ones accept
    >= 90 % with RMS < 0.05 px and p95 < 0.10 px, and the field matches the oracle."""
    from live_dic import ICGNWGPUFullPyramid

    _say("uhd ellipse partial, translation, 2 levels, 16/16")
    ref = load_gray("uhd_reference.png")
    tgt = fractional_shift(ref, *TRUTH)
    mask = _boundary_mask()
    with ICGNWGPUFullPyramid(motion_model="translation", subset_size=16, outlier_policy="median") as eng:
        info = eng.set_reference(ref, roi_mask=mask, roi_policy="partial")
        res = eng.register(tgt)
    assert np.isfinite(res.parameters).all() and not np.any(res.status == -4)
    partial = np.zeros(res.status.shape, bool)
    partial.ravel()[info.roi.partial_indices] = True
    for name, population in (("partial", partial), ("complete", info.roi_selected & ~partial)):
        acc = population & res.accepted_mask
        err = np.linalg.norm(res.displacement[acc] - TRUTH, axis=-1)
        _say(name, int(population.sum()), "accepted", acc.sum() / population.sum(),
             "rms", float(np.sqrt(np.mean(err ** 2))), "p95", float(np.percentile(err, 95)))
        assert acc.sum() / population.sum() >= 0.90
        assert np.sqrt(np.mean(err ** 2)) < 0.05 and np.percentile(err, 95) < 0.10
    rois = oracle_rois(16, 16, mask, 2, policy="partial")
    want = oracle.Pipeline(pyramid_levels=2, motion_model="translation", outlier_policy="median").set_reference(
        ref, rois=rois).register(tgt)
    stats = compare_fields(_packed(res), want, 6)
    _say("vs oracle", stats)
    assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999
