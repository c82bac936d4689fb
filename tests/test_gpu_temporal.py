"""GPU: temporal initialisation (full_pyramid_backend.py:433-881, full_pyramid.wgsl:437-575).

Mirrors the reference's contracts (tests/test_temporal_init.py:43-117): the first temporal frame
equals the zero start bit for bit; a stream equals sequential registration at every depth, with
and without ROI; reset and a new reference clear the history; a corrupted frame keeps it, with
age expiry; and the field matches the oracle's temporal pipeline over a sequence.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import compare_fields

pytestmark = pytest.mark.gpu


def _say(*args):
    print("[case]", *args, flush=True)


def _sequence(n=4, size=256, step=(0.3, -0.2), seed=21):
    from live_dic import synthetic

    w = h = size
    tex = synthetic.SpeckleTexture(w, h, synthetic.blobs(w, h, int(round(synthetic.BLOB_DENSITY * w * h)), seed))
    ref = tex.render(None, noise_sigma=1.0, noise_seed=1)
    targets = []
    for i in range(1, n + 1):
        fld = synthetic.Field(poly=(step[0] * i, 0.0, 0.0, step[1] * i) + (0.0,) * 8)
        targets.append(tex.render(fld, noise_sigma=1.0, noise_seed=1 + i))
    return ref, targets


@pytest.fixture(scope="module")
def seq():
    return _sequence()


def _same(a, b):
    for name in ("parameters", "status", "iterations"):
        np.testing.assert_array_equal(getattr(a, name), getattr(b, name), err_msg=name)


@pytest.mark.parametrize("model", ["translation", "affine", "quadratic"])
@pytest.mark.parametrize("levels", [1, 2, 3])
def test_first_temporal_frame_equals_zero_start(seq, model, levels):
    from live_dic import ICGNWGPUFullPyramid

    _say("first frame", model, levels)
    ref, targets = seq
    with ICGNWGPUFullPyramid(motion_model=model, pyramid_levels=levels) as zero, \
            ICGNWGPUFullPyramid(motion_model=model, pyramid_levels=levels, initialization="temporal") as temporal:
        zero.set_reference(ref)
        temporal.set_reference(ref)
        a = zero.register(targets[0])
        b = temporal.register(targets[0])
        _same(a, b)
        assert b.timing["temporallySeededCoarseSubsets"] == 0
        assert b.timing["temporalAcceptedUpdates"] > 0


@pytest.mark.parametrize("roi", [False, True])
def test_stream_equals_sequential(seq, roi):
    from live_dic import ICGNWGPUFullPyramid

    _say("stream vs sequential, roi", roi)
    ref, targets = seq
    y, x = np.indices(ref.shape)
    mask = (x > 8) & (x < 248) & (y > 8) & (y < 248)
    mask[40:70, 128] = False
    with ICGNWGPUFullPyramid(initialization="temporal") as eng:
        eng.set_reference(ref, roi_mask=mask if roi else None)
        sequential = [eng.register(t) for t in targets]
        assert sequential[-1].timing["temporallySeededCoarseSubsets"] > 0
        for depth in (1, 2, 3):
            eng.reset_temporal_state()
            streamed = list(eng.register_stream(targets, pipeline_depth=depth))
            for a, b in zip(sequential, streamed, strict=True):
                _same(a, b)
                if roi:
                    np.testing.assert_array_equal(b.status[~b.roi_selected], -3)


def test_reset_and_new_reference_clear_history(seq):
    from live_dic import ICGNWGPUFullPyramid

    _say("reset and new reference")
    ref, targets = seq
    with ICGNWGPUFullPyramid(initialization="temporal", pyramid_levels=1) as eng:
        eng.set_reference(ref)
        first = eng.register(targets[0])
        assert eng.register(targets[0]).timing["temporallySeededCoarseSubsets"] > 0
        eng.reset_temporal_state()
        _same(first, eng.register(targets[0]))
        eng.set_reference(ref)
        fresh = eng.register(targets[0])
        _same(first, fresh)
        assert fresh.timing["temporallySeededCoarseSubsets"] == 0


@pytest.mark.parametrize("max_age,expected_seed", [(None, True), (0, False), (1, True)])
def test_corrupted_frame_preserves_history_and_ages_out(seq, max_age, expected_seed):
    from live_dic import ICGNWGPUFullPyramid

    _say("corrupted frame, max_age", max_age)
    ref, targets = seq
    black = np.zeros_like(ref)
    with ICGNWGPUFullPyramid(initialization="temporal", temporal_max_age=max_age) as eng:
        eng.set_reference(ref)
        first = eng.register(targets[0])
        expected = eng.register(targets[0]) if expected_seed else first
        eng.reset_temporal_state()
        eng.register(targets[0])
        corrupted = eng.register(black)
        assert corrupted.accepted == 0
        assert corrupted.timing["temporalAcceptedUpdates"] == 0
        recovered = eng.register(targets[0])
        assert (recovered.timing["temporallySeededCoarseSubsets"] > 0) == expected_seed
        assert recovered.accepted == first.accepted
        _same(recovered, expected)


def test_temporal_sequence_matches_oracle():
    from live_dic import ICGNWGPUFullPyramid

    _say("temporal sequence vs oracle, 1024^2, 2 levels, affine")
    ref, targets = _sequence(n=4, size=1024, step=(1.7, -1.1), seed=33)
    pipe = oracle.Pipeline(pyramid_levels=2, subset_size=16, initialization="temporal")
    pipe.set_reference(ref)
    with ICGNWGPUFullPyramid(pyramid_levels=2, subset_size=16, initialization="temporal") as eng:
        eng.set_reference(ref)
        for i, t in enumerate(targets):
            res = eng.register(t)
            gpu = np.concatenate([res.parameters, res.status[..., None], res.iterations[..., None]], -1)
            ora = pipe.register(t)
            cmp = compare_fields(gpu, ora, 6)
            _say("frame", i, cmp, "counts gpu", res.timing["temporallySeededCoarseSubsets"],
                 res.timing["temporalAcceptedUpdates"], "oracle", pipe.last_counts)
            assert cmp["status_equal"] >= 0.999 and cmp["disp_within_1e3"] >= 0.999
            assert abs(res.timing["temporallySeededCoarseSubsets"] - pipe.last_counts[0]) <= 2
            assert abs(res.timing["temporalAcceptedUpdates"] - pipe.last_counts[1]) <= 0.001 * gpu[..., 0].size + 1


def test_temporal_extends_capture_range():
    """A displacement ramp that ends beyond the zero start's reach: the temporal history carries
    each frame's field to the next (the reference's tensile-sequence contract, on a synthetic ramp)."""
    from live_dic import ICGNWGPUFullPyramid

    _say("capture range, 16 frames up to 24 px translation, 1 level")
    ref, targets = _sequence(n=16, size=384, step=(1.5, -1.0), seed=5)
    with ICGNWGPUFullPyramid(pyramid_levels=1, initialization="temporal") as temporal, \
            ICGNWGPUFullPyramid(pyramid_levels=1) as zero:
        temporal.set_reference(ref)
        zero.set_reference(ref)
        for t in targets:
            a = temporal.register(t)
            b = zero.register(t)
    interior = np.zeros(a.status.shape, bool)
    interior[3:-3, 3:-3] = True
    ok_t = (a.status > 0.5) & interior
    ok_z = (b.status > 0.5) & interior
    _say("accepted temporal", ok_t.mean() / interior.mean(), "zero", ok_z.mean() / interior.mean())
    assert ok_t.sum() > 0.9 * interior.sum() and ok_z.sum() < 0.5 * interior.sum()
    err = np.hypot(a.parameters[..., 0] - 24.0, a.parameters[..., 3] + 16.0)[ok_t]
    assert np.sqrt(np.mean(err ** 2)) < 0.05


@pytest.mark.parametrize("case,levels", [("translation-l2-none-temporal", 2), ("translation-l3-none-temporal", 3)])
def test_temporal_translation_goldens(case, levels):
    """WGSL temporal goldens (before.npz): rolls of the UHD reference through one temporal engine,
    against the capture and against the oracle's temporal pipeline."""
    from helpers import GOLDEN, load_gray
    from live_dic import ICGNWGPUFullPyramid

    _say("temporal golden", case)
    ref = load_gray("uhd_reference.png")
    goldens = np.load(GOLDEN / "uhd_goldens.npz", allow_pickle=False)
    frames = [np.roll(ref, (-i, 2 * i), (0, 1)) for i in range(3)]
    # The capture's "median" policy filtered every subset (median_filter_scope="all").
    pipe = oracle.Pipeline(pyramid_levels=levels, motion_model="translation", outlier_policy="median",
                           median_filter_scope="all", initialization="temporal")
    pipe.set_reference(ref)
    with ICGNWGPUFullPyramid(pyramid_levels=levels, motion_model="translation", outlier_policy="median",
                             median_filter_scope="all", initialization="temporal") as eng:
        eng.set_reference(ref)
        results = list(eng.register_stream(frames, pipeline_depth=2))
    for frame, res in enumerate(results):
        got = np.concatenate([res.parameters, res.status[..., None], res.iterations[..., None]], -1)
        want = pipe.register(frames[frame])
        stats = compare_fields(got, want, 6)
        _say("frame", frame, "vs oracle", stats)
        assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999
        if frame:
            gold = np.concatenate([goldens[f"{case}-f{frame}-parameters"],
                                   goldens[f"{case}-f{frame}-status"][..., None],
                                   goldens[f"{case}-f{frame}-iterations"][..., None]], axis=-1)
            stats = compare_fields(got, gold, 6)
            _say("frame", frame, "vs WGSL golden", stats)
            assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999


LATENCY = dict(initialization="temporal", temporal_prediction="velocity", temporal_fine_correction=True,
               divergence_patience=3, divergence_min_iterations=4)


@pytest.mark.parametrize("case,frames,roi", [("translation-l2-none-latency", (1, 2), "none"),
                                             ("translation-l2-full-latency", (2,), "full"),
                                             ("translation-l2-partial-latency", (2,), "partial")])
def test_latency_translation_goldens(case, frames, roi):
    """WGSL latency-profile goldens (velocity prediction, fine correction, divergence patience 3
    from iteration 4): against the capture and the oracle's pipeline."""
    from helpers import GOLDEN, load_gray, oracle_rois
    from live_dic import ICGNWGPUFullPyramid
    from oracle.roi_plan import seed_permissions

    _say("latency golden", case)
    ref = load_gray("uhd_reference.png")
    goldens = np.load(GOLDEN / "uhd_goldens.npz", allow_pickle=False)
    mask = rois = links = None
    if roi != "none":
        mask = np.zeros(ref.shape, bool)
        mask[160:2000, 160:3680] = True
        mask[880:1200, 1800:2040] = False
        rois = oracle_rois(16, 16, mask, 2, policy=roi)
        links = seed_permissions(rois[0]["plan"], rois[1]["plan"], source_scale=1, destination_scale=2, size=16,
                                 stride=16, reference_mask=mask)
    frames_in = [np.roll(ref, (-i, 2 * i), (0, 1)) for i in range(3)]
    pipe = oracle.Pipeline(pyramid_levels=2, motion_model="translation", outlier_policy="median",
                           median_filter_scope="all", **LATENCY)
    pipe.set_reference(ref, rois=rois, temporal_links=links)
    with ICGNWGPUFullPyramid(pyramid_levels=2, motion_model="translation", outlier_policy="median",
                             median_filter_scope="all", **LATENCY) as eng:
        eng.set_reference(ref, **({} if mask is None else dict(roi_mask=mask, roi_policy=roi)))
        results = list(eng.register_stream(frames_in, pipeline_depth=3))
    for frame, res in enumerate(results):
        got = np.concatenate([res.parameters, res.status[..., None], res.iterations[..., None]], -1)
        want = pipe.register(frames_in[frame])
        stats = compare_fields(got, want, 6)
        _say("frame", frame, "vs oracle", stats, "predicted", res.timing["velocityPredictedSubsets"],
             pipe.last_predicted)
        assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999
        assert res.timing["velocityPredictedSubsets"] == pipe.last_predicted
        if frame in frames:
            key = f"{case}-f{frame}"
            gold = np.concatenate([goldens[f"{key}-parameters"], goldens[f"{key}-status"][..., None],
                                   goldens[f"{key}-iterations"][..., None]], axis=-1)
            stats = compare_fields(got, gold, 6)
            _say("frame", frame, "vs WGSL golden", stats)
            assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999
    assert results[-1].config["temporalPrediction"] == "velocity"
    assert results[-1].config["temporalFineCorrection"] is True
    assert results[-1].config["divergencePatience"] == 3


def test_latency_sequence_matches_oracle():
    """A 6-frame 1024^2 affine ramp with velocity prediction, fine correction and patience."""
    from live_dic import ICGNWGPUFullPyramid

    _say("latency sequence vs oracle, 1024^2, 2 levels, affine")
    ref, targets = _sequence(n=6, size=1024, step=(1.7, -1.1), seed=34)
    pipe = oracle.Pipeline(pyramid_levels=2, subset_size=16, **LATENCY)
    pipe.set_reference(ref)
    with ICGNWGPUFullPyramid(pyramid_levels=2, subset_size=16, **LATENCY) as eng:
        eng.set_reference(ref)
        for i, t in enumerate(targets):
            res = eng.register(t)
            gpu = np.concatenate([res.parameters, res.status[..., None], res.iterations[..., None]], -1)
            ora = pipe.register(t)
            cmp = compare_fields(gpu, ora, 6)
            _say("frame", i, cmp, "predicted gpu", res.timing["velocityPredictedSubsets"], "oracle",
                 pipe.last_predicted)
            assert cmp["status_equal"] >= 0.999 and cmp["disp_within_1e3"] >= 0.999
            assert abs(res.timing["velocityPredictedSubsets"] - pipe.last_predicted) <= 0.001 * gpu[..., 0].size + 1
    assert res.timing["velocityPredictedSubsets"] > 0.9 * gpu[..., 0].size


def test_prediction_requires_two_measurements_and_stops_after_corruption(seq):
    """test_latency_options.py:47-65."""
    from live_dic import ICGNWGPUFullPyramid

    _say("velocity contract")
    ref, targets = seq
    with ICGNWGPUFullPyramid(initialization="temporal", temporal_prediction="velocity",
                             outlier_policy="median") as eng:
        eng.set_reference(ref)
        first = eng.register(targets[0])
        second = eng.register(targets[1])
        assert first.timing["velocityPredictedSubsets"] == 0
        assert second.timing["velocityPredictedSubsets"] == 0
        third = eng.register(targets[2])
        expected = int(np.count_nonzero(first.accepted_mask & second.accepted_mask))
        assert third.timing["velocityPredictedSubsets"] == expected
        assert third.timing["temporalAcceptedUpdates"] == third.accepted
        black = eng.register(np.zeros_like(ref))
        assert black.timing["temporalAcceptedUpdates"] == 0
        recovered = eng.register(targets[2])
        assert recovered.timing["velocityPredictedSubsets"] == 0
        assert recovered.accepted > 0
        eng.reset_temporal_state()
        _same(first, eng.register(targets[0]))


@pytest.mark.parametrize("roi", [False, True])
def test_latency_stream_equals_sequential(seq, roi):
    """The latency profile keeps stream = sequential at depths 1-3 (test_temporal_init.py:61-78 run
    with prediction and correction)."""
    from live_dic import ICGNWGPUFullPyramid

    _say("latency stream vs sequential, roi", roi)
    ref, targets = seq
    y, x = np.indices(ref.shape)
    mask = (x > 8) & (x < 248) & (y > 8) & (y < 248)
    mask[40:70, 128] = False
    with ICGNWGPUFullPyramid(**LATENCY) as eng:
        eng.set_reference(ref, roi_mask=mask if roi else None)
        sequential = [eng.register(t) for t in targets]
        assert sequential[-1].timing["velocityPredictedSubsets"] > 0
        for depth in (1, 2, 3):
            eng.reset_temporal_state()
            for a, b in zip(sequential, eng.register_stream(targets, pipeline_depth=depth), strict=True):
                _same(a, b)


def _shift_pair(size, shift, seed):
    from live_dic import synthetic

    tex = synthetic.SpeckleTexture(size, size, synthetic.blobs(size, size, int(round(synthetic.BLOB_DENSITY * size * size)),
                                                               seed))
    ref = tex.render(None, noise_sigma=1.0, noise_seed=1)
    tgt = tex.render(synthetic.Field(poly=(shift[0], 0.0, 0.0, shift[1]) + (0.0,) * 8), noise_sigma=1.0, noise_seed=2)
    return ref, tgt


def test_finest_abort_does_not_change_the_coarse_field():
    """test_latency_options.py:68-85."""
    from live_dic import ICGNWGPUFullPyramid

    _say("abort: coarse field unchanged")
    ref, tgt = _shift_pair(256, (8.3, -3.7), 7)
    fields = []
    for patience in (None, 1):
        with ICGNWGPUFullPyramid(outlier_policy="disabled", divergence_patience=patience,
                                 divergence_min_iterations=2) as eng:
            eng.set_reference(ref)
            eng.register(tgt)
            fields.append(eng.export_level_state(1))
    np.testing.assert_array_equal(*fields)


def test_abort_preserves_early_convergence_and_matches_oracle():
    """test_latency_options.py:87-105, and the stopped field against the oracle's."""
    from live_dic import ICGNWGPUFullPyramid

    _say("abort: early convergence, disabled path, oracle")
    ref, tgt = _shift_pair(256, (4.3, -2.4), 8)
    fields = []
    for patience in (None, 30, 2):
        with ICGNWGPUFullPyramid(pyramid_levels=1, outlier_policy="disabled", divergence_patience=patience) as eng:
            eng.set_reference(ref)
            fields.append(eng.register(tgt))
    baseline, inactive, stopped = fields
    early = baseline.accepted_mask & (baseline.iterations <= 5)
    for name in ("parameters", "status", "iterations"):
        np.testing.assert_array_equal(getattr(baseline, name), getattr(inactive, name))
        np.testing.assert_array_equal(getattr(baseline, name)[early], getattr(stopped, name)[early])
    assert np.any((stopped.status == 0) & (stopped.iterations < 30))
    assert np.isfinite(stopped.parameters).all()
    want = oracle.Pipeline(pyramid_levels=1, outlier_policy="disabled", divergence_patience=2).set_reference(
        ref).register(tgt)
    got = np.concatenate([stopped.parameters, stopped.status[..., None], stopped.iterations[..., None]], -1)
    stats = compare_fields(got, want, 6)
    _say("vs oracle", stats, "iterations equal", float(np.mean(got[..., 7] == want[..., 7])))
    assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999


def test_adaptive_pyramid_switches_resets_and_is_identical_at_all_depths():
    """test_latency_options.py:107-144, and the level choices and fields against the oracle."""
    from helpers import load_gray
    from live_dic import ICGNWGPUFullPyramid
    from test_oracle_partial import fractional_shift

    _say("adaptive pyramid, 256^2 sequence")
    src = load_gray("frame_000.png")[:256, :256].copy()
    targets = [fractional_shift(src, 0.2 * i, -0.1 * i) for i in range(4)]
    sequence = [*targets, np.zeros_like(src), targets[-1], targets[-1]]
    options = dict(initialization="temporal", temporal_prediction="velocity", pyramid_policy="adaptive",
                   adaptive_acceptance=0.7, temporal_max_age=0, outlier_policy="median",
                   temporal_fine_correction=True)
    pipe = oracle.Pipeline(pyramid_levels=2, **options)
    pipe.set_reference(src)
    with ICGNWGPUFullPyramid(**options) as eng:
        eng.set_reference(src)
        expected = None
        for depth in (1, 2, 3):
            eng.reset_temporal_state()
            outputs = list(eng.register_stream(sequence, pipeline_depth=depth))
            levels = [r.timing["pyramidLevels"] for r in outputs]
            _say("depth", depth, "levels", levels, "expired", [r.timing["temporalExpiredSubsets"] for r in outputs])
            assert levels[:2] == [2, 1] and levels[5] == 2
            assert outputs[4].timing["temporalExpiredSubsets"] > 0
            assert all(r.config["pyramidLevelsUsed"] == lv for r, lv in zip(outputs, levels))
            if expected is None:
                expected = outputs
                for r, t in zip(outputs, sequence):
                    want = pipe.register(t)
                    assert pipe.last_levels == r.timing["pyramidLevels"]
                    got = np.concatenate([r.parameters, r.status[..., None], r.iterations[..., None]], -1)
                    stats = compare_fields(got, want, 6)
                    assert stats["status_equal"] >= 0.999 and stats["disp_within_1e3"] >= 0.999, stats
            for old, new in zip(expected, outputs, strict=True):
                _same(old, new)
                assert old.timing["pyramidLevels"] == new.timing["pyramidLevels"]
        eng.set_reference(src)
        assert eng.register(targets[0]).timing["pyramidLevels"] == 2
