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
