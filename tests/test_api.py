"""Host-side contract of the drop-in API: validation before any device work, result views."""

from __future__ import annotations

import numpy as np
import pytest

import live_dic
from live_dic import ICGNWGPU, ICGNWGPUFullPyramid, RegistrationResult
from live_dic.engine import _as_uint8_gray


@pytest.mark.parametrize("kwargs,message", [
    (dict(subset_size=3), "subset_size"),
    (dict(subset_size=16, subset_stride=17), "subset_stride"),
    (dict(motion_model="rigid"), "motion_model"),
    (dict(interpolation="lanczos"), "interpolation"),
    (dict(outlier_policy="drop"), "outlier_policy"),
    (dict(median_filter_size=5), "median_filter_size > 3"),
    (dict(median_filter_scope="all"), "median_filter_scope='all'"),
    (dict(spatial_weighting="tukey"), "spatial_weighting"),
    (dict(gaussian_sigma=-1.0), "gaussian_sigma"),
    (dict(device_type="tpu"), "device_type"),
    (dict(solver_workgroup_size=24), "solver_workgroup_size"),
])
def test_invalid_options_raise_before_device_creation(kwargs, message):
    for cls in (ICGNWGPU, ICGNWGPUFullPyramid):
        with pytest.raises(ValueError, match=message):
            cls(**kwargs)


def test_pyramid_schedule_validation():
    with pytest.raises(ValueError, match="pyramid_levels"):
        ICGNWGPUFullPyramid(pyramid_levels=5)
    with pytest.raises(ValueError, match="iterations_per_level"):
        ICGNWGPUFullPyramid(pyramid_levels=2, iterations_per_level=(8,))
    with pytest.raises(ValueError, match="iterations"):
        ICGNWGPUFullPyramid(pyramid_levels=2, iterations_per_level=(8, 31))
    with pytest.raises(ValueError, match="initialization"):
        ICGNWGPUFullPyramid(initialization="fourier")
    with pytest.raises(ValueError, match="initialization"):
        ICGNWGPUFullPyramid(initialization="predictive")


def test_temporal_arguments_are_rejected_before_device_creation():
    """test_temporal_init.py:15-27 of the reference."""
    for value in (None, "", "predictive", 1):
        with pytest.raises(ValueError, match="initialization"):
            ICGNWGPUFullPyramid(initialization=value)
    for value in (-1, True, 1.5, "2", 2**32):
        with pytest.raises(ValueError, match="temporal_max_age"):
            ICGNWGPUFullPyramid(temporal_max_age=value)
    with pytest.raises(ValueError, match="temporal_prediction"):
        ICGNWGPUFullPyramid(temporal_prediction="velocity")
    with pytest.raises(ValueError, match="temporal_prediction"):
        ICGNWGPUFullPyramid(initialization="temporal", temporal_prediction="sideways")


def test_public_names():
    for name in ("ICGNWGPU", "ICGNWGPUFullPyramid", "RegistrationResult", "ReferenceInfo", "ICGNError",
                 "PARAMETER_ORDER", "MOTION_MODELS", "OUTLIER_POLICIES", "SEEDED_SOLVE_API_VERSION"):
        assert hasattr(live_dic, name)
    assert live_dic.PARAMETER_ORDER == ("tx", "du_dx", "du_dy", "ty", "dv_dx", "dv_dy")
    assert live_dic.DEFAULT_FULL_PYRAMID_LEVELS == 2


def test_input_normalisation():
    rgb = np.zeros((20, 24, 3), np.uint8)
    rgb[..., 0] = 255
    gray = _as_uint8_gray(rgb)
    assert gray.shape == (20, 24) and gray.dtype == np.uint8
    assert _as_uint8_gray(np.full((16, 16), 0.5)).max() == 127
    with pytest.raises(ValueError, match="at least 16"):
        _as_uint8_gray(np.zeros((8, 32), np.uint8))
    with pytest.raises(ValueError, match=r"\[0, 255\]"):
        _as_uint8_gray(np.full((16, 16), 300, np.int32))


def test_result_views_share_one_packed_buffer(tmp_path):
    raw = bytearray(np.arange(3 * 4 * 8, dtype=np.float32).tobytes())
    packed = np.frombuffer(raw, np.float32).reshape(3, 4, 8)
    packed[..., 6] = [[1, 1, 0, -1], [2, 1, 1, -3], [1, 0, 1, 1]]
    res = RegistrationResult.from_packed(raw, grid=(4, 3), motion_model="affine", timing={},
                                         config={"motionModel": "affine"})
    assert res.parameters.shape == (3, 4, 6)
    assert res.accepted == 8 and res.replaced == 1
    assert np.may_share_memory(res.parameters, res.status)
    assert res.displacement.shape == (3, 4, 2)
    res.parameters[0, 0, 0] = 42.0
    assert packed[0, 0, 0] == 42.0
    path = tmp_path / "field.npz"
    res.save_npz(path)
    with np.load(path, allow_pickle=False) as data:
        assert data["parameters"].shape == (3, 4, 6)
        assert str(data["parameter_order"][0]) == "tx"


def test_latency_arguments_are_rejected_before_device_creation():
    """test_latency_options.py:14-35."""
    for value in (None, True, "linear", "", []):
        with pytest.raises(ValueError, match="prediction"):
            ICGNWGPUFullPyramid(temporal_prediction=value)
    with pytest.raises(ValueError, match="requires temporal"):
        ICGNWGPUFullPyramid(temporal_prediction="velocity")
    for name in ("divergence_patience", "divergence_min_iterations"):
        for value in (0, -1, True, 1.5, "3", 31):
            with pytest.raises(ValueError, match=name):
                ICGNWGPUFullPyramid(**{name: value})
    for options in ({"temporal_fine_correction": 1}, {"temporal_fine_correction": True},
                    {"temporal_fine_correction": True, "initialization": "temporal", "pyramid_levels": 3}):
        with pytest.raises(ValueError, match="temporal_fine_correction"):
            ICGNWGPUFullPyramid(**options)
    for options in ({"pyramid_policy": "unknown"}, {"pyramid_policy": "adaptive"}):
        with pytest.raises(ValueError):
            ICGNWGPUFullPyramid(**options)


def test_reseed_arguments_are_rejected_before_device_creation():
    for value in (-1, 9, True, 1.5, "2"):
        with pytest.raises(ValueError, match="reseed_passes"):
            ICGNWGPUFullPyramid(reseed_passes=value)
    with pytest.raises(ValueError, match="reseed_passes"):
        ICGNWGPUFullPyramid(reseed_passes=2, motion_model="quadratic")


def test_adaptive_pyramid_arguments_are_rejected_before_device_creation():
    """test_latency_options.py:21-35."""
    for options in ({"pyramid_policy": "unknown"}, {"pyramid_policy": "adaptive"},
                    {"pyramid_policy": "adaptive", "initialization": "temporal", "pyramid_levels": 1},
                    {"adaptive_acceptance": float("nan")}, {"adaptive_acceptance": 0},
                    {"adaptive_acceptance": 1.01}, {"adaptive_acceptance": True}):
        with pytest.raises(ValueError):
            ICGNWGPUFullPyramid(**options)


def test_field_summary_host_statistics():
    """FieldSummary.of: counts of status > 0.5 and > 1.5 and float64 sums over the accepted."""
    from live_dic.engine import FieldSummary

    rng = np.random.default_rng(8)
    packed = rng.normal(size=(7, 9, 8)).astype(np.float32)
    packed[..., 6] = rng.choice([-3.0, -1.0, 0.0, 1.0, 2.0], size=(7, 9))
    s = FieldSummary.of(packed, 6)
    ok = packed[..., 6] > 0.5
    assert s.accepted == ok.sum() and s.replaced == (packed[..., 6] > 1.5).sum()
    (mx, my), mi = s.means()
    assert mx == pytest.approx(packed[..., 0][ok].astype(np.float64).mean(), rel=1e-15)
    assert my == pytest.approx(packed[..., 3][ok].astype(np.float64).mean(), rel=1e-15)
    assert mi == pytest.approx(packed[..., 7][ok].astype(np.float64).mean(), rel=1e-15)
    assert FieldSummary(0, 0, 0.0, 0.0, 0.0).means() == (None, None)
