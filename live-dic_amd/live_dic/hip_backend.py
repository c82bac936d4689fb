"""Native HIP engines behind the ``live_dic`` supervisor API.

``ICGNWGPU`` keeps the name and call signatures of the reference's single-scale
engine (wgpu_backend.py:856-3099) and ``ICGNWGPUFullPyramid`` those of its
coarse-to-fine engine (full_pyramid_backend.py:158-1926).  Both drive
``liblivedic.so`` through ctypes: validation happens here, before any device
work, and every numerical step runs in the gfx950 kernels.
"""

from __future__ import annotations

import ctypes
from dataclasses import replace
import operator
import threading
import time
from collections import deque
from collections.abc import Iterable, Iterator
from functools import wraps
from typing import Any

import numpy as np

from . import _native
from .engine import FieldSummary, ICGNError, ImageInput, ReferenceInfo, RegistrationResult, _as_uint8_gray
from .interpolation import INTERPOLATION_CODES, normalize_interpolation
from .motion import MOTION_MODEL_CODES, normalize_motion_model, parameter_count, state_float_count
from .outlier import (
    MEDIAN_FILTER_SCOPE_CODES,
    OUTLIER_POLICY_CODES,
    normalize_median_filter_scope,
    normalize_median_filter_size,
    normalize_outlier_policy,
)
from .roi import immutable, plan_pyramid, validate_roi
from .subset import DEFAULT_SUBSET_SIZE, normalize_subset_size, normalize_subset_stride, subset_grid_shape
from .weighting import SPATIAL_WEIGHTING_CODES, normalize_gaussian_sigma, normalize_spatial_weighting

MAX_ITERATIONS = 30
DEFAULT_IN_FLIGHT_FRAMES = 2
DEFAULT_PYRAMID_LEVELS = 1


def _validated_config(iterations: int, tolerance: float, zncc_min: float) -> dict[str, float | int]:
    """Solver controls (wgpu_backend.py:_validated_config)."""
    count = int(iterations)
    tol = float(tolerance)
    zmin = float(zncc_min)
    if not 1 <= count <= MAX_ITERATIONS:
        raise ValueError(f"iterations must be between 1 and {MAX_ITERATIONS}")
    if not np.isfinite(tol) or tol <= 0.0:
        raise ValueError("tolerance must be a finite positive number")
    if not np.isfinite(zmin) or not 0.0 <= zmin <= 1.0:
        raise ValueError("zncc_min must be between 0 and 1")
    return {"iterations": count, "tolerance": tol, "znccMin": zmin}


def _normalize_device_type(value: str) -> str:
    if not isinstance(value, str) or value.strip().lower() not in ("auto", "gpu", "cpu"):
        raise ValueError("device_type must be 'auto', 'gpu', or 'cpu'")
    return value.strip().lower()


def _normalize_workgroup(value: int) -> int:
    try:
        size = operator.index(value)
    except TypeError as error:
        raise ValueError("solver_workgroup_size must be 16, 32, or 128") from error
    if isinstance(value, bool) or size not in (16, 32, 128):
        raise ValueError("solver_workgroup_size must be 16, 32, or 128")
    return size


def _exclusive(method):
    """Reject a second concurrent registration on one engine (wgpu_backend.py:826-838)."""

    @wraps(method)
    def guarded(self, *args, **kwargs):
        if not self._lock.acquire(blocking=False):
            raise ICGNError("Another registration or registration stream is active.")
        try:
            return method(self, *args, **kwargs)
        finally:
            self._lock.release()

    return guarded


def _exclusive_stream(method):
    """Hold the engine lock until the generator is exhausted or closed (:841-853)."""

    @wraps(method)
    def guarded(self, *args, **kwargs):
        if not self._lock.acquire(blocking=False):
            raise ICGNError("Another registration or registration stream is active.")
        try:
            yield from method(self, *args, **kwargs)
        finally:
            self._lock.release()

    return guarded


class _NativeEngine:
    """Options, the native engine handle and result assembly shared by both engines."""

    def __init__(self, *, levels: int, power_preference: str, device_type: str, solver_workgroup_size: int,
                 interpolation: str, motion_model: str, outlier_policy: str, median_filter_size: int,
                 median_filter_scope: str, subset_size: int, subset_stride: int | None,
                 spatial_weighting: str, gaussian_sigma: float | None) -> None:
        if power_preference not in ("high-performance", "low-power"):
            raise ValueError("power_preference must be 'high-performance' or 'low-power'")
        self.device_type = _normalize_device_type(device_type)
        self.power_preference = power_preference
        self.interpolation = normalize_interpolation(interpolation)
        self.motion_model = normalize_motion_model(motion_model)
        self.outlier_policy = normalize_outlier_policy(outlier_policy)
        self.median_filter_size = normalize_median_filter_size(median_filter_size)
        self.median_filter_scope = normalize_median_filter_scope(median_filter_scope)
        if self.median_filter_scope == "all" and self.outlier_policy != "median":
            raise ValueError("median_filter_scope='all' requires outlier_policy='median'")
        if self.median_filter_size > 3 and self.outlier_policy != "median":
            raise ValueError("median_filter_size > 3 requires outlier_policy='median'")
        self.subset_size = normalize_subset_size(subset_size)
        self.subset_stride = normalize_subset_stride(subset_stride, subset_size=self.subset_size)
        self.spatial_weighting = normalize_spatial_weighting(spatial_weighting)
        self.gaussian_sigma = normalize_gaussian_sigma(gaussian_sigma, subset_size=self.subset_size)
        self.solver_workgroup_size = _normalize_workgroup(solver_workgroup_size)
        self._levels_count = levels
        if self.device_type == "cpu":
            raise ICGNError("device_type='cpu' has no HIP device; the engine runs on the GPU only")
        self._lock = threading.Lock()
        self._batch_holders = 0
        self._temporal_links = None
        self.roi_planning_ms = 0.0
        self._closed = False
        self._reference_ready = False
        self._handle = ctypes.c_void_p()
        self._lib = _native.library()
        opts = _native.Options(
            MOTION_MODEL_CODES[self.motion_model], INTERPOLATION_CODES[self.interpolation], self.subset_size,
            self.subset_stride, SPATIAL_WEIGHTING_CODES[self.spatial_weighting], self.gaussian_sigma,
            OUTLIER_POLICY_CODES[self.outlier_policy], self.median_filter_size,
            MEDIAN_FILTER_SCOPE_CODES[self.median_filter_scope], levels, 0)
        _native.check(self._lib.ldic_create(ctypes.byref(opts), ctypes.byref(self._handle)))
        buf = ctypes.create_string_buffer(256)
        _native.check(self._lib.ldic_device_name(self._handle, buf, 256))
        self.device_name = buf.value.decode()
        self.execution_device_type = "gpu"
        self.adapter_info = {"device": self.device_name, "backend_type": "HIP", "adapter_type": "DiscreteGPU",
                             "vendor": "AMD"}
        self.roi = None
        self.reference_valid = None
        self._level_shapes: list[tuple[int, int, int, int]] = []
        self._roi_level_counts: list[int] = []
        self._config: dict[str, Any] = {}
        self._pyramid_config: dict[str, Any] = {}
        self._slots = 1

    # -- reference -------------------------------------------------------------------------------
    def _prepare_reference(self, image: ImageInput, configs: list[dict], roi_mask, roi_policy, roi_truncation):
        if self._closed:
            raise ICGNError("This engine is closed; create a new engine before set_reference().")
        if self._batch_holders:
            raise ICGNError("The engine belongs to an open BatchRegistration; close the batch first.")
        frame = _as_uint8_gray(image)
        height, width = frame.shape
        mask = validate_roi(roi_mask, roi_policy, frame.shape, self.motion_model, roi_truncation)
        shapes = []
        w, h = width, height
        for k in range(self._levels_count):
            if k:
                w, h = w // 2, h // 2
            gw, gh = subset_grid_shape(w, h, subset_size=self.subset_size, subset_stride=self.subset_stride)
            if gw < 1 or gh < 1:
                raise ValueError(f"Reference dimensions must each be at least {self.subset_size} pixels"
                                 + (f" at pyramid level {k}." if k else "."))
            shapes.append((w, h, gw, gh))
        pyramid = plan_pyramid(mask, self._levels_count, size=self.subset_size, stride=self.subset_stride,
                               policy=roi_policy, truncation=roi_truncation,
                               median_filter_size=self.median_filter_size if self.outlier_policy == "median" else 3,
                               device=0, temporal=getattr(self, "initialization", "") == "temporal")
        plans = pyramid.plans
        started = time.perf_counter()
        keep = []
        rois = (_native.LevelRoi * self._levels_count)()
        for k, plan in enumerate(plans):
            if plan is not None and plan.effective:
                sel = np.ascontiguousarray(plan.selected, dtype=np.uint8)
                words = np.ascontiguousarray(plan.neighbour_words(), dtype=np.float32)
                keep += [sel, words]
                rois[k].selected = sel.ctypes.data
                rois[k].neighbour_words = words.ctypes.data
            if pyramid.seed_links[k] is not None:
                links = np.ascontiguousarray(pyramid.seed_links[k], dtype=np.uint32)
                keep.append(links)
                rois[k].seed_links = links.ctypes.data
        cfgs = (_native.LevelConfig * self._levels_count)(
            *[_native.LevelConfig(c["iterations"], c["tolerance"], c["znccMin"]) for c in configs])
        valid = np.zeros((shapes[0][3], shapes[0][2]), dtype=np.uint8)
        self._reference_ready = False
        _native.check(self._lib.ldic_set_reference(self._handle, frame.ctypes.data, width, height, frame.strides[0],
                                                   cfgs, rois, valid.ctypes.data))
        self._level_shapes = shapes
        for k, plan in enumerate(plans):
            if plan is None or plan.partial_indices is None or plan.partial_indices.size == 0:
                continue
            plans[k], admitted = self._prepare_partial(k, plan)
            if k == 0:
                # reference_valid of a partial subset is its admission (wgpu_backend.py:2428-2433).
                part = np.asarray(plan.partial_indices, np.int64)
                valid.ravel()[part] = 0
                valid.ravel()[admitted] = 1
        build_ms = (time.perf_counter() - started) * 1e3 + pyramid.planning_ms
        self.width, self.height = width, height
        self.grid_width, self.grid_height = shapes[0][2], shapes[0][3]
        self.roi = plans[0]
        self._roi_plans = plans
        self._roi_mask = mask
        self._temporal_links = pyramid.temporal_links
        self.roi_planning_ms = pyramid.planning_ms
        self.reference_valid = valid.astype(bool)
        self.reference_valid.setflags(write=False)
        self._roi_level_counts = [int(p.selected.sum()) if p is not None else s[2] * s[3]
                                  for p, s in zip(plans, shapes)]
        self._slots = 1
        self._reference_ready = True
        return build_ms

    def _prepare_partial(self, level: int, plan):
        """Float64 statistics of the level's partial-support subsets, computed on the device from
        the prepared reference with their support tables (wgpu_backend.py:2412-2450).  Returns the
        plan with its support diagnostics and the flat indices of the admitted subsets."""
        indices = np.ascontiguousarray(plan.partial_indices, dtype=np.uint32)
        words = np.ascontiguousarray(plan.support_words, dtype=np.uint32)
        boxes = np.ascontiguousarray(plan.support_boxes, dtype=np.float32)
        diag = np.empty((indices.size, 4), np.float32)
        _native.check(self._lib.ldic_prepare_partial(self._handle, level, int(indices.size), indices.ctypes.data,
                                                     words.ctypes.data, boxes.ctypes.data, diag.ctypes.data))
        grid = plan.selected.shape
        rank = np.zeros(grid, np.uint32)
        condition = np.zeros(grid, np.float64)
        contrast = np.zeros(grid, np.float64)
        flat = indices.astype(np.int64)
        rank.ravel()[flat] = diag[:, 1].astype(np.uint32)
        condition.ravel()[flat] = diag[:, 2]
        contrast.ravel()[flat] = diag[:, 3]
        plan = replace(plan, support_condition=immutable(condition), support_rank=immutable(rank),
                       support_contrast=immutable(contrast))
        return plan, flat[diag[:, 0] != 0]

    def _reference_info(self, started: float, build_ms: float, config: dict) -> ReferenceInfo:
        total = (time.perf_counter() - started) * 1e3
        return ReferenceInfo(
            preparation_wall_ms=total, runtime_build_wall_ms=total - build_ms, reference_build_wall_ms=build_ms,
            encoding_wall_ms=0.0, total_wall_ms=total, config=config, device=self.device_name,
            grid=(self.grid_width, self.grid_height), width=self.width, height=self.height,
            subset_size=self.subset_size, subset_stride=self.subset_stride, roi=self.roi,
            reference_valid=self.reference_valid)

    def _option_config(self) -> dict[str, Any]:
        return {
            "interpolation": self.interpolation,
            "motionModel": self.motion_model,
            "outlierPolicy": self.outlier_policy,
            "medianFilterSize": self.median_filter_size,
            "medianFilterScope": self.median_filter_scope,
            "subsetStride": self.subset_stride,
            "spatialWeighting": self.spatial_weighting,
            "solverWorkgroupSize": self.solver_workgroup_size,
            "gaussianSigma": self.gaussian_sigma,
        }

    # -- targets ---------------------------------------------------------------------------------
    def _prepare_target(self, target: ImageInput) -> np.ndarray:
        frame = _as_uint8_gray(target)
        if frame.shape != (self.height, self.width):
            raise ValueError(f"Target dimensions must be {self.width}x{self.height}; "
                             f"received {frame.shape[1]}x{frame.shape[0]}.")
        return frame

    def _state_buffer(self) -> np.ndarray:
        """Uninitialised writable bytes for one packed field (the device fills all of it)."""
        return np.empty(self.grid_width * self.grid_height * state_float_count(self.motion_model) * 4, np.uint8)

    @staticmethod
    def _summary(t: _native.Timing) -> FieldSummary | None:
        """The device's summary of the field a call returned (None if it returned none)."""
        if t.accepted < 0:
            return None
        return FieldSummary(int(t.accepted), int(t.replaced), t.sum_translation_x, t.sum_translation_y,
                            t.sum_iterations)

    def _result(self, raw, timing: dict, config: dict, summary: FieldSummary | None = None) -> RegistrationResult:
        full_config = {
            **config,
            **self._pyramid_config,
            "width": self.width,
            "height": self.height,
            "subsetSize": self.subset_size,
            **({"roi": {**self.roi.summary(), "selectedPerLevel": self._roi_level_counts}}
               if self.roi is not None else {}),
            "deviceType": self.device_type,
            "executionDeviceType": self.execution_device_type,
            **self._option_config(),
        }
        return RegistrationResult.from_packed(
            raw, grid=(self.grid_width, self.grid_height), motion_model=self.motion_model, summary=summary,
            timing=timing,
            config=full_config, width=self.width, height=self.height, subset_size=self.subset_size,
            subset_stride=self.subset_stride, roi=self.roi, reference_valid=self.reference_valid)

    def _empty_roi_result(self, started: float, config: dict) -> RegistrationResult:
        """Finite dense field without device work when the ROI selects nothing (:2562-2572)."""
        raw = np.zeros((self.grid_height, self.grid_width, state_float_count(self.motion_model)), np.float32)
        raw[..., parameter_count(self.motion_model)] = -3.0
        result = self._result(raw.reshape(-1).view(np.uint8), {"queueSubmissions": 0.0}, config)
        elapsed = (time.perf_counter() - started) * 1e3
        result.timing = {**result.timing, "totalWallMs": elapsed, "resultAgeWallMs": elapsed}
        return result

    def _timing_dict(self, t: _native.Timing, started: float, prep_ms: float) -> dict[str, float]:
        total = (time.perf_counter() - started) * 1e3
        return {
            "inputPreparationWallMs": prep_ms,
            "uploadWallMs": t.upload_ms,
            "gpuMs": t.gpu_ms,
            "solveGpuMs": t.solve_ms,
            "completionWaitWallMs": t.readback_ms,
            "executeReadbackWallMs": t.wall_ms - t.upload_ms,
            "pipelineLatencyWallMs": total,
            "queueSubmissions": 1.0,
            "totalWallMs": total,
            "resultAgeWallMs": total,
            "temporallySeededCoarseSubsets": float(t.temporal_seeded),
            "temporalAcceptedUpdates": float(t.temporal_updates),
            **({"velocityPredictedSubsets": float(t.velocity_predicted)}
               if getattr(self, "temporal_prediction", "none") == "velocity" else {}),
            **({"temporalExpiredSubsets": float(t.temporal_expired)}
               if getattr(self, "pyramid_policy", "fixed") == "adaptive" else {}),
            **({"levelsUsed": float(t.levels_used)} if hasattr(self, "pyramid_policy") else {}),
        }

    def _run(self, frame: np.ndarray, started: float, prep_ms: float, config: dict) -> RegistrationResult:
        raw = self._state_buffer()
        out = (ctypes.c_char * len(raw)).from_buffer(raw)
        t = _native.Timing()
        _native.check(self._lib.ldic_register(self._handle, frame.ctypes.data, frame.strides[0],
                                              ctypes.addressof(out), ctypes.byref(t)))
        built = time.perf_counter()
        result = self._result(raw, self._timing_dict(t, started, prep_ms), config, self._summary(t))
        result.timing["resultBuildWallMs"] = (time.perf_counter() - built) * 1e3
        return result

    def _stream(self, targets: Iterable[ImageInput], depth: int, config: dict) -> Iterator[RegistrationResult]:
        try:
            depth = operator.index(depth)
        except TypeError as error:
            raise ValueError("pipeline_depth must be a positive integer") from error
        if depth < 1:
            raise ValueError("pipeline_depth must be a positive integer")
        if self.roi is not None and self.roi.empty:
            for target in targets:
                started = time.perf_counter()
                self._prepare_target(target)
                yield self._empty_roi_result(started, config)
            return
        if depth > self._slots:
            _native.check(self._lib.ldic_set_slots(self._handle, depth))
            self._slots = depth
        free = deque(range(depth))
        pending: deque[tuple[int, float, float]] = deque()

        def retire():
            slot, started, prep_ms = pending.popleft()
            raw = self._state_buffer()
            out = (ctypes.c_char * len(raw)).from_buffer(raw)
            t = _native.Timing()
            _native.check(self._lib.ldic_retire(self._handle, slot, ctypes.addressof(out), ctypes.byref(t)))
            free.append(slot)
            return self._result(raw, self._timing_dict(t, started, prep_ms), config, self._summary(t))

        for target in targets:
            if not free:
                yield retire()
            started = time.perf_counter()
            frame = self._prepare_target(target)
            prep_ms = (time.perf_counter() - started) * 1e3
            slot = free.popleft()
            _native.check(self._lib.ldic_submit(self._handle, slot, frame.ctypes.data, frame.strides[0]))
            pending.append((slot, started, prep_ms))
        while pending:
            yield retire()

    def _register_pair(self, reference: ImageInput, target: ImageInput, config: dict) -> RegistrationResult:
        """Replace the reference (same geometry and ROI) and register one target."""
        if not self._reference_ready:
            raise ICGNError("Call set_reference() before register_pair(); it fixes the geometry and ROI.")
        started = time.perf_counter()
        ref = self._prepare_target(reference)
        frame = self._prepare_target(target)
        prep_ms = (time.perf_counter() - started) * 1e3
        raw = self._state_buffer()
        out = (ctypes.c_char * len(raw)).from_buffer(raw)
        valid = np.zeros((self.grid_height, self.grid_width), np.uint8)
        t = _native.Timing()
        _native.check(self._lib.ldic_register_pair(self._handle, ref.ctypes.data, ref.strides[0], frame.ctypes.data,
                                                   frame.strides[0], ctypes.addressof(out), valid.ctypes.data,
                                                   ctypes.byref(t)))
        self.reference_valid = valid.astype(bool)
        self.reference_valid.setflags(write=False)
        return self._result(raw, self._timing_dict(t, started, prep_ms), config, self._summary(t))

    def _set_level_configs(self, configs: list[dict]) -> None:
        for k, c in enumerate(configs):
            cfg = _native.LevelConfig(c["iterations"], c["tolerance"], c["znccMin"])
            _native.check(self._lib.ldic_set_level_config(self._handle, k, ctypes.byref(cfg)))

    def export_reference(self, level: int = 0) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Prepared reference of one level: (H, W, 4) plane, metadata and inverse Hessians."""
        if not self._reference_ready:
            raise ICGNError("Call set_reference() first.")
        w, h, gw, gh = self._level_shapes[level]
        P = parameter_count(self.motion_model)
        M = 28 if self.motion_model == "quadratic" else 16
        ref = np.empty((h, w, 4), np.float32)
        meta = np.empty((gw * gh, M), np.float32)
        inv = np.empty((gw * gh, P * P), np.float32)
        _native.check(self._lib.ldic_export_reference(self._handle, level, ref.ctypes.data, meta.ctypes.data,
                                                      inv.ctypes.data))
        return ref, meta, inv

    def export_coefficients(self, level: int = 0, slot: int = 0) -> np.ndarray:
        """Target B-spline coefficient plane of the slot's last frame."""
        w, h, _, _ = self._level_shapes[level]
        out = np.empty((h, w), np.float32)
        _native.check(self._lib.ldic_export_coefficients(self._handle, slot, level, out.ctypes.data))
        return out

    def export_level_state(self, level: int, slot: int = 0) -> np.ndarray:
        """Output state (after the outlier pass) of one level of the slot's last frame."""
        _, _, gw, gh = self._level_shapes[level]
        out = np.empty((gh, gw, state_float_count(self.motion_model)), np.float32)
        _native.check(self._lib.ldic_export_level_state(self._handle, slot, level, out.ctypes.data))
        return out

    def _set_initialization(self, kind: int, search_radius: int = 0, fallback_zncc: float | None = None,
                            fallback_radius: int = 0, temporal_max_age: int | None = None,
                            temporal_prediction: int = 0, temporal_fine_correction: int = 0,
                            divergence_patience: int = 0, divergence_min_iterations: int = 4,
                            adaptive_pyramid: int = 0, adaptive_acceptance: float = 0.95,
                            reseed_passes: int = 0) -> None:
        init = _native.Initialization(kind, int(search_radius), -2.0 if fallback_zncc is None else float(fallback_zncc),
                                      int(fallback_radius),
                                      0xFFFFFFFE if temporal_max_age is None else int(temporal_max_age),
                                      int(temporal_prediction), int(temporal_fine_correction),
                                      int(divergence_patience), int(divergence_min_iterations),
                                      int(adaptive_pyramid), float(adaptive_acceptance), int(reseed_passes))
        _native.check(self._lib.ldic_set_initialization(self._handle, ctypes.byref(init)))
        self._initialization = init

    def export_search(self, level: int, slot: int = 0) -> tuple[np.ndarray, np.ndarray]:
        """Shift (gh, gw, 2) and ZNCC score (gh, gw) of the slot's last ZNCC search on one level."""
        _, _, gw, gh = self._level_shapes[level]
        shift = np.empty((gh, gw, 2), np.float32)
        score = np.empty((gh, gw), np.float32)
        _native.check(self._lib.ldic_export_search(self._handle, slot, level, shift.ctypes.data, score.ctypes.data))
        return shift, score

    def close(self) -> None:
        """Release device resources permanently; repeated calls are harmless."""
        if self._closed:
            return
        if self._batch_holders:
            raise ICGNError("The engine belongs to an open BatchRegistration; close the batch first.")
        self._reference_ready = False
        self._closed = True
        if self._handle:
            self._lib.ldic_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *_: object) -> None:
        self.close()

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass


class ICGNWGPU(_NativeEngine):
    """Single-scale IC-GN engine (wgpu_backend.py:856-3099).

    Supports ``register``, ``register_stream``, ``register_many`` and the
    version-1 seeded solve ``register_seeded`` with optional final ZNCC.
    ``pyramid_levels`` above one (the reference's selective CPU-pyramid
    recovery) is not built; ``ICGNWGPUFullPyramid`` is the pyramid engine.
    """

    def __init__(self, *, power_preference: str = "high-performance", device_type: str = "auto",
                 prefer_subgroups: bool | None = None, solver_workgroup_size: int = 16,
                 pyramid_levels: int = DEFAULT_PYRAMID_LEVELS, interpolation: str = "bspline",
                 motion_model: str = "affine", outlier_policy: str = "replace", median_filter_size: int = 3,
                 median_filter_scope: str = "rejected", subset_size: int = DEFAULT_SUBSET_SIZE,
                 subset_stride: int | None = None, spatial_weighting: str = "uniform",
                 gaussian_sigma: float | None = None) -> None:
        try:
            levels = operator.index(pyramid_levels)
        except TypeError as error:
            raise ValueError("pyramid_levels must be an integer from 1 to 3") from error
        if isinstance(pyramid_levels, bool) or not 1 <= levels <= 3:
            raise ValueError("pyramid_levels must be an integer from 1 to 3")
        if levels != 1:
            raise ValueError("selective pyramid recovery is not built; use ICGNWGPUFullPyramid for pyramids")
        if prefer_subgroups is not None and not isinstance(prefer_subgroups, bool):
            raise ValueError("prefer_subgroups must be a boolean or None")
        super().__init__(levels=1, power_preference=power_preference, device_type=device_type,
                         solver_workgroup_size=solver_workgroup_size, interpolation=interpolation,
                         motion_model=motion_model, outlier_policy=outlier_policy,
                         median_filter_size=median_filter_size, median_filter_scope=median_filter_scope,
                         subset_size=subset_size, subset_stride=subset_stride,
                         spatial_weighting=spatial_weighting, gaussian_sigma=gaussian_sigma)
        self.pyramid_levels = 1
        self.solver = "hip-wave64"

    def set_reference(self, image: ImageInput, *, iterations: int = 30, tolerance: float = 1e-3,
                      zncc_min: float = 0.8, roi_mask: np.ndarray | None = None, roi_policy: str = "full",
                      roi_truncation: str = "none") -> ReferenceInfo:
        started = time.perf_counter()
        config = _validated_config(iterations, tolerance, zncc_min)
        build_ms = self._prepare_reference(image, [config], roi_mask, roi_policy, roi_truncation)
        self._config = config
        self._pyramid_config = {"pyramidLevels": 1, "pyramidStrategy": "single"}
        return self._reference_info(started, build_ms, {
            **config,
            **({"roi": {**self.roi.summary(), "selectedPerLevel": self._roi_level_counts}} if self.roi else {}),
            **self._option_config()})

    def _resolve_config(self, iterations, tolerance, zncc_min) -> dict:
        config = _validated_config(
            self._config["iterations"] if iterations is None else iterations,
            self._config["tolerance"] if tolerance is None else tolerance,
            self._config["znccMin"] if zncc_min is None else zncc_min)
        if config != self._config:
            self._set_level_configs([config])
            self._config = config
        return config

    @_exclusive
    def register(self, target: ImageInput, *, iterations: int | None = None, tolerance: float | None = None,
                 zncc_min: float | None = None, include_zncc: bool = False) -> RegistrationResult:
        if not self._reference_ready:
            raise ICGNError("Call set_reference() before register().")
        started = time.perf_counter()
        frame = self._prepare_target(target)
        config = self._resolve_config(iterations, tolerance, zncc_min)
        if include_zncc:
            self._check_final_zncc()
        if self.roi is not None and self.roi.empty:
            return self._empty_roi_result(started, config)
        prep_ms = (time.perf_counter() - started) * 1e3
        if include_zncc:
            shape = (self.grid_height, self.grid_width)
            return self._seeded(frame, np.zeros((*shape, 6), np.float32), np.ones(shape, bool), True, started,
                                prep_ms, config)
        return self._run(frame, started, prep_ms, config)

    def _check_final_zncc(self) -> None:
        partial = self.roi is not None and self.roi.partial_indices is not None and self.roi.partial_indices.size
        if self.motion_model != "affine" or self.spatial_weighting != "uniform" or partial:
            raise ValueError("include_zncc requires affine motion and uniform full windows")

    def _seeded(self, frame, values, mask, include_zncc, started, prep_ms, config) -> RegistrationResult:
        n = self.grid_width * self.grid_height
        packed = np.zeros((n, 8), np.float32)
        packed[:, :6] = np.where(mask.reshape(-1, 1), values.reshape(n, 6), 0)
        active = np.ascontiguousarray(mask.reshape(-1), dtype=np.uint8)
        raw = self._state_buffer()
        out = (ctypes.c_char * len(raw)).from_buffer(raw)
        zncc = np.empty((self.grid_height, self.grid_width), np.float32)
        t = _native.Timing()
        _native.check(self._lib.ldic_register_seeded(
            self._handle, frame.ctypes.data, frame.strides[0], packed.ctypes.data, active.ctypes.data,
            1 if include_zncc else 0, ctypes.addressof(out), zncc.ctypes.data, ctypes.byref(t)))
        result = self._result(raw, self._timing_dict(t, started, prep_ms), config, self._summary(t))
        if include_zncc:
            zncc.setflags(write=False)
            result._zncc = zncc
        return result

    @_exclusive
    def register_seeded(self, target: ImageInput, parameters, active, *, iterations: int | None = None,
                        tolerance: float | None = None, zncc_min: float | None = None,
                        include_zncc: bool = False, fallback_zncc: float | None = None,
                        fallback_radius: int = 8, reseed_passes: int = 0) -> RegistrationResult:
        """Solve active affine subsets from supplied initial states (wgpu_backend.py:2773-2819).

        ``fallback_zncc`` (an addition; None keeps the reference behaviour) re-solves every
        selected subset that did not converge or whose final ZNCC is below it from the best
        integer ZNCC shift within ``fallback_radius`` pixels of its seed.  ``reseed_passes``
        (an addition, 0-8) then re-solves the subsets still unaccepted from the warp of their
        nearest accepted subset, extrapolated to their centre, once per pass.
        """
        if not self._reference_ready:
            raise ICGNError("Call set_reference() before register_seeded().")
        if self.motion_model != "affine":
            raise ValueError("register_seeded currently requires motion_model='affine'")
        if include_zncc:
            self._check_final_zncc()
        shape = (self.grid_height, self.grid_width)
        values = np.asarray(parameters, dtype=np.float32)
        mask = np.asarray(active)
        if values.shape != (*shape, 6):
            raise ValueError(f"parameters must have shape {(*shape, 6)}")
        if mask.shape != shape or mask.dtype != np.bool_:
            raise ValueError(f"active must be a Boolean array with shape {shape}")
        if not np.isfinite(values[mask]).all():
            raise ValueError("active parameters must be finite")
        if fallback_zncc is not None and not (-1.0 < float(fallback_zncc) <= 1.0):
            raise ValueError("fallback_zncc must be in (-1, 1] or None")
        if isinstance(reseed_passes, bool) or not isinstance(reseed_passes, int) or not 0 <= reseed_passes <= 8:
            raise ValueError("reseed_passes must be an integer from 0 to 8")
        started = time.perf_counter()
        frame = self._prepare_target(target)
        config = self._resolve_config(iterations, tolerance, zncc_min)
        prep_ms = (time.perf_counter() - started) * 1e3
        self._set_initialization(_native.INIT_ZERO, 0, fallback_zncc, fallback_radius if fallback_zncc is not None else 0,
                                 reseed_passes=reseed_passes)
        try:
            return self._seeded(frame, values, mask, include_zncc, started, prep_ms, config)
        finally:
            if fallback_zncc is not None or reseed_passes:
                self._set_initialization(_native.INIT_ZERO)

    @_exclusive_stream
    def register_stream(self, targets: Iterable[ImageInput], *, pipeline_depth: int = DEFAULT_IN_FLIGHT_FRAMES,
                        iterations: int | None = None, tolerance: float | None = None,
                        zncc_min: float | None = None) -> Iterator[RegistrationResult]:
        if not self._reference_ready:
            raise ICGNError("Call set_reference() before register_stream().")
        config = self._resolve_config(iterations, tolerance, zncc_min)
        yield from self._stream(targets, pipeline_depth, config)

    @_exclusive
    def register_pair(self, reference: ImageInput, target: ImageInput) -> RegistrationResult:
        """Consecutive-pair mode: prepare `reference` on the GPU and register `target` against it,
        keeping the geometry, ROI and solver controls of the last set_reference()."""
        return self._register_pair(reference, target, self._config)

    def register_many(self, targets: Iterable[ImageInput], *, pipeline_depth: int = DEFAULT_IN_FLIGHT_FRAMES,
                      iterations: int | None = None, tolerance: float | None = None,
                      zncc_min: float | None = None) -> list[RegistrationResult]:
        return list(self.register_stream(targets, pipeline_depth=pipeline_depth, iterations=iterations,
                                         tolerance=tolerance, zncc_min=zncc_min))
