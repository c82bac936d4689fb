"""Coarse-to-fine engine (reference: live_dic/full_pyramid_backend.py).

Each target is uploaded once; its pyramid, per-level prefilters, solves,
outlier passes and seed transfers run as one queued sequence on a HIP stream,
and only the finest field is read back.  ``register_stream`` keeps
``pipeline_depth`` slots in flight and returns results in input order.
"""

from __future__ import annotations

import operator
import time
from collections.abc import Iterable, Iterator

import numpy as np

from . import _native
from .engine import ICGNError, ImageInput, ReferenceInfo, RegistrationResult
from .hip_backend import (
    DEFAULT_IN_FLIGHT_FRAMES,
    _exclusive,
    _exclusive_stream,
    _NativeEngine,
    _validated_config,
)
from .subset import DEFAULT_SUBSET_SIZE

DEFAULT_FULL_PYRAMID_LEVELS = 2
# Iteration ceilings and tolerances per level, coarsest first
# (full_pyramid_backend.py:51-62).
_DEFAULT_ITERATION_SCHEDULES = {1: (30,), 2: (8, 16), 3: (6, 8, 16), 4: (12, 8, 8, 16)}
_DEFAULT_TOLERANCES = (1e-2, 5e-3, 5e-3)
_FOUR_LEVEL_TOLERANCES = (1e-2, *_DEFAULT_TOLERANCES)


def _level_count(value: int) -> int:
    try:
        count = operator.index(value)
    except TypeError as error:
        raise ValueError("pyramid_levels must be an integer from 1 to 4") from error
    if isinstance(value, bool) or not 1 <= count <= 4:
        raise ValueError("pyramid_levels must be an integer from 1 to 4")
    return count


def _iteration_count(value, name: str) -> int:
    """An integer from 1 to 30 (full_pyramid_backend.py:135-142)."""
    try:
        count = operator.index(value)
    except TypeError as error:
        raise ValueError(f"{name} must be an integer from 1 to 30") from error
    if isinstance(value, bool) or not 1 <= count <= 30:
        raise ValueError(f"{name} must be an integer from 1 to 30")
    return count


def _schedule(values, *, level_count: int, defaults, name: str):
    selected = tuple(defaults[-level_count:] if values is None else values)
    if len(selected) != level_count:
        raise ValueError(f"{name} must contain one value per pyramid level")
    return selected


class ICGNWGPUFullPyramid(_NativeEngine):
    """GPU-resident image pyramid with a coarse-to-fine IC-GN solve.

    ``iterations_per_level`` and ``tolerance_per_level`` are ordered from the
    coarsest level to the finest.  Initialisers: ``"zero"`` (the reference
    default); ``"temporal"`` (full_pyramid_backend.py:433-881): the last accepted
    finest field, projected onto the coarsest grid, seeds every frame; history
    entries expire after more than ``temporal_max_age`` rejected frames; targets
    must come in chronological order; ``temporal_prediction="velocity"``
    extrapolates subsets with two consecutive fresh accepted measurements, and
    ``temporal_fine_correction`` (two levels) seeds the finest level with its
    history plus the coarse solve's correction; and ``"zncc"``, an addition:
    the coarsest level is seeded by an exhaustive integer ZNCC search covering
    ``seed_search_radius`` pixels of the finest level.  ``divergence_patience``
    stops unconverged finest-level solves whose change stopped decreasing
    (icgn.wgsl:103-111).  ``pyramid_policy="adaptive"`` (temporal, two levels) solves only
    the finest level after a frame whose merged subsets reach ``adaptive_acceptance`` of the
    reference-valid ones with no history expiry, and both levels otherwise
    (full_pyramid_backend.py:1775-1791).  ``reseed_passes`` (an addition, 0-8) re-solves finest-level
    subsets still unaccepted from the warp of their nearest accepted subset.  Not built: the
    Fourier and CPM-HOG matching initialisers.
    """

    def __init__(self, *, pyramid_levels: int = DEFAULT_FULL_PYRAMID_LEVELS,
                 iterations_per_level: tuple[int, ...] | None = None,
                 tolerance_per_level: tuple[float, ...] | None = None,
                 power_preference: str = "high-performance", device_type: str = "auto",
                 solver_workgroup_size: int = 16, interpolation: str = "bspline", motion_model: str = "affine",
                 outlier_policy: str = "replace", median_filter_size: int = 3,
                 median_filter_scope: str = "rejected", subset_size: int = DEFAULT_SUBSET_SIZE,
                 subset_stride: int | None = None, spatial_weighting: str = "uniform",
                 gaussian_sigma: float | None = None, initialization: str = "zero",
                 temporal_max_age: int | None = None, temporal_fine_correction: bool = False,
                 temporal_prediction: str = "none", pyramid_policy: str = "fixed", seed_search_radius: int = 32,
                 divergence_patience: int | None = None, divergence_min_iterations: int = 4,
                 reseed_passes: int = 0, adaptive_acceptance: float = 0.95, **unsupported) -> None:
        if unsupported:
            raise ValueError(f"unsupported options for the HIP engine: {sorted(unsupported)}")
        if not isinstance(pyramid_policy, str) or pyramid_policy not in ("fixed", "adaptive"):
            raise ValueError("pyramid_policy must be 'fixed' or 'adaptive'")
        if pyramid_policy == "adaptive" and (initialization != "temporal" or pyramid_levels != 2):
            raise ValueError("adaptive pyramid requires temporal initialization and two levels")
        if (isinstance(adaptive_acceptance, bool) or not isinstance(adaptive_acceptance, (int, float))
                or not np.isfinite(adaptive_acceptance) or not 0 < adaptive_acceptance <= 1):
            raise ValueError("adaptive_acceptance must be finite and in (0, 1]")
        self.adaptive_acceptance = float(adaptive_acceptance)
        if not isinstance(initialization, str) or initialization not in (
                "zero", "zncc", "temporal", "fourier", "wgpu-matching"):
            raise ValueError("initialization must be 'zero', 'zncc', 'temporal', 'fourier' or 'wgpu-matching'")
        if initialization not in ("zero", "zncc", "temporal"):
            raise ValueError(f"initialization={initialization!r} is not built; use 'zero', 'temporal' or 'zncc'")
        if not isinstance(temporal_prediction, str) or temporal_prediction not in ("none", "velocity"):
            raise ValueError("temporal_prediction must be 'none' or 'velocity'")
        if temporal_prediction != "none" and initialization != "temporal":
            raise ValueError("temporal_prediction requires temporal initialization")
        self.divergence_patience = (None if divergence_patience is None
                                    else _iteration_count(divergence_patience, "divergence_patience"))
        self.divergence_min_iterations = _iteration_count(divergence_min_iterations, "divergence_min_iterations")
        if isinstance(reseed_passes, bool) or not isinstance(reseed_passes, int) or not 0 <= reseed_passes <= 8:
            raise ValueError("reseed_passes must be an integer from 0 to 8")
        if reseed_passes and motion_model == "quadratic":
            raise ValueError("reseed_passes supports the translation and affine models")
        self.reseed_passes = reseed_passes
        if not isinstance(temporal_fine_correction, bool):
            raise ValueError("temporal_fine_correction must be a boolean")
        if temporal_fine_correction and (initialization != "temporal" or _level_count(pyramid_levels) != 2):
            raise ValueError("temporal_fine_correction requires temporal initialization and two levels")
        if temporal_max_age is not None:
            try:
                age = operator.index(temporal_max_age)
            except TypeError as error:
                raise ValueError("temporal_max_age must be a nonnegative integer or None") from error
            if isinstance(temporal_max_age, bool) or not 0 <= age <= 0xFFFFFFFE:
                raise ValueError("temporal_max_age must be between 0 and 4294967294 or None")
            temporal_max_age = age
        self.temporal_max_age = temporal_max_age
        self.temporal_prediction = temporal_prediction
        self.temporal_fine_correction = temporal_fine_correction
        self.pyramid_levels = _level_count(pyramid_levels)
        try:
            radius = operator.index(seed_search_radius)
        except TypeError as error:
            raise ValueError("seed_search_radius must be a non-negative integer") from error
        if isinstance(seed_search_radius, bool) or radius < 0:
            raise ValueError("seed_search_radius must be a non-negative integer")
        self.seed_search_radius = radius
        # Coarsest-level radius: the finest-level range scaled down, plus one
        # pixel for the rounding of the scaled displacement.
        self.coarse_search_radius = -(-radius // (1 << (self.pyramid_levels - 1))) + 1
        self.pyramid_policy = pyramid_policy
        self.initialization = initialization
        L = self.pyramid_levels
        tol_defaults = _FOUR_LEVEL_TOLERANCES if L == 4 else _DEFAULT_TOLERANCES
        self.iterations_per_level = tuple(int(v) for v in _schedule(
            iterations_per_level, level_count=L, defaults=_DEFAULT_ITERATION_SCHEDULES[L],
            name="iterations_per_level"))
        self.tolerance_per_level = tuple(float(v) for v in _schedule(
            tolerance_per_level, level_count=L, defaults=tol_defaults, name="tolerance_per_level"))
        for it, tol in zip(self.iterations_per_level, self.tolerance_per_level, strict=True):
            _validated_config(it, tol, 0.8)
        super().__init__(levels=L, power_preference=power_preference, device_type=device_type,
                         solver_workgroup_size=solver_workgroup_size, interpolation=interpolation,
                         motion_model=motion_model, outlier_policy=outlier_policy,
                         median_filter_size=median_filter_size, median_filter_scope=median_filter_scope,
                         subset_size=subset_size, subset_stride=subset_stride,
                         spatial_weighting=spatial_weighting, gaussian_sigma=gaussian_sigma)
        self._zncc_min = 0.8
        sequence = dict(temporal_prediction=1 if temporal_prediction == "velocity" else 0,
                        temporal_fine_correction=int(temporal_fine_correction),
                        divergence_patience=self.divergence_patience or 0,
                        divergence_min_iterations=self.divergence_min_iterations,
                        reseed_passes=reseed_passes, adaptive_pyramid=int(pyramid_policy == "adaptive"),
                        adaptive_acceptance=self.adaptive_acceptance)
        if initialization == "zncc":
            self._set_initialization(_native.INIT_ZNCC, self.coarse_search_radius, **sequence)
        elif initialization == "temporal":
            self._set_initialization(_native.INIT_TEMPORAL, temporal_max_age=temporal_max_age, **sequence)
        elif self.divergence_patience is not None or reseed_passes:
            self._set_initialization(_native.INIT_ZERO, **sequence)

    def set_reference(self, image: ImageInput, *, iterations_per_level: tuple[int, ...] | None = None,
                      tolerance_per_level: tuple[float, ...] | None = None, zncc_min: float = 0.8,
                      roi_mask: np.ndarray | None = None, roi_policy: str = "full",
                      roi_truncation: str = "none") -> ReferenceInfo:
        """Prepare every level's reference on the device (full_pyramid_backend.py:1135-1368)."""
        started = time.perf_counter()
        L = self.pyramid_levels
        tol_defaults = _FOUR_LEVEL_TOLERANCES if L == 4 else _DEFAULT_TOLERANCES
        iterations = self.iterations_per_level if iterations_per_level is None else tuple(
            int(v) for v in _schedule(iterations_per_level, level_count=L,
                                      defaults=_DEFAULT_ITERATION_SCHEDULES[L], name="iterations_per_level"))
        tolerances = self.tolerance_per_level if tolerance_per_level is None else tuple(
            float(v) for v in _schedule(tolerance_per_level, level_count=L, defaults=tol_defaults,
                                        name="tolerance_per_level"))
        configs_coarse_first = [_validated_config(i, t, zncc_min) for i, t in zip(iterations, tolerances)]
        configs = configs_coarse_first[::-1]  # finest first for the native engine
        build_ms = self._prepare_reference(image, configs, roi_mask, roi_policy, roi_truncation)
        if self.initialization == "temporal":
            self._set_temporal_links()
        self.iterations_per_level = iterations
        self.tolerance_per_level = tolerances
        self._zncc_min = float(zncc_min)
        self._config = configs[0]
        self._pyramid_config = {
            "pyramidLevels": L,
            "pyramidStrategy": "full",
            "iterationsPerLevel": list(iterations),
            "tolerancePerLevel": list(tolerances),
        }
        if self.initialization == "zncc":
            self._pyramid_config.update(initialization="zncc", seedSearchRadius=self.seed_search_radius,
                                        coarseSearchRadius=self.coarse_search_radius)
        if self.reseed_passes:
            self._pyramid_config.update(reseedPasses=self.reseed_passes)
        if self.pyramid_policy == "adaptive":
            self._pyramid_config.update(pyramidPolicy="adaptive", adaptiveAcceptance=self.adaptive_acceptance,
                                        adaptiveFeedback="ordered-single-slot")
        if self.divergence_patience is not None:
            self._pyramid_config.update(divergencePatience=self.divergence_patience,
                                        divergenceMinIterations=self.divergence_min_iterations,
                                        divergenceLevels="finest")
        if self.initialization == "temporal":
            filtered = self.outlier_policy == "replace" and self.temporal_prediction == "none"
            self._pyramid_config.update(initialization="temporal", temporalMaxAge=self.temporal_max_age,
                                        temporalSeedLevel="coarsest",
                                        temporalSeedState="filtered" if filtered else "raw",
                                        **({"temporalFineCorrection": True} if self.temporal_fine_correction else {}),
                                        **({"temporalPrediction": "velocity"}
                                           if self.temporal_prediction == "velocity" else {}))
        return self._reference_info(started, build_ms, {
            **configs[0],
            **({"roi": {**self.roi.summary(), "selectedPerLevel": self._roi_level_counts}} if self.roi else {}),
            **self._option_config()})

    def _set_temporal_links(self) -> None:
        """Projection links finest -> coarsest under an effective ROI (full_pyramid_backend.py:608-620),
        planned on the device with the reference's tables."""
        links = None
        if self.roi is not None and self.roi.effective and self._temporal_links is not None:
            links = np.ascontiguousarray(self._temporal_links, dtype=np.uint32)
        _native.check(self._lib.ldic_set_temporal_links(self._handle, None if links is None else links.ctypes.data))

    def reset_temporal_state(self) -> None:
        """Forget the temporal history (full_pyramid_backend.py:660-677); a no-op before
        set_reference or in other initialisation modes.  Finish consuming a stream first."""
        if self._reference_ready and self.initialization == "temporal":
            _native.check(self._lib.ldic_reset_temporal(self._handle))

    @_exclusive
    def register(self, target: ImageInput) -> RegistrationResult:
        """One upload, one queued coarse-to-fine sequence and one readback."""
        if not self._reference_ready:
            raise ICGNError("Call set_reference() before register().")
        started = time.perf_counter()
        frame = self._prepare_target(target)
        if self.roi is not None and self.roi.empty:
            return self._empty_roi_result(started, self._config)
        prep_ms = (time.perf_counter() - started) * 1e3
        return self._levels_used(self._run(frame, started, prep_ms, self._config))

    @_exclusive
    def register_pair(self, reference: ImageInput, target: ImageInput) -> RegistrationResult:
        """Consecutive-pair mode: the reference pyramid and every level's preparation run on the
        GPU in the same queued sequence as the target's coarse-to-fine solve."""
        return self._levels_used(self._register_pair(reference, target, self._config))

    @_exclusive_stream
    def register_stream(self, targets: Iterable[ImageInput], *,
                        pipeline_depth: int = DEFAULT_IN_FLIGHT_FRAMES) -> Iterator[RegistrationResult]:
        """Ordered results with at most ``pipeline_depth`` frames in flight."""
        if not self._reference_ready:
            raise ICGNError("Call set_reference() before register_stream().")
        if isinstance(pipeline_depth, bool):
            raise ValueError("pipeline_depth must be a positive integer")
        if self.pyramid_policy == "adaptive":
            # Each frame's level count depends on the previous frame's result: retire
            # every frame before submitting the next (full_pyramid_backend.py:1865-1870).
            operator.index(pipeline_depth)
            pipeline_depth = 1
        for result in self._stream(targets, pipeline_depth, self._config):
            yield self._levels_used(result)

    def _levels_used(self, result: RegistrationResult) -> RegistrationResult:
        """Levels the frame solved, as timing (and, for the adaptive pyramid, config)."""
        used = int(result.timing.pop("levelsUsed", 0)) or self.pyramid_levels
        result.timing["pyramidLevels"] = float(used)
        if self.pyramid_policy == "adaptive":
            result.config["pyramidLevelsUsed"] = used
        return result

    def register_many(self, targets: Iterable[ImageInput], *,
                      pipeline_depth: int = DEFAULT_IN_FLIGHT_FRAMES) -> list[RegistrationResult]:
        return list(self.register_stream(targets, pipeline_depth=pipeline_depth))
