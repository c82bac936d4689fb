"""Result and input contract of the native supervisor (reference: live_dic/engine.py:56-362, 778-828)."""

from __future__ import annotations

import io
import json
import os
from collections.abc import Mapping, Sequence
from dataclasses import dataclass
from typing import Any, Final, TypeAlias

import numpy as np
import numpy.typing as npt

from .motion import AFFINE_PARAMETER_ORDER, parameter_order
from .roi import RoiPlan

try:  # Pillow is optional: NumPy arrays never need it.
    from PIL import Image
except ImportError:  # pragma: no cover - depends on the installation
    Image = None

MIN_FRAME_DIMENSION: Final = 16
PARAMETER_ORDER: Final = AFFINE_PARAMETER_ORDER
ImageInput: TypeAlias = Any


class ICGNError(RuntimeError):
    """Raised when the native runtime or the IC-GN computation fails."""


@dataclass(frozen=True, slots=True)
class ReferenceInfo:
    """Dimensions and timings of reference preparation (engine.py:82-103)."""

    preparation_wall_ms: float
    runtime_build_wall_ms: float
    reference_build_wall_ms: float
    encoding_wall_ms: float
    total_wall_ms: float
    config: Mapping[str, Any]
    device: str
    grid: tuple[int, int]
    width: int
    height: int
    subset_size: int
    subset_stride: int
    roi: RoiPlan | None = None
    reference_valid: np.ndarray | None = None

    @property
    def roi_selected(self) -> np.ndarray | None:
        return None if self.roi is None else self.roi.selected


@dataclass(slots=True)
class RegistrationResult:
    """Dense field over the subset grid (engine.py:106-273).

    ``parameters[..., :]`` follows ``[tx, du/dx, du/dy, ty, dv/dx, dv/dy]``;
    the quadratic model appends the six second derivatives.  Status is 1 for
    accepted, 2 for displacement outliers, 0 for non-converged and negative for
    rejected subsets.  All arrays are views of one packed float32 buffer.
    """

    parameters: np.ndarray
    status: np.ndarray
    iterations: np.ndarray
    accepted: int
    replaced: int
    mean_translation: tuple[float, float] | None
    mean_iterations: float | None
    timing: Mapping[str, float]
    config: Mapping[str, Any]
    width: int = 512
    height: int = 512
    subset_size: int = 16
    subset_stride: int = 16
    roi: RoiPlan | None = None
    reference_valid: np.ndarray | None = None
    _zncc: np.ndarray | None = None

    @property
    def zncc(self) -> np.ndarray | None:
        """Optional final normalised correlation; NaN means unavailable."""
        return self._zncc

    @property
    def roi_selected(self) -> np.ndarray | None:
        return None if self.roi is None else self.roi.selected

    @property
    def populations(self) -> dict[str, int | float | None]:
        selected = self.status.size if self.roi is None else self.roi.selected_count
        admissible = None
        if self.reference_valid is not None:
            valid = self.reference_valid if self.roi is None else self.reference_valid & self.roi.selected
            admissible = int(np.count_nonzero(valid))
        return {
            "gridTotal": self.status.size,
            "roiSelected": selected,
            "roiExcluded": self.status.size - selected,
            "roiPartial": 0,
            "referenceAdmissible": admissible,
            "accepted": self.accepted,
            "rejectedSelected": selected - self.accepted,
            "acceptedFractionSelected": self.accepted / selected if selected else None,
        }

    @property
    def displacement(self) -> np.ndarray:
        """(tx, ty) as a zero-copy view."""
        return self.parameters[..., 0:4:3]

    @property
    def accepted_mask(self) -> np.ndarray:
        mask = self.status > 0.5
        return mask if self.roi is None else mask & self.roi.selected

    @property
    def flagged(self) -> int:
        return self.replaced

    @property
    def outlier_mask(self) -> np.ndarray:
        return self.status > 1.5

    @property
    def replaced_mask(self) -> np.ndarray:
        return self.outlier_mask

    @property
    def parameter_order(self) -> tuple[str, ...]:
        return parameter_order(str(self.config.get("motionModel", "affine")))

    def rmse(self, truth: Sequence[float]) -> float | None:
        """Vector RMSE of accepted displacements against a constant (dx, dy)."""
        expected = np.asarray(truth, dtype=np.float64)
        if expected.shape != (2,):
            raise ValueError("truth must contain exactly two values: (dx, dy)")
        mask = self.accepted_mask
        if not np.any(mask):
            return None
        error = self.displacement[mask].astype(np.float64) - expected
        return float(np.sqrt(np.mean(np.sum(error * error, axis=1))))

    def save_npz(self, path: str | os.PathLike[str]) -> None:
        """Self-contained field archive, including the static ROI (engine.py:180-193)."""
        np.savez_compressed(
            path,
            parameters=self.parameters,
            displacement=self.displacement,
            status=self.status,
            iterations=self.iterations,
            parameter_order=np.asarray(self.parameter_order),
            config_json=np.asarray(json.dumps(dict(self.config), sort_keys=True, allow_nan=False)),
            populations_json=np.asarray(json.dumps(self.populations, allow_nan=False)),
            **({} if self.roi is None else self.roi.archive_fields()),
            **({} if self.reference_valid is None else {"reference_valid": self.reference_valid}),
        )

    @classmethod
    def from_packed(cls, raw, *, grid: tuple[int, int], motion_model: str, **fields) -> RegistrationResult:
        """Views of one packed state buffer (gh, gw, 8|16) as a result."""
        gw, gh = grid
        order = parameter_order(motion_model)
        stride = 16 if len(order) == 12 else 8
        packed = np.frombuffer(raw, dtype=np.float32).reshape(gh, gw, stride)
        P = len(order)
        parameters = packed[..., :P]
        status = packed[..., P]
        iterations = packed[..., P + 1]
        accepted_mask = status > 0.5
        accepted = int(np.count_nonzero(accepted_mask))
        mean_translation = mean_iterations = None
        if accepted:
            mean_translation = (
                float(np.mean(parameters[..., 0][accepted_mask], dtype=np.float64)),
                float(np.mean(parameters[..., 3][accepted_mask], dtype=np.float64)),
            )
            mean_iterations = float(np.mean(iterations[accepted_mask], dtype=np.float64))
        return cls(parameters=parameters, status=status, iterations=iterations, accepted=accepted,
                   replaced=int(np.count_nonzero(status > 1.5)), mean_translation=mean_translation,
                   mean_iterations=mean_iterations, **fields)

    @classmethod
    def from_packed_browser_result(cls, raw: bytes, payload: Mapping[str, Any]) -> RegistrationResult:
        """Interleaved little-endian float32 state plus its JSON header (engine.py:275-327)."""
        grid = tuple(int(v) for v in payload["grid"])
        config = dict(payload.get("config", {}))
        model = str(config.get("motionModel", "affine"))
        stride = int(payload.get("stateStride", 16 if model == "quadratic" else 8))
        expected = grid[0] * grid[1] * stride * 4
        if len(raw) != expected:
            raise ICGNError(f"Browser binary result size mismatch: expected {expected} bytes, received {len(raw)}")
        return cls.from_packed(
            bytearray(raw), grid=grid, motion_model=model,
            timing={k: float(v) for k, v in payload.get("timing", {}).items() if v is not None},
            config=config, width=int(payload["width"]), height=int(payload["height"]),
            subset_size=int(payload["subsetSize"]),
            subset_stride=int(payload.get("subsetStride", payload["subsetSize"])))


def _as_uint8_gray(image: ImageInput) -> np.ndarray:
    """Contiguous 8-bit grayscale frame from an array, PIL image, path or bytes (engine.py:778-828)."""
    if isinstance(image, (str, os.PathLike, bytes)) or (Image is not None and isinstance(image, Image.Image)):
        if Image is None:
            raise ValueError("Pillow is required to read image files")
        opened = image if isinstance(image, Image.Image) else Image.open(
            io.BytesIO(image) if isinstance(image, bytes) else image)
        try:
            if opened.mode == "F" or opened.mode.startswith("I"):
                return _as_uint8_gray(np.asarray(opened))
            array = np.asarray(opened.convert("L"), dtype=np.uint8)
        finally:
            if opened is not image:
                opened.close()
    else:
        source = np.asarray(image)
        if source.ndim not in (2, 3):
            raise ValueError("Expected a 2D, RGB, or RGBA image.")
        if source.ndim == 3 and source.shape[2] not in (3, 4):
            raise ValueError("Expected an RGB or RGBA image with 3 or 4 channels.")
        if np.issubdtype(source.dtype, np.floating):
            source = np.nan_to_num(source, nan=0.0, posinf=255.0, neginf=0.0)
            scale = 255.0 if source.size and float(np.max(source)) <= 1.0 else 1.0
            source = np.clip(source * scale, 0, 255).astype(np.uint8)
        elif source.dtype != np.uint8:
            if not np.issubdtype(source.dtype, np.integer):
                raise ValueError(f"Unsupported image dtype {source.dtype}.")
            if source.size and (int(np.min(source)) < 0 or int(np.max(source)) > 255):
                raise ValueError(
                    "Integer image values must be in [0, 255]. Convert higher-bit-depth images explicitly "
                    "to uint8 with the same calibrated intensity range for the reference and every target."
                )
            source = source.astype(np.uint8)
        if source.ndim == 2:
            array = source
        else:
            if Image is None:
                raise ValueError("Pillow is required to convert colour images")
            array = np.asarray(Image.fromarray(source).convert("L"), dtype=np.uint8)
    height, width = array.shape
    if height < MIN_FRAME_DIMENSION or width < MIN_FRAME_DIMENSION:
        raise ValueError(
            f"Image dimensions must each be at least {MIN_FRAME_DIMENSION} pixels; received {array.shape}."
        )
    return np.ascontiguousarray(array)


def empty_timing() -> dict[str, float]:
    return {"queueSubmissions": 0.0}


NDArrayF32 = npt.NDArray[np.float32]
