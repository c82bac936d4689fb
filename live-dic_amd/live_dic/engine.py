"""Result and input contract of the native supervisor.

The public shapes follow the reference (live_dic/engine.py:56-362 for the result
fields and archive keys, :778-828 for the accepted image inputs and their error
messages); the code is this package's own.  A result is a set of zero-copy views
of one packed float32 field ``(gh, gw, 8|16)`` as the device writes it, plus the
summary statistics the device computes beside it.
"""

from __future__ import annotations

import io
import json
import os
from collections.abc import Mapping, Sequence
from dataclasses import dataclass
from typing import Any, Final, NamedTuple, TypeAlias

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

# Status codes of the packed field: > 0.5 accepted (1) or a displacement outlier (2).
_ACCEPTED_ABOVE: Final = 0.5
_OUTLIER_ABOVE: Final = 1.5


class ICGNError(RuntimeError):
    """Raised when the native runtime or the IC-GN computation fails."""


class FieldSummary(NamedTuple):
    """Counts and float64 sums of one packed field: subsets with status > 0.5 and > 1.5, and
    the sums of tx, ty and iterations over the former."""

    accepted: int
    replaced: int
    sum_tx: float
    sum_ty: float
    sum_iterations: float

    @classmethod
    def of(cls, packed: np.ndarray, params: int) -> FieldSummary:
        """Host computation, for fields that did not come from the device with a summary."""
        status = packed[..., params]
        keep = status > _ACCEPTED_ABOVE
        picked = packed[keep]
        return cls(int(picked.shape[0]), int(np.count_nonzero(status > _OUTLIER_ABOVE)),
                   float(picked[:, 0].sum(dtype=np.float64)), float(picked[:, 3].sum(dtype=np.float64)),
                   float(picked[:, params + 1].sum(dtype=np.float64)))

    def means(self) -> tuple[tuple[float, float] | None, float | None]:
        if self.accepted == 0:
            return None, None
        n = float(self.accepted)
        return (self.sum_tx / n, self.sum_ty / n), self.sum_iterations / n


@dataclass(frozen=True, slots=True)
class ReferenceInfo:
    """Dimensions and timings of one reference preparation (fields: engine.py:82-103)."""

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
        return self.roi.selected if self.roi is not None else None


@dataclass(slots=True)
class RegistrationResult:
    """Dense field over the subset grid (fields and properties: engine.py:106-273).

    ``parameters[..., :]`` is ``[tx, du/dx, du/dy, ty, dv/dx, dv/dy]``, followed for the
    quadratic model by the six second derivatives.  Status 1 is accepted, 2 a displacement
    outlier, 0 not converged, negative rejected.
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
        """Final normalised correlation when it was requested; NaN where undefined."""
        return self._zncc

    @property
    def roi_selected(self) -> np.ndarray | None:
        return self.roi.selected if self.roi is not None else None

    def _selected_count(self) -> int:
        return self.status.size if self.roi is None else self.roi.selected_count

    @property
    def populations(self) -> dict[str, int | float | None]:
        total = self.status.size
        chosen = self._selected_count()
        partial = self.roi.partial_indices if self.roi is not None else None
        admissible = None
        if self.reference_valid is not None:
            usable = self.reference_valid if self.roi is None else np.logical_and(self.reference_valid,
                                                                                self.roi.selected)
            admissible = int(np.count_nonzero(usable))
        return {
            "gridTotal": total,
            "roiSelected": chosen,
            "roiExcluded": total - chosen,
            "roiPartial": 0 if partial is None else int(np.size(partial)),
            "referenceAdmissible": admissible,
            "accepted": self.accepted,
            "rejectedSelected": chosen - self.accepted,
            "acceptedFractionSelected": (self.accepted / chosen) if chosen else None,
        }

    @property
    def displacement(self) -> np.ndarray:
        """(tx, ty) per subset, a view of the parameters."""
        return self.parameters[..., 0:4:3]

    @property
    def accepted_mask(self) -> np.ndarray:
        """Accepted and flagged subsets (status > 0.5) inside the ROI selection."""
        keep = self.status > _ACCEPTED_ABOVE
        if self.roi is not None:
            keep &= self.roi.selected
        return keep

    @property
    def flagged(self) -> int:
        """Displacement outliers found by the outlier pass (status 2); same as ``replaced``."""
        return self.replaced

    @property
    def outlier_mask(self) -> np.ndarray:
        return self.status > _OUTLIER_ABOVE

    @property
    def replaced_mask(self) -> np.ndarray:
        return self.outlier_mask

    @property
    def parameter_order(self) -> tuple[str, ...]:
        return parameter_order(str(self.config.get("motionModel", "affine")))

    def rmse(self, truth: Sequence[float]) -> float | None:
        """Root mean square of the displacement error vector against a constant (dx, dy)."""
        target = np.asarray(truth, dtype=np.float64)
        if target.shape != (2,):
            raise ValueError("truth must contain exactly two values: (dx, dy)")
        picked = self.displacement[self.accepted_mask]
        if picked.shape[0] == 0:
            return None
        squared = np.square(picked.astype(np.float64) - target).sum(axis=1)
        return float(np.sqrt(squared.mean()))

    def archive(self) -> dict[str, np.ndarray]:
        """Arrays of the self-contained field archive (keys: engine.py:180-193)."""
        out = {
            "parameters": self.parameters,
            "displacement": self.displacement,
            "status": self.status,
            "iterations": self.iterations,
            "parameter_order": np.asarray(self.parameter_order),
            "config_json": np.asarray(json.dumps(dict(self.config), sort_keys=True, allow_nan=False)),
            "populations_json": np.asarray(json.dumps(self.populations, allow_nan=False)),
        }
        if self.roi is not None:
            out.update(self.roi.archive_fields())
        if self.reference_valid is not None:
            out["reference_valid"] = self.reference_valid
        return out

    def save_npz(self, path: str | os.PathLike[str]) -> None:
        """Write :meth:`archive` as one compressed NPZ file."""
        np.savez_compressed(path, **self.archive())

    @classmethod
    def from_packed(cls, raw, *, grid: tuple[int, int], motion_model: str, summary: FieldSummary | None = None,
                    **fields) -> RegistrationResult:
        """Views of one packed state buffer (gh, gw, 8|16).  ``summary`` carries the device's
        statistics of the same field; without it they are computed here."""
        gw, gh = grid
        params = len(parameter_order(motion_model))
        packed = np.frombuffer(raw, dtype=np.float32).reshape(gh, gw, 16 if params == 12 else 8)
        stats = summary if summary is not None else FieldSummary.of(packed, params)
        mean_translation, mean_iterations = stats.means()
        return cls(parameters=packed[..., :params], status=packed[..., params], iterations=packed[..., params + 1],
                   accepted=stats.accepted, replaced=stats.replaced, mean_translation=mean_translation,
                   mean_iterations=mean_iterations, **fields)

    @classmethod
    def from_packed_browser_result(cls, raw: bytes, payload: Mapping[str, Any]) -> RegistrationResult:
        """The browser engine's interleaved little-endian float32 state and its JSON header
        (engine.py:275-327)."""
        gw, gh = (int(v) for v in payload["grid"])
        config = dict(payload.get("config", {}))
        model = str(config.get("motionModel", "affine"))
        stride = int(payload.get("stateStride", 16 if model == "quadratic" else 8))
        want = gw * gh * stride * 4
        if len(raw) != want:
            raise ICGNError(f"Browser binary result size mismatch: expected {want} bytes, received {len(raw)}")
        size = int(payload["subsetSize"])
        timing = {k: float(v) for k, v in payload.get("timing", {}).items() if v is not None}
        return cls.from_packed(bytearray(raw), grid=(gw, gh), motion_model=model, timing=timing, config=config,
                               width=int(payload["width"]), height=int(payload["height"]), subset_size=size,
                               subset_stride=int(payload.get("subsetStride", size)))


# -- image inputs (engine.py:778-828) ------------------------------------------------------------
def _from_pil(picture) -> np.ndarray:
    """Float and 32-bit integer modes go through the array rules; others convert to "L"."""
    if picture.mode == "F" or picture.mode.startswith("I"):
        return _from_array(np.asarray(picture))
    return np.asarray(picture.convert("L"), dtype=np.uint8)


def _from_array(source: np.ndarray) -> np.ndarray:
    if source.ndim not in (2, 3):
        raise ValueError("Expected a 2D, RGB, or RGBA image.")
    if source.ndim == 3 and source.shape[2] not in (3, 4):
        raise ValueError("Expected an RGB or RGBA image with 3 or 4 channels.")
    kind = source.dtype
    if np.issubdtype(kind, np.floating):
        finite = np.nan_to_num(source, nan=0.0, posinf=255.0, neginf=0.0)
        if finite.size and float(finite.max()) <= 1.0:  # normalised intensities
            finite = finite * 255.0
        values = np.clip(finite, 0, 255).astype(np.uint8)
    elif kind == np.uint8:
        values = source
    elif np.issubdtype(kind, np.integer):
        if source.size and (int(source.min()) < 0 or int(source.max()) > 255):
            raise ValueError(
                "Integer image values must be in [0, 255]. Convert higher-bit-depth images explicitly "
                "to uint8 with the same calibrated intensity range for the reference and every target."
            )
        values = source.astype(np.uint8)
    else:
        raise ValueError(f"Unsupported image dtype {kind}.")
    if values.ndim == 2:
        return values
    if Image is None:
        raise ValueError("Pillow is required to convert colour images")
    return np.asarray(Image.fromarray(values).convert("L"), dtype=np.uint8)


def _as_uint8_gray(image: ImageInput) -> np.ndarray:
    """Contiguous 8-bit grayscale frame from an array, a PIL image, a path or encoded bytes."""
    is_pil = Image is not None and isinstance(image, Image.Image)
    if isinstance(image, (str, os.PathLike, bytes)) or is_pil:
        if Image is None:
            raise ValueError("Pillow is required to read image files")
        if is_pil:
            frame = _from_pil(image)
        else:
            with Image.open(io.BytesIO(image) if isinstance(image, bytes) else image) as opened:
                frame = _from_pil(opened)
    else:
        frame = _from_array(np.asarray(image))
    if min(frame.shape) < MIN_FRAME_DIMENSION:
        raise ValueError(
            f"Image dimensions must each be at least {MIN_FRAME_DIMENSION} pixels; received {frame.shape}."
        )
    return np.ascontiguousarray(frame)


def empty_timing() -> dict[str, float]:
    return {"queueSubmissions": 0.0}


NDArrayF32 = npt.NDArray[np.float32]
