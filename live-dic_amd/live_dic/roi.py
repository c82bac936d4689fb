"""Static reference-coordinate ROI planning, once per reference.

The tables the device reads (window selection, 3x3 neighbour links, the 81-bit
link words of wider medians, partial supports, transfer links between pyramid
levels and the reduced level masks) are computed on the device by the
library's planner (``ldic_roi_planner_*``, csrc/ldic_roi.hip).  This module
validates the mask with the reference's rules and messages (roi.py:29-60),
drives the planner and keeps its output as immutable :class:`RoiPlan` tables,
whose fields, summary keys and archive keys follow the reference's RoiPlan.

A subset is selected when its whole S x S window is allowed (``full``).  Under
``partial`` (translation and affine, as in the reference) an incomplete window
is also selected when the 4-connected component of allowed pixels grown from
its centre is not empty; that component, optionally cut to the pixels visible
from the centre (``roi_truncation="visibility"``), is its support.  The
reference rejects ROIs for the quadratic model (roi.py:45-46); this build
supports them under the full policy.
"""

from __future__ import annotations

import ctypes
import hashlib
import struct
from dataclasses import dataclass

import numpy as np

from .outlier import normalize_median_filter_size

# Partial-support admission limits (roi.py:20-21): normalised Hessian condition and
# weighted RMS contrast of the support.
PARTIAL_MAX_NORMALIZED_CONDITION = 1e8
PARTIAL_MIN_WEIGHTED_CONTRAST = 1.0
# The device planner grows partial supports one 64-bit row per lane.
MAX_PARTIAL_SUBSET = 64


def immutable(array: np.ndarray) -> np.ndarray:
    """Read-only copy owning its bytes."""
    out = np.array(array, copy=True, order="C")
    out.setflags(write=False)
    return out


def validate_roi(mask, policy: str, shape: tuple[int, int], motion_model: str, truncation: str = "none"):
    """The reference's checks (policy, truncation, dtype, shape) before any device work."""
    if not isinstance(policy, str) or policy not in ("full", "partial"):
        raise ValueError("roi_policy must be 'full' or 'partial'")
    if not isinstance(truncation, str) or truncation not in ("none", "visibility"):
        raise ValueError("roi_truncation must be 'none' or 'visibility'")
    if truncation != "none" and (policy != "partial" or mask is None):
        raise ValueError("roi_truncation requires roi_policy='partial' and a mask")
    if mask is None:
        return None
    if policy == "partial" and motion_model not in ("translation", "affine"):
        raise ValueError("roi_policy='partial' supports only translation and affine motion models")
    if not isinstance(mask, np.ndarray) or mask.dtype != np.bool_:
        raise TypeError("roi_mask must be a two-dimensional boolean NumPy array")
    if mask.ndim != 2 or mask.shape != tuple(shape):
        raise ValueError(f"roi_mask must have reference shape {tuple(shape)}")
    return immutable(mask)


def mask_digest(mask: np.ndarray) -> str:
    """SHA-256 of "live-dic-roi-v1\\0", the uint64le height and width, and the C-order bytes."""
    head = b"live-dic-roi-v1\0" + struct.pack("<QQ", *mask.shape)
    return hashlib.sha256(head + np.ascontiguousarray(mask, dtype=np.bool_).tobytes()).hexdigest()


@dataclass(frozen=True, slots=True)
class RoiPlan:
    """Immutable per-level ROI tables shared by a reference and its results."""

    mask: np.ndarray
    selected: np.ndarray
    neighbour_bits: np.ndarray
    mask_hash: str
    effective: bool
    selected_count: int
    policy: str = "full"
    truncation: str = "none"
    partial_indices: np.ndarray | None = None
    median_neighbour_words: np.ndarray | None = None
    support_words: np.ndarray | None = None     # (partial, ceil(S*S/32)) uint32, row-major, pixel 0 in bit 0
    support_boxes: np.ndarray | None = None     # (partial, 4) float32 xmin, ymin, xmax, ymax (local pixels)
    coverage: np.ndarray | None = None          # (gh, gw) support fraction of each selected subset
    support_condition: np.ndarray | None = None
    support_rank: np.ndarray | None = None
    support_contrast: np.ndarray | None = None

    @property
    def empty(self) -> bool:
        return self.selected_count == 0

    def summary(self) -> dict:
        out = {
            "policy": self.policy,
            "truncation": self.truncation,
            "partial": 0 if self.partial_indices is None else int(self.partial_indices.size),
        }
        if self.policy == "partial":
            out.update(supportConditionLimit=PARTIAL_MAX_NORMALIZED_CONDITION,
                       supportContrastRmsMinimum=PARTIAL_MIN_WEIGHTED_CONTRAST, supportConnectivity=4,
                       evenCenter="all four central pixels required")
        total = int(self.selected.size)
        out.update(maskShape=list(self.mask.shape), maskSha256=self.mask_hash, gridTotal=total,
                   selected=self.selected_count, excluded=total - self.selected_count,
                   hashFormat="live-dic-roi-v1\\0 + uint64le(height,width) + C-order uint8 bool")
        return out

    def archive_fields(self) -> dict:
        fields = {"roi_mask": self.mask, "roi_selected": self.selected}
        if self.partial_indices is not None:
            fields.update(roi_partial_indices=self.partial_indices, roi_support_words=self.support_words,
                          roi_support_boxes=self.support_boxes, roi_coverage=self.coverage)
            if self.support_condition is not None:
                fields.update(roi_support_condition=self.support_condition, roi_support_rank=self.support_rank,
                              roi_support_contrast=self.support_contrast)
        return fields

    def neighbour_words(self) -> np.ndarray:
        """(grid, 4) float32 link words as the device reads them after the metadata."""
        if self.median_neighbour_words is not None:
            return np.ascontiguousarray(self.median_neighbour_words, dtype=np.float32)
        words = np.zeros((self.selected.size, 4), dtype=np.float32)
        words[:, 0] = self.neighbour_bits.ravel()
        return words


@dataclass(frozen=True)
class PyramidPlan:
    """Plans of every pyramid level (finest first; None without a mask), the transfer links into
    each finer level (``seed_links[k]``: from level k + 1 into k) and the projection links from the
    finest level onto the coarsest (temporal initialisation), when the finest plan is effective."""

    plans: list
    seed_links: list
    temporal_links: np.ndarray | None
    planning_ms: float


def plan_pyramid(mask, levels: int, *, size: int, stride: int, policy: str = "full", truncation: str = "none",
                 median_filter_size: int = 3, device: int = 0, temporal: bool = False) -> PyramidPlan:
    """Plan every level of a validated mask on the device (one upload of the mask)."""
    import time

    from . import _native

    median_filter_size = normalize_median_filter_size(median_filter_size)
    if mask is None:
        return PyramidPlan([None] * levels, [None] * levels, None, 0.0)
    if policy == "partial" and size > MAX_PARTIAL_SUBSET:
        raise ValueError(f"roi_policy='partial' supports subset sizes up to {MAX_PARTIAL_SUBSET} pixels")
    lib = _native.library()
    started = time.perf_counter()
    h, w = mask.shape
    raw = np.ascontiguousarray(mask, dtype=np.uint8)
    handle = ctypes.c_void_p()
    _native.check(lib.ldic_roi_planner_create(raw.ctypes.data, w, h, levels, device, ctypes.byref(handle)))
    try:
        opts = _native.RoiPlanOptions(size, stride, median_filter_size, 1 if policy == "partial" else 0,
                                      1 if truncation == "visibility" else 0)
        plans = [_plan_level(lib, handle, k, opts, policy, truncation) for k in range(levels)]
        links = [None] * levels
        temporal_links = None
        if plans[0].effective:
            for k in range(levels - 1):
                links[k] = _links(lib, handle, k + 1, k, plans[k].selected.shape)
            if temporal:
                temporal_links = _links(lib, handle, 0, levels - 1, plans[-1].selected.shape)
    finally:
        lib.ldic_roi_planner_destroy(handle)
    return PyramidPlan(plans, links, temporal_links, (time.perf_counter() - started) * 1e3)


def _links(lib, handle, source: int, destination: int, shape) -> np.ndarray:
    from . import _native

    out = np.empty(shape, np.uint32)
    _native.check(lib.ldic_roi_planner_seed_links(handle, source, destination, out.ctypes.data))
    return immutable(out)


def _plan_level(lib, handle, level: int, opts, policy: str, truncation: str) -> RoiPlan:
    from . import _native

    w, h, every = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    _native.check(lib.ldic_roi_planner_mask(handle, level, ctypes.byref(w), ctypes.byref(h), None, ctypes.byref(every)))
    mask = np.empty((h.value, w.value), np.uint8)
    _native.check(lib.ldic_roi_planner_mask(handle, level, None, None, mask.ctypes.data, None))
    mask = mask.astype(bool)
    S, D = opts.subset_size, opts.subset_stride
    gw, gh = 1 + (w.value - S) // D, 1 + (h.value - S) // D
    selected = np.empty((gh, gw), np.uint8)
    bits = np.empty((gh, gw), np.uint32)
    wide = opts.median_filter_size > 3
    words = np.empty((gh * gw, 4), np.float32) if wide else None
    coverage = np.empty((gh, gw), np.float32)
    count = ctypes.c_int32()
    _native.check(lib.ldic_roi_planner_plan(handle, level, ctypes.byref(opts), selected.ctypes.data, bits.ctypes.data,
                                            None if words is None else words.ctypes.data, coverage.ctypes.data,
                                            ctypes.byref(count)))
    partial = (None,) * 4
    if policy == "partial":
        n, nwords = count.value, (S * S + 31) // 32
        indices = np.empty(n, np.uint32)
        support = np.empty((n, nwords), np.uint32)
        boxes = np.empty((n, 4), np.float32)
        if n:
            _native.check(lib.ldic_roi_planner_partial(handle, level, indices.ctypes.data, support.ctypes.data,
                                                       boxes.ctypes.data))
        partial = (immutable(indices), immutable(support), immutable(boxes), immutable(coverage))
    chosen = selected.astype(bool)
    return RoiPlan(immutable(mask), immutable(chosen), immutable(bits), mask_digest(mask), not every.value,
                   int(chosen.sum()), policy, truncation, partial[0], None if words is None else immutable(words),
                   partial[1], partial[2], partial[3])
