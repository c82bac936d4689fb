"""ctypes binding of liblivedic.so (C ABI declared in include/livedic.h).

The library is built in-tree by ``live-dic_amd/build.py``.  There is no
fallback: a missing library or device is an ``ICGNError``.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

from .engine import ICGNError

LIBRARY_PATH = Path(os.environ.get("LIVEDIC_LIBRARY", Path(__file__).resolve().parent / "liblivedic.so"))
MAX_LEVELS = 4


class Options(ctypes.Structure):
    _fields_ = [
        ("motion_model", ctypes.c_int32),
        ("interpolation", ctypes.c_int32),
        ("subset_size", ctypes.c_int32),
        ("subset_stride", ctypes.c_int32),
        ("spatial_weighting", ctypes.c_int32),
        ("gaussian_sigma", ctypes.c_float),
        ("outlier_policy", ctypes.c_int32),
        ("median_filter_size", ctypes.c_int32),
        ("median_filter_scope", ctypes.c_int32),
        ("pyramid_levels", ctypes.c_int32),
        ("device_ordinal", ctypes.c_int32),
    ]


class LevelConfig(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("tolerance", ctypes.c_float), ("zncc_min", ctypes.c_float)]


class LevelRoi(ctypes.Structure):
    _fields_ = [
        ("selected", ctypes.c_void_p),
        ("neighbour_words", ctypes.c_void_p),
        ("seed_links", ctypes.c_void_p),
    ]


class Field(ctypes.Structure):
    _fields_ = [
        ("enabled", ctypes.c_int32),
        ("origin_x", ctypes.c_double),
        ("origin_y", ctypes.c_double),
        ("poly", ctypes.c_double * 12),
        ("wave", (ctypes.c_double * 4) * 2),
    ]


class RoiPlanOptions(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int32) for name in (
        "subset_size", "subset_stride", "median_filter_size", "partial", "visibility")]


class MedianOp(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int32) for name in (
        "grid_width", "grid_height", "subset_stride", "motion_model", "outlier_policy", "median_filter_size",
        "median_filter_scope", "seed_available")]


class Speckle(ctypes.Structure):
    _fields_ = [
        ("device_blobs", ctypes.c_void_p),
        ("device_cell_start", ctypes.c_void_p),
        ("cells_x", ctypes.c_int32),
        ("cells_y", ctypes.c_int32),
        ("cell", ctypes.c_int32),
        ("cell_origin", ctypes.c_float),
        ("sigma", ctypes.c_float),
        ("background", ctypes.c_float),
    ]


class Timing(ctypes.Structure):
    _fields_ = [
        ("upload_ms", ctypes.c_double),
        ("gpu_ms", ctypes.c_double),
        ("solve_ms", ctypes.c_double),
        ("readback_ms", ctypes.c_double),
        ("wall_ms", ctypes.c_double),
        ("frames", ctypes.c_int32),
        ("temporal_seeded", ctypes.c_int32),
        ("temporal_updates", ctypes.c_int32),
        ("velocity_predicted", ctypes.c_int32),
        ("levels_used", ctypes.c_int32),
        ("temporal_expired", ctypes.c_int32),
        ("accepted", ctypes.c_int64),
        ("replaced", ctypes.c_int64),
        ("sum_translation_x", ctypes.c_double),
        ("sum_translation_y", ctypes.c_double),
        ("sum_iterations", ctypes.c_double),
    ]


class Initialization(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("search_radius", ctypes.c_int32),
        ("fallback_zncc", ctypes.c_float),
        ("fallback_radius", ctypes.c_int32),
        ("temporal_max_age", ctypes.c_uint32),
        ("temporal_prediction", ctypes.c_int32),
        ("temporal_fine_correction", ctypes.c_int32),
        ("divergence_patience", ctypes.c_int32),
        ("divergence_min_iterations", ctypes.c_int32),
        ("adaptive_pyramid", ctypes.c_int32),
        ("adaptive_acceptance", ctypes.c_float),
        ("reseed_passes", ctypes.c_int32),
    ]


class NNConv(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "batch", "height", "width", "in_channels", "in_stride", "in_offset", "out_height", "out_width",
        "out_channels", "out_stride", "out_offset", "stride", "out_channel_tile")]


class NNCorr(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "batch", "height", "width", "channels", "radius", "f1_stride", "f1_offset", "f2_stride", "f2_offset",
        "out_stride", "out_offset")] + [("scale", ctypes.c_float)]


class NNSeed(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "grid_width", "grid_height", "subset_size", "subset_stride", "flow_width", "flow_height")] + [
        ("scale", ctypes.c_float), ("state_stride", ctypes.c_int32), ("gradients", ctypes.c_int32),
        ("selected", ctypes.c_void_p)]


INIT_ZERO = 0
INIT_ZNCC = 1
INIT_TEMPORAL = 2


# Every symbol include/livedic.h declares, with its argument types.
_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
SIGNATURES = {
    "ldic_abi_version": (ctypes.c_int, []),
    "ldic_last_error": (ctypes.c_char_p, []),
    "ldic_create": (ctypes.c_int, [ctypes.POINTER(Options), ctypes.POINTER(_P)]),
    "ldic_destroy": (None, [_P]),
    "ldic_device_name": (ctypes.c_int, [_P, ctypes.c_char_p, _I32]),
    "ldic_set_reference": (ctypes.c_int, [_P, _P, _I32, _I32, _I64, _P, _P, _P]),
    "ldic_level_shape": (ctypes.c_int, [_P, _I32, _P, _P, _P, _P]),
    "ldic_set_level_config": (ctypes.c_int, [_P, _I32, ctypes.POINTER(LevelConfig)]),
    "ldic_register": (ctypes.c_int, [_P, _P, _I64, _P, ctypes.POINTER(Timing)]),
    "ldic_register_pair": (ctypes.c_int, [_P, _P, _I64, _P, _I64, _P, _P, ctypes.POINTER(Timing)]),
    "ldic_register_pair_resident": (ctypes.c_int, [_P, _I32, _P, _P, _I64]),
    "ldic_set_slots": (ctypes.c_int, [_P, _I32]),
    "ldic_submit": (ctypes.c_int, [_P, _I32, _P, _I64]),
    "ldic_retire": (ctypes.c_int, [_P, _I32, _P, ctypes.POINTER(Timing)]),
    "ldic_register_seeded": (ctypes.c_int, [_P, _P, _I64, _P, _P, _I32, _P, _P, ctypes.POINTER(Timing)]),
    "ldic_register_resident": (ctypes.c_int, [_P, _I32, _P, _I64]),
    "ldic_set_initialization": (ctypes.c_int, [_P, ctypes.POINTER(Initialization)]),
    "ldic_register_seeded_resident": (ctypes.c_int, [_P, _I32, _P, _I64, _P, _P]),
    "ldic_export_search": (ctypes.c_int, [_P, _I32, _I32, _P, _P]),
    "ldic_reset_temporal": (ctypes.c_int, [_P]),
    "ldic_prepare_partial": (ctypes.c_int, [_P, _I32, _I32, _P, _P, _P, _P]),
    "ldic_roi_planner_create": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, ctypes.POINTER(ctypes.c_void_p)]),
    "ldic_roi_planner_mask": (ctypes.c_int, [_P, _I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32), _P,
                                             ctypes.POINTER(_I32)]),
    "ldic_roi_planner_plan": (ctypes.c_int, [_P, _I32, ctypes.POINTER(RoiPlanOptions), _P, _P, _P, _P,
                                             ctypes.POINTER(_I32)]),
    "ldic_roi_planner_partial": (ctypes.c_int, [_P, _I32, _P, _P, _P]),
    "ldic_roi_planner_seed_links": (ctypes.c_int, [_P, _I32, _I32, _P]),
    "ldic_roi_planner_destroy": (None, [_P]),
    "ldic_set_temporal_links": (ctypes.c_int, [_P, _P]),
    "ldic_prepare_reference_resident": (ctypes.c_int, [_P, _I32, _P, _I64]),
    "ldic_batch_create": (ctypes.c_int, [_P, _I32, _P, ctypes.POINTER(_P)]),
    "ldic_batch_tune": (ctypes.c_int, [_P]),
    "ldic_batch_launch": (ctypes.c_int, [_P, _I32]),
    "ldic_batch_wait": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double)]),
    "ldic_batch_describe": (ctypes.c_int, [_P, ctypes.c_char_p, _I32]),
    "ldic_batch_destroy": (None, [_P]),
    "ldic_op_zncc_search": (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _I32, _P, _I32, _P, _P, _P]),
    "ldic_slot_wait": (ctypes.c_int, [_P, _I32, ctypes.POINTER(Timing)]),
    "ldic_slot_output": (ctypes.c_int, [_P, _I32, _P]),
    "ldic_export_reference": (ctypes.c_int, [_P, _I32, _P, _P, _P]),
    "ldic_export_coefficients": (ctypes.c_int, [_P, _I32, _I32, _P]),
    "ldic_export_level_state": (ctypes.c_int, [_P, _I32, _I32, _P]),
    "ldic_op_reduce": (ctypes.c_int, [_P, _I32, _I32, _P]),
    "ldic_op_prefilter": (ctypes.c_int, [_P, _I32, _I32, _P]),
    "ldic_op_median": (ctypes.c_int, [ctypes.POINTER(MedianOp), _P, _P, _P, _P]),
    "ldic_render_speckle": (ctypes.c_int, [ctypes.POINTER(Speckle), ctypes.POINTER(Field), _I32, _I32,
                                           ctypes.c_float, ctypes.c_uint64, _P, _P, _P]),
    "ldic_render_speckle_batch": (ctypes.c_int, [ctypes.POINTER(Speckle), _P, _P, _I32, _I32, _I32,
                                                 ctypes.c_float, ctypes.c_uint64, _P, _P]),
    "ldic_reduce_device": (ctypes.c_int, [_P, _I32, _I32, _I32, _P, _P]),
    "ldic_slot_stream": (ctypes.c_int, [_P, _I32, ctypes.POINTER(_P)]),
    "ldic_nn_conv_forward": (ctypes.c_int, [ctypes.POINTER(NNConv), _P, _P, _P, _I32, _P, _P, _P]),
    "ldic_nn_conv_forward_split": (ctypes.c_int, [ctypes.POINTER(NNConv), _P, _P, _P, _I32, _P, _P, _P, _I32, _P]),
    "ldic_nn_conv_forward_split_static": (ctypes.c_int, [ctypes.POINTER(NNConv), _P, _P, _P, _I32, _P, _P, _I32,
                                                         _P]),
    "ldic_nn_conv_backward": (ctypes.c_int, [ctypes.POINTER(NNConv), _P, _P, _P, _P, _I32, _I32, _I32, _P, _P,
                                             _P]),
    "ldic_nn_transpose_weights": (ctypes.c_int, [_I32, _I32, _P, _P, _P]),
    "ldic_nn_leaky_backward": (ctypes.c_int, [ctypes.c_size_t, _P, _P, _P]),
    "ldic_nn_corr_forward": (ctypes.c_int, [ctypes.POINTER(NNCorr), _P, _P, _P, _P]),
    "ldic_nn_corr_backward": (ctypes.c_int, [ctypes.POINTER(NNCorr), _P, _P, _P, _I32, _I32, _P, _I32, _I32, _P,
                                             _I32, _I32, _P]),
    "ldic_nn_warp_forward": (ctypes.c_int, [_I32, _I32, _I32, _I32, _P, _P, _P, _P]),
    "ldic_nn_warp_backward": (ctypes.c_int, [_I32, _I32, _I32, _I32, _P, _P, _P, _P]),
    "ldic_nn_upsample_flow": (ctypes.c_int, [_I32, _I32, _I32, _I32, _I32, _P, _P, _P, _I32, _I32, _P]),
    "ldic_nn_copy_slice": (ctypes.c_int, [ctypes.c_size_t, _I32, _P, _I32, _I32, _P, _I32, _I32, _P]),
    "ldic_nn_add_slice": (ctypes.c_int, [ctypes.c_size_t, _I32, _P, _I32, _I32, _P, _I32, _I32, _P]),
    "ldic_nn_images_to_bf16": (ctypes.c_int, [ctypes.c_size_t, _P, _P, _P]),
    "ldic_nn_to_bf16": (ctypes.c_int, [ctypes.c_size_t, _P, _P, _P]),
    "ldic_nn_flow_loss": (ctypes.c_int, [_I32, _P, _P, ctypes.c_float, _P, _P, _P]),
    "ldic_nn_adam": (ctypes.c_int, [ctypes.c_size_t, _P, _P, _P, _P, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float, _I32, _P, _P]),
    "ldic_nn_seed_from_flow": (ctypes.c_int, [ctypes.POINTER(NNSeed), _P, _P, _P, _P]),
    "ldic_device_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_P)]),
    "ldic_device_free": (ctypes.c_int, [_P]),
    "ldic_copy_to_device": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "ldic_copy_to_host": (ctypes.c_int, [_P, _P, ctypes.c_size_t]),
    "ldic_device_synchronize": (ctypes.c_int, []),
    "ldic_device_zero": (ctypes.c_int, [_P, ctypes.c_size_t, _P]),
    "ldic_device_copy_async": (ctypes.c_int, [_P, _P, ctypes.c_size_t, _P]),
    "ldic_stream_synchronize": (ctypes.c_int, [_P]),
}

_lib = None


def library() -> ctypes.CDLL:
    """Load liblivedic.so once and declare its signatures."""
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            raise ICGNError(
                f"{LIBRARY_PATH} is missing; build it with `python live-dic_amd/build.py` (hipcc, gfx950)"
            )
        try:
            lib = ctypes.CDLL(str(LIBRARY_PATH))
        except OSError as error:
            raise ICGNError(f"cannot load {LIBRARY_PATH}: {error}") from error
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = lib
    return _lib


class DeviceBuffer:
    """Device allocation owned by Python (freed on close or collection)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(library().ldic_device_alloc(self.nbytes, ctypes.byref(self.ptr)))

    @property
    def address(self) -> int:
        return int(self.ptr.value or 0)

    def upload(self, array) -> None:
        if array.nbytes > self.nbytes:
            raise ValueError("array larger than the device buffer")
        check(library().ldic_copy_to_device(self.ptr, array.ctypes.data, array.nbytes))

    def download(self, array) -> None:
        if array.nbytes > self.nbytes:
            raise ValueError("array larger than the device buffer")
        check(library().ldic_copy_to_host(array.ctypes.data, self.ptr, array.nbytes))

    def close(self) -> None:
        if self.ptr.value:
            library().ldic_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


BF16 = "bfloat16"


class DeviceArray(DeviceBuffer):
    """A C-ordered device array of a fixed shape and element type (``BF16`` is stored as 16-bit
    patterns and read back as float32).  Transfers run on a HIP stream handle (0 or None: the
    null stream); host reads wait for the device first."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(n) for n in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = dtype if dtype == BF16 else np.dtype(dtype)
        self.itemsize = 2 if dtype == BF16 else self.dtype.itemsize
        self.size = int(np.prod(self.shape)) if self.shape else 1
        super().__init__(self.size * self.itemsize)

    @classmethod
    def zeros(cls, shape, dtype, stream: int | None = 0) -> "DeviceArray":
        """A cleared array; the clear has completed on return, so any stream may use it."""
        out = cls(shape, dtype)
        out.zero(stream)
        check(library().ldic_stream_synchronize(stream or None))
        return out

    @classmethod
    def from_numpy(cls, array, dtype=None) -> "DeviceArray":
        host = np.ascontiguousarray(array)
        out = cls(host.shape, dtype or host.dtype)
        out.upload(host)
        return out

    def numel(self) -> int:
        return self.size

    def data_ptr(self) -> int:
        return self.address

    def zero(self, stream: int | None = 0) -> None:
        check(library().ldic_device_zero(self.ptr, self.nbytes, stream or None))

    def copy_from(self, source, stream: int | None = 0) -> None:
        """Copy a same-sized device array (or any object with ``data_ptr()``) or a host array."""
        if isinstance(source, np.ndarray):
            self.upload(source)
            return
        check(library().ldic_device_copy_async(self.ptr, source.data_ptr(), self.nbytes, stream or None))

    def upload(self, array) -> None:
        host = np.ascontiguousarray(array)
        if self.dtype == BF16:
            if host.dtype != np.uint16:
                # float32 -> bfloat16 bit patterns, rounded to nearest even.
                u = np.ascontiguousarray(host, np.float32).view(np.uint32).astype(np.uint64)
                host = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
        elif host.dtype != self.dtype:
            host = host.astype(self.dtype)
        if host.nbytes != self.nbytes:
            raise ValueError(f"host array of {host.nbytes} bytes for a device array of {self.nbytes}")
        super().upload(host)

    def numpy(self) -> np.ndarray:
        """Wait for the device, then copy to a host array (float32 for ``BF16``)."""
        check(library().ldic_device_synchronize())
        if self.dtype == BF16:
            bits = np.empty(self.shape, np.uint16)
            self.download(bits)
            return (bits.astype(np.uint32) << 16).view(np.float32)
        out = np.empty(self.shape, self.dtype)
        self.download(out)
        return out

    def item(self) -> float:
        return float(self.numpy().reshape(-1)[0])


def check(code: int) -> None:
    if code != 0:
        message = library().ldic_last_error()
        raise ICGNError((message or b"").decode("utf-8", "replace") or f"native error {code}")


def ptr(array) -> int | None:
    return None if array is None else array.ctypes.data
