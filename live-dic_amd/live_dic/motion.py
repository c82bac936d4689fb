"""Motion models and parameter orders (reference: live_dic/motion.py:1-56)."""

from __future__ import annotations

from typing import Final

MOTION_MODELS: Final = ("affine", "translation", "quadratic")
MOTION_MODEL_CODES: Final = {"affine": 0, "translation": 1, "quadratic": 2}
AFFINE_PARAMETER_ORDER: Final = ("tx", "du_dx", "du_dy", "ty", "dv_dx", "dv_dy")
QUADRATIC_PARAMETER_ORDER: Final = (
    *AFFINE_PARAMETER_ORDER,
    "d2u_dx2",
    "d2u_dxdy",
    "d2u_dy2",
    "d2v_dx2",
    "d2v_dxdy",
    "d2v_dy2",
)
_ERROR = "motion_model must be 'affine', 'translation', or 'quadratic'"


def normalize_motion_model(value: str) -> str:
    if not isinstance(value, str) or value.strip().lower() not in MOTION_MODELS:
        raise ValueError(_ERROR)
    return value.strip().lower()


def parameter_order(motion_model: str) -> tuple[str, ...]:
    if normalize_motion_model(motion_model) == "quadratic":
        return QUADRATIC_PARAMETER_ORDER
    return AFFINE_PARAMETER_ORDER


def parameter_count(motion_model: str) -> int:
    return len(parameter_order(motion_model))


def state_float_count(motion_model: str) -> int:
    """Packed per-subset state stride: 8 floats, or 16 for quadratic."""
    return 16 if normalize_motion_model(motion_model) == "quadratic" else 8
