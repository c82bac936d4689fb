"""Target reconstruction modes (reference: live_dic/interpolation.py:1-27).

``bspline`` is the Gaussian plus cubic B-spline reconstruction; ``linear`` and
``cubic`` (Keys, a = -0.5) interpolate the unfiltered image directly.
"""

from __future__ import annotations

INTERPOLATION_MODES = ("bspline", "linear", "cubic")
INTERPOLATION_CODES = {"bspline": 0, "linear": 1, "cubic": 2}


def normalize_interpolation(value: str) -> str:
    if not isinstance(value, str) or value.strip().lower() not in INTERPOLATION_MODES:
        raise ValueError("interpolation must be 'bspline', 'linear', or 'cubic'")
    return value.strip().lower()
