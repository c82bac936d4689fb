"""MI355X-native IC-GN digital image correlation with the ``live_dic`` supervisor API.

NumPy reference/target ``uint8`` images and an optional ROI mask go in; dense
NumPy displacement, gradient, curvature, status and iteration fields come out.
All numerical work runs in hand-written gfx950 HIP kernels (``liblivedic.so``)
called through ctypes.  Public names follow the reference package
(live_dic/__init__.py:1-41); the browser engine is not part of this build.
"""

from .engine import PARAMETER_ORDER, ICGNError, ImageInput, ReferenceInfo, RegistrationResult
from .full_pyramid_backend import DEFAULT_FULL_PYRAMID_LEVELS, ICGNWGPUFullPyramid
from .hip_backend import DEFAULT_IN_FLIGHT_FRAMES, DEFAULT_PYRAMID_LEVELS, ICGNWGPU
from .motion import AFFINE_PARAMETER_ORDER, MOTION_MODELS, QUADRATIC_PARAMETER_ORDER
from .outlier import OUTLIER_POLICIES

SEEDED_SOLVE_API_VERSION = 1

__all__ = [
    "AFFINE_PARAMETER_ORDER",
    "DEFAULT_FULL_PYRAMID_LEVELS",
    "DEFAULT_IN_FLIGHT_FRAMES",
    "DEFAULT_PYRAMID_LEVELS",
    "ICGNError",
    "ICGNWGPU",
    "ICGNWGPUFullPyramid",
    "ImageInput",
    "MOTION_MODELS",
    "OUTLIER_POLICIES",
    "PARAMETER_ORDER",
    "QUADRATIC_PARAMETER_ORDER",
    "ReferenceInfo",
    "RegistrationResult",
    "SEEDED_SOLVE_API_VERSION",
]
