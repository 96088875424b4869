"""MI355X-native Green--Laplace bound-wave reconstruction (gfx950 HIP kernels).

Drop-in for the reference API: ``gl_spectral_coefficients`` /
``gl_spectral_surface`` (MATLAB two-stage API) and
``reconstruct_directional_field`` (Python package).
"""

from .api import (
    DirectionalFieldResult,
    GreenLaplaceError,
    gl_spectral_coefficients,
    gl_pure_sum_order4,
    gl_spectral_surface,
    reconstruct_directional_field,
)
from .plan import build_plan, gauss_laguerre_rule

__all__ = [
    "DirectionalFieldResult",
    "GreenLaplaceError",
    "gl_spectral_coefficients",
    "gl_pure_sum_order4",
    "gl_spectral_surface",
    "reconstruct_directional_field",
    "build_plan",
    "gauss_laguerre_rule",
]
__version__ = "0.1.0"
