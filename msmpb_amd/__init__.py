"""MI355X-native B-spline multilevel summation (MSM) for periodic Coulomb
energy and forces: a drop-in for the FF_energy path of hkayabilisim/msmpb.

The computation lives in libmsmpb_hip.so (HIP kernels for gfx950 behind the
reference's forcefield.h C ABI); this package binds it for Python and
PyTorch.
"""
from .msm import MSM, MSMError, msm_energy  # noqa: F401

__all__ = ["MSM", "MSMError", "msm_energy"]
