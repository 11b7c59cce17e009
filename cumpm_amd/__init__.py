"""cumpm_amd — MI355X-native mutual-polarization (MPM) dipole solver.

Python API of the reference package cuMPM (``dipole_solver``,
``drude_dielectric``, ``Near_Field``, ``EELS``, ``dev`` operator API) over libcumpm_amd, a C-ABI
library of hand-written HIP kernels for gfx950 and rocFFT.
"""
from .solver import dipole_solver, drude_dielectric, PrecisionMode
from .near_field import Near_Field
from .eels import EELS
from . import dev


__all__ = ["dipole_solver", "drude_dielectric", "PrecisionMode", "Near_Field", "EELS", "dev"]
