"""Near field: the intensity |E0 + E_ind|^2 that given dipoles produce at arbitrary points.

Same interface as the reference's ``cuMPM.Near_Field`` (src/cuMPM/near_field.py:5-126) over
the C ABI (``mpm_near_field_*``); the field is the periodic Ewald field of the operator, with
the field points taking part as sites that carry no dipole.  Addition: ``precision``
("auto" / None as the reference's default FP32 reciprocal grid, "mixed", or "double").
"""
import ctypes

import numpy as np

from . import _capi as C
from ._capi import check

_FIELD = {"auto": C.FIELD_AUTO, "monodisperse": C.FIELD_MONODISPERSE, "polydisperse": C.FIELD_POLYDISPERSE,
          "direct": C.FIELD_DIRECT}


def _coords(arr, what):
    if arr is None:
        raise ValueError(f"{what} can't be set to None")
    a = np.asarray(arr, dtype=float)
    return [np.ascontiguousarray(a[:, d]) for d in range(3)]


class Near_Field:
    r"""Local field intensity :math:`\lvert \mathbf{E}_0 + \mathbf{E}_{\text{ind}} \rvert^2` at target
    points from complex point dipoles under periodic boundary conditions."""

    def __init__(self, box, E0, radius=1.0, dip=None, dip_pos=None, field_points=None, xi=0.5, errortol=1e-3,
                 field_type="auto", *, precision=None):
        self.box = [float(x) for x in box]
        self.E0 = [complex(x) for x in E0]
        if len(self.box) != 3:
            raise ValueError("Near_Field box must have length 3.")
        if len(self.E0) != 3:
            raise ValueError("Near_Field E0 must have length 3.")
        self.radius_list = [float(r) for r in radius] if np.iterable(radius) else [float(radius)]
        self.xi = float(xi)
        self.errortol = float(errortol)
        if field_type not in _FIELD:
            raise RuntimeError(f"Unknown field_type: {field_type}")
        self.field_type = field_type
        box_a = np.ascontiguousarray(self.box, np.float64)
        e0 = np.ascontiguousarray(self.E0, np.complex128)
        rad = np.ascontiguousarray(self.radius_list, np.float64)
        h = ctypes.c_void_p()
        check(C.load().mpm_near_field_create(C.dptr(box_a), C.dptr(e0.view(np.float64)), C.dptr(rad), rad.size,
                                             self.xi, self.errortol, _FIELD[field_type], ctypes.byref(h)))
        self._h = h
        self._M = 0
        prec = {None: C.PREC_AUTO, "auto": C.PREC_AUTO, "mixed": C.PREC_MIXED, "double": C.PREC_DOUBLE}
        if precision not in prec:
            raise ValueError(f"precision must be None, 'auto', 'mixed' or 'double', got {precision!r}")
        check(C.load().mpm_near_field_set_precision(h, prec[precision]))
        if dip is not None:
            self.set_dipoles(dip)
        if dip_pos is not None:
            self.set_dipole_positions(dip_pos)
        if field_points is not None:
            self.set_field_points(field_points)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and C._lib is not None:
            C._lib.mpm_near_field_destroy(h)
            self._h = None

    def set_dipoles(self, dip):
        """Complex dipoles, shape (num_particles, 3)."""
        if dip is None:
            raise ValueError("dipoles can't be set to None")
        d = np.ascontiguousarray(np.asarray(dip, dtype=complex).reshape(-1))
        if d.size % 3:
            raise ValueError("Dipoles vector size must be a multiple of 3.")
        check(C.load().mpm_near_field_set_dipoles(self._h, C.dptr(d.view(np.float64)), d.size // 3))

    def set_dipole_positions(self, dip_pos):
        """Dipole positions, shape (num_particles, 3)."""
        x, y, z = _coords(dip_pos, "dipole positions")
        check(C.load().mpm_near_field_set_dipole_positions(self._h, C.dptr(x), C.dptr(y), C.dptr(z), x.size))

    def set_field_points(self, field_points):
        """Points at which the field is evaluated, shape (num_field_points, 3)."""
        x, y, z = _coords(field_points, "field points")
        check(C.load().mpm_near_field_set_field_points(self._h, C.dptr(x), C.dptr(y), C.dptr(z), x.size))
        self._M = x.size

    def calculate(self):
        """Intensity |E0 + E_ind|^2 at the field points, shape (num_field_points,)."""
        out = np.zeros(max(self._M, 1))
        check(C.load().mpm_near_field_calculate(self._h, C.dptr(out)))
        return out[:self._M]
