"""Operator-level API mirroring src/cuMPM/dev/electric_field.py of the reference.

``MonodisperseEwaldElectricField`` applies the Ewald operator once
(``calculate``) to dipoles uploaded from the host and reads the field back,
with the reference's split helpers ``compute_real_field`` and
``compute_inverse_field`` (:764-778).  Device memory is owned by
libcumpm_amd; dipoles and fields are [particle][component] complex128.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _capi as C
from .._capi import check
from ..solver import PrecisionMode


class FieldCalcMode:
    SOLVER_AX = 0
    INTERACTION_FIELD = 1


def _parse_precision(precision):
    if precision is None:
        return PrecisionMode.AUTO
    if isinstance(precision, str):
        p = precision.lower()
        if p in ("auto",):
            return PrecisionMode.AUTO
        if p in ("mixed", "fp32", "single"):
            return PrecisionMode.MIXED
        if p in ("double", "fp64"):
            return PrecisionMode.DOUBLE
        raise ValueError(f"Unknown precision mode string: '{precision}'. Must be 'auto', 'mixed', or 'double'.")
    return precision


def _vec(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))


class MonodisperseEwaldElectricField:
    """Monodisperse spectral-Ewald operator (dev/electric_field.py:430-842 of the reference)."""

    def __init__(self, box_x, box_y, box_z, errortol, xi, calc_inter_dipole, radius=1.0,
                 solve_quadrupoles=False, quad_idxs=None, precision=None, *, grid=None, support=0):
        mode = calc_inter_dipole
        if isinstance(calc_inter_dipole, bool):
            mode = FieldCalcMode.SOLVER_AX if calc_inter_dipole else FieldCalcMode.INTERACTION_FIELD
        if mode != FieldCalcMode.SOLVER_AX:
            raise NotImplementedError("INTERACTION_FIELD mode (near fields) is not provided by this library")
        self._prec = _parse_precision(precision)
        d = C.OpDesc()
        d.box[:] = [float(box_x), float(box_y), float(box_z)]
        d.errortol = float(errortol)
        d.xi = float(xi)
        if grid is not None:
            d.grid[:] = [int(g) for g in grid]
        d.support = int(support)
        self._h = self._create(d)
        self._solve_quads = bool(solve_quadrupoles)
        self._Q = 0
        self._quad = None     # host copy [Q][5] complex
        if self._solve_quads:
            if quad_idxs is None or len(quad_idxs) == 0:
                check(C.load().mpm_op_set_quads(self._h, None, -1))
            else:
                idx = (ctypes.c_int * len(quad_idxs))(*[int(i) for i in quad_idxs])
                check(C.load().mpm_op_set_quads(self._h, idx, len(quad_idxs)))
        self._box = (float(box_x), float(box_y), float(box_z))
        self._errortol = float(errortol)
        self._xi = float(xi)
        self._N = 0
        self._dip = None      # host copy [N][3] complex
        self._sc = None
        self._epoint = None
        self._xbuf = self._ybuf = None

    def _create(self, d):
        h = ctypes.c_void_p()
        check(C.load().mpm_op_create(ctypes.byref(d), ctypes.byref(h)))
        return h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and C._lib is not None:
            C._lib.mpm_op_destroy(h)
            self._h = None

    # ---- properties ----
    def _info(self):
        info = C.OpInfo()
        check(C.load().mpm_op_get_info(self._h, ctypes.byref(info)))
        return info

    @property
    def is_using_recip_fp32(self) -> bool:
        return self._prec != PrecisionMode.DOUBLE

    @property
    def recip_precision_mode(self):
        return self._prec

    @property
    def box_x(self):
        return self._box[0]

    @property
    def box_y(self):
        return self._box[1]

    @property
    def box_z(self):
        return self._box[2]

    @property
    def errortol(self):
        return self._errortol

    @property
    def xi(self):
        return self._xi

    @property
    def rc(self):
        return self._info().rc

    @property
    def mode(self):
        return FieldCalcMode.SOLVER_AX

    @property
    def calc_inter_dipole(self):
        return True

    @property
    def self_coef(self):
        return self._info().self_coef_ewald

    @property
    def num_particles(self):
        return self._N

    @property
    def num_field_points(self):
        return self._N

    @property
    def num_pairs(self):
        return int(self._info().num_pairs)

    @property
    def num_offsets(self):
        return int(self._info().stencil_width) ** 3

    @property
    def num_spread(self):
        return self._N * self.num_offsets

    @property
    def num_contract(self):
        return self._N * self.num_offsets

    @property
    def table_size(self):
        return 9001

    @property
    def solve_quadrupoles(self):
        return self._solve_quads

    @property
    def num_quads(self):
        return self._Q

    @property
    def num_grid(self):
        return np.array(list(self._info().grid), dtype=np.int32)

    @property
    def grid_spacing(self):
        return np.array(list(self._info().grid_spacing))

    @property
    def spectral_split(self):
        return np.array(list(self._info().spectral_split))

    # ---- updates ----
    def update_particle_coordinates(self, x_part, y_part, z_part):
        x, y, z = _vec(x_part), _vec(y_part), _vec(z_part)
        if not (x.size == y.size == z.size):
            raise ValueError("Base_Electric_Field: coordinate vectors must have equal size.")
        xyz = np.ascontiguousarray(np.stack([x, y, z], axis=1))
        check(C.load().mpm_op_set_positions(self._h, C.dptr(xyz), x.size))
        if x.size != self._N:
            self._N = x.size
            self._Q = int(self._info().num_quads)
            self._dip = np.zeros((self._N, 3), np.complex128)
            self._quad = np.zeros((self._Q, 5), np.complex128)
            self._sc = np.zeros(self._N, np.complex128)
            self._epoint = np.zeros(3 * self._N + 5 * self._Q, np.complex128)
            self._xbuf = C.DeviceBuffer(self._epoint.nbytes)
            self._ybuf = C.DeviceBuffer(self._epoint.nbytes)
        self._push_self()

    def update_dipoles(self, dip_x, dip_y, dip_z):
        self.update_dipoles_complex(dip_x, np.zeros(self._N), dip_y, np.zeros(self._N), dip_z, np.zeros(self._N))

    def update_dipoles_complex(self, dip_xr, dip_xi, dip_yr, dip_yi, dip_zr, dip_zi):
        comps = [_vec(v) for v in (dip_xr, dip_xi, dip_yr, dip_yi, dip_zr, dip_zi)]
        if any(c.size != self._N for c in comps):
            raise ValueError(
                "Base_Electric_Field: complex dipole component vectors must match the allocated number of particles.")
        self._dip = np.stack([comps[0] + 1j * comps[1], comps[2] + 1j * comps[3], comps[4] + 1j * comps[5]], axis=1)

    def set_dipoles(self, p):
        """Convenience: p of shape (N, 3) complex."""
        p = np.asarray(p, np.complex128).reshape(self._N, 3)
        self._dip = p.copy()

    def update_quadrupoles(self, quad_1, quad_2, quad_3, quad_4, quad_5):
        z = np.zeros(self._Q)
        self.update_quadrupoles_complex(quad_1, z, quad_2, z, quad_3, z, quad_4, z, quad_5, z)

    def update_quadrupoles_complex(self, quad_1r, quad_1i, quad_2r, quad_2i, quad_3r, quad_3i, quad_4r, quad_4i,
                                   quad_5r, quad_5i):
        """Components (Qxx, Qxy, Qxz, Qyy, Qyz) of the quadrupole subset, Qzz = -(Qxx + Qyy)."""
        comps = [_vec(v) for v in (quad_1r, quad_1i, quad_2r, quad_2i, quad_3r, quad_3i, quad_4r, quad_4i,
                                   quad_5r, quad_5i)]
        if any(c.size != self._Q for c in comps):
            raise ValueError("quadrupole component vectors must match the number of quadrupoles")
        self._quad = np.stack([comps[2 * i] + 1j * comps[2 * i + 1] for i in range(5)], axis=1)

    def set_quadrupoles(self, q):
        """Convenience: q of shape (Q, 5) complex."""
        self._quad = np.asarray(q, np.complex128).reshape(self._Q, 5).copy()

    def set_self_coef(self, self_coef_r, self_coef_i=None):
        if self_coef_i is None:
            if np.iterable(self_coef_r):
                sc = _vec(self_coef_r).astype(np.complex128)
            else:
                sc = np.full(self._N, complex(float(self_coef_r), 0.0))
        else:
            r = np.broadcast_to(np.asarray(self_coef_r, np.float64), (self._N,))
            i = np.broadcast_to(np.asarray(self_coef_i, np.float64), (self._N,))
            sc = (r + 1j * i).astype(np.complex128)
        if sc.size != self._N:
            raise ValueError("self coefficients must match the number of particles")
        self._sc = np.ascontiguousarray(sc)
        self._push_self()

    def _push_self(self):
        if self._N and self._sc is not None:
            check(C.load().mpm_op_set_self_coef(self._h, C.dptr(self._sc.view(np.float64)), 1))

    def set_bloch_wavevector(self, kx: float, ky: float):
        check(C.load().mpm_op_set_bloch(self._h, float(kx), float(ky)))

    # ---- evaluation ----
    def _apply(self, parts):
        if not self._N:
            raise RuntimeError("electricField: particle coordinates have not been set")
        x = np.concatenate([self._dip.reshape(-1), self._quad.reshape(-1)])
        self._xbuf.upload(np.ascontiguousarray(x))
        check(C.load().mpm_op_apply(self._h, self._xbuf.ptr, self._ybuf.ptr, 1, int(self._prec), int(parts), None))
        out = np.zeros_like(x)
        self._ybuf.download(out)
        self._epoint = out

    def calculate(self):
        self._apply(C.PART_ALL)

    def clear_epoint(self):
        if self._epoint is not None:
            self._epoint[...] = 0

    def real_space(self, E_point_ptr=None):
        self._apply(C.PART_REAL)

    def compute_real_field(self):
        self.clear_epoint()
        self.real_space(None)
        return self.get_epoint_host()

    def compute_inverse_field(self):
        self.calculate()
        total = self.get_epoint_host()
        real = self.compute_real_field()
        if isinstance(total, tuple):
            return total[0] - real[0], total[1] - real[1]
        return total - real

    def get_epoint_host(self, reshape=True):
        """Field at the particles, (N, 3); with quadrupoles the pair ((N, 3), (Q, 5)) of field and
        field gradient (dev/electric_field.py:816-842 of the reference)."""
        e = self._epoint.copy()
        if not reshape:
            return e
        if self._solve_quads and self._Q > 0:
            return e[:3 * self._N].reshape(self._N, 3), e[3 * self._N:].reshape(self._Q, 5)
        return e.reshape(self._N, 3)

    # ---- host copies of the operator's tables and per-frame precalculations ----
    def _table(self, which):
        size = C._i64()
        lib = C.load()
        check(lib.mpm_op_get_pair_table(self._h, int(which), None, None, 0, ctypes.byref(size)))
        r = np.zeros(size.value)
        v = np.zeros(size.value)
        check(lib.mpm_op_get_pair_table(self._h, int(which), C.dptr(r), C.dptr(v), size.value, ctypes.byref(size)))
        return r, v

    def get_real_space_tables_host(self):
        """(r_table, field_dip_1, field_dip_2): samples 0, 1.000 ... 9.999 and the perpendicular /
        parallel pair functions (entry 0 the self term)."""
        r, perp = self._table(0)
        return r, perp, self._table(1)[1]

    def _neighbors(self):
        if not self._N:
            raise RuntimeError("electricField: particle coordinates have not been set")
        m = C._i64()
        lib = C.load()
        check(lib.mpm_op_get_neighbors(self._h, None, None, None, None, 0, ctypes.byref(m)))
        M = m.value
        start = np.zeros(self._N + 1, np.int64)
        j = np.zeros(max(M, 1), np.int32)
        perp = np.zeros(max(M, 1))
        para = np.zeros(max(M, 1))
        check(lib.mpm_op_get_neighbors(self._h, start.ctypes.data_as(ctypes.POINTER(C._i64)),
                                       j.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), C.dptr(perp), C.dptr(para),
                                       M, ctypes.byref(m)))
        return start, j[:M], perp[:M], para[:M]

    @property
    def max_neighbors(self):
        start = self._neighbors()[0]
        return int(np.diff(start).max()) if self._N else 0

    def get_neighbor_list_host(self):
        """(list, counts): row i of list (length max_neighbors, padded with -1) holds the
        neighbours j != i of particle i within rc (the same pairs as the reference; the order
        within a row may differ)."""
        start, j, _, _ = self._neighbors()
        counts = np.diff(start).astype(np.int32)
        width = int(counts.max()) if counts.size else 0
        out = np.full((self._N, width), -1, np.int32)
        for i in range(self._N):
            out[i, :counts[i]] = j[start[i]:start[i + 1]]
        return out.reshape(-1), counts

    def get_real_space_precalcs_host(self):
        """(self_perp, perp, para): the self term and each neighbour pair's perpendicular /
        parallel values at d_eff = max(d, 1), in get_neighbor_list_host's row order."""
        _, _, perp, para = self._neighbors()
        return float(self._info().self_perp), perp, para

    def compute_neighbor_list(self):
        """The neighbour list, stencils and pair values are built by update_particle_coordinates."""

    compute_precalculations = compute_real_space_tables = compute_neighbor_list
    spread_precalcs = contract_precalcs = real_space_precalcs = compute_neighbor_list

    def get_dipoles_complex_host(self):
        p = self._dip
        return (p[:, 0].real.copy(), p[:, 0].imag.copy(), p[:, 1].real.copy(), p[:, 1].imag.copy(),
                p[:, 2].real.copy(), p[:, 2].imag.copy())


class PolydisperseEwaldElectricField(MonodisperseEwaldElectricField):
    """Polydisperse spectral-Ewald operator: one radius per particle (nondimensional).
    Reference: dev/electric_field.py (PolydisperseEwaldElectricField) over
    polydisperse_ewald_electric_field.cu; radii within 1e-5 share a class."""

    def __init__(self, box_x, box_y, box_z, errortol, xi, calc_inter_dipole, particle_radii,
                 solve_quadrupoles=False, quad_idxs=None, precision=None, *, grid=None, support=0):
        self._radii = _vec(particle_radii)
        if self._radii.size == 0:
            raise ValueError("particle_radii is empty")
        super().__init__(box_x, box_y, box_z, errortol, xi, calc_inter_dipole, 1.0, solve_quadrupoles,
                         quad_idxs, precision, grid=grid, support=support)

    def _create(self, d):
        h = ctypes.c_void_p()
        check(C.load().mpm_op_create_poly(ctypes.byref(d), C.dptr(self._radii), self._radii.size,
                                          ctypes.byref(h)))
        return h

    @property
    def radii(self):
        return self._radii.copy()

    @property
    def num_unique_radii(self):
        return int(self._info().num_unique_radii)

    @property
    def num_offsets(self):
        return int(self._info().support) ** 3

    @property
    def table_size(self):
        return int(np.ceil((max(self.rc, 1.001) - 1.0) / 0.001)) + 1

    def update_particle_coordinates(self, x_part, y_part, z_part):
        if _vec(x_part).size != self._radii.size:
            raise ValueError("the polydisperse operator needs one radius per particle")
        super().update_particle_coordinates(x_part, y_part, z_part)
