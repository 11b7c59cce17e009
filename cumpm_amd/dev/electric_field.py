"""Operator-level API mirroring src/cuMPM/dev/electric_field.py of the reference.

``MonodisperseEwaldElectricField``, ``PolydisperseEwaldElectricField`` and
``DirectElectricField`` apply one operator of libcumpm_amd to dipoles (and
quadrupoles) uploaded from the host and read the field back: ``calculate`` (the
fused apply), the reference's pipeline ``spread`` / ``scale`` / ``contract`` /
``real_space`` with optional device pointers, the split helpers
``compute_real_field`` / ``compute_inverse_field`` (:764-778), the host getters and
the ``dev_*`` device addresses (:142-164 of electric_field_bindings.cpp).  Device
memory is owned by libcumpm_amd; dipoles and fields are [particle][component]
complex128.

Modes: SOLVER_AX (calc_inter_dipole=True) returns A p at the particles (field of the
other particles with the solver's sign plus the self terms); INTERACTION_FIELD
(False) returns the field E of the dipoles at separate field points
(update_field_coordinates): the field points join the operator as sites without
dipoles and with zero self coefficients, and E is minus their rows of A [p; 0]
(monodisperse_ewald_electric_field.cu:2385-2411, direct_electric_field.cu:37-300).
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


def _mode(calc_inter_dipole):
    if calc_inter_dipole is None:
        return FieldCalcMode.SOLVER_AX
    if isinstance(calc_inter_dipole, bool):
        return FieldCalcMode.SOLVER_AX if calc_inter_dipole else FieldCalcMode.INTERACTION_FIELD
    return int(calc_inter_dipole)


class _Field:
    """Shared host logic: coordinates, dipoles, quadrupoles, self coefficients, evaluation."""

    def _init_common(self, mode, precision, solve_quadrupoles, quad_idxs):
        self._mode = _mode(mode)
        self._prec = _parse_precision(precision)
        self._solve_quads = bool(solve_quadrupoles)
        self._quad_idxs = [int(i) for i in (quad_idxs or [])]
        if self._mode == FieldCalcMode.INTERACTION_FIELD and self._solve_quads:
            raise NotImplementedError("INTERACTION_FIELD mode evaluates dipole fields only in this library")
        self._N = 0
        self._M = 0
        self._Q = 0
        self._xyz = np.zeros((0, 3))
        self._fxyz = np.zeros((0, 3))
        self._dip = np.zeros((0, 3), np.complex128)
        self._quad = np.zeros((0, 5), np.complex128)
        self._sc = np.zeros(0, np.complex128)
        self._xbuf = self._ybuf = self._grid = None
        self._pending = True
        self._particles_updated = self._field_points_updated = self._dipoles_updated = False

    def _set_quads(self):
        if not self._solve_quads:
            return
        lib = C.load()
        if not self._quad_idxs:
            check(lib.mpm_op_set_quads(self._h, None, -1))
        else:
            idx = (ctypes.c_int * len(self._quad_idxs))(*self._quad_idxs)
            check(lib.mpm_op_set_quads(self._h, idx, len(self._quad_idxs)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and C._lib is not None:
            C._lib.mpm_op_destroy(h)
            self._h = None

    # ---- geometry of the operator's sites ----
    def _interaction(self):
        return self._mode == FieldCalcMode.INTERACTION_FIELD

    def _site_count(self):
        return self._N + (self._M if self._interaction() else 0)

    def _n(self):
        return 3 * self._site_count() + 5 * self._Q

    def _sync(self):
        """Positions, buffers and self coefficients of the operator for the current inputs."""
        if not self._N:
            raise RuntimeError("electricField: particle coordinates have not been set")
        if self._interaction() and not self._M:
            raise RuntimeError("electricField: field coordinates have not been set")
        if self._pending:
            xyz = self._xyz if not self._interaction() else np.concatenate([self._xyz, self._fxyz])
            xyz = np.ascontiguousarray(xyz)
            self._push_positions(xyz)
            self._Q = int(self._info().num_quads)
            if self._quad.shape[0] != self._Q:
                self._quad = np.zeros((self._Q, 5), np.complex128)
            nbytes = self._n() * 16
            if self._xbuf is None or self._xbuf.nbytes < nbytes:
                self._xbuf = C.DeviceBuffer(nbytes)
                self._ybuf = C.DeviceBuffer(nbytes)
                self._ybuf.upload(np.zeros(self._n(), np.complex128))
            self._pending = False
        sites = self._site_count()
        sc = np.zeros(sites, np.complex128)
        sc[:self._N] = self._sc if not self._interaction() else 0.0
        check(C.load().mpm_op_set_self_coef(self._h, C.dptr(sc.view(np.float64)), 1))
        x = np.zeros(self._n(), np.complex128)
        x[:3 * self._N] = self._dip.reshape(-1)
        x[3 * sites:] = self._quad.reshape(-1)
        self._xbuf.upload(x)

    def _push_positions(self, xyz):
        check(C.load().mpm_op_set_positions(self._h, C.dptr(xyz), xyz.shape[0]))

    def _info(self):
        info = C.OpInfo()
        check(C.load().mpm_op_get_info(self._h, ctypes.byref(info)))
        return info

    # ---- updates ----
    def update_particle_coordinates(self, x_part, y_part, z_part):
        x, y, z = _vec(x_part), _vec(y_part), _vec(z_part)
        if not (x.size == y.size == z.size):
            raise ValueError("Base_Electric_Field: coordinate vectors must have equal size.")
        if x.size != self._N:
            self._dip = np.zeros((x.size, 3), np.complex128)
            self._sc = np.zeros(x.size, np.complex128)
        self._N = x.size
        self._xyz = np.stack([x, y, z], axis=1)
        self._pending = True
        self._particles_updated = True
        if not self._interaction():
            self._sync()

    def update_field_coordinates(self, x_field, y_field, z_field):
        x, y, z = _vec(x_field), _vec(y_field), _vec(z_field)
        if not (x.size == y.size == z.size):
            raise ValueError("Base_Electric_Field: coordinate vectors must have equal size.")
        self._M = x.size
        self._fxyz = np.stack([x, y, z], axis=1)
        self._pending = True
        self._field_points_updated = True

    def update_dipoles(self, dip_x, dip_y, dip_z):
        z = np.zeros(self._N)
        self.update_dipoles_complex(dip_x, z, dip_y, z, dip_z, z)

    def update_dipoles_complex(self, dip_xr, dip_xi, dip_yr, dip_yi, dip_zr, dip_zi):
        comps = [_vec(v) for v in (dip_xr, dip_xi, dip_yr, dip_yi, dip_zr, dip_zi)]
        if any(c.size != self._N for c in comps):
            raise ValueError(
                "Base_Electric_Field: complex dipole component vectors must match the allocated number of particles.")
        self._dip = np.stack([comps[0] + 1j * comps[1], comps[2] + 1j * comps[3], comps[4] + 1j * comps[5]], axis=1)
        self._dipoles_updated = True

    def set_dipoles(self, p):
        """Convenience: p of shape (N, 3) complex."""
        self._dip = np.asarray(p, np.complex128).reshape(self._N, 3).copy()
        self._dipoles_updated = True

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
        self._dipoles_updated = True

    def set_quadrupoles(self, q):
        """Convenience: q of shape (Q, 5) complex."""
        self._quad = np.asarray(q, np.complex128).reshape(self._Q, 5).copy()

    def set_self_coef(self, self_coef_r, self_coef_i=None):
        if self_coef_i is None:
            if np.iterable(self_coef_r):
                sc = np.asarray(self_coef_r).astype(np.complex128).reshape(-1)
            else:
                sc = np.full(self._N, complex(self_coef_r))
        else:
            r = np.broadcast_to(np.asarray(self_coef_r, np.float64), (self._N,))
            i = np.broadcast_to(np.asarray(self_coef_i, np.float64), (self._N,))
            sc = (r + 1j * i).astype(np.complex128)
        if sc.size != self._N:
            raise ValueError("self coefficients must match the number of particles")
        self._sc = np.ascontiguousarray(sc)

    def clear_particles_updated(self):
        self._particles_updated = False

    def clear_field_points_updated(self):
        self._field_points_updated = False

    def clear_dipoles_updated(self):
        self._dipoles_updated = False

    @property
    def particles_updated(self):
        return self._particles_updated

    @property
    def field_points_updated(self):
        return self._field_points_updated

    @property
    def dipoles_updated(self):
        return self._dipoles_updated

    # ---- evaluation ----
    def calculate(self):
        """The field at the particles (SOLVER_AX) or at the field points (INTERACTION_FIELD)."""
        self._sync()
        check(C.load().mpm_op_apply(self._h, self._xbuf.ptr, self._ybuf.ptr, 1, int(self._prec), C.PART_ALL, None))

    def clear_epoint(self):
        if self._ybuf is not None:
            self._ybuf.upload(np.zeros(self._n(), np.complex128))

    def _grid_buffer(self):
        b = ctypes.c_size_t()
        check(C.load().mpm_op_grid_bytes(self._h, int(self._prec), ctypes.byref(b)))
        if self._grid is None or self._grid.nbytes < b.value:
            self._grid = C.DeviceBuffer(max(b.value, 16))
        return self._grid.ptr

    def _stage(self, which, grid_ptr, y_ptr):
        self._sync()
        g = ctypes.c_void_p(int(grid_ptr)) if grid_ptr is not None else (
            self._grid_buffer() if which != C.STAGE_REAL_SPACE else None)
        y = ctypes.c_void_p(int(y_ptr)) if y_ptr is not None else self._ybuf.ptr
        check(C.load().mpm_op_stage(self._h, which, self._xbuf.ptr, g, y, int(self._prec), None))

    def spread(self, grid_ptr=None):
        """Spread the dipoles onto the grid (grid_ptr: a device buffer of mpm_op_grid_bytes)."""
        self._stage(C.STAGE_SPREAD, grid_ptr, None)

    def scale(self, grid_ptr=None):
        """Forward transform, Green's tensor and inverse transform of the grid, in place."""
        self._stage(C.STAGE_SCALE, grid_ptr, None)

    def contract(self, E_point_ptr=None, Es_grid_ptr=None):
        """Add the contraction of the grid at the particles to E_point."""
        self._stage(C.STAGE_CONTRACT, Es_grid_ptr, E_point_ptr)

    def real_space(self, E_point_ptr=None):
        """Add the real-space pair terms and self terms to E_point."""
        self._stage(C.STAGE_REAL_SPACE, None, E_point_ptr)

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
        """SOLVER_AX: A p at the particles, (N, 3), with quadrupoles the pair ((N, 3), (Q, 5));
        INTERACTION_FIELD: the field at the field points, (M, 3)."""
        if self._ybuf is None:
            raise RuntimeError("electricField: particle coordinates have not been set")
        e = np.zeros(self._n(), np.complex128)
        self._ybuf.download(e)
        if self._interaction():
            out = -e[3 * self._N:3 * (self._N + self._M)]
            return out.reshape(self._M, 3) if reshape else out
        if not reshape:
            return e
        if self._solve_quads and self._Q > 0:
            return e[:3 * self._N].reshape(self._N, 3), e[3 * self._N:].reshape(self._Q, 5)
        return e.reshape(self._N, 3)

    def get_dipoles_host(self):
        return self._dip[:, 0].real.copy(), self._dip[:, 1].real.copy(), self._dip[:, 2].real.copy()

    def get_dipoles_complex_host(self):
        p = self._dip
        return (p[:, 0].real.copy(), p[:, 0].imag.copy(), p[:, 1].real.copy(), p[:, 1].imag.copy(),
                p[:, 2].real.copy(), p[:, 2].imag.copy())

    # ---- device addresses (getDev* of the reference; 0 where this layout has none) ----
    def _dev(self, name, offset=0):
        p = ctypes.c_void_p()
        check(C.load().mpm_op_device_pointer(self._h, name.encode(), ctypes.byref(p), None))
        return (p.value + offset) if p.value else 0

    @property
    def dev_dipoles(self):
        return self._xbuf.ptr.value if self._xbuf is not None else 0

    @property
    def dev_epoint(self):
        return self._ybuf.ptr.value if self._ybuf is not None else 0

    @property
    def dev_x_part(self):
        return self._dev("x_part")

    @property
    def dev_y_part(self):
        return self._dev("y_part")

    @property
    def dev_z_part(self):
        return self._dev("z_part")

    # field points follow the particles in the operator's coordinate arrays
    @property
    def dev_x_field(self):
        return self._dev("x_part", 8 * self._N) if self._interaction() else self._dev("x_part")

    @property
    def dev_y_field(self):
        return self._dev("y_part", 8 * self._N) if self._interaction() else self._dev("y_part")

    @property
    def dev_z_field(self):
        return self._dev("z_part", 8 * self._N) if self._interaction() else self._dev("z_part")

    # self coefficients are interleaved complex doubles: real at +0, imaginary at +8, stride 16
    @property
    def dev_self_coef_real(self):
        return self._dev("self_coef")

    @property
    def dev_self_coef_imag(self):
        v = self._dev("self_coef")
        return v + 8 if v else 0


class _EwaldField(_Field):
    """Properties shared by the Ewald operators."""

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
        return self._mode

    @property
    def calc_inter_dipole(self):
        return self._mode == FieldCalcMode.SOLVER_AX

    @property
    def self_coef(self):
        return self._info().self_coef_ewald

    @property
    def num_particles(self):
        return self._N

    @property
    def num_field_points(self):
        return self._M if self._interaction() else self._N

    @property
    def num_pairs(self):
        return int(self._info().num_pairs)

    @property
    def num_offsets(self):
        return int(self._info().stencil_width) ** 3

    @property
    def num_spread(self):
        return self._site_count() * self.num_offsets

    @property
    def num_contract(self):
        return self._site_count() * self.num_offsets

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

    def set_bloch_wavevector(self, kx: float, ky: float):
        check(C.load().mpm_op_set_bloch(self._h, float(kx), float(ky)))

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
        self._sync()
        m = C._i64()
        lib = C.load()
        check(lib.mpm_op_get_neighbors(self._h, None, None, None, None, 0, ctypes.byref(m)))
        M = m.value
        n = self._site_count()
        start = np.zeros(n + 1, np.int64)
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
        out = np.full((counts.size, width), -1, np.int32)
        for i in range(counts.size):
            out[i, :counts[i]] = j[start[i]:start[i + 1]]
        return out.reshape(-1), counts

    def get_real_space_precalcs_host(self):
        """(self_perp, perp, para): the self term and each neighbour pair's perpendicular /
        parallel values at d_eff = max(d, 1), in get_neighbor_list_host's row order."""
        _, _, perp, para = self._neighbors()
        return float(self._info().self_perp), perp, para

    def compute_neighbor_list(self, max_neighbors_per_particle=128):
        """The neighbour list, stencils and pair values are built with the coordinates."""
        self._sync()

    compute_precalculations = compute_real_space_tables = compute_neighbor_list
    spread_precalcs = contract_precalcs = real_space_precalcs = compute_neighbor_list

    # ---- device addresses of the Ewald buffers ----
    @property
    def dev_spread_coef(self):
        return self._dev("spread_coef_fp32" if self.is_using_recip_fp32 else "spread_coef")

    @property
    def dev_contract_coef(self):
        return self.dev_spread_coef

    @property
    def dev_spread_idxs(self):
        return self._dev("spread_idxs")

    @property
    def dev_contract_idxs(self):
        return self._dev("contract_idxs")

    @property
    def dev_particle_index(self):
        return self._dev("particle_index")

    @property
    def dev_self_perp(self):
        return self._dev("self_perp")

    # per-pair (perp, para) as interleaved doubles: perp at +0, para at +8, stride 16
    @property
    def dev_perp(self):
        return self._dev("perp_para")

    @property
    def dev_para(self):
        v = self._dev("perp_para")
        return v + 8 if v else 0

    @property
    def dev_fe_grid(self):
        return self._grid.ptr.value if self._grid is not None else self._dev("fe_grid")

    @property
    def dev_fes_grid(self):
        return self._dev("fes_grid")

    @property
    def dev_neighbor_list(self):
        return self._dev("neighbor_list")

    @property
    def dev_neighbor_counts(self):
        """Row starts of the CSR neighbour list ([N + 1]; counts are their differences)."""
        return self._dev("neighbor_start")

    @property
    def dev_scale_coef(self):
        return self._dev("scale_coef")

    @property
    def dev_khat(self):
        return self._dev("scale_coef")

    # tables stay on the host and stencil offsets are implicit in this library
    dev_rtable = dev_field_dip1 = dev_field_dip2 = dev_offset = dev_offsetxyz = property(lambda self: 0)


class MonodisperseEwaldElectricField(_EwaldField):
    """Monodisperse spectral-Ewald operator (dev/electric_field.py:430-842 of the reference)."""

    def __init__(self, box_x, box_y, box_z, errortol, xi, calc_inter_dipole, radius=1.0,
                 solve_quadrupoles=False, quad_idxs=None, precision=None, *, grid=None, support=0):
        self._init_common(calc_inter_dipole, precision, solve_quadrupoles, quad_idxs)
        d = C.OpDesc()
        d.box[:] = [float(box_x), float(box_y), float(box_z)]
        d.errortol = float(errortol)
        d.xi = float(xi)
        if grid is not None:
            d.grid[:] = [int(g) for g in grid]
        d.support = int(support)
        self._box = (float(box_x), float(box_y), float(box_z))
        self._errortol = float(errortol)
        self._xi = float(xi)
        self._h = self._create(d)
        self._set_quads()

    def _create(self, d):
        h = ctypes.c_void_p()
        check(C.load().mpm_op_create(ctypes.byref(d), ctypes.byref(h)))
        return h


class PolydisperseEwaldElectricField(MonodisperseEwaldElectricField):
    """Polydisperse spectral-Ewald operator: one radius per particle (nondimensional).
    Reference: dev/electric_field.py (PolydisperseEwaldElectricField) over
    polydisperse_ewald_electric_field.cu; radii within 1e-5 share a class."""

    def __init__(self, box_x, box_y, box_z, errortol, xi, calc_inter_dipole, particle_radii,
                 solve_quadrupoles=False, quad_idxs=None, precision=None, *, grid=None, support=0):
        self._radii = _vec(particle_radii)
        if self._radii.size == 0:
            raise ValueError("particle_radii is empty")
        if _mode(calc_inter_dipole) == FieldCalcMode.INTERACTION_FIELD:
            raise NotImplementedError("INTERACTION_FIELD mode of the polydisperse operator is not provided by this library")
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


class DirectElectricField(_Field):
    """Direct (open-boundary) field, O(N^2) per evaluation (dev/electric_field.py:29-427 of the
    reference over direct_electric_field.cu).  radius: one per particle (or one for all)."""

    def __init__(self, radius=None, mode=None, solve_quadrupoles=False, quad_idxs=None, precision=None):
        self._init_common(mode, precision, solve_quadrupoles, quad_idxs)
        self._radius = _vec(radius if radius is not None and np.size(radius) else [1.0])
        self._h = None
        self._made_for = -1

    def _push_positions(self, xyz):
        n = xyz.shape[0]
        if self._h is None or self._made_for != n:
            if self._h is not None:
                C.load().mpm_op_destroy(self._h)
                self._h = None
            r = self._radius
            if r.size not in (1, self._N):
                raise ValueError("Direct_Electric_Field: host_radius size must match num_particles.")
            rr = np.ones(n)
            rr[:self._N] = r if r.size == self._N else r[0]
            h = ctypes.c_void_p()
            check(C.load().mpm_op_create_direct(C.dptr(np.ascontiguousarray(rr)), n, ctypes.byref(h)))
            self._h = h
            self._set_quads()
            self._made_for = n
        super()._push_positions(xyz)

    @property
    def is_using_fp32(self) -> bool:
        return self._prec != PrecisionMode.DOUBLE

    @property
    def precision_mode(self):
        return self._prec

    box_x = box_y = box_z = errortol = rc = xi = self_coef = property(lambda self: 0.0)
    num_spread = num_contract = num_pairs = max_neighbors = table_size = num_offsets = property(lambda self: 0)

    @property
    def mode(self):
        return self._mode

    @property
    def calc_inter_dipole(self):
        return self._mode == FieldCalcMode.SOLVER_AX

    @property
    def num_particles(self):
        return self._N

    @property
    def num_field_points(self):
        return self._M if self._interaction() else self._N

    @property
    def solve_quadrupoles(self):
        return self._solve_quads

    @property
    def num_quads(self):
        return self._Q

    @property
    def num_grid(self):
        return np.zeros(3, np.int32)

    @property
    def grid_spacing(self):
        return np.zeros(3)

    @property
    def spectral_split(self):
        return np.zeros(3)
