"""User API: ``EELS``, the electron energy-loss driver.

Mirrors src/cuMPM/eels.py of the reference (constructor arguments, their
validation and error messages, sub-dipole splitting of particles near the beam,
output shapes and the physical rescaling of the getters) over the C ABI
``mpm_eels_solver_*`` of libcumpm_amd (include/cumpm_amd.h).  Additions, all
keyword-only and off by default: ``batched`` (every frequency of a Bloch point
in one batched GPU solve), ``precision="refined"``, ``restart``/``maxiter``,
and ``get_reports()``/``get_incident_field()``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C
from ._capi import check
from .solver import _FIELD, _GUESS, _parse_precision


class _NativeEELS:
    """One mpm_eels_solver handle (the reference's _cuMPM.EELS_Solver)."""

    def __init__(self, box, eps_p_2d, omega, v, radius, eps_m, xi, tol, quiet, guess_type, solver_type,
                 field_type, integration_step, solve_quadrupoles, quad_idxs, asm_flag, asm_Nx, asm_Ny,
                 precision, refined, batched, restart, maxiter):
        if guess_type not in _GUESS:
            raise RuntimeError(f"Guess type {guess_type} not supported.")
        if solver_type not in ("gmres", "bicgstab"):
            raise ValueError(f"solver_type must be 'gmres' or 'bicgstab', got {solver_type}")
        eps = np.ascontiguousarray(np.asarray(eps_p_2d, np.complex128))
        om = np.ascontiguousarray(np.asarray(omega, np.float64).reshape(-1))
        rad = np.ascontiguousarray(np.asarray(radius, np.float64).reshape(-1))
        self._keep = [eps, om, rad]
        d = C.EelsDesc()
        d.box[:] = [float(b) for b in box]
        d.eps_p = C.dptr(eps.view(np.float64))
        d.num_wavevectors, d.eps_cols = eps.shape
        d.omega = C.dptr(om)
        d.num_omega = om.size
        d.v = float(v)
        d.radius = C.dptr(rad)
        d.num_radii = rad.size
        d.eps_m = float(eps_m)
        d.xi = float(xi)
        d.tol = float(tol)
        d.quiet = int(bool(quiet))
        d.guess = _GUESS[guess_type]
        d.method = C.GMRES if solver_type == "gmres" else C.BICGSTAB
        d.field_type = _FIELD[field_type]
        d.integration_step = float(integration_step)
        d.solve_quadrupoles = int(bool(solve_quadrupoles))
        qi = [int(i) for i in (quad_idxs or [])]
        self._qi = (ctypes.c_int * max(len(qi), 1))(*qi)
        d.quad_idxs = self._qi
        d.num_quad_idxs = len(qi)
        d.asm_flag = int(bool(asm_flag))
        d.asm_nx = int(asm_Nx)
        d.asm_ny = int(asm_Ny)
        d.precision = int(precision)
        d.refined = int(bool(refined))
        d.batched = int(bool(batched))
        d.restart = int(restart)
        d.maxiter = int(maxiter)
        self.h = ctypes.c_void_p()
        check(C.load().mpm_eels_solver_create(ctypes.byref(d), ctypes.byref(self.h)))

    def __del__(self):
        h = getattr(self, "h", None)
        if h and C._lib is not None:
            C._lib.mpm_eels_solver_destroy(h)
            self.h = None

    def compute(self, epos, x, y, z):
        e = np.ascontiguousarray(np.asarray(epos, np.float64).reshape(2))
        x, y, z = (np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)) for a in (x, y, z))
        if not (x.size == y.size == z.size):
            raise ValueError("Input coordinate vectors must have the exact same size.")
        check(C.load().mpm_eels_solver_compute(self.h, C.dptr(e), C.dptr(x), C.dptr(y), C.dptr(z), x.size))

    def dims(self):
        d = (C._i64 * 4)()
        check(C.load().mpm_eels_solver_get_dims(self.h, d))
        return [int(v) for v in d]

    def get_eels(self):
        F, W, _, _ = self.dims()
        out = np.zeros((F, W))
        if out.size:
            check(C.load().mpm_eels_solver_get_eels(self.h, C.dptr(out)))
        return out

    def get_dipoles(self, physical=True):
        F, W, N, _ = self.dims()
        out = np.zeros((F, W, N, 3), np.complex128)
        if out.size:
            check(C.load().mpm_eels_solver_get_dipoles(self.h, int(bool(physical)), C.dptr(out.view(np.float64))))
        return out

    def get_quadrupoles(self, physical=True):
        F, W, _, Q = self.dims()
        out = np.zeros((F, W, Q, 5), np.complex128)
        if out.size:
            check(C.load().mpm_eels_solver_get_quadrupoles(self.h, int(bool(physical)),
                                                          C.dptr(out.view(np.float64))))
        return out

    def get_incident_field(self):
        F, W, N, Q = self.dims()
        out = np.zeros((W, 3 * N + 5 * Q), np.complex128)
        check(C.load().mpm_eels_solver_get_incident_field(self.h, C.dptr(out.view(np.float64))))
        return out

    def get_reports(self):
        n = C._i64()
        check(C.load().mpm_eels_solver_get_reports(self.h, None, 0, ctypes.byref(n)))
        reps = (C.SolveReport * max(n.value, 1))()
        check(C.load().mpm_eels_solver_get_reports(self.h, reps, n.value, ctypes.byref(n)))
        return [reps[i].as_dict() for i in range(n.value)]

    def get_timing(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(C.load().mpm_eels_solver_get_timing(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"solve_ms": a.value, "setup_ms": b.value, "total_ms": c.value}

    def isUsingJacobiPrecond(self):
        v = ctypes.c_int()
        check(C.load().mpm_eels_solver_get_use_jacobi(self.h, ctypes.byref(v)))
        return bool(v.value)

    def setUseJacobiPrecond(self, enable):
        check(C.load().mpm_eels_solver_set_use_jacobi(self.h, int(bool(enable))))


def _eps_rows(eps_p):
    """eps_p as the native solver takes it: (W, cols) complex (eels.py:124-147 of the reference:
    a scalar is one wavelength, a 1-D array one value per wavelength, a 2-D array per particle)."""
    a = np.asarray(eps_p)
    if a.ndim == 0:
        return np.array([[complex(a.item())]])
    if a.ndim == 1:
        if a.size == 1:
            return np.array([[complex(a.item())]])
        return a.astype(np.complex128).reshape(-1, 1)
    if a.ndim == 2:
        return a.astype(np.complex128)
    raise ValueError("eps_p must be a scalar, 1D array, or 2D array")


class EELS:
    """Electron energy-loss probability and induced dipoles of particles in the field of a
    fast electron (the reference's cuMPM.EELS, src/cuMPM/eels.py:21-609).

    box: [Lx, Ly, Lz] (nm); eps_p: scalar, per wavelength, or (W, N); omega: wavenumbers
    (1/nm); v: beam speed / c.  compute(epos, positions) takes the beam's (x, y) impact
    coordinates, (2,) or (num_epos, 2), and the particles (N, 3); every beam position is one
    frame.  Particles within split_dist of the beam can be replaced by N_split sub-dipoles
    (split_method "cubic", "fibonacci" or "fcc") of equal total volume."""

    def __init__(self, box, eps_p, omega, v, eps_m=1.0, radius=1.0, xi=0.5, tol=1e-3, quiet=False,
                 guess_type="derivative", solver_type="gmres", field_type="auto", split_dist=None, N_split=None,
                 integration_step=0.05, split_method="cubic", solve_quadrupoles=False, quad_idxs=None,
                 asm_flag=False, asm_Nx=1, asm_Ny=1, precision=None, *, batched=False, restart=50,
                 maxiter=1000):
        if split_method not in ("cubic", "fibonacci", "fcc"):
            raise ValueError("split_method must be 'cubic', 'fibonacci', or 'fcc', got " + repr(split_method))
        self.split_method = split_method
        if (split_dist is None) != (N_split is None):
            raise ValueError("Both split_dist and N_split must be specified, or both left as None.")
        if split_dist is None:
            self.split_dist, self.N_split = 0.0, 0
        else:
            self.split_dist, self.N_split = float(split_dist), int(N_split)
        self._prec, self._refined = _parse_precision(precision)
        self.precision = self._prec
        if field_type not in ("auto", "monodisperse", "polydisperse", "direct"):
            raise ValueError(f"field_type must be 'auto', 'monodisperse', 'polydisperse', or 'direct', got {field_type}")
        box_list = [float(x) for x in box]
        if len(box_list) != 3:
            raise ValueError("box must have length 3")
        self.box = box_list
        self.eps_p = eps_p
        self.omega = [float(o) for o in omega]
        self.v = float(v)
        self.eps_m = float(eps_m)
        self.radius = radius
        self.xi = float(xi)
        self.tol = float(tol)
        self.quiet = bool(quiet)
        self.guess_type = guess_type
        self.solver_type = solver_type
        self.field_type = field_type
        self.integration_step = float(integration_step)
        self.solve_quadrupoles = bool(solve_quadrupoles)
        self.quad_idxs = [int(i) for i in quad_idxs] if quad_idxs is not None else []
        self.asm_flag = bool(asm_flag)
        self.asm_Nx = int(asm_Nx)
        self.asm_Ny = int(asm_Ny)
        self._opts = dict(batched=batched, restart=restart, maxiter=maxiter)
        self._use_jacobi = True
        self.positions = None
        self._eels_results, self._dips_results, self._quads_results, self._split_poss = [], [], [], []
        self._dips_scale_factors, self._quads_scale_factors = [], []
        radius_list = [float(r) for r in radius] if np.iterable(radius) else [float(radius)]
        if self._splitting():
            self._solver = None
        else:
            self._solver = self._native(box_list, _eps_rows(eps_p), radius_list, self.quad_idxs, self.asm_flag,
                                        self.asm_Nx, self.asm_Ny)

    def _splitting(self):
        return self.split_dist > 0.0 and self.N_split > 1

    def _native(self, box, eps_2d, radii, quad_idxs, asm_flag, asm_Nx, asm_Ny):
        s = _NativeEELS(box, eps_2d, self.omega, self.v, radii, self.eps_m, self.xi, self.tol, self.quiet,
                        self.guess_type, self.solver_type, self.field_type, self.integration_step,
                        self.solve_quadrupoles, quad_idxs, asm_flag, asm_Nx, asm_Ny, self._prec, self._refined,
                        **self._opts)
        if not self._use_jacobi:
            s.setUseJacobiPrecond(False)
        return s

    # ---- sub-dipole lattices on the unit ball (eels.py:176-264 of the reference) ----
    def _make_cubic_offsets(self):
        """Cubic lattice of n^3 >= N_split cell centres in [-1, 1]^3, kept inside the unit ball."""
        n = int(np.ceil(self.N_split ** (1.0 / 3.0)))
        while n ** 3 < self.N_split:
            n += 1
        c = -1.0 + (2.0 * np.arange(n) + 1.0) / n
        g = np.stack(np.meshgrid(c, c, c, indexing="ij"), axis=-1).reshape(-1, 3)
        return g[np.linalg.norm(g, axis=1) <= 1.0]

    def _make_fibonacci_offsets(self):
        """N_split golden-angle directions (y from 1 to -1) at radii ((i + 1/2)/N)^(1/3), the
        radii shuffled by a fixed generator (seed 0) so that shells and latitudes decouple."""
        N = self.N_split
        i = np.arange(N, dtype=np.float64)
        yv = 1.0 - 2.0 * i / float(N - 1)
        th = np.pi * (3.0 - np.sqrt(5.0)) * i
        rs = np.sqrt(np.maximum(0.0, 1.0 - yv * yv))
        dirs = np.stack([np.cos(th) * rs, yv, np.sin(th) * rs], axis=1)
        radii = ((i + 0.5) / N) ** (1.0 / 3.0)
        np.random.default_rng(0).shuffle(radii)
        return radii[:, None] * dirs

    def _make_fcc_offsets(self):
        """FCC points (integer triples of even sum) times a step h inside the unit ball, h from
        the point density 2 / h^3 and shrunk by 2% until at least N_split points fit."""
        N = max(1, self.N_split)
        h = (2.0 * np.pi / (3.0 * N)) ** (1.0 / 3.0)

        def pts(step):
            m = int(np.ceil(1.0 / step))
            k = np.arange(-m, m + 1)
            g = np.stack(np.meshgrid(k, k, k, indexing="ij"), axis=-1).reshape(-1, 3)
            g = g[g.sum(axis=1) % 2 == 0] * step
            return g[np.linalg.norm(g, axis=1) <= 1.0]

        p = pts(h)
        while len(p) < N and h > 1e-4:
            h *= 0.98
            p = pts(h)
        return p

    def _get_unit_offsets(self):
        return {"cubic": self._make_cubic_offsets, "fibonacci": self._make_fibonacci_offsets,
                "fcc": self._make_fcc_offsets}[self.split_method]()

    def _split_close_dipoles(self, positions, e_pos, radii, eps_p_2d):
        """Particles whose in-plane distance to the beam is below split_dist become M
        sub-dipoles at x + R u_k of radius R / M^(1/3) (same total volume), appended after the
        unsplit particles, which keep their order (eels.py:281-345)."""
        close = np.linalg.norm(positions[:, :2] - e_pos[None, :], axis=1) < self.split_dist
        if not np.any(close) or self.N_split <= 1:
            return positions.copy(), radii.copy(), eps_p_2d.copy()
        unit = self._get_unit_offsets()
        M = len(unit)
        keep = np.where(~close)[0]
        split = np.where(close)[0]
        new_pos = [positions[keep]] + [positions[i] + unit * radii[i] for i in split]
        new_rad = [radii[keep]] + [np.full(M, radii[i] / M ** (1.0 / 3.0)) for i in split]
        new_eps = [eps_p_2d[:, keep]] + [np.repeat(eps_p_2d[:, i:i + 1], M, axis=1) for i in split]
        return np.concatenate(new_pos), np.concatenate(new_rad), np.concatenate(new_eps, axis=1)

    def compute(self, epos, positions):
        """EELS spectrum and induced dipoles for each beam position (one frame each)."""
        positions = np.asarray(positions)
        if positions.ndim != 2 or positions.shape[1] != 3:
            raise ValueError("positions must be of shape (num_particles, 3)")
        epos = np.asarray(epos)
        if epos.ndim == 1:
            if epos.shape[0] != 2:
                raise ValueError("epos must be of shape (2,) or (num_epos, 2)")
            epos = epos[np.newaxis, :]
        elif epos.ndim != 2 or epos.shape[1] != 2:
            raise ValueError("epos must be of shape (2,) or (num_epos, 2)")
        num_particles = positions.shape[0]
        self.positions = positions
        if not self._splitting():
            for i in range(epos.shape[0]):
                if not self.quiet:
                    print(f"  [EELS] Processing impact parameter {i + 1}/{epos.shape[0]}...")
                self._solver.compute(epos[i], positions[:, 0], positions[:, 1], positions[:, 2])
            return
        self._eels_results, self._dips_results, self._split_poss = [], [], []
        self._quads_results = []
        self._dips_scale_factors, self._quads_scale_factors = [], []
        radii = np.asarray(self.radius, np.float64).reshape(-1)
        if radii.size == 1:
            radii = np.repeat(radii.item(), num_particles)
        elif radii.size != num_particles:
            raise ValueError("The number of particles is inconsistent with the number of radii provided.")
        W = len(self.omega)
        e = np.asarray(self.eps_p)
        if e.ndim == 0 or (e.ndim == 1 and e.size == 1):
            eps_2d = np.full((W, num_particles), complex(np.ravel(e)[0]))
        elif e.ndim == 1:
            if e.size != W:
                raise ValueError("1D eps_p must have length equal to num_wavevectors")
            eps_2d = np.tile(e.astype(np.complex128)[:, None], (1, num_particles))
        elif e.ndim == 2:
            if e.shape[0] != W:
                raise ValueError("First dimension of 2D eps_p must be equal to num_wavevectors")
            if e.shape[1] == 1:
                eps_2d = np.tile(e.astype(np.complex128), (1, num_particles))
            elif e.shape[1] != num_particles:
                raise ValueError("Second dimension of 2D eps_p must be equal to num_particles")
            else:
                eps_2d = e.astype(np.complex128).copy()
        else:
            raise ValueError("eps_p must be a scalar, 1D array, or 2D array")
        for i in range(epos.shape[0]):
            e_pos = np.asarray(epos[i], np.float64)
            if not self.quiet:
                print(f"  [EELS] Processing impact parameter {i + 1}/{epos.shape[0]}...")
            sp, sr, se = self._split_close_dipoles(positions, e_pos, radii, eps_2d)
            qi = self.quad_idxs
            if self.solve_quadrupoles and qi:
                close = np.linalg.norm(positions[:, :2] - e_pos[None, :], axis=1) < self.split_dist
                keep = list(np.where(~close)[0])
                split = list(np.where(close)[0])
                M = len(self._get_unit_offsets())
                mapped = []
                for q in qi:
                    if q in keep:
                        mapped.append(keep.index(q))
                    elif q in split:
                        start = len(keep) + split.index(q) * M
                        mapped.extend(range(start, start + M))
                qi = mapped
            s = self._native(self.box, se, sr.tolist(), qi, False, 1, 1)
            s.compute(e_pos, sp[:, 0], sp[:, 1], sp[:, 2])
            fe = np.atleast_1d(np.squeeze(s.get_eels()))
            fd = s.get_dipoles(physical=False)[0]  # (W, M_particles, 3)
            self._eels_results.append(fe)
            self._dips_results.append(fd)
            self._split_poss.append(sp)
            # physical units follow the smallest radius of the split system, the solver's
            # length scale (the reference scales by the first radius, eels.py:436-437)
            self._dips_scale_factors.append(self.eps_m * sr.min() ** 3)
            if self.solve_quadrupoles:
                self._quads_results.append(s.get_quadrupoles(physical=False)[0])
                self._quads_scale_factors.append(self.eps_m * sr.min() ** 4)

    @staticmethod
    def _stack(results, factors=None):
        if not results:
            return []
        out = [r * f for r, f in zip(results, factors)] if factors is not None else list(results)
        if len(out) == 1:
            return np.squeeze(out[0])
        try:
            return np.squeeze(np.stack(out))
        except ValueError:
            return [np.squeeze(r) for r in out]

    def get_eels(self):
        """Loss probability, (num_epos, W) squeezed."""
        if self._splitting():
            if not self._eels_results:
                return np.array([])
            return np.squeeze(np.array(self._eels_results))
        return np.squeeze(self._solver.get_eels())

    def get_dipoles(self, physical=True):
        """Induced dipoles (num_epos, W, N, 3) squeezed; physical units scale by eps_m a_min^3."""
        if self._splitting():
            return self._stack(self._dips_results, self._dips_scale_factors if physical else None)
        return np.squeeze(self._solver.get_dipoles(physical))

    def get_quadrupoles(self, physical=True):
        """Induced quadrupoles (num_epos, W, Q, 5) squeezed; physical units scale by eps_m a_min^4."""
        if self._splitting():
            return self._stack(self._quads_results, self._quads_scale_factors if physical else None)
        return np.squeeze(self._solver.get_quadrupoles(physical))

    def get_positions(self):
        if self._splitting():
            if not self._split_poss:
                return []
            if len(self._split_poss) == 1:
                return np.squeeze(self._split_poss[0])
            try:
                return np.squeeze(np.stack(self._split_poss))
            except ValueError:
                return [np.squeeze(p) for p in self._split_poss]
        return self.positions

    def get_incident_field(self):
        """Nondimensional incident field of the last frame, (W, 3N + 5Q) (an addition)."""
        return self._solver.get_incident_field()

    def get_reports(self):
        """Per-solve reports of every frame and Bloch point (an addition)."""
        return self._solver.get_reports()

    def get_timing(self):
        return self._solver.get_timing()

    @property
    def use_jacobi_precond(self) -> bool:
        if self._solver is None:
            return self._use_jacobi
        return self._solver.isUsingJacobiPrecond()

    @use_jacobi_precond.setter
    def use_jacobi_precond(self, enable: bool):
        self._use_jacobi = bool(enable)
        if self._solver is not None:
            self._solver.setUseJacobiPrecond(enable)
