"""User API: ``dipole_solver`` and ``drude_dielectric``.

Mirrors src/cuMPM/solver.py of the reference (same constructor arguments,
argument normalisation, error types and output shapes) on top of the C ABI
of libcumpm_amd (include/cumpm_amd.h).  Additions, all keyword-only and off
by default: ``batched`` (all frequency x polarisation systems in one batched
GPU solve), ``precision="refined"`` (FP32 reciprocal space inside the Krylov
iterations, FP64 residual at every restart), ``restart``/``maxiter``,
Ewald overrides ``grid``/``support``/``ewald_tol``, ``autotune`` (time the
reciprocal-space layouts and FFT back-ends on the first frame and keep the
fastest; ``autotune="grid"`` may also grow the grid to 7-smooth lengths), and
``get_reports()``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C
from ._capi import check


class PrecisionMode:
    """Reference enum _cuMPM.PrecisionMode (ewald_electric_field_base.h; AUTO == MIXED; FP32 and
    FP64 are the direct field's names for MIXED and DOUBLE)."""
    AUTO = C.PREC_AUTO
    MIXED = C.PREC_MIXED
    DOUBLE = C.PREC_DOUBLE
    FP32 = C.PREC_MIXED
    FP64 = C.PREC_DOUBLE


def drude_dielectric(omega, gamma, omega_p, eps_inf):
    """Complex relative permittivity of a Drude metal (src/cuMPM/solver.py:4-51).

    eps = eps_inf - omega_p^2 / (omega^2 + i omega gamma); shape (num_omega, num_particles),
    squeezed."""
    if not np.iterable(omega):
        omega = np.asarray([omega])
    omega = np.asarray(omega)
    if len(omega.shape) != 1:
        raise ValueError("omega must be a scalar or a 1-D array of wavenumbers")
    if not np.iterable(gamma):
        gamma = np.asarray([gamma])
    gamma = np.asarray(gamma)
    if len(gamma.shape) != 1:
        raise ValueError("gamma must be a scalar or a 1-D array of damping coefficients")
    if not np.iterable(omega_p):
        omega_p = np.asarray([omega_p])
    omega_p = np.asarray(omega_p)
    if len(omega_p.shape) != 1:
        raise ValueError("omega_p must be a scalar or a 1-D array of plasma frequencies")
    if not np.iterable(eps_inf):
        eps_inf = np.asarray([eps_inf])
    eps_inf = np.asarray(eps_inf)
    if len(eps_inf.shape) != 1:
        raise ValueError("eps_inf must be a scalar or a 1-D array of high frequency permitivities")
    if np.any(len(eps_inf) != np.array([len(gamma), len(omega_p)])):
        raise ValueError("eps_inf, gamma, and omega_p must all have the same number of values applied")
    eps_p = eps_inf[None, :] - omega_p[None, :] ** 2 / (
        omega[:, None] ** 2 + 1j * omega[:, None] * gamma[None, :])
    return np.squeeze(eps_p.astype(np.complex128))


def _parse_precision(precision):
    """-> (mpm_precision, refined).  src/cuMPM/solver.py:97-115 plus "refined"."""
    if precision is None:
        return C.PREC_AUTO, False
    if isinstance(precision, (int, np.integer)) and precision in (C.PREC_AUTO, C.PREC_MIXED, C.PREC_DOUBLE):
        return int(precision), False
    if isinstance(precision, str):
        p = precision.lower()
        if p in ("mixed", "fp32", "float32", "single"):
            return C.PREC_MIXED, False
        if p in ("double", "fp64", "double64"):
            return C.PREC_DOUBLE, False
        if p in ("auto",):
            return C.PREC_AUTO, False
        if p in ("refined", "mixed_refined", "fp32_refined"):
            return C.PREC_MIXED, True
        raise ValueError(f"Unknown precision mode: {precision}")
    raise ValueError(f"Invalid precision type: {type(precision)}")


_GUESS = {"derivative": C.GUESS_DERIVATIVE, "mean_field": C.GUESS_MEAN_FIELD,
          "mean-field": C.GUESS_MEAN_FIELD, "previous": C.GUESS_PREVIOUS}
_FIELD = {"auto": C.FIELD_AUTO, "monodisperse": C.FIELD_MONODISPERSE,
          "polydisperse": C.FIELD_POLYDISPERSE, "direct": C.FIELD_DIRECT}


class dipole_solver:
    """Self-consistent induced dipoles of a periodic particle assembly (MPM).

    Same interface as the reference's ``cuMPM.dipole_solver``; the GPU path is
    libcumpm_amd (HIP for gfx950 + rocFFT)."""

    def __init__(self, box, eps_p, radius=1.0, eps_m=1.0, xi=0.5, tol=1e-3, quiet=False,
                 guess_type="derivative", solver_type="gmres", field_type="auto", E0=None,
                 quadrupoles=False, precision=None, *, batched=False, restart=50, maxiter=1000,
                 grid=None, support=0, ewald_tol=0.0, autotune=False):
        if autotune not in (False, True, "grid"):
            raise ValueError(f"autotune must be False, True or 'grid', got {autotune!r}")
        if field_type not in ("auto", "monodisperse", "polydisperse", "direct"):
            raise ValueError(
                f"field_type must be 'auto', 'monodisperse', 'polydisperse', or 'direct', got {field_type}")
        prec, refined = _parse_precision(precision)
        box_list = [float(x) for x in box]
        if len(box_list) != 3:
            raise ValueError("box must have length 3")
        radius_list = [float(r) for r in radius] if np.iterable(radius) else [float(radius)]

        E0_list = []
        if E0 is not None:
            E0_arr = np.asarray(E0)
            if E0_arr.ndim == 1:
                if E0_arr.size % 3 != 0:
                    raise ValueError("E0 size must be divisible by 3")
                E0_list = [[complex(x) for x in E0_arr[i:i + 3]] for i in range(0, E0_arr.size, 3)]
            elif E0_arr.ndim == 2:
                if E0_arr.shape[1] == 3 or E0_arr.shape[1] % 3 == 0:
                    E0_list = [[complex(x) for x in row] for row in E0_arr]
                else:
                    raise ValueError("Each vector in E0 must be 3-dimensional or have a size that is a multiple of 3")
            elif E0_arr.ndim == 3:
                if E0_arr.shape[2] != 3:
                    raise ValueError("The last dimension of E0 must be 3")
                E0_list = [[complex(x) for x in block.ravel()] for block in E0_arr]
            else:
                raise ValueError("E0 must be a 1D, 2D, or 3D array")
        # solver.py:145-154 of the reference: True = every particle, an iterable = that subset
        quad_idxs_list = []
        if isinstance(quadrupoles, bool):
            solve_quadrupoles = quadrupoles
        elif np.iterable(quadrupoles):
            solve_quadrupoles = True
            quad_idxs_list = [int(i) for i in quadrupoles]
        else:
            raise TypeError("quadrupoles must be a boolean or a list/iterable of particle indices")

        eps_arr = np.asarray(eps_p)
        if eps_arr.ndim == 0 or (eps_arr.ndim == 1 and eps_arr.size == 1):
            eps2 = np.array([[complex(eps_arr.reshape(-1)[0])]], np.complex128)
        elif eps_arr.ndim == 1:
            eps2 = eps_arr.astype(np.complex128).reshape(-1, 1)
        elif eps_arr.ndim == 2:
            eps2 = eps_arr.astype(np.complex128)
        else:
            raise ValueError("eps_p must be a scalar, 1D array, or 2D array")
        if guess_type not in _GUESS:
            raise RuntimeError(f"Guess type {guess_type} not supported.")

        # keep the host arrays alive until create() has copied them
        self._eps = np.ascontiguousarray(eps2)
        self._rad = np.ascontiguousarray(radius_list, np.float64)
        self._qidx = (ctypes.c_int * max(1, len(quad_idxs_list)))(*quad_idxs_list)
        if E0_list:
            lens = {len(r) for r in E0_list}
            if len(lens) != 1:
                raise RuntimeError("all incident fields must have the same length")
            self._E0 = np.ascontiguousarray(np.array(E0_list, np.complex128))
        else:
            self._E0 = None
        d = C.DipoleDesc()
        d.box[:] = box_list
        d.eps_p = C.dptr(self._eps.view(np.float64))
        d.num_wavevectors, d.eps_cols = self._eps.shape
        d.radius = C.dptr(self._rad)
        d.num_radii = self._rad.size
        d.eps_m = float(eps_m)
        d.xi = float(xi)
        d.tol = float(tol)
        d.quiet = int(bool(quiet))
        d.guess = _GUESS[guess_type]
        d.method = C.BICGSTAB if solver_type == "bicgstab" else C.GMRES
        d.field_type = _FIELD[field_type]
        if self._E0 is not None:
            d.E0 = C.dptr(self._E0.view(np.float64))
            d.num_E0, d.E0_len = self._E0.shape
        d.solve_quadrupoles = int(bool(solve_quadrupoles))
        d.quad_idxs = self._qidx
        d.num_quad_idxs = len(quad_idxs_list)
        d.precision = prec
        d.refined = int(refined)
        d.batched = int(bool(batched))
        d.restart = int(restart)
        d.maxiter = int(maxiter)
        if grid is not None:
            d.grid[:] = [int(g) for g in grid]
        d.support = int(support)
        d.ewald_tol = float(ewald_tol)
        d.autotune = 2 if autotune == "grid" else int(bool(autotune))
        h = ctypes.c_void_p()
        check(C.load().mpm_dipole_solver_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and C._lib is not None:
            C._lib.mpm_dipole_solver_destroy(h)
            self._h = None

    def _dims(self):
        dims = (C._i64 * 4)()
        check(C.load().mpm_dipole_solver_get_dims(self._h, dims))
        return tuple(int(v) for v in dims)

    def compute(self, positions):
        """Solve for the dipoles of one or more frames, positions (N,3) or (F,N,3)."""
        positions = np.asarray(positions)
        if positions.ndim == 2:
            positions = positions[np.newaxis, ...]
        elif positions.ndim != 3 or positions.shape[-1] != 3:
            raise ValueError("positions must be of shape (num_particles, 3) or (num_frames, num_particles, 3)")
        for f in range(positions.shape[0]):
            x = np.ascontiguousarray(positions[f, :, 0], np.float64)
            y = np.ascontiguousarray(positions[f, :, 1], np.float64)
            z = np.ascontiguousarray(positions[f, :, 2], np.float64)
            check(C.load().mpm_dipole_solver_compute(self._h, C.dptr(x), C.dptr(y), C.dptr(z), x.size))

    def get_eff_polarizability(self, physical=True):
        """Average dipole per (frequency, polarisation): (W, K, 3), squeezed."""
        F, W, N, K = self._dims()
        out = np.zeros((W, K, 3), np.complex128)
        check(C.load().mpm_dipole_solver_get_eff_polarizability(self._h, int(bool(physical)),
                                                                C.dptr(out.view(np.float64))))
        return np.squeeze(out)

    def get_dipoles(self, physical=True):
        """Dipoles (F, W, N, K, 3), squeezed."""
        F, W, N, K = self._dims()
        out = np.zeros((F, W, N, K, 3), np.complex128)
        check(C.load().mpm_dipole_solver_get_dipoles(self._h, int(bool(physical)), C.dptr(out.view(np.float64))))
        return np.squeeze(out)

    def get_num_quads(self):
        v = C._i64()
        check(C.load().mpm_dipole_solver_get_num_quads(self._h, ctypes.byref(v)))
        return int(v.value)

    def get_quadrupoles(self, physical=True):
        """Quadrupoles (F, W, Q, K, 5) of the quadrupole subset, squeezed; the five
        components are (Qxx, Qxy, Qxz, Qyy, Qyz) of the traceless symmetric tensor,
        Qzz = -(Qxx + Qyy)."""
        F, W, N, K = self._dims()
        out = np.zeros((F, W, self.get_num_quads(), K, 5), np.complex128)
        check(C.load().mpm_dipole_solver_get_quadrupoles(self._h, int(bool(physical)),
                                                         C.dptr(out.view(np.float64))))
        return np.squeeze(out)

    def get_cap_dip(self):
        """Deprecated. Use get_dipoles and get_eff_polarizability instead."""
        return self.get_eff_polarizability(), self.get_dipoles()

    def get_reports(self):
        """Per-solve statistics [F][W][K] (iterations, applies, residuals, convergence)."""
        F, W, N, K = self._dims()
        arr = (C.SolveReport * max(1, F * W * K))()
        check(C.load().mpm_dipole_solver_get_reports(self._h, arr))
        flat = [arr[i].as_dict() for i in range(F * W * K)]
        return [[[flat[(f * W + w) * K + k] for k in range(K)] for w in range(W)] for f in range(F)]

    def get_timing(self):
        """(solve_ms, setup_ms, total_ms) of the last compute(); solve_ms covers the device
        solves only (inputs already in device memory)."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(C.load().mpm_dipole_solver_get_timing(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def get_tune_info(self):
        """The autotuner's choice on the first frame: layout, fft_backend, grid, best_ms and
        default_ms per apply of one solve's batch, candidates, cached (zeros when autotune is
        off)."""
        info = C.TuneInfo()
        check(C.load().mpm_dipole_solver_get_tune_info(self._h, ctypes.byref(info)))
        return info.as_dict()

    def reset(self):
        """Drop stored frames (keeps the operator and FFT plans); the next compute() is a
        first frame again."""
        check(C.load().mpm_dipole_solver_reset(self._h))

    def get_op_info(self):
        info = C.OpInfo()
        check(C.load().mpm_dipole_solver_get_op_info(self._h, ctypes.byref(info)))
        return info

    @property
    def use_jacobi_precond(self) -> bool:
        v = ctypes.c_int()
        check(C.load().mpm_dipole_solver_get_use_jacobi(self._h, ctypes.byref(v)))
        return bool(v.value)

    @use_jacobi_precond.setter
    def use_jacobi_precond(self, enable: bool):
        check(C.load().mpm_dipole_solver_set_use_jacobi(self._h, int(bool(enable))))
