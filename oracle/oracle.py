"""CPU oracle for the cuMPM hot path — TEST INFRASTRUCTURE ONLY.

Python harness over ``oracle/lib/liboracle_mpm.so`` (built from
``oracle/mpm_oracle.cpp`` by ``oracle/Makefile``).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module, and only as the checker of the HIP product path.

Besides thin wrappers over the C++ operator and Krylov solvers, this module
restates the reference's ``Dipole_Solver`` driver
(src/cpp/dipole_solver.cu:131-607): nondimensionalisation (:184-205), volume
fraction (:207-221), the mean-field / previous / derivative initial guesses
(:245-346), the frequency x polarisation loops (:361-489), the quadrupole
unknowns, right-hand sides and guesses (:131-182, :233-345, :361-435) and the
physical rescaling of the outputs (:609-647).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liboracle_mpm.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle library with its Makefile (g++ + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_fft3d.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_sizing.argtypes = [_dp, ctypes.c_double, ctypes.c_double, _ip, ctypes.c_int,
                                 _ip, _dp, _dp, _ip, _ip, _dp]
        L.orc_tables.argtypes = [ctypes.c_double, _dp, _dp, _dp, _dp]
        L.orc_pair_functions.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.orc_op_create.restype = ctypes.c_void_p
        L.orc_op_create.argtypes = [_dp, ctypes.c_double, ctypes.c_double, _ip, ctypes.c_int,
                                    ctypes.c_char_p, ctypes.c_int]
        L.orc_op_destroy.argtypes = [ctypes.c_void_p]
        L.orc_op_set_bloch.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_op_set_positions.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        L.orc_op_info.argtypes = [ctypes.c_void_p, _ip, _dp, _dp, _ip, _ip, _dp, _dp,
                                  ctypes.POINTER(ctypes.c_long)]
        L.orc_op_apply.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_int]
        L.orc_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_op_set_quads.argtypes = [ctypes.c_void_p, _ip, ctypes.c_int]
        L.orc_quad_pair_functions.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_quad_self_grad.restype = ctypes.c_double
        L.orc_quad_self_grad.argtypes = [ctypes.c_double]
        L.orc_poly_create.restype = ctypes.c_void_p
        L.orc_poly_create.argtypes = [_dp, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int, _ip,
                                      ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.orc_poly_pair_functions.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, _dp, _dp]
        L.orc_poly_self.restype = ctypes.c_double
        L.orc_poly_self.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_poly_weight.argtypes = [_dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def num_threads():
    return lib().orc_num_threads()


def fft3d(a, sign=-1):
    """Unnormalised 3-D DFT of a complex (nx, ny, nz) array (sign -1 forward)."""
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    nx, ny, nz = a.shape
    lib().orc_fft3d(_d(a.view(np.float64)), nx, ny, nz, int(sign))
    return a


def sizing(box, errortol, xi, grid=None, P=None):
    """Ewald sizing (monodisperse_ewald_electric_field.cu:546-601)."""
    L = np.asarray(box, dtype=np.float64)
    n = np.zeros(3, np.int32)
    h = np.zeros(3)
    eta = np.zeros(3)
    Pout = ctypes.c_int()
    width = ctypes.c_int()
    rc = ctypes.c_double()
    g = np.zeros(3, np.int32) if grid is None else np.asarray(grid, np.int32)
    rcode = lib().orc_sizing(_d(L), float(errortol), float(xi), g.ctypes.data_as(_ip),
                             int(P or 0), n.ctypes.data_as(_ip), _d(h), _d(eta),
                             ctypes.byref(Pout), ctypes.byref(width), ctypes.byref(rc))
    if rcode != 0:
        raise RuntimeError("Real space cutoff larger than half the box length.")
    return dict(grid=tuple(int(v) for v in n), h=h, eta=eta, P=Pout.value, width=width.value,
                rc=rc.value)


def tables(xi):
    """Dipole real-space tables (r, perp, para) and the self term."""
    n = 9001
    r = np.zeros(n)
    pe = np.zeros(n)
    pa = np.zeros(n)
    sp = ctypes.c_double()
    lib().orc_tables(float(xi), _d(r), _d(pe), _d(pa), ctypes.byref(sp))
    return r, pe, pa, sp.value


def pair_functions(r, xi):
    r = np.ascontiguousarray(r, dtype=np.float64)
    pe = np.zeros_like(r)
    pa = np.zeros_like(r)
    lib().orc_pair_functions(_d(r), r.size, float(xi), _d(pe), _d(pa))
    return pe, pa


def quad_pair_functions(r, xi):
    """Quadrupole pair functions (field_quad_1..3, grad_quad_1..4) at distances r: (len(r), 7)."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros((r.size, 7))
    lib().orc_quad_pair_functions(_d(r), r.size, float(xi), _d(out))
    return out


def quad_self_grad(xi):
    return lib().orc_quad_self_grad(float(xi))


def poly_pair_functions(r, a, b, xi):
    """Polydisperse real-space pair functions (perp, para) of radii a, b at distances r."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    pe = np.zeros_like(r)
    pa = np.zeros_like(r)
    lib().orc_poly_pair_functions(_d(r), r.size, float(a), float(b), float(xi), _d(pe), _d(pa))
    return pe, pa


def poly_self(a, b, xi):
    return lib().orc_poly_self(float(a), float(b), float(xi))


def poly_weight(gd, a, xi, eta):
    """Finite-sphere spread weight vector of radius a at grid offset gd (3,)."""
    gd = np.ascontiguousarray(gd, dtype=np.float64)
    w = np.zeros(3)
    lib().orc_poly_weight(_d(gd), float(a), float(xi), float(eta), _d(w))
    return w


class Operator:
    """E = A p for the monodisperse Ewald operator (radius 1, SOLVER_AX mode)."""

    def __init__(self, box, errortol, xi, grid=None, P=None):
        L = np.asarray(box, dtype=np.float64)
        g = np.zeros(3, np.int32) if grid is None else np.asarray(grid, np.int32)
        err = ctypes.create_string_buffer(256)
        self._h = lib().orc_op_create(_d(L), float(errortol), float(xi), g.ctypes.data_as(_ip),
                                      int(P or 0), err, 256)
        if not self._h:
            raise RuntimeError(err.value.decode())
        self.N = 0
        self.Q = 0

    def set_quads(self, quad_idxs=None):
        """Quadrupoles on the particles quad_idxs (all when None); after set_positions.
        Unknowns then are [3N dipole components, 5Q quadrupole components]."""
        if quad_idxs is None:
            q = lib().orc_op_set_quads(self._h, None, -1)
        else:
            idx = np.ascontiguousarray(quad_idxs, np.int32)
            q = lib().orc_op_set_quads(self._h, idx.ctypes.data_as(_ip), idx.size)
        if q < 0:
            raise ValueError("bad quadrupole indices")
        self.Q = q
        return q

    @property
    def n(self):
        return 3 * self.N + 5 * self.Q

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_op_destroy(self._h)
            self._h = None

    def set_bloch(self, kx, ky):
        lib().orc_op_set_bloch(self._h, float(kx), float(ky))

    def set_positions(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        self.N = pos.shape[0]
        return lib().orc_op_set_positions(self._h, _d(pos), self.N)

    def info(self):
        n = np.zeros(3, np.int32)
        h = np.zeros(3)
        eta = np.zeros(3)
        P = ctypes.c_int()
        w = ctypes.c_int()
        rc = ctypes.c_double()
        sp = ctypes.c_double()
        npairs = ctypes.c_long()
        lib().orc_op_info(self._h, n.ctypes.data_as(_ip), _d(h), _d(eta), ctypes.byref(P),
                          ctypes.byref(w), ctypes.byref(rc), ctypes.byref(sp),
                          ctypes.byref(npairs))
        return dict(grid=tuple(int(v) for v in n), h=h, eta=eta, P=P.value, width=w.value,
                    rc=rc.value, self_perp=sp.value, num_pairs=npairs.value)

    def apply(self, p, self_coef, parts=3):
        """p: (N, 3) complex; self_coef: (N,) complex. parts: 1 real+self, 2 recip, 3 both."""
        shape = (self.N, 3) if self.Q == 0 else (self.n,)
        p = np.ascontiguousarray(p, dtype=np.complex128).reshape(shape)
        sc = np.ascontiguousarray(np.broadcast_to(self_coef, (self.N,)), dtype=np.complex128)
        E = np.zeros(shape, np.complex128)
        lib().orc_op_apply(self._h, _d(sc.view(np.float64)), _d(p.view(np.float64)),
                           _d(E.view(np.float64)), int(parts))
        return E

    def solve(self, b, self_coef, x0=None, tol=1e-6, method="gmres", restart=50, maxiter=1000,
              jacobi=True):
        b = np.ascontiguousarray(b, dtype=np.complex128).reshape(-1)
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.complex128).reshape(-1).copy()
        sc = np.ascontiguousarray(np.broadcast_to(self_coef, (self.N,)), dtype=np.complex128)
        rep = np.zeros(5)
        lib().orc_solve(self._h, 0 if method == "gmres" else 1, _d(sc.view(np.float64)),
                        _d(b.view(np.float64)), _d(x.view(np.float64)), float(tol), int(restart),
                        int(maxiter), int(bool(jacobi)), _d(rep))
        report = dict(iters=int(rep[0]), matvecs=int(rep[1]), converged=bool(rep[2]),
                      relres_pre=float(rep[3]), relres_true=float(rep[4]))
        return x, report


def self_coefficient(eps, radius=1.0):
    """-3 / (4 pi (1 - eps) a^3) with the |1-eps| < 1e-12 guard (dipole_solver.cu:453-460)."""
    eps = np.asarray(eps, dtype=np.complex128)
    diff = 1.0 - eps
    diff = np.where(np.abs(diff) < 1e-12, 1e-12, diff)
    return -3.0 / (4.0 * np.pi * diff) / np.asarray(radius, dtype=np.float64) ** 3


def near_field_intensity(box, E0, radius, dip, dip_pos, field_points, xi=0.5, errortol=1e-3):
    """|E0 + E_ind|^2 at the field points (src/cpp/near_field.cu:105-213, monodisperse):
    lengths over the first radius; E_ind(f) = -(row f of A [p; 0]) with the field points as
    sites without dipoles and zero self coefficients, i.e. the Gaussian contraction of the
    reciprocal grid at f plus the pair terms of every particle within rc at d_eff = max(d, 1)
    (the INTERACTION_FIELD evaluation, monodisperse_ewald_electric_field.cu:1631-1869,
    2385-2411, with its final negation)."""
    a = float(np.atleast_1d(radius)[0])
    dip = np.asarray(dip, np.complex128).reshape(-1, 3)
    N = dip.shape[0]
    fp = np.asarray(field_points, np.float64).reshape(-1, 3)
    op = Operator(np.asarray(box, np.float64) / a, errortol, xi)
    op.set_positions(np.concatenate([np.asarray(dip_pos, np.float64).reshape(-1, 3), fp]) / a)
    x = np.concatenate([dip, np.zeros_like(fp, dtype=np.complex128)])
    y = op.apply(x, 0.0)[N:]
    return np.sum(np.abs(np.asarray(E0, np.complex128)[None, :] - y) ** 2, axis=1)


class DipoleSolver:
    """Restatement of Dipole_Solver for one or more frames (Ewald field, dipoles and optional
    quadrupoles).

    eps_p: (W, N) complex in the medium's units (divided by eps_m here);
    E0: (K, 3), (K, 3N) or (K, 3N + 5Q) incident fields; quadrupoles: False, True (every
    particle) or the indices of the particles that carry one (dipole_solver.cu:131-182).
    Physical outputs as the reference.
    """

    def __init__(self, box, eps_p, radius, eps_m=1.0, xi=0.5, tol=1e-3, guess_type="derivative",
                 solver_type="gmres", E0=None, restart=50, maxiter=1000, jacobi=True,
                 grid=None, P=None, ewald_tol=None, field_type="auto", quadrupoles=False):
        self.box = np.asarray(box, np.float64)
        self.eps_p = np.asarray(eps_p, np.complex128)
        self.radius = np.asarray(radius, np.float64)  # scalar or one per particle
        self.field_type = field_type
        self.eps_m = float(eps_m)
        self.xi = xi
        self.tol = tol
        self.guess_type = guess_type
        self.solver_type = solver_type
        self.E0 = np.eye(3, dtype=np.complex128) if E0 is None else np.asarray(E0, np.complex128)
        self.restart = restart
        self.maxiter = maxiter
        self.jacobi = jacobi
        self.grid = grid
        self.P = P
        self.ewald_tol = ewald_tol
        self.op = None
        if isinstance(quadrupoles, bool):
            self.quad_idxs = None if quadrupoles else []
        else:
            self.quad_idxs = [int(i) for i in quadrupoles]
        self.quads = []      # per frame (W, Q, K, 5)
        self.dips = []       # per frame (W, N, K, 3)
        self.alpha = []      # per frame (W, K, 3)
        self.reports = []

    def _setup(self, N):
        W = self.eps_p.shape[0] if self.eps_p.ndim >= 1 else 1
        eps = self.eps_p.reshape(W, -1) if self.eps_p.ndim else self.eps_p.reshape(1, 1)
        if eps.shape[1] == 1:
            eps = np.repeat(eps, N, axis=1)
        self.eps = eps / self.eps_m
        self.W = eps.shape[0]
        rad = np.broadcast_to(self.radius, (N,)).astype(np.float64)
        # dipole_solver.cu:184-221 (lengths over the smallest radius) and :527-536 (auto:
        # polydisperse iff the radii differ)
        self.length_scale = rad.min()
        self.L = self.box / self.length_scale
        self.a = rad / self.length_scale
        self.vol_frac = 4.0 / 3.0 * np.pi * np.sum(self.a ** 3) / np.prod(self.L)
        poly = self.field_type == "polydisperse" or (self.field_type == "auto" and np.any(rad != rad[0]))
        if poly:
            self.op = PolyOperator(self.L, self.ewald_tol or self.tol, self.xi, self.a, self.grid, self.P)
        else:
            self.op = Operator(self.L, self.ewald_tol or self.tol, self.xi, self.grid, self.P)
        if self.quad_idxs is None or self.quad_idxs:
            if poly:
                raise NotImplementedError("quadrupoles with the polydisperse field")
            self.quad_owner = np.arange(N) if self.quad_idxs is None else np.asarray(self.quad_idxs)
            self.quad_set = True

    def _rhs(self, N):
        """(K, n) right-hand sides; dipole_solver.cu:368-390 (quadrupole rows zero unless given)."""
        K = self.E0.shape[0]
        n = self.op.n
        out = np.zeros((K, n), np.complex128)
        for k in range(K):
            e = self.E0[k].reshape(-1)
            if e.size == 3:
                out[k, :3 * N] = np.tile(e, N)
            elif e.size == 3 * N:
                out[k, :3 * N] = e
            elif e.size == n:
                out[k] = e
            else:
                raise RuntimeError(f"Incident field E0[{k}] has size {e.size}")
        return out

    def _mean_field(self, w, N):
        """dipole_solver.cu:244-261: alpha_MF E0 on the dipoles, zero quadrupoles."""
        beta = (self.eps[w] - 1.0) / (self.eps[w] + 2.0)
        val = 4.0 * np.pi * self.a ** 3 * beta / (1.0 - beta * self.vol_frac)  # per particle
        g = self._rhs(N).copy()
        g[:, :3 * N] *= np.repeat(val, 3)[None, :]
        g[:, 3 * N:] = 0.0
        return g

    def _guess(self, w, prev, N):
        """dipole_solver.cu:233-345 on (K, n) vectors; a quadrupole follows its particle's eps."""
        if self.guess_type in ("mean_field", "mean-field") or w < 1:
            return self._mean_field(w, N)
        if self.guess_type == "previous" or w < 2:
            return prev[w - 1].copy()
        if self.guess_type != "derivative":
            raise RuntimeError(f"Guess type {self.guess_type} not supported.")
        X = np.zeros_like(self.eps)
        ok = self.eps != 1.0
        X[ok] = (self.eps[ok] + 2.0) / (self.eps[ok] - 1.0)
        run = X[w - 1] - X[w - 2]
        new_run = X[w] - X[w - 1]
        small = np.abs(run) < 1e-12
        run = np.where(small, 1.0, run)
        new_run = np.where(small, 0.0, new_run)
        ratio = new_run / run  # per particle
        per_row = np.repeat(ratio, 3)
        if self.op.Q:
            per_row = np.concatenate([per_row, np.repeat(ratio[self.quad_owner], 5)])
        rise = prev[w - 1] - prev[w - 2]
        return prev[w - 1] + per_row[None, :] * rise

    def compute(self, positions):
        pos = np.asarray(positions, np.float64).reshape(-1, 3)
        N = pos.shape[0]
        if self.op is None:
            self._setup(N)
            if getattr(self, "quad_set", False):
                self.op.set_positions(pos / self.length_scale)
                self.op.set_quads(self.quad_idxs)
        self.op.set_positions(pos / self.length_scale)
        K = self.E0.shape[0]
        rhs = self._rhs(N)
        n = self.op.n
        frame = np.zeros((self.W, K, n), np.complex128)
        reps = []
        for w in range(self.W):
            sc = self_coefficient(self.eps[w], self.a)
            if w == 0 and self.dips:
                guess = self._last_first  # previous frame, omega 0
            else:
                guess = self._guess(w, frame, N)
            for k in range(K):
                x, rep = self.op.solve(rhs[k], sc, guess[k], self.tol,
                                       self.solver_type, self.restart, self.maxiter, self.jacobi)
                frame[w, k] = x.reshape(-1)
                reps.append(rep)
        self._last_first = frame[0].copy()
        dips = np.transpose(frame[:, :, :3 * N].reshape(self.W, K, N, 3), (0, 2, 1, 3))  # (W, N, K, 3)
        Q = self.op.Q
        self.quads.append(np.transpose(frame[:, :, 3 * N:].reshape(self.W, K, Q, 5), (0, 2, 1, 3)))
        self.dips.append(dips)
        self.alpha.append(dips.mean(axis=1))
        self.reports.append(reps)

    def get_eff_polarizability(self, physical=True):
        f = self.eps_m * self.length_scale ** 3 if physical else 1.0
        return np.squeeze(np.mean(np.stack(self.alpha), axis=0) * f)

    def get_dipoles(self, physical=True):
        f = self.eps_m * self.length_scale ** 3 if physical else 1.0
        return np.squeeze(np.stack(self.dips) * f)

    def get_quadrupoles(self, physical=True):
        """(F, W, Q, K, 5), squeezed; physical units scale by eps_m a_min^4 (dipole_solver.cu:640-647)."""
        f = self.eps_m * self.length_scale ** 4 if physical else 1.0
        return np.squeeze(np.stack(self.quads) * f)


class PolyOperator(Operator):
    """E = A p for the polydisperse Ewald operator (per-particle radii, nondimensional)."""

    def __init__(self, box, errortol, xi, radii, grid=None, P=None):
        L = np.asarray(box, dtype=np.float64)
        g = np.zeros(3, np.int32) if grid is None else np.asarray(grid, np.int32)
        r = np.ascontiguousarray(radii, dtype=np.float64)
        err = ctypes.create_string_buffer(256)
        self._h = lib().orc_poly_create(_d(L), float(errortol), float(xi), _d(r), int(r.size),
                                        g.ctypes.data_as(_ip), int(P or 0), err, 256)
        if not self._h:
            raise RuntimeError(err.value.decode())
        self.N = 0
        self.Q = 0
        self.radii = r

    def set_positions(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        if pos.shape[0] != self.radii.size:
            raise ValueError("one radius per particle")
        return super().set_positions(pos)
