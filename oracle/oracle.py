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


def near_field_intensity(box, E0, radius, dip, dip_pos, field_points, xi=0.5, errortol=1e-3, field_type="auto"):
    """|E0 + E_ind|^2 at the field points (src/cpp/near_field.cu:105-213): lengths over the
    first radius; E_ind(f) = -(row f of A [p; 0]) with the field points as sites without dipoles
    and zero self coefficients, i.e. the contraction of the reciprocal grid at f plus the pair
    terms of every particle within rc (the INTERACTION_FIELD evaluation,
    monodisperse_ewald_electric_field.cu:1631-1869, 2385-2411, with its final negation).
    Polydisperse (field_type "polydisperse", or "auto" with unequal radii,
    near_field.cu:121-147): field point f is a site of particle f's radius for f < N; beyond,
    the reference's kernels take the first radius for its spread weight
    (polydisperse_ewald_electric_field.cu:297) and the smallest radius class for its pair
    functions (:402); this restatement covers inputs where the two agree (M <= N, or the first
    radius the smallest) and raises otherwise."""
    r = np.atleast_1d(np.asarray(radius, np.float64))
    a = float(r[0])
    dip = np.asarray(dip, np.complex128).reshape(-1, 3)
    N = dip.shape[0]
    fp = np.asarray(field_points, np.float64).reshape(-1, 3)
    M = fp.shape[0]
    if r.size == 1:
        r = np.full(N, a)
    poly = field_type == "polydisperse" or (field_type == "auto" and np.any(r != r[0]))
    if poly:
        if M > N and r[0] != r.min():
            raise NotImplementedError("field points beyond the N-th take two different radii in the reference")
        fr = np.array([r[f] if f < N else r.min() for f in range(M)])
        op = PolyOperator(np.asarray(box, np.float64) / a, errortol, xi, np.concatenate([r, fr]) / a)
        op.set_positions(np.concatenate([np.asarray(dip_pos, np.float64).reshape(-1, 3), fp]) / a)
        x = np.concatenate([dip, np.zeros_like(fp, dtype=np.complex128)])
        y = op.apply(x, 0.0)[N:]
        return np.sum(np.abs(np.asarray(E0, np.complex128)[None, :] - y) ** 2, axis=1)
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


class DirectOperator:
    """A p for the direct (open-boundary) field in the solver's mode, dense numpy restatement
    of src/cpp/electric_field/direct_electric_field.cu:37-300 (direct_field_kernel, mode
    SOLVER_AX): for every target j and source i != j, r = x_i - x_j clamped to the source
    radius when shorter (a zero-length r keeps its direction, i.e. none), the point dipole
    field c (3 (p.u) u - p) / r^3 with c = -1/(4 pi), with quadrupoles the point quadrupole
    field, and for quadrupole targets the dipole and quadrupole field gradients; the self
    terms (sc + 1/(4 pi a^3)) p and (2.5 sc / a^2 + 3/(4 pi a^5)) Q.

    Quadrupole couplings follow the Ewald operator's conventions, not the direct kernel's:
    the components (Q0..Q4) = (xx, xy, xz, yy, yz) enter the fields as the traceless matrix
    Qm = -[[Q0, Q1, Q2], [Q1, Q3, Q4], [Q2, Q4, -Q0 - Q3]] (the kernel doubles the off-diagonal
    entries, :206-212), and the gradient of a quadrupole's field is the derivative of that
    field, (-3c / r^5) [5 (v u^T + u v^T) + 2.5 s I - 17.5 s u u^T - Qm] with v = Qm u,
    s = u.v, in the components (xx, xy, xz, yy, yz) of d_target E (the kernel's :253-272
    differs in sign and in the Qm term).  With these the direct field is the large-box limit
    of the Ewald operator for every coupling, which the reference's kernel is not.
    """

    def __init__(self, radii, quad_idxs=None):
        self.radii = np.asarray(radii, np.float64).reshape(-1)
        self.quad_idxs = quad_idxs  # None: no quadrupoles; True: every particle; or indices
        self.N = 0
        self.Q = 0

    @property
    def n(self):
        return 3 * self.N + 5 * self.Q

    def set_positions(self, pos):
        self.pos = np.asarray(pos, np.float64).reshape(-1, 3)
        self.N = self.pos.shape[0]
        a = np.broadcast_to(self.radii, (self.N,)) if self.radii.size == 1 else self.radii
        if a.size != self.N:
            raise ValueError("one radius per particle")
        self.a = np.asarray(a, np.float64)
        if self.quad_idxs is None or self.quad_idxs is False:
            self.owner = np.zeros(0, int)
        elif self.quad_idxs is True:
            self.owner = np.arange(self.N)
        else:
            self.owner = np.asarray(self.quad_idxs, int)
        self.Q = self.owner.size
        self.qmap = -np.ones(self.N, int)
        self.qmap[self.owner] = np.arange(self.Q)

    @staticmethod
    def _qmat(q):
        return -np.array([[q[0], q[1], q[2]],
                          [q[1], q[3], q[4]],
                          [q[2], q[4], -q[0] - q[3]]])

    def apply(self, x, self_coef):
        N, Q = self.N, self.Q
        x = np.asarray(x, np.complex128).reshape(-1)
        sc = np.broadcast_to(np.asarray(self_coef, np.complex128), (N,))
        p = x[:3 * N].reshape(N, 3)
        qv = x[3 * N:].reshape(Q, 5)
        E = np.zeros((N, 3), np.complex128)
        G = np.zeros((Q, 5), np.complex128)
        c = -1.0 / (4.0 * np.pi)
        I5 = np.array([1.0, 0, 0, 1.0, 0])
        for j in range(N):
            qt = self.qmap[j]
            for i in range(N):
                if i == j:
                    continue
                r = self.pos[i] - self.pos[j]
                r2 = r @ r
                ai = self.a[i]
                if r2 < ai * ai:
                    if np.sqrt(r2) > 1e-9:
                        r = r * np.sqrt(ai * ai / r2)
                    r2 = ai * ai
                if r2 < 1e-18:
                    continue
                d = np.sqrt(r2)
                u = r / d
                pu = p[i] @ u
                E[j] += c / d ** 3 * (3.0 * pu * u - p[i])
                if Q == 0:
                    continue
                qs = self.qmap[i]
                Qm = self._qmat(qv[qs]) if qs >= 0 else np.zeros((3, 3), np.complex128)
                v = Qm @ u
                s = u @ v
                E[j] += 3.0 * c / d ** 4 * (2.5 * s * u - v)
                if qt < 0:
                    continue
                rr = np.array([u[0] * u[0], u[0] * u[1], u[0] * u[2], u[1] * u[1], u[1] * u[2]])
                sym = lambda a_, b_: np.array([2 * a_[0] * b_[0], a_[0] * b_[1] + a_[1] * b_[0],
                                               a_[0] * b_[2] + a_[2] * b_[0], 2 * a_[1] * b_[1],
                                               a_[1] * b_[2] + a_[2] * b_[1]])
                Gd = pu * (-5.0 * rr + I5) + sym(p[i], u)
                G[qt] += -3.0 * c / d ** 4 * Gd
                Qc = np.array([Qm[0, 0], Qm[0, 1], Qm[0, 2], Qm[1, 1], Qm[1, 2]])
                Gq = 5.0 * sym(v, u) + 2.5 * s * I5 - 17.5 * s * rr - Qc
                G[qt] += -3.0 * c / d ** 5 * Gq
        a3 = self.a ** 3
        E += ((sc + 1.0 / (4.0 * np.pi * a3))[:, None]) * p
        if Q:
            ao = self.a[self.owner]
            G += ((2.5 * sc[self.owner] / ao ** 2 + 3.0 / (4.0 * np.pi * ao ** 5))[:, None]) * qv
        return np.concatenate([E.reshape(-1), G.reshape(-1)])

    def matrix(self, self_coef):
        n = self.n
        A = np.zeros((n, n), np.complex128)
        e = np.zeros(n, np.complex128)
        for k in range(n):
            e[:] = 0.0
            e[k] = 1.0
            A[:, k] = self.apply(e, self_coef)
        return A

    def solve(self, b, self_coef, **_):
        """Exact solution of A x = b (dense LU; sizes where the direct field is used are small)."""
        return np.linalg.solve(self.matrix(self_coef), np.asarray(b, np.complex128).reshape(-1))


# ---------------------------------------------------------------------------
# EELS (src/cpp/eels_solver.cu:241-880)
# ---------------------------------------------------------------------------
def surface_normals24():
    """The 24 surface points of the quadrupole gradient rule: a Fibonacci sphere with
    z_k = 1 - (2k + 1)/24 and azimuth -(k + 1/2) pi (3 - sqrt 5) (eels_solver.cu:1-27)."""
    k = np.arange(24)
    z = 1.0 - (2.0 * k + 1.0) / 24.0
    r = np.sqrt(1.0 - z * z)
    a = -(k + 0.5) * np.pi * (3.0 - np.sqrt(5.0))
    return np.stack([r * np.cos(a), r * np.sin(a), z], axis=1)


def beam_field(points, epos, omega, v, sigma, gamma=None):
    """Relativistic field of a charge moving along z through (epos[0], epos[1]) at speed v
    (units of c), frequency omega (wavenumber), at points (M, 3) (eels_solver.cu:610-640):
    E_perp = -A K1(xi) r/r_s, E_z = A (i/gamma) K0(xi), A = omega/(v^2 gamma) e^{2 pi i omega
    z / v}, xi = 2 pi omega r_s / (v gamma), r_s = sqrt(r_perp^2 + sigma^2)."""
    import scipy.special as sp
    if gamma is None:
        gamma = 1.0 / np.sqrt(1.0 - v * v)
    p = np.asarray(points, np.float64).reshape(-1, 3)
    dx = p[:, 0] - epos[0]
    dy = p[:, 1] - epos[1]
    rs = np.sqrt(dx * dx + dy * dy + sigma * sigma)
    xi = 2.0 * np.pi * omega * rs / (v * gamma)
    A = omega / (v * v * gamma) * np.exp(1j * 2.0 * np.pi * omega * p[:, 2] / v)
    E = np.zeros((p.shape[0], 3), np.complex128)
    E[:, 0] = -A * sp.kv(1, xi) * dx / rs
    E[:, 1] = -A * sp.kv(1, xi) * dy / rs
    E[:, 2] = A * 1j / gamma * sp.kv(0, xi)
    return E


def beam_gradient(centres, epos, omega, v, sigma, R=1.0):
    """Volume-averaged gradient <d_i E_j> of the beam field over spheres of radius R by the
    24-point surface rule 3/(24 R) sum_k E(x + R n_k) n_k^T, as the five components
    (Gxx - Gzz, Gxy, Gxz, Gyy - Gzz, Gyz) of its symmetric part (eels_solver.cu:646-706)."""
    n = surface_normals24()
    c = np.asarray(centres, np.float64).reshape(-1, 3)
    out = np.zeros((c.shape[0], 5), np.complex128)
    for i in range(c.shape[0]):
        E = beam_field(c[i][None, :] + R * n, epos, omega, v, sigma)  # (24, 3)
        T = 3.0 / (24.0 * R) * (n.T @ E)  # T[i][j] = sum_k n_i E_j
        out[i] = [T[0, 0] - T[2, 2], 0.5 * (T[0, 1] + T[1, 0]), 0.5 * (T[0, 2] + T[2, 0]),
                  T[1, 1] - T[2, 2], 0.5 * (T[1, 2] + T[2, 1])]
    return out


class EELSSolver:
    """Restatement of EELS_Solver for one frame per beam position (no aperiodic sampling).

    box, positions, radius, epos in physical units; omega wavenumbers; eps_p (W, N) or (W, 1)
    relative to vacuum (divided by eps_m here).  field: "direct", "monodisperse" or
    "polydisperse".  Solutions are exact (dense LU) for the direct field and GMRES(50) to tol
    for the Ewald fields; the loss is the reciprocity sum (eels_solver.cu:844-880)."""

    def __init__(self, box, eps_p, omega, v, radius=1.0, eps_m=1.0, xi=0.5, tol=1e-8, field="direct",
                 quad_idxs=None, asm=None):
        self.box = np.asarray(box, np.float64)
        self.eps_p = np.asarray(eps_p, np.complex128)
        if self.eps_p.ndim == 1:
            self.eps_p = self.eps_p[:, None]
        self.omega = np.asarray(omega, np.float64)
        self.v = float(v)
        self.radius = radius
        self.eps_m = float(eps_m)
        self.xi = xi
        self.tol = tol
        self.field = field
        self.quad_idxs = quad_idxs  # None: no quadrupoles, True: every particle, or indices
        self.asm = asm              # None, or (Nx, Ny) Bloch points of the aperiodic sampling
        self.eels, self.dips, self.quads, self.einc = [], [], [], []

    def compute(self, epos, positions):
        pos = np.asarray(positions, np.float64).reshape(-1, 3)
        N = pos.shape[0]
        rad = np.broadcast_to(np.asarray(self.radius, np.float64), (N,)).astype(np.float64)
        ls = rad.min()
        a = rad / ls
        L = self.box / ls
        x = pos / ls
        om = self.omega * ls
        eps = np.broadcast_to(self.eps_p, (self.eps_p.shape[0], N)) / self.eps_m
        W = eps.shape[0]
        e = np.asarray(epos, np.float64) / ls
        sigma = 0.5 / ls
        if self.field == "direct":
            op = DirectOperator(a, self.quad_idxs if self.quad_idxs is not None else None)
            op.set_positions(x)
        elif self.field == "polydisperse":
            op = PolyOperator(L, self.tol, self.xi, a)
            op.set_positions(x)
        else:
            op = Operator(L, self.tol, self.xi)
            op.set_positions(x)
            if self.quad_idxs is not None:
                op.set_quads(None if self.quad_idxs is True else self.quad_idxs)
        Q = op.Q
        owner = (np.arange(N) if self.quad_idxs is True else np.asarray(self.quad_idxs, int)) if Q else np.zeros(0, int)
        n = 3 * N + 5 * Q
        gamma = 1.0 / np.sqrt(1.0 - self.v ** 2)
        # aperiodic sampling (eels_solver.cu:521-537, 584-607, 858-872): Bloch points k = -pi/L +
        # (i + 1/2) dk, the beam's images within 12 v gamma / (2 pi w) summed with the phase
        # e^{i k.(x - R)}, the loss and dipoles averaged over the points
        Nx, Ny = self.asm if self.asm else (1, 1)
        dkx, dky = (2 * np.pi / (L[0] * Nx), 2 * np.pi / (L[1] * Ny)) if self.asm else (0.0, 0.0)
        fe = np.zeros(W)
        fd = np.zeros((W, N, 3), np.complex128)
        fq = np.zeros((W, Q, 5), np.complex128)
        fi = []
        for ix in range(Nx):
            for iy in range(Ny):
                kx = -np.pi / L[0] + (ix + 0.5) * dkx if self.asm else 0.0
                ky = -np.pi / L[1] + (iy + 0.5) * dky if self.asm else 0.0
                if self.asm:
                    op.set_bloch(kx, ky)
                fi = []
                for w in range(W):
                    E = np.zeros(n, np.complex128)
                    if self.asm:
                        rd = 12.0 * abs(self.v * gamma / (2 * np.pi * om[w]))
                        rx, ry = int(np.ceil(rd / L[0])), int(np.ceil(rd / L[1]))
                    else:
                        rx = ry = 0
                    for jx in range(-rx, rx + 1):
                        for jy in range(-ry, ry + 1):
                            R = np.array([jx * L[0], jy * L[1]])
                            ph = np.exp(1j * (kx * (x[:, 0] - R[0]) + ky * (x[:, 1] - R[1])))
                            E[:3 * N] += (beam_field(x, e + R, om[w], self.v, sigma) * ph[:, None]).reshape(-1)
                            if Q:
                                E[3 * N:] += (beam_gradient(x[owner], e + R, om[w], self.v, sigma)
                                              * ph[owner][:, None]).reshape(-1)
                    sc = -3.0 / (4.0 * np.pi * (1.0 - eps[w]))  # without a^-3 (eels_solver.cu:730-733)
                    if isinstance(op, DirectOperator):
                        p = op.solve(E, sc)
                    else:
                        p, _ = op.solve(E, sc, tol=self.tol)
                    sm = np.sum((p[:3 * N] * np.conj(E[:3 * N])).imag)
                    if Q:
                        wq = np.tile([1.0, 2.0, 2.0, 1.0, 2.0], Q)
                        sm += np.sum((wq * p[3 * N:] * np.conj(E[3 * N:])).imag) / 6.0
                    fe[w] += sm / (2.0 * np.pi ** 2)
                    fd[w] += p[:3 * N].reshape(N, 3)
                    fq[w] += p[3 * N:].reshape(Q, 5)
                    fi.append(E)
        f = 1.0 / (Nx * Ny)
        fe, fd, fq = fe * f, fd * f, fq * f
        self.eels.append(np.array(fe))
        self.dips.append(np.array(fd))
        self.quads.append(np.array(fq))
        self.einc.append(np.array(fi))
        self.length_scale = ls

    def get_eels(self):
        return np.squeeze(np.array(self.eels))

    def get_dipoles(self, physical=True):
        f = self.eps_m * self.length_scale ** 3 if physical else 1.0
        return np.squeeze(np.array(self.dips) * f)

    def get_quadrupoles(self, physical=True):
        f = self.eps_m * self.length_scale ** 4 if physical else 1.0
        return np.squeeze(np.array(self.quads) * f)
