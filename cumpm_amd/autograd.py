"""Differentiable solves on the GPU: ``torch.autograd.Function`` around mpm_solve /
mpm_solve_adjoint, with Drude parameters, extinction and the EELS loss as torch expressions.

No counterpart in the reference (its solvers return dipoles only).  The forward model is its
own: per-particle Drude permittivities eps_i(w) = eps_inf_i - wp_i^2 / (w^2 + i w gamma_i)
relative to the medium, self coefficients -3 / (4 pi (1 - eps) a^3) (dipole_solver.cu:453-461;
the EELS driver's drop the a^-3, eels_solver.cu:730-733, equal for unit radii), the dipoles
(and quadrupoles) x = A(sc)^-1 b, extinction C_ext = 4 pi k_w Im(E0^H p) with k_w = 2 pi w
sqrt(eps_m) (examples/mpm_hex_monolayer_sweep.py:117-119) and the EELS loss
sum Im(x . conj E_inc) / (2 pi^2) with the quadrupole rows weighted (1, 2, 2, 1, 2) / 6
(eels_solver.cu:844-880).

Backward of x = A(sc)^-1 b (torch's convention: the gradient of a real loss is conj-linear,
grad_in = J^H grad_out): grad_b = A^-H grad_x by mpm_solve_adjoint, and, since dA/dsc_i is
the identity on particle i's dipole rows and 2.5 times it on its quadrupole rows,
grad_sc_i = -sum_c w_c conj(x_ic) grad_b_ic.  Everything else (Drude, losses) is ordinary
torch autograd.

torch must own the process's HIP runtime: import this module (which imports torch) before
any other cumpm_amd call loads the library; otherwise two HIP runtimes would coexist, which
``load_with_torch`` detects and refuses.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi as C
from ._capi import check


def load_with_torch():
    """Load libcumpm_amd into a process whose HIP runtime is torch's; raise if the library
    was loaded before torch brought its own (two runtimes in one process)."""
    lib = C.load()
    maps = open("/proc/self/maps").read().split("\n")
    hips = sorted({ln.split()[-1] for ln in maps if "libamdhip64" in ln and "/" in ln})
    if len(hips) > 1:
        raise RuntimeError("two HIP runtimes are mapped (%s): import torch (or cumpm_amd.autograd) before "
                           "the first cumpm_amd call" % ", ".join(hips))
    return lib


class Operator:
    """An Ewald operator (monodisperse, radius 1 after scaling lengths by `radius`) at fixed
    positions, for differentiable solves.  quadrupoles: False, True or particle indices."""

    def __init__(self, box, positions, radius=1.0, quadrupoles=False, xi=0.5, ewald_tol=1e-6, grid=None,
                 support=0, tol=1e-10, restart=50, maxiter=2000, precision=C.SOLVE_DOUBLE):
        lib = load_with_torch()
        self.radius = float(radius)
        pos = np.ascontiguousarray(np.asarray(positions, np.float64).reshape(-1, 3) / self.radius)
        self.N = pos.shape[0]
        d = C.OpDesc()
        d.box[:] = [float(v) / self.radius for v in box]
        d.errortol = float(ewald_tol)
        d.xi = float(xi)
        if grid is not None:
            d.grid[:] = [int(g) for g in grid]
        d.support = int(support)
        self.h = ctypes.c_void_p()
        check(lib.mpm_op_create(ctypes.byref(d), ctypes.byref(self.h)))
        if quadrupoles is True:
            check(lib.mpm_op_set_quads(self.h, None, -1))
        elif quadrupoles is not False and quadrupoles is not None:
            idx = [int(i) for i in quadrupoles]
            check(lib.mpm_op_set_quads(self.h, (ctypes.c_int * len(idx))(*idx), len(idx)))
        check(lib.mpm_op_set_positions(self.h, C.dptr(pos), self.N))
        info = C.OpInfo()
        check(lib.mpm_op_get_info(self.h, ctypes.byref(info)))
        self.Q = int(info.num_quads)
        self.n = 3 * self.N + 5 * self.Q
        owners = (np.arange(self.N) if quadrupoles is True else np.asarray(quadrupoles, int)) if self.Q else []
        self.quad_owner = np.asarray(owners, np.int64)
        self.positions = pos
        self.desc = C.SolveDesc(method=C.GMRES, restart=int(restart), maxiter=int(maxiter), tol=float(tol), jacobi=1,
                                precision=int(precision))
        self.reports = []

    def __del__(self):
        h = getattr(self, "h", None)
        if h and C._lib is not None:
            C._lib.mpm_op_destroy(h)
            self.h = None

    def _solve(self, entry, b, sc):
        b = b.detach().contiguous()
        sc = sc.detach().contiguous()
        B = b.shape[0]
        stream = ctypes.c_void_p(torch.cuda.current_stream(b.device).cuda_stream)
        check(C.load().mpm_op_set_self_coef_device(self.h, ctypes.c_void_p(sc.data_ptr()), B, stream))
        x = torch.zeros_like(b)
        reps = (C.SolveReport * B)()
        check(getattr(C.load(), entry)(self.h, ctypes.byref(self.desc), ctypes.c_void_p(b.data_ptr()),
                                      ctypes.c_void_p(x.data_ptr()), B, reps, stream))
        self.reports.append([reps[i].as_dict() for i in range(B)])
        return x

    def solve(self, b, sc):
        """x = A(sc)^-1 b for B systems: b (B, 3N + 5Q), sc (B, N), complex128 CUDA tensors."""
        if b.dtype != torch.complex128 or sc.dtype != torch.complex128:
            raise TypeError("b and sc must be complex128")
        if b.shape[-1] != self.n or sc.shape != (b.shape[0], self.N):
            raise ValueError(f"b must be (B, {self.n}) and sc (B, {self.N})")
        if not (b.is_cuda and sc.is_cuda):
            raise ValueError("b and sc must live on the GPU")
        return _Solve.apply(b, sc, self)

    def dipoles(self, x):
        return x[:, :3 * self.N].reshape(x.shape[0], self.N, 3)

    def quadrupoles(self, x):
        return x[:, 3 * self.N:].reshape(x.shape[0], self.Q, 5)


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, sc, op):
        x = op._solve("mpm_solve", b, sc)
        ctx.op = op
        ctx.save_for_backward(x, sc)
        return x

    @staticmethod
    def backward(ctx, gx):
        x, sc = ctx.saved_tensors
        op = ctx.op
        gb = op._solve("mpm_solve_adjoint", gx.contiguous(), sc)
        N, Q = op.N, op.Q
        gsc = -(torch.conj(x[:, :3 * N]) * gb[:, :3 * N]).reshape(-1, N, 3).sum(-1)
        if Q:
            tq = -2.5 * (torch.conj(x[:, 3 * N:]) * gb[:, 3 * N:]).reshape(-1, Q, 5).sum(-1)
            own = torch.as_tensor(op.quad_owner, device=gsc.device)
            gsc = gsc.index_add(1, own, tq)
        return gb, gsc, None


class PolyOperator:
    """The polydisperse Ewald operator at fixed positions with the particle radii as an input
    of the differentiable solve (`solve(b, sc, radii)`), in the units of box, positions and
    radii as given (no rescaling: pick units where the radii are of order one).  The grid
    and support are fixed at construction (from the initial radii), so that radius changes
    move the operator smoothly; the real-space cutoff is max(sqrt(-ln ewald_tol) / xi,
    2 a_max), which moves with the largest radius only when 2 a_max sets it.  Dipoles only
    (the polydisperse operator carries no quadrupoles), no Bloch wave vector."""

    def __init__(self, box, positions, radii, xi=0.5, ewald_tol=1e-6, grid=None, support=0, tol=1e-10, restart=50,
                 maxiter=2000, precision=C.SOLVE_DOUBLE):
        load_with_torch()
        self.radius = 1.0
        self.positions = np.ascontiguousarray(np.asarray(positions, np.float64).reshape(-1, 3))
        self.N = self.positions.shape[0]
        self.Q = 0
        self.n = 3 * self.N
        self.quad_owner = np.zeros(0, np.int64)
        self.box = [float(v) for v in box]
        self.xi = float(xi)
        self.ewald_tol = float(ewald_tol)
        self.grid = None if grid is None else [int(g) for g in grid]
        self.support = int(support)
        self.desc = C.SolveDesc(method=C.GMRES, restart=int(restart), maxiter=int(maxiter), tol=float(tol), jacobi=1,
                                precision=int(precision))
        self.reports = []
        self.h = None
        self._radii = None
        self._build(np.asarray(radii, np.float64).reshape(-1))
        info = C.OpInfo()
        check(C.load().mpm_op_get_info(self.h, ctypes.byref(info)))
        self.grid = [int(g) for g in info.grid]
        self.support = int(info.support)

    def _build(self, radii):
        if radii.shape != (self.N,):
            raise ValueError(f"need {self.N} radii")
        if self._radii is not None and np.array_equal(radii, self._radii):
            return
        lib = C.load()
        d = C.OpDesc()
        d.box[:] = self.box
        d.errortol = self.ewald_tol
        d.xi = self.xi
        if self.grid is not None:
            d.grid[:] = self.grid
        d.support = self.support
        r = np.ascontiguousarray(radii, np.float64)
        h = ctypes.c_void_p()
        check(lib.mpm_op_create_poly(ctypes.byref(d), C.dptr(r), self.N, ctypes.byref(h)))
        try:
            check(lib.mpm_op_set_positions(h, C.dptr(self.positions), self.N))
        except Exception:
            lib.mpm_op_destroy(h)
            raise
        self.__del__()
        self.h = h
        self._radii = r.copy()

    __del__ = Operator.__del__
    _solve = Operator._solve
    dipoles = Operator.dipoles

    def solve(self, b, sc, radii):
        """x = A(sc, radii)^-1 b for B systems: b (B, 3N), sc (B, N) complex128 and radii (N,)
        float64, CUDA tensors; differentiable in all three."""
        if b.dtype != torch.complex128 or sc.dtype != torch.complex128 or radii.dtype != torch.float64:
            raise TypeError("b and sc must be complex128, radii float64")
        if b.shape[-1] != self.n or sc.shape != (b.shape[0], self.N) or radii.shape != (self.N,):
            raise ValueError(f"b must be (B, {self.n}), sc (B, {self.N}) and radii ({self.N},)")
        if not (b.is_cuda and sc.is_cuda and radii.is_cuda):
            raise ValueError("b, sc and radii must live on the GPU")
        return _PolySolve.apply(b, sc, radii, self)


class _PolySolve(torch.autograd.Function):
    """Backward: grad_b and grad_sc as for _Solve; the operator part of the radius gradient
    is dL/da_i = -Re(sum_b grad_b^H (dA/da_i) x_b) (mpm_op_radius_gradient)."""

    @staticmethod
    def forward(ctx, b, sc, radii, op):
        op._build(radii.detach().cpu().numpy())
        x = op._solve("mpm_solve", b, sc)
        ctx.op = op
        ctx.save_for_backward(x, sc, radii)
        return x

    @staticmethod
    def backward(ctx, gx):
        x, sc, radii = ctx.saved_tensors
        op = ctx.op
        op._build(radii.detach().cpu().numpy())  # another forward may have moved the radii
        gb = op._solve("mpm_solve_adjoint", gx.contiguous(), sc)
        N = op.N
        gsc = -(torch.conj(x) * gb).reshape(-1, N, 3).sum(-1)
        out = torch.empty(N, dtype=torch.float64, device=x.device)
        xc = x.detach().contiguous()
        stream = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
        check(C.load().mpm_op_radius_gradient(op.h, ctypes.c_void_p(xc.data_ptr()), ctypes.c_void_p(gb.data_ptr()),
                                              xc.shape[0], ctypes.c_void_p(out.data_ptr()), 0, stream))
        return gb, gsc, -out, None


def drude(omega, eps_inf, wp, gamma, eps_m=1.0):
    """(W, N) permittivities relative to the medium: eps_inf - wp^2 / (w^2 + i w gamma), every
    parameter a (N,) tensor (or scalar), omega (W,)."""
    w = torch.as_tensor(omega, dtype=torch.float64, device=wp.device)[:, None]
    den = w * w + 1j * w * gamma
    return (eps_inf - wp * wp / den) / eps_m


def self_coefficient(eps, radius=1.0):
    """-3 / (4 pi (1 - eps) a^3) (dipole_solver.cu:453-461); radius a scalar or a tensor
    broadcasting against eps (per-particle radii: (N,))."""
    return -3.0 / (4.0 * np.pi * (1.0 - eps)) / radius ** 3


def extinction(p, E0, omega, eps_m=1.0):
    """Sum over systems of C_ext = 4 pi k_w Im(E0^H p): p (W, K, N, 3), E0 (K, 3)."""
    w = torch.as_tensor(omega, dtype=torch.float64, device=p.device)
    k = 2.0 * np.pi * w * np.sqrt(eps_m)
    e0 = torch.as_tensor(E0, dtype=torch.complex128, device=p.device)
    s = torch.einsum("kc,wknc->wk", torch.conj(e0), p)
    return (4.0 * np.pi * k[:, None] * s.imag).sum()


def eels_loss(x, einc, N, Q):
    """Sum over systems of the EELS reciprocity sum: x, einc (W, 3N + 5Q)."""
    s = (x[:, :3 * N] * torch.conj(einc[:, :3 * N])).imag.sum()
    if Q:
        wq = torch.tensor([1.0, 2.0, 2.0, 1.0, 2.0], dtype=torch.float64, device=x.device).repeat(Q)
        s = s + (wq * (x[:, 3 * N:] * torch.conj(einc[:, 3 * N:])).imag).sum() / 6.0
    return s / (2.0 * np.pi ** 2)


def ebeam_field(op, omega, v, epos, sigma=0.5):
    """The EELS driver's incident field at the operator's particles (lengths in the units of
    `op`'s box and radius): (W, 3N + 5Q) complex128 on the GPU, for wavenumbers omega, speed
    v / c, beam position epos (x, y) and probe softening sigma."""
    a = op.radius
    om = np.ascontiguousarray(np.asarray(omega, np.float64).reshape(-1) * a)
    W = om.size
    out = np.zeros((W, op.n), np.complex128)
    q = np.ascontiguousarray(op.quad_owner, np.int32)
    e = np.ascontiguousarray(np.asarray(epos, np.float64).reshape(2) / a)
    check(C.load().mpm_ebeam_field(C.dptr(op.positions), op.N, q.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), op.Q,
                                   C.dptr(om), W, float(v), C.dptr(e), float(sigma) / a, C.dptr(out.view(np.float64))))
    return torch.from_numpy(out).cuda()


def uniform_rhs(op, E0, W):
    """(W K, 3N + 5Q) right-hand sides of uniform incident fields E0 (K, 3), system w K + k."""
    e0 = torch.as_tensor(np.asarray(E0, np.complex128).reshape(-1, 3))
    K = e0.shape[0]
    b = torch.zeros((W * K, op.n), dtype=torch.complex128)
    b[:, :3 * op.N] = torch.stack([e0[k].repeat(op.N) for k in range(K)]).repeat(W, 1)
    return b.cuda()
