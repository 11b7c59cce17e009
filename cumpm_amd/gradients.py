"""Extinction of a frequency sweep and its gradient with respect to each particle's plasma
frequency, through one batched forward solve and one batched adjoint solve on the GPU.

No counterpart in the reference code (its solvers return dipoles only); the forward model
is its own: Drude permittivities eps_i(w) = eps_inf - wp_i^2 / (w^2 + i w gamma) relative to
the medium, self coefficients -3 / (4 pi (1 - eps_i)) (dipole_solver.cu:453-461), dipoles
(and optionally quadrupoles) from A p = E0, and the extinction of polarisation k at
frequency w as C_ext = 4 pi k_w Im(sum_i conj(E0_k) . p_i) with k_w = 2 pi w sqrt(eps_m)
(the reference's examples/mpm_hex_monolayer_sweep.py:117-119).

Gradient: with c the right-hand side conj(E0) on the dipole rows, dC/dtheta =
4 pi k Im(-lambda^T (dA/dtheta) p), where lambda = conj(y) and y solves A^H y = E0
(mpm_solve_adjoint); dA/dwp_i is the derivative of particle i's self coefficient times
the identity on its dipole rows and 2.5 times it on its quadrupole rows.
"""
import ctypes

import numpy as np

from . import _capi as C
from ._capi import check


def _self_coef(eps):
    diff = 1.0 - eps
    diff = np.where(np.abs(diff) < 1e-12, 1e-12, diff)
    return -3.0 / (4.0 * np.pi * diff)


def extinction_and_wp_gradient(box, positions, omega, eps_inf, wp, gamma, eps_m=1.0, radius=1.0, E0=None,
                               quadrupoles=False, xi=0.5, tol=1e-10, ewald_tol=1e-6, restart=50,
                               maxiter=2000, gradient=True):
    """(L, dL/dwp, reports) for L = sum over frequencies and polarisations of C_ext.

    box, positions: physical lengths (scaled here by the common radius); omega (W,), wp (N,)
    per-particle plasma frequencies, eps_inf and gamma scalars; E0 (K, 3) uniform incident
    fields (identity when None); quadrupoles: False, True or particle indices.  The solves
    run in FP64 to `tol` (both the forward and the adjoint), so that the gradient is exact to
    that tolerance for the operator at `ewald_tol`.  gradient=False skips the adjoint solve
    (the gradient is then None)."""
    lib = C.load()
    pos = np.ascontiguousarray(np.asarray(positions, np.float64).reshape(-1, 3) / radius)
    N = pos.shape[0]
    omega = np.atleast_1d(np.asarray(omega, np.float64))
    wp = np.asarray(wp, np.float64).reshape(-1)
    if wp.size != N:
        raise ValueError("one plasma frequency per particle")
    E0 = np.eye(3) if E0 is None else np.asarray(E0, np.complex128).reshape(-1, 3)
    W, K = omega.size, E0.shape[0]
    d = C.OpDesc()
    d.box[:] = [float(v) / radius for v in box]
    d.errortol = float(ewald_tol)
    d.xi = float(xi)
    h = ctypes.c_void_p()
    check(lib.mpm_op_create(ctypes.byref(d), ctypes.byref(h)))
    try:
        if quadrupoles is True:
            check(lib.mpm_op_set_quads(h, None, -1))
        elif quadrupoles is not False and quadrupoles is not None:
            idx = [int(i) for i in quadrupoles]
            check(lib.mpm_op_set_quads(h, (ctypes.c_int * len(idx))(*idx), len(idx)))
        check(lib.mpm_op_set_positions(h, C.dptr(pos), N))
        info = C.OpInfo()
        check(lib.mpm_op_get_info(h, ctypes.byref(info)))
        Q = int(info.num_quads)
        qidx = (np.arange(N) if quadrupoles is True else np.asarray(quadrupoles, int)) if Q else np.zeros(0, int)
        n = 3 * N + 5 * Q
        B = W * K
        den = omega ** 2 + 1j * omega * gamma                     # (W,)
        eps = (eps_inf - wp[None, :] ** 2 / den[:, None]) / eps_m  # (W, N)
        sc = _self_coef(eps)
        rows = np.ascontiguousarray(np.repeat(sc, K, axis=0))   # (B, N): system w K + k
        check(lib.mpm_op_set_self_coef(h, C.dptr(rows.view(np.float64)), B))
        b = np.zeros((B, n), np.complex128)
        for w in range(W):
            for k in range(K):
                b[w * K + k, :3 * N] = np.tile(E0[k], N)
        desc = C.SolveDesc(method=C.GMRES, restart=restart, maxiter=maxiter, tol=tol, jacobi=1,
                           precision=C.SOLVE_DOUBLE)
        bb = C.DeviceBuffer(b.nbytes)
        xb = C.DeviceBuffer(b.nbytes)
        bb.upload(b)
        xb.upload(np.zeros_like(b))
        reps = (C.SolveReport * B)()
        check(lib.mpm_solve(h, ctypes.byref(desc), bb.ptr, xb.ptr, B, reps, None))
        p = np.zeros_like(b)
        xb.download(p)
        reps_adj = (C.SolveReport * B)()
        y = np.zeros_like(b)
        if gradient:
            xb.upload(y)
            check(lib.mpm_solve_adjoint(h, ctypes.byref(desc), bb.ptr, xb.ptr, B, reps_adj, None))
            xb.download(y)
    finally:
        lib.mpm_op_destroy(h)
    lam = np.conj(y)
    kw = np.repeat(2.0 * np.pi * omega * np.sqrt(eps_m), K)   # (B,)
    # C_ext per system and their sum
    cext = 4.0 * np.pi * kw * np.imag(np.einsum("bc,bic->b", np.conj(np.repeat(E0[None], W, 0).reshape(B, 3)),
                                                p[:, :3 * N].reshape(B, N, 3)))
    L = float(cext.sum())
    if not gradient:
        return L, None, ([reps[i].as_dict() for i in range(B)], [])
    # lambda^T (dA/dsc_i) p per system and particle
    lp = np.einsum("bic,bic->bi", lam[:, :3 * N].reshape(B, N, 3), p[:, :3 * N].reshape(B, N, 3))
    if Q:
        lq = np.einsum("bqc,bqc->bq", lam[:, 3 * N:].reshape(B, Q, 5), p[:, 3 * N:].reshape(B, Q, 5))
        np.add.at(lp, (slice(None), qidx), 2.5 * lq)
    dsc_deps = -3.0 / (4.0 * np.pi * (1.0 - eps) ** 2)          # (W, N)
    deps_dwp = -2.0 * wp[None, :] / den[:, None] / eps_m        # (W, N)
    dsc = np.repeat(dsc_deps * deps_dwp, K, axis=0)            # (B, N)
    grad = np.sum(4.0 * np.pi * kw[:, None] * np.imag(-dsc * lp), axis=0)
    return L, grad, ([reps[i].as_dict() for i in range(B)], [reps_adj[i].as_dict() for i in range(B)])
