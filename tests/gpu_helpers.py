"""Thin helpers that drive the C ABI (include/cumpm_amd.h) from the tests."""
import ctypes

import numpy as np

from cumpm_amd import _capi as C


class Op:
    """Owns an mpm_op plus device buffers; every call goes through the C ABI."""

    def __init__(self, box, errortol, xi, grid=None, support=0, layout=0, fft_backend=0, radii=None, quads=None):
        d = C.OpDesc()
        d.layout = int(layout)
        d.fft_backend = int(fft_backend)
        d.box[:] = [float(v) for v in box]
        d.errortol = float(errortol)
        d.xi = float(xi)
        if grid is not None:
            d.grid[:] = [int(g) for g in grid]
        d.support = int(support)
        self.h = ctypes.c_void_p()
        if radii is None:
            C.check(C.load().mpm_op_create(ctypes.byref(d), ctypes.byref(self.h)))
        else:
            r = np.ascontiguousarray(radii, np.float64)
            C.check(C.load().mpm_op_create_poly(ctypes.byref(d), C.dptr(r), r.size, ctypes.byref(self.h)))
        self.N = 0
        self.Q = 0
        self._quads = quads
        if quads is True:
            C.check(C.load().mpm_op_set_quads(self.h, None, -1))
        elif quads is not None:
            idx = (ctypes.c_int * len(quads))(*[int(v) for v in quads])
            C.check(C.load().mpm_op_set_quads(self.h, idx, len(quads)))

    def __del__(self):
        if getattr(self, "h", None) and C._lib is not None:
            C._lib.mpm_op_destroy(self.h)

    def info(self):
        i = C.OpInfo()
        C.check(C.load().mpm_op_get_info(self.h, ctypes.byref(i)))
        return i

    def set_positions(self, pos):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        self.N = pos.shape[0]
        C.check(C.load().mpm_op_set_positions(self.h, C.dptr(pos), self.N))
        self.Q = int(self.info().num_quads)

    @property
    def n(self):
        return 3 * self.N + 5 * self.Q

    def _shape(self, B):
        return (B, self.N, 3) if self.Q == 0 else (B, self.n)

    def set_bloch(self, kx, ky):
        C.check(C.load().mpm_op_set_bloch(self.h, float(kx), float(ky)))

    def set_self_coef(self, rows):
        rows = np.ascontiguousarray(np.atleast_2d(rows), np.complex128)
        C.check(C.load().mpm_op_set_self_coef(self.h, C.dptr(rows.view(np.float64)), rows.shape[0]))

    def apply(self, p, precision=C.PREC_DOUBLE, parts=C.PART_ALL):
        p = np.ascontiguousarray(p, np.complex128)
        p = p.reshape(self._shape(p.size // self.n))
        B = p.shape[0]
        xb = C.DeviceBuffer(p.nbytes)
        yb = C.DeviceBuffer(p.nbytes)
        xb.upload(p)
        C.check(C.load().mpm_op_apply(self.h, xb.ptr, yb.ptr, B, int(precision), int(parts), None))
        out = np.zeros_like(p)
        yb.download(out)
        return out

    def solve_adjoint(self, g, y0=None, **kw):
        """A^H y = g through mpm_solve_adjoint."""
        return self.solve(g, y0, _entry="mpm_solve_adjoint", **kw)

    def solve(self, b, x0=None, tol=1e-8, method=C.GMRES, restart=50, maxiter=1000, jacobi=True,
              precision=C.SOLVE_DOUBLE, _entry="mpm_solve"):
        b = np.ascontiguousarray(b, np.complex128)
        b = b.reshape(self._shape(b.size // self.n))
        B = b.shape[0]
        x0 = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.complex128).reshape(b.shape)
        bb = C.DeviceBuffer(b.nbytes)
        xb = C.DeviceBuffer(b.nbytes)
        bb.upload(b)
        xb.upload(x0)
        d = C.SolveDesc(method=method, restart=restart, maxiter=maxiter, tol=tol, jacobi=int(jacobi),
                        precision=precision)
        reps = (C.SolveReport * B)()
        C.check(getattr(C.load(), _entry)(self.h, ctypes.byref(d), bb.ptr, xb.ptr, B, reps, None))
        x = np.zeros_like(b)
        xb.download(x)
        return x, [reps[i].as_dict() for i in range(B)]


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def hex_monolayer(cols, rows, spacing, jitter=0.0, seed=0):
    """Hexagonal monolayer: x = (i + (j mod 2)/2) s, y = j s sqrt(3)/2, z = 0, plus jitter."""
    j, i = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    x = (i + 0.5 * (j % 2)) * spacing
    y = j * spacing * np.sqrt(3) / 2
    pos = np.stack([x.ravel(), y.ravel(), np.zeros(x.size)], axis=1)
    if jitter:
        pos = pos + np.random.default_rng(seed).normal(0.0, jitter, pos.shape)
    box = np.array([cols * spacing, rows * spacing * np.sqrt(3) / 2, 0.0])
    return pos, box
