"""Independent closed-form checks used to pin the oracle (not a restatement of it).

* ``point_dipole_field`` — the classic Ewald lattice sum of point dipoles with
  tin-foil (conducting) boundary conditions, written from the textbook
  formulas (erfc real-space kernel, Gaussian-screened reciprocal sum, dipole
  self-correction 4 alpha^3 / (3 sqrt(pi))).  For non-overlapping unit spheres
  the reference operator's off-diagonal blocks are minus this field, and its
  diagonal is self_coef + 1/(4 pi) minus the field of a particle's own images.
* ``overlap_correction`` — the volume-averaged field of a uniformly polarised
  unit sphere on an overlapping unit sphere (centre distance 1 < r < 2), minus
  the point-dipole field, by 1-D quadrature of the analytic angular integrals.

Units: the field of a dipole p at distance r is (3 r^(r^.p) - p) / (4 pi r^3).
"""
import itertools

import numpy as np
from scipy.integrate import quad
from scipy.special import erfc


def point_dipole_field(pos, p, box, alpha=None, real_images=3, kmax_factor=6.5):
    """Field at every particle from all *other* dipoles and every periodic image
    (own images included, own dipole excluded).  pos (N,3), p (N,3) complex."""
    pos = np.asarray(pos, float)
    p = np.asarray(p, complex)
    L = np.asarray(box, float)
    N = len(pos)
    V = np.prod(L)
    if alpha is None:
        alpha = 5.5 / L.min()
    E = np.zeros((N, 3), complex)
    rng = range(-real_images, real_images + 1)
    shifts = np.array(list(itertools.product(rng, rng, rng)), float) * L  # (S, 3)
    for i in range(N):
        R = (pos[i] - pos)[:, None, :] + shifts[None, :, :]  # (N, S, 3)
        d = np.linalg.norm(R, axis=-1)
        keep = d > 0.0  # drops only the own dipole (i == j, zero shift)
        ds = np.where(keep, d, 1.0)
        ar = alpha * ds
        g = 2.0 * ar / np.sqrt(np.pi) * np.exp(-ar * ar)
        B = np.where(keep, (erfc(ar) + g) / ds ** 3, 0.0)
        C = np.where(keep, (3.0 * erfc(ar) + g * (3.0 + 2.0 * ar * ar)) / ds ** 5, 0.0)
        Rp = np.einsum("jsc,jc->js", R, p)
        E[i] += -(B.sum(axis=1)[:, None] * p).sum(axis=0) + np.einsum("js,jsc->c", C * Rp, R)
    # reciprocal space, k != 0
    kmax = kmax_factor * alpha * 2.0
    m = [int(np.ceil(kmax * L[d] / (2 * np.pi))) for d in range(3)]
    ks = []
    for a in range(-m[0], m[0] + 1):
        for b in range(-m[1], m[1] + 1):
            for c in range(-m[2], m[2] + 1):
                if a == 0 and b == 0 and c == 0:
                    continue
                k = 2 * np.pi * np.array([a, b, c]) / L
                k2 = k @ k
                if k2 > kmax * kmax:
                    continue
                ks.append(k)
    ks = np.array(ks)
    k2 = np.sum(ks * ks, axis=1)
    w = 4 * np.pi / V * np.exp(-k2 / (4 * alpha * alpha)) / k2
    for i in range(N):
        for j in range(N):
            ph = np.cos(ks @ (pos[i] - pos[j]))
            kp = ks @ p[j]
            E[i] += -np.sum((w * ph * kp)[:, None] * ks, axis=0)
        E[i] += 4.0 * alpha ** 3 / (3.0 * np.sqrt(np.pi)) * p[i]
    return E / (4 * np.pi)


def overlap_correction(r):
    """(E_avg - E_dip) for unit spheres at centre distance r in (1, 2).

    Returns (perp, para): the field component along p for p perpendicular /
    parallel to the line of centres, per unit dipole moment."""
    Vs = 4.0 * np.pi / 3.0

    def lens_u0(rho):
        return (rho * rho + r * r - 1.0) / (2.0 * rho * r)

    def integrand(rho, which):
        u0 = max(-1.0, min(1.0, lens_u0(rho)))
        du = 1.0 - u0
        # int_{u0}^{1} u^2 du
        u2 = (1.0 - u0 ** 3) / 3.0
        uniform = -du / (4 * np.pi)                     # interior field -p/(4 pi)
        if which == "para":                              # 3u^2 - 1
            ang = (3.0 * u2 - du)
        else:                                            # (3/2)(1-u^2) - 1
            ang = 1.5 * (du - u2) - du
        dip = -ang / (4 * np.pi * rho ** 3)
        return 2 * np.pi * rho * rho * (uniform + dip)

    lo = r - 1.0
    perp = quad(integrand, lo, 1.0, args=("perp",), epsabs=1e-14, epsrel=1e-13, limit=200)[0] / Vs
    para = quad(integrand, lo, 1.0, args=("para",), epsabs=1e-14, epsrel=1e-13, limit=200)[0] / Vs
    return perp, para


def overlap_correction_radii(r, a, b):
    """(E_avg - E_dip) on a sphere of radius b from a uniformly polarised sphere of
    radius a (unit dipole), centre distance r < a + b: the interior field -p/(4 pi a^3)
    minus the point-dipole field, integrated over the part of sphere a inside sphere b
    (shells of radius rho about a's centre, caps u >= u0(rho)), divided by b's volume.

    Returns (perp, para) as ``overlap_correction``."""
    Vb = 4.0 * np.pi * b ** 3 / 3.0

    def integrand(rho, which):
        u0 = (rho * rho + r * r - b * b) / (2.0 * rho * r)
        u0 = max(-1.0, min(1.0, u0))
        du = 1.0 - u0
        u2 = (1.0 - u0 ** 3) / 3.0
        uniform = -du / (4 * np.pi * a ** 3)
        ang = (3.0 * u2 - du) if which == "para" else (1.5 * (du - u2) - du)
        dip = -ang / (4 * np.pi * rho ** 3)
        return 2 * np.pi * rho * rho * (uniform + dip)

    lo, hi = max(r - b, 0.0), min(a, r + b)
    pts = [p for p in (b - r,) if lo < p < hi]
    perp = quad(integrand, lo, hi, args=("perp",), points=pts or None, epsabs=1e-14, epsrel=1e-13,
                limit=400)[0] / Vb
    para = quad(integrand, lo, hi, args=("para",), points=pts or None, epsabs=1e-14, epsrel=1e-13,
                limit=400)[0] / Vb
    return perp, para
