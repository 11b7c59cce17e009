"""Independent pin of the oracle's quadrupole real-space pair functions, including the
overlap polynomials (1 < r < 2) of the seven tables and the gradient self term.

The reference's quadrupole model (read off its k-space couplings, which the golden
fixture pins): a unit sphere with dipole p carries the uniform polarisation
P = p / V (V = 4 pi / 3); a sphere with traceless quadrupole Q carries the linear
polarisation P(y) = c Q y, c = 15 / (8 pi), y measured from its centre.  A target sphere
reads the volume-averaged field E_t = (1/V) int E dV and the first moment
G_t = (15 / 4 pi) int y (x) E dV (which is grad E at the centre for a harmonic field),
its traceless part in the components (xx, xy, xz, yy, yz).  The operator's coupling is
minus these.  The field of each source is written from electrostatics: outside the
sphere the point dipole / point quadrupole field, inside -p / (4 pi) and -(2c/5) Q y
(a surface charge of pure l = 1 or l = 2 angular dependence).

The real-space part at splitting xi is the bare coupling minus the Gaussian-screened
reciprocal part.  Here
  * the bare coupling is a 3-D quadrature over the target sphere in shells about the
    source centre, split at the source surface (no closed-form overlap formula used);
  * the screened reciprocal part is a 3-D quadrature in k-space of the same model:
    form factors 3 j1(k)/k (uniform sphere) and 4 pi i j2(k)/k (first moment), the
    longitudinal Green's function -k^ k^., the window exp(-k^2 / 4 xi^2).
Both are compared with the oracle operator's real-space output for two spheres
(every coupling block: dipole/quadrupole -> field/gradient) at r on the table nodes,
overlapping and not, and with the single-sphere self terms (self_perp and the
gradient self term, the tables' entry 0).  A wrong sign, a dropped term or a wrong
power in any overlap polynomial changes these by far more than the bound.
Value and slope continuity of every table at r = 2 are checked as well.
"""
import numpy as np
import pytest
from scipy.special import spherical_jn

from oracle import oracle as O

C_Q = 15.0 / (8.0 * np.pi)
IDX5 = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2)]


def q_from5(q5):
    xx, xy, xz, yy, yz = q5
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, -(xx + yy)]])


def to5(M):
    S = 0.5 * (M + M.T)
    S = S - np.trace(S) / 3.0 * np.eye(3)
    return np.array([S[a, b] for a, b in IDX5])


def source_field(y, p, Q):
    """Field at points y (..., 3) (relative to the source centre) of a unit sphere carrying
    dipole p and quadrupole Q (3x3), from its uniform / linear polarisation."""
    rho = np.linalg.norm(y, axis=-1)[..., None]
    inside = rho < 1.0
    rs = np.where(inside, 1.0, rho)
    py = np.einsum("...c,c->...", y, p)[..., None]
    Qy = np.einsum("ab,...b->...a", Q, y)
    yQy = np.einsum("...a,...a->...", y, Qy)[..., None]
    E_out = (3.0 * py * y / rs ** 5 - p / rs ** 3) / (4 * np.pi) \
        - (C_Q / 5.0) * (2.0 * Qy / rs ** 5 - 5.0 * yQy * y / rs ** 7)
    E_in = -p / (4 * np.pi) * np.ones_like(y) - (2.0 * C_Q / 5.0) * Qy
    return np.where(inside, E_in, E_out)


def bare_coupling(r_vec, p, Q, n_rho=48, n_u=48, n_phi=24):
    """(E_t, G_t) on a unit sphere at r_vec from a source sphere at the origin, by quadrature
    over the target sphere in shells of radius rho about the source centre (the
    integrand is smooth on each side of the source surface rho = 1)."""
    r = np.linalg.norm(r_vec)
    ez = r_vec / r
    tmp = np.array([1.0, 0.0, 0.0]) if abs(ez[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    ex = tmp - ez * (tmp @ ez)
    ex /= np.linalg.norm(ex)
    ey = np.cross(ez, ex)
    lo, hi = max(r - 1.0, 0.0), r + 1.0
    pieces = [(lo, 1.0), (1.0, hi)] if lo < 1.0 < hi else [(lo, hi)]
    xr, wr = np.polynomial.legendre.leggauss(n_rho)
    xu, wu = np.polynomial.legendre.leggauss(n_u)
    phi = 2 * np.pi * np.arange(n_phi) / n_phi
    E = np.zeros(3)
    G = np.zeros((3, 3))
    for a, b in pieces:
        rho = 0.5 * (b - a) * xr + 0.5 * (b + a)
        wrho = 0.5 * (b - a) * wr
        u0 = np.clip((rho ** 2 + r * r - 1.0) / (2.0 * rho * r), -1.0, 1.0)
        u = 0.5 * (1.0 - u0)[:, None] * xu[None, :] + 0.5 * (1.0 + u0)[:, None]
        wu2 = 0.5 * (1.0 - u0)[:, None] * wu[None, :]
        s = np.sqrt(np.clip(1.0 - u * u, 0.0, None))
        # points (rho, u, phi) in the frame (ex, ey, ez)
        d = (s[..., None] * np.cos(phi))[..., None] * ex + (s[..., None] * np.sin(phi))[..., None] * ey \
            + u[..., None, None] * ez
        y = rho[:, None, None, None] * d
        w = (wrho[:, None] * rho[:, None] ** 2 * wu2)[..., None] * (2 * np.pi / n_phi)
        F = source_field(y, p, Q)
        E += np.einsum("ijk,ijkc->c", w, F)
        G += np.einsum("ijk,ijka,ijkb->ab", w, y - r_vec, F)
    V = 4 * np.pi / 3
    return E / V, (15.0 / (4 * np.pi)) * G.T  # G[j][k] = int y_k E_j


def screened_reciprocal(r_vec, p, Q, xi, n_k=96, n_u=64, n_phi=24):
    """(E_t, G_t) of the Gaussian-screened reciprocal part, open boundary:
    (2 pi)^-3 int d^3k exp(-k^2/4xi^2) e^{i k.r} [target form factor] (-k^ k^ . P~(k))."""
    kmax = 2.0 * xi * np.sqrt(-np.log(1e-18))
    xk, wk = np.polynomial.legendre.leggauss(n_k)
    k = 0.5 * kmax * (xk + 1.0)
    wkk = 0.5 * kmax * wk
    xu, wu = np.polynomial.legendre.leggauss(n_u)
    phi = 2 * np.pi * np.arange(n_phi) / n_phi
    st = np.sqrt(1.0 - xu * xu)
    kh = np.stack([st[:, None] * np.cos(phi)[None, :], st[:, None] * np.sin(phi)[None, :],
                   np.repeat(xu[:, None], n_phi, 1)], -1)  # (u, phi, 3)
    j1 = spherical_jn(1, k)
    j2 = spherical_jn(2, k)
    F1 = 3.0 * j1 / k                      # (1/V) int_ball e^{i k.y}
    F2 = 4.0 * np.pi * j2 / k              # int_ball y e^{i k.y} = i F2 k^
    win = np.exp(-k * k / (4 * xi * xi))
    # source: P~ = p F1 + c Q (-i F2 k^);  k^.P~
    kp = np.einsum("uvc,c->uv", kh, p)
    kQk = np.einsum("uva,ab,uvb->uv", kh, Q, kh)
    kP = kp[None] * F1[:, None, None] - 1j * C_Q * kQk[None] * F2[:, None, None]  # (k, u, v)
    phase = np.exp(1j * k[:, None, None] * np.einsum("uvc,c->uv", kh, r_vec)[None])
    base = (wkk * k * k * win)[:, None, None] * wu[None, :, None] * (2 * np.pi / n_phi) * phase * (-kP)
    Et = np.einsum("kuv,uvc->c", base * F1[:, None, None], kh) / (2 * np.pi) ** 3
    Gt = np.einsum("kuv,uva,uvb->ab", base * (1j * F2[:, None, None]) * (15.0 / (4 * np.pi)), kh, kh) \
        / (2 * np.pi) ** 3
    assert np.abs(Et.imag).max() < 1e-12 and np.abs(Gt.imag).max() < 1e-12
    return Et.real, Gt.real


def oracle_real_coupling(r_vec, p, Q5, xi):
    """Real-space part of the oracle operator for two spheres with quadrupoles (source 0,
    target 1): the target's (field, gradient) rows."""
    L = 40.0
    op = O.Operator([L, L, L], 1e-8, xi)
    x0 = np.array([L / 2, L / 2, L / 2])
    op.set_positions(np.stack([x0, x0 + r_vec]))
    op.set_quads(None)
    v = np.zeros(op.n, complex)
    v[0:3] = p
    v[6:11] = Q5
    y = op.apply(v, np.zeros(2, complex), 1).real
    return y[3:6], y[11:16]


@pytest.mark.parametrize("xi", [0.5, 0.7])
@pytest.mark.parametrize("r", [1.05, 1.3, 1.5, 1.75, 1.95, 2.4, 3.1])
def test_quadrupole_pair_couplings_match_model(r, xi):
    rng = np.random.default_rng(int(r * 1000) + int(xi * 10))
    d = rng.standard_normal(3)
    r_vec = r * d / np.linalg.norm(d)
    for src in ("dipole", "quadrupole"):
        p = rng.standard_normal(3) if src == "dipole" else np.zeros(3)
        Q5 = rng.standard_normal(5) if src == "quadrupole" else np.zeros(5)
        Q = q_from5(Q5)
        Eb, Gb = bare_coupling(r_vec, p, Q)
        Er, Gr = screened_reciprocal(r_vec, p, Q, xi)
        E_model = -(Eb - Er)
        G_model = -to5(Gb - Gr)
        E_orc, G_orc = oracle_real_coupling(r_vec, p, Q5, xi)
        scale = max(np.abs(E_model).max(), np.abs(G_model).max())
        assert np.abs(E_orc - E_model).max() < 2e-9 * max(scale, 1.0), (src, E_orc, E_model)
        assert np.abs(G_orc - G_model).max() < 2e-9 * max(scale, 1.0), (src, G_orc, G_model)


@pytest.mark.parametrize("xi", [0.4, 0.5, 0.7])
def test_self_terms_match_model(xi):
    """Entry 0 of the tables: the real-space self terms are the bare self coupling of one
    sphere minus the screened reciprocal part at r = 0 (dipole: 1/(4 pi); quadrupole
    gradient: (2c/5) Q)."""
    rng = np.random.default_rng(11)
    p = rng.standard_normal(3)
    Q5 = rng.standard_normal(5)
    Q = q_from5(Q5)
    Er, _ = screened_reciprocal(np.zeros(3), p, np.zeros((3, 3)), xi)
    _, Gr = screened_reciprocal(np.zeros(3), np.zeros(3), Q, xi)
    info = O.Operator([40.0] * 3, 1e-8, xi)
    info.set_positions(np.array([[20.0, 20.0, 20.0]]))
    sp = info.info()["self_perp"]
    np.testing.assert_allclose(sp * p, -(-p / (4 * np.pi) - Er), rtol=0, atol=1e-11)
    sg = O.quad_self_grad(xi)
    np.testing.assert_allclose(0.5 * sg * Q5, -to5(-(2 * C_Q / 5.0) * Q - Gr), rtol=0, atol=1e-11)


@pytest.mark.parametrize("xi", [0.3, 0.5, 0.8])
def test_tables_continuous_at_contact(xi):
    """The overlap polynomials switch off at r = 2, where the lens vanishes with its volume:
    value and slope of all nine pair functions are continuous there."""
    h = 1e-5
    rs = np.array([2 - 2 * h, 2 - h, 2 + h, 2 + 2 * h])
    q = O.quad_pair_functions(rs, xi)       # (4, 7)
    perp, para = O.pair_functions(rs, xi)
    f = np.concatenate([q, np.stack([perp, para], 1)], 1)
    scale = np.abs(f).max(0)
    jump = np.abs((2 * f[1] - f[0]) - (2 * f[2] - f[3]))   # linear extrapolations to r = 2
    slope_l = (f[1] - f[0]) / h
    slope_r = (f[3] - f[2]) / h
    assert np.all(jump < 1e-8 * np.maximum(scale, 1.0)), jump
    assert np.all(np.abs(slope_l - slope_r) < 1e-3 * np.maximum(scale, 1.0)), slope_l - slope_r
