"""Pins for the oracle's polydisperse operator (polydisperse_ewald_electric_field.cu).

The pair functions, the finite-sphere spread weights, the scalar k-space
coefficient and the sizing are checked against what fixes them independently:
the monodisperse closed forms (radii 1), a textbook point-dipole Ewald lattice
sum (non-overlapping spheres of different radii: the total operator does not
depend on the radii there), quadrature of the field of overlapping spheres of
different radii (lens and one-inside-the-other), xi-independence of real +
reciprocal, and the reference's own mono-vs-poly equivalence tests.
"""
import numpy as np
import pytest

from oracle import oracle as O
import ewald_textbook as T


def test_pair_functions_reduce_to_monodisperse():
    r = np.linspace(1.0, 9.5, 60)
    for xi in (0.4, 0.5, 0.7):
        pe, pa = O.pair_functions(r, xi)
        qe, qa = O.poly_pair_functions(r, 1.0, 1.0, xi)
        np.testing.assert_allclose(qe, pe, rtol=1e-11, atol=1e-14)
        np.testing.assert_allclose(qa, pa, rtol=1e-11, atol=1e-14)
        assert O.poly_self(1.0, 1.0, xi) == pytest.approx(O.tables(xi)[3], rel=1e-13)


def test_pair_functions_symmetric_in_radii():
    r = np.linspace(0.3, 8.0, 40)
    pe, pa = O.poly_pair_functions(r, 1.0, 1.7, 0.5)
    qe, qa = O.poly_pair_functions(r, 1.7, 1.0, 0.5)
    np.testing.assert_allclose(pe, qe, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(pa, qa, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("xi,tol,bound", [(0.5, 1e-10, 2e-7), (0.6, 1e-12, 2e-7), (0.45, 1e-8, 2e-7)])
def test_textbook_lattice_sum_different_radii(xi, tol, bound):
    """Non-overlapping spheres: off-diagonal = -(periodic point-dipole field), diagonal =
    1/(4 pi a^3) minus the own-image field, whatever the radii."""
    L = np.array([25.0, 26.0, 27.0])
    pos = np.array([[1.0, 2.0, 3.0], [5.5, 2.5, 3.2], [2.0, 7.0, 9.0], [20.0, 20.0, 20.0]])
    rad = np.array([1.0, 1.5, 1.2, 2.0])
    rng = np.random.default_rng(1)
    p = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
    ref = -T.point_dipole_field(pos, p, L) + p / (4 * np.pi * rad[:, None] ** 3)
    op = O.PolyOperator(L, tol, xi, rad)
    op.set_positions(pos)
    A = op.apply(p, 0.0)
    assert np.abs(A - ref).max() / np.abs(ref).max() < bound


@pytest.mark.parametrize("a,b,r", [(1.0, 1.6, 2.2), (1.6, 1.0, 1.5), (1.0, 1.6, 1.5), (2.5, 1.0, 1.2),
                                   (1.0, 2.5, 1.2)])
def test_overlapping_spheres_of_different_radii_quadrature(a, b, r):
    """Field on sphere b (radius b) from sphere a at distance r < a + b: point-dipole
    lattice field plus the quadrature of the overlap; when b contains a's centre the
    direct dipole field drops out of b's volume average (its principal value is zero)."""
    L = np.array([20.0, 20.0, 20.0])
    pos = np.array([[5.0, 5.0, 5.0], [5.0, 5.0, 5.0 + r]])
    op = O.PolyOperator(L, 1e-10, 0.5, [a, b])
    op.set_positions(pos)
    corr = T.overlap_correction_radii(r, a, b)
    u = np.array([0.0, 0.0, 1.0])
    for c, k in ((0, 0), (2, 1)):
        p = np.zeros((2, 3), complex)
        p[0, c] = 1.0
        A = op.apply(p, 0.0)
        E = T.point_dipole_field(pos, p, L)[1, c]
        if r < b:
            E -= ((3 * u * (u @ p[0]) - p[0]) / (4 * np.pi * r ** 3))[c]
        assert abs(A[1, c] - (-E - corr[k])) < 5e-8


def test_total_operator_independent_of_xi():
    L = np.array([22.0, 23.0, 24.0])
    rng = np.random.default_rng(4)
    pos = rng.uniform(0, 20, (12, 3))
    rad = rng.choice([1.0, 1.3, 1.8], 12)
    p = rng.standard_normal((12, 3)) + 1j * rng.standard_normal((12, 3))
    sc = rng.standard_normal(12) + 1j * rng.standard_normal(12)
    out = []
    for xi in (0.45, 0.6):
        op = O.PolyOperator(L, 1e-10, xi, rad)
        op.set_positions(pos)
        out.append(op.apply(p, sc))
    assert np.abs(out[0] - out[1]).max() / np.abs(out[0]).max() < 1e-7


def test_equals_monodisperse_for_identical_radii():
    """Reference tests/test_electric_fields.py:30-82 (1e-5) and the split additivity (1e-12)."""
    pos = np.array([[0.0, 0.0, 0.0], [5.0, 0.0, 0.0], [10.0, 0.0, 0.0], [15.0, 0.0, 0.0]])
    p = np.zeros((4, 3), complex)
    p[:, 0] = 1.0
    mono = O.Operator([20.0] * 3, 1e-6, 0.5)
    mono.set_positions(pos)
    poly = O.PolyOperator([20.0] * 3, 1e-6, 0.5, [1.0] * 4)
    poly.set_positions(pos)
    Em, Ep = mono.apply(p, 1.0), poly.apply(p, 1.0)
    np.testing.assert_allclose(Ep, Em, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(Ep, poly.apply(p, 1.0, 1) + poly.apply(p, 1.0, 2), rtol=1e-12, atol=1e-12)


def test_bloch_equals_monodisperse_for_identical_radii():
    """Reference tests/test_electric_fields.py:226-285 (dipole part, 1e-3).  The polydisperse
    k-space coefficient takes the window correction at the Bloch-shifted k (the
    monodisperse one at the grid wave vector), so the two differ by O(eta k_b G) here."""
    pos = np.array([[0.0, 0.0, 0.0], [4.5, 0.0, 0.0], [9.0, 0.0, 0.0], [13.5, 0.0, 0.0]])
    rs = np.random.RandomState(123)  # the reference test's inputs
    P = rs.rand(4, 3) + 1j * rs.rand(4, 3)
    mono = O.Operator([18.0] * 3, 1e-6, 0.5)
    poly = O.PolyOperator([18.0] * 3, 1e-6, 0.5, [1.0] * 4)
    for op in (mono, poly):
        op.set_bloch(0.15, -0.22)
        op.set_positions(pos)
    np.testing.assert_allclose(poly.apply(P, 1.0), mono.apply(P, 1.0), rtol=1e-3, atol=1e-3)


def test_sizing_rule():
    """rc = max(sqrt(-ln tol)/xi, 2 a_max); P = ceil(2/h_mean - 2 ln tol/pi); polydisperse.cu:2039-2069."""
    op = O.PolyOperator([30.0, 31.0, 32.0], 1e-6, 0.5, [1.0, 4.0])
    info = op.info()
    assert info["rc"] == pytest.approx(8.0)
    kcut = 2 * 0.25 * 8.0
    n = [int(np.ceil(1 + Lx * kcut / np.pi)) for Lx in (30.0, 31.0, 32.0)]
    assert info["grid"] == tuple(n)
    h = np.array([30.0, 31.0, 32.0]) / n
    assert info["P"] == int(np.ceil(2 / h.mean() - 2 * np.log(1e-6) / np.pi))
    assert info["eta"][0] == pytest.approx(info["P"] * (h.mean() * 0.5) ** 2 / np.pi)


def _continuous_radii_lattice(n=70, seed=4):
    """n spheres of n distinct radii in [1, 2] on a jittered 5 x 5 x 4 lattice of spacing 5
    (no overlaps): more radius classes than pair tables are built for."""
    rng = np.random.default_rng(seed)
    sites = np.stack(np.meshgrid(np.arange(5), np.arange(5), np.arange(4), indexing="ij"), -1).reshape(-1, 3)
    sites = sites[rng.permutation(len(sites))[:n]] * 5.0 + 2.5
    pos = sites + rng.uniform(-0.4, 0.4, sites.shape)
    rad = 1.0 + rng.random(n)
    rad[0] = 1.0
    return np.array([25.0, 25.0, 20.0]), pos, rad


def test_textbook_lattice_sum_continuous_radii():
    """70 distinct radii (no tables; each pair's closed forms at its own radii): the
    non-overlapping operator equals the textbook lattice sum without the tables'
    interpolation error (the same bound as the tables' case: what remains is the Ewald
    truncation of both sums)."""
    L, pos, rad = _continuous_radii_lattice()
    assert len(np.unique(np.round(rad, 5))) > 64
    rng = np.random.default_rng(2)
    p = rng.standard_normal((len(rad), 3)) + 1j * rng.standard_normal((len(rad), 3))
    ref = -T.point_dipole_field(pos, p, L) + p / (4 * np.pi * rad[:, None] ** 3)
    op = O.PolyOperator(L, 1e-10, 0.5, rad)
    op.set_positions(pos)
    A = op.apply(p, 0.0)
    assert np.abs(A - ref).max() / np.abs(ref).max() < 2e-7


def test_total_operator_independent_of_xi_continuous_radii():
    """Overlapping pairs of 70 distinct radii (closed forms at r >= 1): real + reciprocal
    does not depend on xi."""
    L, pos, rad = _continuous_radii_lattice(seed=6)
    pos[1] = pos[0] + [1.3, 0.9, 0.4]     # overlapping (d = 1.63 < a0 + a1)
    pos[2] = pos[0] + [-2.1, 0.3, -0.2]   # close
    rng = np.random.default_rng(3)
    p = rng.standard_normal((len(rad), 3)) + 1j * rng.standard_normal((len(rad), 3))
    sc = rng.standard_normal(len(rad)) + 1j * rng.standard_normal(len(rad))
    out = []
    for xi in (0.5, 0.6):
        op = O.PolyOperator(L, 1e-10, xi, rad)
        op.set_positions(pos)
        out.append(op.apply(p, sc))
    assert np.abs(out[0] - out[1]).max() / np.abs(out[0]).max() < 1e-8


def test_polydisperse_near_field_reduces_to_monodisperse_and_is_xi_free():
    """The polydisperse near field (field points as sites of the reference's radii) with equal
    radii equals the monodisperse near field, and with unequal radii it does not depend on the
    Ewald splitting (both to the Ewald tolerance)."""
    rng = np.random.default_rng(31)
    box = np.array([22.0, 23.0, 24.0])
    N, M = 8, 5
    pos = rng.uniform(0, 20, (N, 3))
    dip = rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))
    fp = rng.uniform(0, 20, (M, 3))
    fp[0] = pos[2] + [1.3, 0.4, 0.0]  # overlapping the particle's sphere
    E0 = [0.0, 0.0, 0.0]
    mono = O.near_field_intensity(box, E0, 1.0, dip, pos, fp, errortol=1e-8)
    poly = O.near_field_intensity(box, E0, np.ones(N), dip, pos, fp, errortol=1e-8, field_type="polydisperse")
    np.testing.assert_allclose(poly, mono, rtol=2e-6)
    r = rng.uniform(0.9, 1.3, N)
    r[0] = r.min() - 0.05
    a = O.near_field_intensity(box, E0, r, dip, pos, fp, xi=0.45, errortol=1e-9)
    b = O.near_field_intensity(box, E0, r, dip, pos, fp, xi=0.6, errortol=1e-9)
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-7 * np.abs(b).max())
    with pytest.raises(NotImplementedError):
        r2 = r.copy()
        r2[0] = r.max() + 0.1
        O.near_field_intensity(box, E0, r2, dip, pos, np.concatenate([fp, fp + 0.5]), errortol=1e-6)
