"""Pins for the CPU oracle (oracle/): golden fixture, closed forms, invariants.

None of these re-types the oracle's formulas: each compares it with something
computed independently (numpy.fft, a textbook Ewald lattice sum, quadrature,
np.linalg.solve, the reference's own fixture, or integers fixed by the
reference's sizing rule worked out by hand).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
import ewald_textbook as T

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_sizing_matches_reference_rule():
    # Config 2 (hex monolayer, Lz = 250 nm, a = 5 nm): 210 x 114 x 61, P = 9 -> 9^3 stencil,
    # rc = sqrt(-ln 1e-6)/0.5.
    Ly = 50 * 11 * np.sqrt(3) / 2
    s = O.sizing(np.array([880.0, Ly, 250.0]) / 5.0, 1e-6, 0.5)
    assert s["grid"] == (210, 114, 61)
    assert s["P"] == 9 and s["width"] == 9
    assert s["rc"] == pytest.approx(np.sqrt(-np.log(1e-6)) / 0.5, rel=1e-15)
    # Config 1 (SC 8^3, L = 24, tol 1e-8): 34^3 and a 13^3 stencil.
    s = O.sizing([24.0] * 3, 1e-8, 0.5)
    assert s["grid"] == (34, 34, 34) and s["width"] == 13
    # Config 5 cutoff 4 radii at tol 1e-6 -> xi = 0.9292; reference grid 247^3.
    xi5 = np.sqrt(-np.log(1e-6)) / 4.0
    s = O.sizing([111.769] * 3, 1e-6, xi5)
    assert s["grid"] == (247, 247, 247) and s["width"] == 9
    # overrides
    s = O.sizing([24.0] * 3, 1e-8, 0.5, grid=(32, 32, 32), P=8)
    assert s["grid"] == (32, 32, 32) and s["width"] == 9
    assert s["eta"][0] == pytest.approx(8 * (24.0 / 32 * 0.5) ** 2 / np.pi, rel=1e-14)
    with pytest.raises(RuntimeError):
        O.sizing([10.0, 10.0, 10.0], 1e-6, 0.5)  # rc > L/2


@pytest.mark.parametrize("shape", [(4, 6, 5), (8, 7, 61), (3, 19, 114), (5, 1, 9), (2, 113, 3)])
def test_fft_matches_numpy(shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f = O.fft3d(a, -1)
    np.testing.assert_allclose(f, np.fft.fftn(a), rtol=0, atol=1e-12 * np.abs(f).max())
    g = O.fft3d(a, +1)
    np.testing.assert_allclose(g, np.fft.ifftn(a) * a.size, rtol=0, atol=1e-12 * np.abs(g).max())


def _drude(omega, gamma, omega_p, eps_inf):
    return eps_inf - omega_p ** 2 / (omega ** 2 + 1j * omega * gamma)


@pytest.mark.parametrize("method", ["bicgstab", "gmres"])
def test_golden_pympm_fixture(method):
    """tests/test_dipole_solvers.py:8-60 of the reference (pyMPM fixture, rtol=atol=1e-3)."""
    ref = np.load(os.path.join(GOLDEN, "ref_dipole_monodisperse.npz"))
    d, a, eps_m = 12.0, 4.0, 2.13
    eps = _drude(np.linspace(2000, 6000, 3), 700, 12000, 4.0)
    pos = np.array([[0, 0, 0], [d, 0, 0], [0, d, 0], [d, d, 0]], float)
    s = O.DipoleSolver([5 * d] * 3, np.repeat(eps[:, None], 4, 1), a, eps_m, tol=1e-4,
                       solver_type=method)
    s.compute(pos)
    np.testing.assert_allclose(s.get_eff_polarizability(False), ref["alpha_py"], rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(s.get_dipoles(False), ref["dips_py"], rtol=1e-3, atol=1e-3)


@pytest.mark.parametrize("xi,tol,bound", [(0.5, 1e-10, 2e-8), (0.6, 1e-12, 2e-8), (0.5, 1e-6, 2e-5)])
def test_textbook_point_dipole_lattice_sum(xi, tol, bound):
    """Non-overlapping spheres: off-diagonal = -(periodic dipole field), diagonal = 1/(4 pi)
    minus the own-image field.  Holds for any xi to the Ewald tolerance."""
    L = np.array([25.0, 26.0, 27.0])
    pos = np.array([[1.0, 2.0, 3.0], [4.5, 2.5, 3.2], [2.0, 6.0, 9.0], [20.0, 20.0, 20.0]])
    rng = np.random.default_rng(1)
    p = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
    E = T.point_dipole_field(pos, p, L)
    op = O.Operator(L, tol, xi)
    op.set_positions(pos)
    A = op.apply(p, 0.0)
    ref = -E + p / (4 * np.pi)
    assert np.abs(A - ref).max() / np.abs(ref).max() < bound


@pytest.mark.parametrize("r", [1.2, 1.5, 1.9])
def test_overlapping_spheres_quadrature(r):
    """1 < r < 2: the pair kernel differs from the point dipole by the volume-averaged field of
    the overlap lens, computed here by quadrature."""
    L = np.array([20.0, 20.0, 20.0])
    pos = np.array([[5.0, 5.0, 5.0], [5.0, 5.0, 5.0 + r]])
    op = O.Operator(L, 1e-10, 0.5)
    op.set_positions(pos)
    corr = T.overlap_correction(r)
    for c, k in ((0, 0), (2, 1)):  # perpendicular (x) and parallel (z) to the line of centres
        p = np.zeros((2, 3), complex)
        p[0, c] = 1.0
        A = op.apply(p, 0.0)
        E = T.point_dipole_field(pos, p, L)
        assert abs(A[1, c] - (-E[1, c] - corr[k])) < 1e-9


def test_real_plus_reciprocal_is_total():
    op = O.Operator([20.0] * 3, 1e-6, 0.5)
    pos = np.array([[0, 0, 0], [5.0, 0, 0], [10.0, 0, 0], [15.0, 0, 0]])
    op.set_positions(pos)
    rng = np.random.default_rng(3)
    p = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
    sc = rng.standard_normal(4) + 1j * rng.standard_normal(4)
    tot = op.apply(p, sc, 3)
    np.testing.assert_allclose(tot, op.apply(p, sc, 1) + op.apply(p, sc, 2), rtol=1e-12, atol=1e-12)


def test_table_clamps_beyond_last_sample():
    """Beyond r = 9.999 the reference interpolation returns the last table entry (quirk kept)."""
    r, pe, pa, sp = O.tables(0.4)
    assert r[0] == 0.0 and r[1] == 1.0 and r[-1] == pytest.approx(9.999)
    assert pe[0] == sp and pa[0] == sp
    f_perp, f_para = O.pair_functions(np.array([9.999]), 0.4)
    assert pe[-1] == pytest.approx(f_perp[0], rel=1e-14)


@pytest.mark.parametrize("method,tol,bound", [("gmres", 1e-12, 1e-9), ("bicgstab", 1e-8, 1e-6)])
def test_krylov_matches_dense_solve(method, tol, bound):
    """Both solvers reach np.linalg.solve.  BiCGSTAB keeps the reference's absolute breakdown
    thresholds (|rho| < 1e-16), so it is run at a tolerance it can reach."""
    L = np.array([16.0, 17.0, 18.0])
    rng = np.random.default_rng(7)
    pos = rng.uniform(0, 16, (6, 3))
    op = O.Operator(L, 1e-6, 0.5)
    op.set_positions(pos)
    N = 6
    eps = -2.5 + 0.6j
    sc = O.self_coefficient(np.full(N, eps))
    A = np.zeros((3 * N, 3 * N), complex)
    for c in range(3 * N):
        e = np.zeros(3 * N, complex)
        e[c] = 1.0
        A[:, c] = op.apply(e.reshape(N, 3), sc).ravel()
    b = np.zeros((N, 3), complex)
    b[:, 0] = 1.0
    x_ref = np.linalg.solve(A, b.ravel())
    x, rep = op.solve(b, sc, tol=tol, method=method, restart=8)
    assert rep["converged"]
    assert np.abs(x - x_ref).max() / np.abs(x_ref).max() < bound


def test_near_field_golden_fixture():
    """Reference tests/test_near_fields.py:31-80: the oracle's near field against the pyMPM
    fixture (its bar 1e-4; the restatement agrees to 1e-9)."""
    ref = np.load(os.path.join(GOLDEN, "ref_near_field.npz"))["intensity_py"]
    box = [50.0] * 3
    dip_pos = np.array([[0.0, 0, 0], [10.0, 0, 0], [0, 10.0, 0], [0, 0, 10.0]])
    dip = np.array([[0.1 + 0.05j, 0, 0], [-0.1, 0.02j, 0], [0, 0.05, -0.01j], [0.01, 0, 0.1j]])
    fp = np.array([[5.0, 0, 0], [0, 5.0, 0], [2.5, 2.5, 2.5]])
    I = O.near_field_intensity(box, [1.0, 0, 0], 1.0, dip, dip_pos, fp)
    np.testing.assert_allclose(I, ref, rtol=1e-9, atol=1e-12)


def test_near_field_is_the_periodic_point_dipole_field():
    """Field points at least 2 from every sphere centre: the volume average over a unit sphere
    of a uniformly polarised sphere's field is its point-dipole field at the centre, so the
    near field equals the textbook periodic point-dipole sum there (1e-7 at errortol 1e-10)."""
    L = np.array([24.0, 25.0, 26.0])
    rng = np.random.default_rng(3)
    dip_pos = np.array([[2.0, 3.0, 4.0], [9.0, 3.5, 4.2], [4.0, 11.0, 13.0], [18.0, 18.0, 19.0]])
    p = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
    fp = np.array([[5.5, 3.2, 4.1], [12.0, 12.0, 12.0], [20.0, 6.0, 2.0]])
    sites = np.concatenate([dip_pos, fp])
    ref = T.point_dipole_field(sites, np.concatenate([p, np.zeros((3, 3))]), L)[4:]
    op = O.Operator(L, 1e-10, 0.5)
    op.set_positions(sites)
    E_ind = -op.apply(np.concatenate([p, np.zeros((3, 3), complex)]), 0.0)[4:]
    assert np.abs(E_ind - ref).max() / np.abs(ref).max() < 1e-7
    E0 = np.array([0.3, 1.0j, 0.2])
    I = O.near_field_intensity(L, E0, 1.0, p, dip_pos, fp, errortol=1e-10)
    np.testing.assert_allclose(I, np.sum(np.abs(E0 + ref) ** 2, axis=1), rtol=1e-9)
