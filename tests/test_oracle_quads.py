"""Pins for the oracle's quadrupole terms (monodisperse operator).

* The reference's own golden fixture tests/golden/ref_quadrupoles_monodisperse.npz
  (a dense solve of its Python formulation, tests/quadrupole_field.py; two spheres
  farther apart than rc, so reciprocal space, the projectors and the self terms)
  through a dense solve of the oracle operator.
* xi-independence of real + reciprocal for every coupling (dipole -> field,
  quadrupole -> field, dipole -> gradient, quadrupole -> gradient) with pairs
  inside the cutoff, overlapping ones included: this pins the seven quadrupole
  real-space tables' screened parts against the reciprocal sum.  Their
  overlap polynomials (r < 2) do not depend on xi; tests/test_oracle_quad_overlap.py
  pins them against an electrostatic model by quadrature.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _dense(op, sc):
    n = op.n
    A = np.zeros((n, n), complex)
    for c in range(n):
        e = np.zeros(n, complex)
        e[c] = 1.0
        A[:, c] = op.apply(e, sc)
    return A


def test_golden_quadrupole_fixture_dense_solve():
    """Reference tests/test_dipole_solvers.py:63-148 (its bar 1e-5)."""
    ref = np.load(os.path.join(GOLDEN, "ref_quadrupoles_monodisperse.npz"))
    op = O.Operator([100.0] * 3, 1e-6, 0.4)
    op.set_positions(np.array([[10.0, 10.0, 10.0], [20.0, 20.0, 20.0]]))
    assert op.set_quads() == 2 and op.n == 16
    A = _dense(op, O.self_coefficient(np.full(2, 4.5 + 0.2j)))
    for dim in range(3):
        b = np.zeros(16, complex)
        b[dim:6:3] = 1.0
        x = np.linalg.solve(A, b)
        np.testing.assert_allclose(x[:6].reshape(2, 3), ref["dips_py_all"][dim], rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(x[6:].reshape(2, 5), ref["quads_py_all"][dim], rtol=1e-9, atol=1e-12)


def _xi_pair(mask_fn, quad_idxs=None):
    L = np.array([22.0, 23.0, 24.0])
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, 20, (10, 3))
    pos[1] = pos[0] + [1.5, 0.4, 0.3]   # overlapping pair
    pos[2] = pos[0] + [0.0, 2.6, 0.2]   # close, not overlapping
    out = []
    for xi in (0.5, 0.6):
        op = O.Operator(L, 1e-10, xi)
        op.set_positions(pos)
        op.set_quads(quad_idxs)
        r2 = np.random.default_rng(9)
        v = (r2.standard_normal(op.n) + 1j * r2.standard_normal(op.n)) * mask_fn(op.n)
        sc = r2.standard_normal(10) + 1j * r2.standard_normal(10)
        out.append(op.apply(v, sc))
    return out


@pytest.mark.parametrize("src", ["all", "dipoles", "quadrupoles"])
def test_quadrupole_operator_independent_of_xi(src):
    mask = {"all": lambda n: np.ones(n), "dipoles": lambda n: np.r_[np.ones(30), np.zeros(n - 30)],
            "quadrupoles": lambda n: np.r_[np.zeros(30), np.ones(n - 30)]}[src]
    a, b = _xi_pair(mask)
    assert np.abs(a[:30] - b[:30]).max() / np.abs(a[:30]).max() < 2e-8
    assert np.abs(a[30:] - b[30:]).max() / np.abs(a[30:]).max() < 2e-8


def test_quadrupole_subset_and_split():
    """Quadrupoles on a subset of particles; real + reciprocal = total (1e-12)."""
    a, b = _xi_pair(lambda n: np.ones(n), quad_idxs=[0, 3, 7])
    assert a.shape == (30 + 15,)
    assert np.abs(a - b).max() / np.abs(a).max() < 2e-8
    op = O.Operator([22.0, 23.0, 24.0], 1e-8, 0.5)
    op.set_positions(np.random.default_rng(5).uniform(0, 20, (10, 3)))
    op.set_quads([0, 3, 7])
    v = np.random.default_rng(1).standard_normal(op.n) + 0j
    np.testing.assert_allclose(op.apply(v, 1.0), op.apply(v, 1.0, 1) + op.apply(v, 1.0, 2), rtol=1e-12, atol=1e-12)


def test_oracle_driver_quadrupole_fixture():
    """The oracle's driver (quadrupoles=True, xi 0.4, tol 1e-8) against the same fixture,
    as reference tests/test_dipole_solvers.py:63-148 drives cuMPM.dipole_solver (its bar 1e-5).
    The Ewald accuracy is the fixture's 1e-6 (a coarser grid than tol would give, for time)."""
    ref = np.load(os.path.join(GOLDEN, "ref_quadrupoles_monodisperse.npz"))
    s = O.DipoleSolver([100.0] * 3, 4.5 + 0.2j, 1.0, 1.0, xi=0.4, tol=1e-8, field_type="monodisperse",
                       quadrupoles=True, ewald_tol=1e-6)
    s.compute(np.array([[10.0, 10.0, 10.0], [20.0, 20.0, 20.0]]))
    dips, quads = s.get_dipoles(), s.get_quadrupoles()
    assert dips.shape == (2, 3, 3) and quads.shape == (2, 3, 5)
    for dim in range(3):
        np.testing.assert_allclose(dips[:, dim], ref["dips_py_all"][dim], rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(quads[:, dim], ref["quads_py_all"][dim], rtol=1e-5, atol=1e-5)


def test_oracle_driver_quadrupole_rhs_layouts():
    """E0 of length 3, 3N and 3N + 5Q give the same solution when they describe the same field
    (dipole_solver.cu:368-390: quadrupole rows are zero unless given); a full-length E0 with
    quadrupole rows solves A x = b for that b."""
    L = [22.0, 23.0, 24.0]
    pos = np.random.default_rng(5).uniform(0, 20, (6, 3))
    e3 = np.array([[1.0, 0.5j, 0.2]])
    outs = []
    for E0 in (e3, np.tile(e3, (1, 6)), np.concatenate([np.tile(e3, (1, 6)), np.zeros((1, 10))], axis=1)):
        s = O.DipoleSolver(L, [-2.0 + 0.5j], 1.0, tol=1e-10, quadrupoles=[1, 4], E0=E0)
        s.compute(pos)
        outs.append((s.get_dipoles(), s.get_quadrupoles()))
    for d, q in outs[1:]:
        np.testing.assert_allclose(d, outs[0][0], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(q, outs[0][1], rtol=1e-12, atol=1e-14)
    b = np.random.default_rng(3).standard_normal(28) + 0j
    s = O.DipoleSolver(L, [-2.0 + 0.5j], 1.0, tol=1e-12, quadrupoles=[1, 4], E0=b[None, :])
    s.compute(pos)
    x = np.concatenate([s.get_dipoles(physical=False).reshape(-1), s.get_quadrupoles(physical=False).reshape(-1)])
    op = O.Operator(L, 1e-12, 0.5)
    op.set_positions(pos)
    op.set_quads([1, 4])
    assert np.abs(op.apply(x, O.self_coefficient(np.full(6, -2.0 + 0.5j))) - b).max() < 1e-9 * np.abs(b).max()
