"""CPU pins of the oracle's direct field and EELS driver (no GPU).

Direct field (oracle.DirectOperator): the textbook point-dipole field, field gradients that are
the numerical derivatives of the fields, the isolated-sphere polarisabilities (Clausius-Mossotti
for l = 1 and l = 2), and the large-box limit of the Ewald operator (pinned elsewhere) for all
four couplings.  EELS (oracle.EELSSolver, beam_field, beam_gradient): the reference's fixture
(an external Python EELS module), the reciprocity identity between the loss sum and the line
integral of the induced field along the beam, and the 24-point gradient rule against the
derivative of the beam field.
"""
import os

import numpy as np
import pytest
from scipy.integrate import simpson

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _num_grad(f, x, h=1e-5):
    """(d_x f_x, d_x f_y, d_x f_z, d_y f_y, d_y f_z) of a 3-vector field at x."""
    J = np.zeros((3, 3), complex)
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        J[a] = (f(x + e) - f(x - e)) / (2 * h)
    return np.array([J[0, 0], J[0, 1], J[0, 2], J[1, 1], J[1, 2]])


def _two_site(src_pos, tgt_pos, x_src, quads):
    """Operator output at a target site from one source site with the given unknowns."""
    D = O.DirectOperator(1.0, True if quads else None)
    D.set_positions(np.array([src_pos, tgt_pos]))
    x = np.zeros(D.n, complex)
    if quads:
        x[0:3] = x_src[:3]
        x[6:11] = x_src[3:]
    else:
        x[0:3] = x_src
    y = D.apply(x, np.zeros(2))
    return (y[3:6], y[11:16]) if quads else (y[3:6], None)


def test_direct_dipole_field_is_textbook_and_gradients_are_derivatives():
    rng = np.random.default_rng(0)
    src = np.array([0.3, -0.2, 0.1])
    p = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    q = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    for tgt in (np.array([4.0, 1.0, -2.0]), np.array([-1.5, 2.5, 3.0])):
        # dipole -> field: minus the textbook field (3 (p.n) n - p) / (4 pi r^3), n = (t - s)/r
        r = tgt - src
        d = np.linalg.norm(r)
        n = r / d
        E_phys = (3 * (p @ n) * n - p) / (4 * np.pi * d ** 3)
        f, _ = _two_site(src, tgt, np.concatenate([p, np.zeros(5)]), True)
        np.testing.assert_allclose(f, -E_phys, rtol=1e-12, atol=1e-15)
        # dipole -> gradient and quadrupole -> gradient: derivatives of the fields at the target
        for xs in (np.concatenate([p, np.zeros(5)]), np.concatenate([np.zeros(3), q])):
            fld = lambda t: _two_site(src, t, xs, True)[0]
            g_num = _num_grad(fld, tgt)
            _, g = _two_site(src, tgt, xs, True)
            np.testing.assert_allclose(g, g_num, rtol=1e-7, atol=1e-12 * np.abs(g_num).max())


def test_direct_self_terms_are_isolated_sphere_polarisabilities():
    # an isolated sphere of radius a: p = 4 pi a^3 (eps - 1)/(eps + 2) E (Clausius-Mossotti) and
    # Q = G / self_Q with self_Q = (3 / (4 pi a^5)) (eps + 3/2)/(eps - 1) (the l = 2 mode)
    for a, eps in ((1.0, 4.5 + 0.2j), (1.7, -2.0 + 0.3j)):
        D = O.DirectOperator(a, True)
        D.set_positions(np.zeros((1, 3)))
        sc = O.self_coefficient(eps, a)
        A = D.matrix(np.array([sc]))
        np.testing.assert_allclose(np.diag(A)[:3], (eps + 2) / (4 * np.pi * a ** 3 * (eps - 1)), rtol=1e-13)
        np.testing.assert_allclose(np.diag(A)[3:], 3 * (eps + 1.5) / (4 * np.pi * a ** 5 * (eps - 1)), rtol=1e-13)
        assert np.allclose(A - np.diag(np.diag(A)), 0.0)


def test_direct_field_is_large_box_limit_of_ewald():
    # every coupling (dipole/quadrupole -> field/gradient) of the direct field against the Ewald
    # operator in a box of 80 radii; the periodic images contribute O((d/L)^3) and shrink with L
    L = 80.0
    pos = np.array([[0, 0, 0], [12.0, 0.0, 0.0], [0.0, 11.0, 3.0], [6, 6, -3]]) + L / 2
    N = 4
    sc = O.self_coefficient(np.full(N, 4.5 + 0.2j))
    op = O.Operator([L, L, L], 1e-8, 0.6)
    op.set_positions(pos)
    op.set_quads(None)
    D = O.DirectOperator(1.0, True)
    D.set_positions(pos)
    rng = np.random.default_rng(1)
    bars = {("dip", "field"): 1e-2, ("dip", "grad"): 1e-3, ("quad", "field"): 1e-3, ("quad", "grad"): 2e-4}
    for src in ("dip", "quad"):
        x = np.zeros(op.n, complex)
        if src == "dip":
            x[0:3] = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        else:
            x[3 * N:3 * N + 5] = rng.standard_normal(5) + 1j * rng.standard_normal(5)
        ye, yd = op.apply(x, sc), D.apply(x, sc)
        for tgt, sl in (("field", slice(3, 3 * N)), ("grad", slice(3 * N + 5, None))):
            err = np.abs(ye[sl] - yd[sl]).max() / np.abs(ye[sl]).max()
            assert err < bars[(src, tgt)], (src, tgt, err)


def _fixture_case():
    d = 25.0 + 2.6
    R = 12.5
    box = np.array([10.0, 10.0, 10.0]) * d
    points = np.array([[-0.5, 0.0, 0.0], [0.5, 0.0, 0.0]]) * d
    e_pos = np.array([[0.0, 0.0], [1.0, 0.0]]) * d
    omega = np.linspace(2000.0, 9000.0, 5)
    eps_p = 2.25 - 11886.0 ** 2 / (omega ** 2 + 1j * omega * 845.0)
    return box, points, e_pos, eps_p, omega / 1e7, 0.45, R


def test_oracle_eels_matches_reference_fixture():
    # reference tests/test_eels_solver.py:114-173 (rtol 2e-2); the oracle solves exactly
    box, points, e_pos, eps_p, omega_nm, v, R = _fixture_case()
    ref = np.load(os.path.join(GOLDEN, "ref_eels_vs_python_module.npz"))
    s = O.EELSSolver(box, eps_p, omega_nm, v, radius=R, field="direct")
    for ep in e_pos:
        s.compute(ep, points)
    eels = s.get_eels()
    np.testing.assert_allclose(eels / (omega_nm * R) * (4 * np.pi) ** 2, ref["ref_eels"], rtol=5e-3)
    np.testing.assert_allclose(s.get_dipoles(physical=False), ref["ref_dips_stacked"], rtol=2e-2, atol=1e-5)
    np.testing.assert_allclose(s.get_dipoles(physical=False), ref["ref_dips_stacked"], rtol=0, atol=2e-3 * 15)


@pytest.mark.parametrize("v,om", [(0.3, 0.3), (0.45, 0.12), (0.7, 0.05)])
def test_loss_sum_equals_line_integral_of_induced_field(v, om):
    # For the beam field with gamma = 1 (a line charge -e^{2 pi i w z / v}/v of the static
    # Coulomb kernel) reciprocity makes the loss sum the work of the induced field on the
    # beam: sum_p Im(p . conj E(x_p)) = Re int e^{-2 pi i w z / v} E_ind,z(beam, z) dz / (2 pi w).
    pos = np.array([[0.0, 0.0, -1.5], [0.0, 0.0, 1.5], [0.5, 1.0, 0.25]])
    epos = np.array([2.5, 0.0])
    N = 3
    D = O.DirectOperator(1.0)
    D.set_positions(pos)
    E = O.beam_field(pos, epos, om, v, 0.0, gamma=1.0)
    p = D.solve(E.reshape(-1), O.self_coefficient(np.full(N, 4.0 + 0.1j))).reshape(N, 3)
    rec = np.sum((p * np.conj(E)).imag)
    z = np.arange(-3000.0, 3000.0, 0.01)
    tp = np.stack([np.full_like(z, epos[0]), np.full_like(z, epos[1]), z], axis=1)
    Ez = np.zeros(z.size, complex)
    for i in range(N):
        r = tp - pos[i]
        d = np.linalg.norm(r, axis=1)
        u = r / d[:, None]
        Ez += (3 * (u @ p[i]) * u[:, 2] - p[i][2]) / (4 * np.pi * d ** 3)
    lin = simpson(np.real(Ez * np.exp(-1j * 2 * np.pi * om * z / v)), x=z) / (2 * np.pi * om)
    assert abs(rec - lin) < 2e-6 * abs(lin)


def test_beam_gradient_rule_approximates_the_field_derivative():
    # the 24-point surface rule averages the gradient over a unit sphere; far from the beam the
    # field varies slowly and the average is the derivative at the centre to O(a^2 / r^2)
    n = O.surface_normals24()
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, rtol=1e-15)
    assert np.abs(n.sum(axis=0)).max() < 0.2
    np.testing.assert_allclose(n[:, 2], 1 - (2 * np.arange(24) + 1) / 24, rtol=1e-15)
    epos = np.array([0.0, 0.0])
    for centre in (np.array([9.0, 4.0, 0.3]), np.array([-6.0, 10.0, -2.0])):
        G = O.beam_gradient(centre[None, :], epos, 0.02, 0.6, 0.1)[0]
        f = lambda x: O.beam_field(x[None, :], epos, 0.02, 0.6, 0.1)[0]
        g = _num_grad(f, centre)
        # rows are (Gxx - Gzz, Gxy, Gxz, Gyy - Gzz, Gyz) with G_ij = d_i E_j
        J = np.zeros((3, 3), complex)
        for a in range(3):
            e = np.zeros(3)
            e[a] = 1e-5
            J[a] = (f(centre + e) - f(centre - e)) / 2e-5
        want = np.array([J[0, 0] - J[2, 2], 0.5 * (J[0, 1] + J[1, 0]), 0.5 * (J[0, 2] + J[2, 0]),
                         J[1, 1] - J[2, 2], 0.5 * (J[1, 2] + J[2, 1])])
        assert np.abs(G - want).max() < 2e-2 * np.abs(want).max(), (G, want, g)
