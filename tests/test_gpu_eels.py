"""GPU parity of the direct (open-boundary) field and the EELS driver, through the C ABI.

Oracles: oracle.DirectOperator (dense, pinned in test_oracle_eels.py) and oracle.EELSSolver
(beam field with scipy's Bessel functions, exact or GMRES solves, reciprocity sum); the
reference's EELS fixture and its direct-vs-Ewald equivalence tests with their own inputs.
"""
import ctypes
import os

import numpy as np
import pytest

import cumpm_amd
from cumpm_amd import _capi as C
from gpu_helpers import rel
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


class DirectOp:
    """mpm_op_create_direct through the C ABI."""

    def __init__(self, radii, quads=None):
        r = np.ascontiguousarray(np.atleast_1d(radii), np.float64)
        self.h = ctypes.c_void_p()
        C.check(C.load().mpm_op_create_direct(C.dptr(r), r.size, ctypes.byref(self.h)))
        if quads is True:
            C.check(C.load().mpm_op_set_quads(self.h, None, -1))
        elif quads is not None:
            idx = (ctypes.c_int * len(quads))(*quads)
            C.check(C.load().mpm_op_set_quads(self.h, idx, len(quads)))

    def __del__(self):
        if getattr(self, "h", None) and C._lib is not None:
            C._lib.mpm_op_destroy(self.h)

    def set_positions(self, pos):
        pos = np.ascontiguousarray(pos, np.float64)
        C.check(C.load().mpm_op_set_positions(self.h, C.dptr(pos), pos.shape[0]))
        info = C.OpInfo()
        C.check(C.load().mpm_op_get_info(self.h, ctypes.byref(info)))
        self.n = 3 * pos.shape[0] + 5 * int(info.num_quads)

    def set_self_coef(self, rows):
        rows = np.ascontiguousarray(np.atleast_2d(rows), np.complex128)
        C.check(C.load().mpm_op_set_self_coef(self.h, C.dptr(rows.view(np.float64)), rows.shape[0]))

    def apply(self, x, precision=C.PREC_DOUBLE):
        x = np.ascontiguousarray(x, np.complex128).reshape(-1, self.n)
        xb, yb = C.DeviceBuffer(x.nbytes), C.DeviceBuffer(x.nbytes)
        xb.upload(x)
        C.check(C.load().mpm_op_apply(self.h, xb.ptr, yb.ptr, x.shape[0], int(precision), C.PART_ALL, None))
        y = np.zeros_like(x)
        yb.download(y)
        return y


def _cluster(seed, N=37):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 12, (N, 3))
    pos[1] = pos[0] + [0.5, 0.1, -0.2]   # overlapping pair: evaluated at the source radius
    radii = rng.uniform(0.8, 1.3, N)
    return pos, radii


@pytest.mark.parametrize("quads", [None, [0, 1, 5, 17, 30]])
def test_direct_apply_matches_oracle(quads):
    pos, radii = _cluster(3)
    N = pos.shape[0]
    op = DirectOp(radii, quads)
    op.set_positions(pos)
    D = O.DirectOperator(radii, quads)
    D.set_positions(pos)
    rng = np.random.default_rng(4)
    B = 5
    eps = -2.0 + 0.4j + 0.3 * rng.standard_normal(B)
    rows = np.array([O.self_coefficient(np.full(N, e), radii) for e in eps])
    op.set_self_coef(rows)
    x = rng.standard_normal((B, op.n)) + 1j * rng.standard_normal((B, op.n))
    print(f"direct apply N={N} quads={quads} B={B}", flush=True)
    y = op.apply(x)
    ref = np.array([D.apply(x[b], rows[b]) for b in range(B)])
    assert rel(y, ref) < 1e-12
    y32 = op.apply(x, C.PREC_MIXED)
    assert rel(y32, ref) < 1e-5
    y1 = op.apply(x[:1])
    assert rel(y1, ref[:1]) < 1e-12


def test_dipole_solver_direct_matches_oracle_and_ewald_large_box():
    # reference tests/test_dipole_solvers.py:151-187 with its inputs (rtol 1e-3), and the direct
    # solve against the oracle's dense solution
    box = np.array([120.0, 120.0, 120.0])
    eps_p = 4.5 + 0.2j
    pos = np.array([[0.0, 0.0, 0.0], [12.0, 0.0, 0.0], [-12.0, 0.0, 0.0], [0.0, 12.0, 0.0], [0.0, -12.0, 0.0]])
    sd = cumpm_amd.dipole_solver(box, eps_p, radius=1.0, tol=1e-5, field_type="direct", quiet=True,
                                 quadrupoles=True, precision="double")
    sd.compute(pos)
    se = cumpm_amd.dipole_solver(box, eps_p, radius=1.0, xi=0.4, tol=1e-5, field_type="monodisperse",
                                 quiet=True, quadrupoles=True, precision="double")
    se.compute(pos)
    np.testing.assert_allclose(sd.get_eff_polarizability(), se.get_eff_polarizability(), rtol=1e-3, atol=1e-5)
    np.testing.assert_allclose(sd.get_dipoles(), se.get_dipoles(), rtol=1e-3, atol=1e-5)
    np.testing.assert_allclose(sd.get_quadrupoles(), se.get_quadrupoles(), rtol=1e-3, atol=5e-5)
    # tight: the direct solve at tol 1e-10 against the oracle's exact solution
    s2 = cumpm_amd.dipole_solver(box, eps_p, radius=1.0, tol=1e-10, field_type="direct", quiet=True,
                                 quadrupoles=True, precision="double")
    s2.compute(pos)
    D = O.DirectOperator(1.0, True)
    D.set_positions(pos)
    sc = O.self_coefficient(np.full(5, eps_p))
    dips, quads = s2.get_dipoles(), s2.get_quadrupoles()
    for k in range(3):
        b = np.zeros(D.n, complex)
        b[k:15:3] = 1.0
        x = D.solve(b, sc)
        assert rel(dips[:, k, :], x[:15].reshape(5, 3)) < 1e-8
        assert rel(quads[:, k, :], x[15:].reshape(5, 5)) < 1e-7


def _fixture_case():
    d = 25.0 + 2.6
    R = 12.5
    box = np.array([10.0, 10.0, 10.0]) * d
    points = np.array([[-0.5, 0.0, 0.0], [0.5, 0.0, 0.0]]) * d
    e_pos = np.array([[0.0, 0.0], [1.0, 0.0]]) * d
    omega = np.linspace(2000.0, 9000.0, 5)
    eps_p = cumpm_amd.drude_dielectric(omega, 845.0, 11886.0, 2.25)
    return box, points, e_pos, eps_p, omega / 1e7, 0.45, R


@pytest.mark.parametrize("batched", [False, True])
def test_eels_reference_fixture_and_oracle(batched):
    # reference tests/test_eels_solver.py:114-173: BiCGSTAB at tol 1e-3, direct field, rtol 2e-2
    box, points, e_pos, eps_p, omega_nm, v, R = _fixture_case()
    ref = np.load(os.path.join(GOLDEN, "ref_eels_vs_python_module.npz"))
    s = cumpm_amd.EELS(box=box, eps_p=eps_p, omega=omega_nm, v=v, eps_m=1.0, radius=R, xi=0.5, tol=1e-3,
                       field_type="direct", solver_type="bicgstab", integration_step=0.01 * R, precision="double",
                       quiet=True, batched=batched)
    s.compute(e_pos, points)
    eels, dips = s.get_eels(), s.get_dipoles(physical=False)
    assert eels.shape == (2, 5) and dips.shape == (2, 5, 2, 3)
    np.testing.assert_allclose(eels / (omega_nm * R) * (4.0 * np.pi) ** 2, ref["ref_eels"], rtol=2e-2, atol=1e-5)
    np.testing.assert_allclose(dips, ref["ref_dips_stacked"], rtol=2e-2, atol=1e-5)
    # tight: GMRES at tol 1e-11 against the oracle (exact solve, scipy Bessel functions)
    s2 = cumpm_amd.EELS(box, eps_p, omega_nm, v, radius=R, tol=1e-11, field_type="direct", precision="double",
                        quiet=True, batched=batched)
    s2.compute(e_pos, points)
    o = O.EELSSolver(box, eps_p, omega_nm, v, radius=R, field="direct")
    for ep in e_pos:
        o.compute(ep, points)
    assert rel(s2.get_eels(), o.get_eels()) < 1e-9
    assert rel(s2.get_dipoles(), o.get_dipoles()) < 1e-9


def test_eels_incident_field_matches_oracle():
    # the device Bessel functions and 24-point gradient rule against scipy through the oracle
    rng = np.random.default_rng(7)
    pos = rng.uniform(-6, 6, (12, 3))
    om = np.array([0.002, 0.01, 0.05, 0.2])  # 1/nm
    s = cumpm_amd.EELS([200.0] * 3, np.full(4, 3.0 + 0.2j), om, 0.6, radius=1.5, tol=1e-8, field_type="direct",
                       solve_quadrupoles=True, quad_idxs=[0, 3, 11], precision="double", quiet=True)
    ep = np.array([0.7, -1.1])
    s.compute(ep, pos)
    Einc = s.get_incident_field()  # nondimensional (lengths / 1.5)
    a = 1.5
    for w in range(4):
        want = O.beam_field(pos / a, ep / a, om[w] * a, 0.6, 0.5 / a).reshape(-1)
        assert rel(Einc[w, :36], want) < 1e-12
        wq = O.beam_gradient(pos[[0, 3, 11]] / a, ep / a, om[w] * a, 0.6, 0.5 / a).reshape(-1)
        assert rel(Einc[w, 36:], wq) < 1e-12


def test_eels_quadrupoles_direct_vs_ewald_and_oracle():
    # reference tests/test_eels_quadrupoles.py:5-49 with its inputs (rtol 1e-3)
    box = np.array([150.0, 150.0, 150.0])
    pos = np.array([[0.0, 0.0, -10.0], [0.0, 0.0, 0.0], [0.0, 0.0, 10.0]])
    epos = np.array([12.0, 0.0])
    kw = dict(eps_m=1.0, radius=1.2, tol=1e-6, quiet=True, solve_quadrupoles=True, precision="double")
    sd = cumpm_amd.EELS(box, [4.5 + 0.2j], [0.15], 0.35, field_type="direct", **kw)
    sd.compute(epos, pos)
    se = cumpm_amd.EELS(box, [4.5 + 0.2j], [0.15], 0.35, field_type="monodisperse", xi=0.35, **kw)
    se.compute(epos, pos)
    ed, ee = sd.get_eels(), se.get_eels()
    assert np.ndim(ed) == 0 or np.shape(ed) == (1,)
    assert sd.get_dipoles().shape == (3, 3) and sd.get_quadrupoles().shape == (3, 5)
    np.testing.assert_allclose(ed, ee, rtol=1e-3, atol=1e-5)
    np.testing.assert_allclose(sd.get_dipoles(), se.get_dipoles(), rtol=1e-3, atol=1e-5)
    np.testing.assert_allclose(sd.get_quadrupoles(), se.get_quadrupoles(), rtol=1e-3, atol=1e-5)
    o = O.EELSSolver(box, [4.5 + 0.2j], [0.15], 0.35, radius=1.2, field="direct", quad_idxs=True)
    o.compute(epos, pos)
    s3 = cumpm_amd.EELS(box, [4.5 + 0.2j], [0.15], 0.35, field_type="direct",
                        **dict(kw, tol=1e-11))
    s3.compute(epos, pos)
    assert abs(s3.get_eels() - o.get_eels()) < 1e-9 * abs(o.get_eels())
    assert rel(s3.get_quadrupoles(), o.get_quadrupoles()) < 1e-8
    # the quadrupole subset (reference :51-79): shapes
    sub = cumpm_amd.EELS([100.0] * 3, [3.0 + 0.1j], [0.2], 0.4, radius=1.0, tol=1e-5, field_type="direct",
                         quiet=True, solve_quadrupoles=True, quad_idxs=[0, 2], precision="double")
    sub.compute(np.array([6.0, 0.0]), pos / 2)
    assert sub.get_dipoles().shape == (3, 3) and sub.get_quadrupoles().shape == (2, 5)


def test_eels_splitting_like_reference():
    # reference tests/test_eels_solver.py:176-217
    s = cumpm_amd.EELS(box=[100, 100, 100], eps_p=[4.0], omega=[0.15], v=0.3, radius=2.0, split_dist=4.0,
                       N_split=27, field_type="direct", precision="double", quiet=True)
    pos = np.array([[0.0, 0.0, 0.0], [10.0, 0.0, 0.0]])
    s.compute(np.array([1.0, 0.0]), pos)
    sp = s.get_positions()
    assert len(sp) > 2
    assert np.isfinite(s.get_eels()).all()
    f = cumpm_amd.EELS(box=[100, 100, 100], eps_p=[4.0], omega=[0.15], v=0.3, radius=2.0, split_dist=4.0,
                       N_split=50, split_method="fcc", field_type="direct", precision="double", quiet=True)
    f.compute(np.array([1.0, 0.0]), pos)
    sp = f.get_positions()
    assert len(sp) > 2
    assert np.linalg.norm(np.mean(sp[1:], axis=0)) < 1e-12
    # a lossy material at a tight tolerance: the loss of the split system equals the oracle's on
    # the split particles (the unsplit particle first, then the sub-dipoles of equal volume)
    g = cumpm_amd.EELS(box=[100, 100, 100], eps_p=[4.0 + 0.5j], omega=[0.15], v=0.3, radius=2.0, split_dist=4.0,
                       N_split=50, split_method="fcc", field_type="direct", precision="double", quiet=True, tol=1e-11)
    g.compute(np.array([1.0, 0.0]), pos)
    sp = g.get_positions()
    M = len(sp) - 1
    o = O.EELSSolver([100.0] * 3, [4.0 + 0.5j], [0.15], 0.3, radius=np.r_[2.0, np.full(M, 2.0 / M ** (1 / 3))],
                     field="direct")
    o.compute([1.0, 0.0], sp)
    assert abs(g.get_eels() - o.get_eels()) < 1e-8 * abs(o.get_eels())


def _small_monolayer():
    rng = np.random.default_rng(11)
    g = np.stack(np.meshgrid(np.arange(3), np.arange(3), indexing="ij"), -1).reshape(-1, 2) * 8.0 + 4.0
    pos = np.concatenate([g, np.full((9, 1), 12.0)], axis=1) + rng.normal(0, 0.2, (9, 3))
    return pos, np.array([24.0, 24.0, 24.0])


@pytest.mark.parametrize("batched,precision", [(False, "double"), (True, "double"), (True, "refined")])
def test_eels_ewald_orders_and_oracle(batched, precision):
    pos, box = _small_monolayer()
    om = np.linspace(0.08, 0.2, 5)
    eps = cumpm_amd.drude_dielectric(om, 0.01, 0.25, 4.0)
    s = cumpm_amd.EELS(box, eps, om, 0.5, radius=1.0, tol=1e-8, field_type="monodisperse", quiet=True,
                       precision=precision, batched=batched)
    s.compute(np.array([4.5, 4.0]), pos)
    o = O.EELSSolver(box, eps, om, 0.5, radius=1.0, tol=1e-8, field="monodisperse")
    o.compute([4.5, 4.0], pos)
    assert rel(s.get_eels(), o.get_eels()) < 1e-6
    assert rel(s.get_dipoles(), o.get_dipoles()) < 1e-6
    assert all(r["converged"] for r in s.get_reports())


def test_eels_aperiodic_sampling_matches_oracle():
    pos, box = _small_monolayer()
    om = np.array([0.05, 0.12])
    eps = cumpm_amd.drude_dielectric(om, 0.01, 0.25, 4.0)
    s = cumpm_amd.EELS(box, eps, om, 0.5, radius=1.0, tol=1e-8, field_type="monodisperse", quiet=True,
                       precision="double", asm_flag=True, asm_Nx=2, asm_Ny=2, batched=True)
    s.compute(np.array([4.5, 4.0]), pos)
    o = O.EELSSolver(box, eps, om, 0.5, radius=1.0, tol=1e-8, field="monodisperse", asm=(2, 2))
    o.compute([4.5, 4.0], pos)
    assert rel(s.get_eels(), o.get_eels()) < 1e-6
    assert rel(s.get_dipoles(), o.get_dipoles()) < 1e-6
    assert len(s.get_reports()) == 4 * 2


def test_near_field_direct_matches_oracle():
    rng = np.random.default_rng(5)
    dpos = rng.uniform(-5, 5, (6, 3))
    dip = rng.standard_normal((6, 3)) + 1j * rng.standard_normal((6, 3))
    fpts = rng.uniform(-6, 6, (9, 3))
    fpts[0] = dpos[2] + [0.3, 0.0, 0.1]  # inside a sphere: the source-radius clamp
    E0 = np.array([1.0, 0.5j, 0.0])
    nf = cumpm_amd.Near_Field([60.0] * 3, E0, radius=1.5, field_type="direct")
    nf.set_dipoles(dip)
    nf.set_dipole_positions(dpos)
    nf.set_field_points(fpts)
    got = nf.calculate()
    D = O.DirectOperator(np.r_[np.full(6, 1.0), np.ones(9)])
    D.set_positions(np.concatenate([dpos, fpts]) / 1.5)
    y = D.apply(np.concatenate([dip.reshape(-1), np.zeros(27)]), np.zeros(15))[18:].reshape(9, 3)
    want = np.sum(np.abs(E0[None, :] - y) ** 2, axis=1)
    assert rel(got, want) < 1e-12
