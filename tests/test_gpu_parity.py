"""GPU parity tests: the HIP path (through the C ABI) against the CPU oracle.

Bars: operator applies bit-for-bit up to summation order (relative 1e-11 in
FP64); FP32-reciprocal applies within 1e-4 of FP64 (the reference's own
tolerance, tests/test_electric_fields.py:324-357); solves within 1e-6 of the
oracle's solution at tol 1e-8 (north-star parity bar), and the reference's
pyMPM fixture within its 1e-3.
"""
import os
import sys

import numpy as np
import pytest

from cumpm_amd import _capi as C
from oracle import oracle as O
from gpu_helpers import Op, rel, hex_monolayer

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


GEOMS = [
    # name, box, errortol, xi, positions, grid override, support override
    ("random3d", [16.0, 17.0, 18.0], 1e-6, 0.5, lambda: np.random.default_rng(1).uniform(0, 16, (40, 3)), None, 0),
    ("outside_box", [16.0, 16.0, 16.0], 1e-6, 0.5, lambda: np.random.default_rng(2).uniform(-9, 30, (25, 3)), None, 0),
    ("hex_slab", None, 1e-6, 0.5, None, None, 0),
    # odd nz (21): the slab Green's-tensor table has real / imaginary entries
    ("hex_slab_odd", None, 1e-6, 0.5, None, None, 0),
    ("sc_tight", [24.0] * 3, 1e-8, 0.5,
     lambda: np.stack(np.meshgrid(*[np.arange(4) * 3.0 + 1.5] * 3, indexing="ij"), -1).reshape(-1, 3), None, 0),
    ("tiny_grid_wrap", [15.0, 15.0, 15.0], 1e-4, 0.5,
     lambda: np.random.default_rng(3).uniform(0, 15, (12, 3)), (6, 7, 5), 0),
    ("override_32_P8", [24.0] * 3, 1e-8, 0.5, lambda: np.random.default_rng(4).uniform(0, 24, (30, 3)), (32, 32, 32), 8),
    ("single_particle", [20.0] * 3, 1e-6, 0.5, lambda: np.array([[3.0, 4.0, 5.0]]), None, 0),
    ("bilayer", None, 1e-4, 0.5, None, None, 0),
    # FFT lengths covering the register radices 3..17, 23 and the generic stage (29, 61 = prime)
    ("radix_mix_a", [24.0, 22.0, 26.0], 1e-6, 0.5, lambda: np.random.default_rng(5).uniform(0, 22, (30, 3)),
     (39, 34, 46), 0),
    ("radix_mix_b", [24.0, 21.0, 30.0], 1e-6, 0.5, lambda: np.random.default_rng(6).uniform(0, 21, (30, 3)),
     (58, 27, 121), 0),
    ("radix_mix_c", [24.0, 24.0, 24.0], 1e-6, 0.5, lambda: np.random.default_rng(7).uniform(0, 24, (30, 3)),
     (61, 44, 45), 0),
]


def _geom(name):
    for g in GEOMS:
        if g[0] == name:
            nm, box, tol, xi, pf, grid, P = g
            if nm in ("hex_slab", "hex_slab_odd"):
                pos, box = hex_monolayer(10, 8, 3.0, jitter=0.04, seed=5)
                box = np.array([box[0], box[1], 16.0 if nm == "hex_slab" else 16.5])
            elif nm == "bilayer":  # two layers: the occupied z-planes are not contiguous
                p1, box = hex_monolayer(8, 6, 2.5, jitter=0.05, seed=6)
                p2 = p1 + np.array([1.25, 0.7, 12.0])
                pos = np.concatenate([p1, p2])
                box = np.array([box[0], box[1], 40.0])
            else:
                pos = pf()
            return np.asarray(box, float), tol, xi, pos, grid, P
    raise KeyError(name)


SLAB_GEOMS = ("hex_slab", "hex_slab_odd", "bilayer")


@pytest.mark.parametrize("name,layout,fft", [(g[0], 0, 0) for g in GEOMS] + [(g, 1, 0) for g in SLAB_GEOMS] +
                         [("random3d", 0, 1), ("hex_slab", 0, 1), ("hex_slab", 1, 1), ("sc_tight", 0, 1)])
def test_apply_fp64_matches_oracle(name, layout, fft):
    box, tol, xi, pos, grid, P = _geom(name)
    _case(f"apply fp64 {name} layout={layout} fft={fft} N={len(pos)}")
    N = len(pos)
    rng = np.random.default_rng(11)
    # 12 right-hand sides: wide-batch kernels (four column groups, thread-per-RHS
    # finish); the first 3 again on their own: narrow-batch kernels
    B = 12
    p = rng.standard_normal((B, N, 3)) + 1j * rng.standard_normal((B, N, 3))
    rows = rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N))
    op = Op(box, tol, xi, grid, P, layout=layout, fft_backend=fft)
    op.set_positions(pos)
    op.set_self_coef(rows)
    ref_op = O.Operator(box, tol, xi, grid, P)
    ref_op.set_positions(pos)
    info = op.info()
    assert info.fft_backend == fft
    if layout == 1:
        assert info.slab_planes == 0
    elif name in SLAB_GEOMS:
        assert info.slab_planes > 0
    assert tuple(info.grid) == ref_op.info()["grid"]
    assert info.num_pairs == ref_op.info()["num_pairs"]
    for parts in (C.PART_REAL, C.PART_RECIP, C.PART_ALL):
        got = op.apply(p, C.PREC_DOUBLE, parts)
        got3 = op.apply(p[:3], C.PREC_DOUBLE, parts)
        for b in range(B):
            ref = ref_op.apply(p[b], rows[b], parts)
            if np.abs(ref).max() == 0:
                assert np.abs(got[b]).max() == 0
                continue
            assert rel(got[b], ref) < 1e-11, (name, parts, b, rel(got[b], ref))
            if b < 3:
                assert rel(got3[b], ref) < 1e-11, (name, parts, b, rel(got3[b], ref))


@pytest.mark.parametrize("name,layout,fft", [("random3d", 0, 0), ("hex_slab", 0, 0), ("hex_slab", 1, 0),
                                             ("sc_tight", 0, 0), ("bilayer", 0, 0), ("hex_slab", 0, 1),
                                             ("radix_mix_a", 0, 0), ("radix_mix_b", 0, 0), ("hex_slab_odd", 0, 0)])
def test_apply_fp32_reciprocal_within_reference_tolerance(name, layout, fft):
    box, tol, xi, pos, grid, P = _geom(name)
    _case(f"apply fp32 {name} layout={layout} fft={fft}")
    N = len(pos)
    rng = np.random.default_rng(12)
    p = rng.standard_normal((2, N, 3)) + 1j * rng.standard_normal((2, N, 3))
    rows = rng.standard_normal((2, N)) + 1j * rng.standard_normal((2, N))
    op = Op(box, tol, xi, grid, P, layout=layout, fft_backend=fft)
    op.set_positions(pos)
    op.set_self_coef(rows)
    ref_op = O.Operator(box, tol, xi, grid, P)
    ref_op.set_positions(pos)
    got = op.apply(p, C.PREC_MIXED)
    for b in range(2):
        ref = ref_op.apply(p[b], rows[b])
        assert rel(got[b], ref) < 1e-4


@pytest.mark.parametrize("name", ["hex_slab", "hex_slab_odd"])
def test_fused_x_transform_slab_fp32_matches_fp64(name):
    """FP32 reciprocal part (x transform, plane coupling and inverse x transform in one
    kernel) against the FP64 one (separate passes) at FP32 rounding, with a Bloch phase."""
    box, tol, xi, pos, grid, P = _geom(name)
    grid = (40, 26, 20 if name == "hex_slab" else 21)  # x = 10 x 4: two register stages
    _case(f"fused x transform {name} grid={grid}")
    N = len(pos)
    rng = np.random.default_rng(21)
    p = rng.standard_normal((5, N, 3)) + 1j * rng.standard_normal((5, N, 3))
    op = Op(box, tol, xi, grid, P)
    op.set_positions(pos)
    op.set_bloch(0.11, -0.07)
    op.set_self_coef(np.ones(N))
    C.profile_reset()
    C.profile_enable(True)
    got = op.apply(p, C.PREC_MIXED, C.PART_RECIP)
    C.profile_enable(False)
    names = C.profile_names()
    assert any(n.startswith("xtoeplitz_kernel<float") for n in names), names
    ref = op.apply(p, C.PREC_DOUBLE, C.PART_RECIP)
    assert rel(got, ref) < 1e-5
    ref_op = O.Operator(box, tol, xi, grid, P)
    ref_op.set_positions(pos)
    ref_op.set_bloch(0.11, -0.07)
    assert rel(ref[0], ref_op.apply(p[0], np.ones(N), parts=2)) < 1e-11


def test_bloch_wavevector_matches_oracle():
    box, tol, xi, pos, grid, P = _geom("hex_slab")
    _case("bloch")
    N = len(pos)
    rng = np.random.default_rng(13)
    p = rng.standard_normal((6, N, 3)) + 1j * rng.standard_normal((6, N, 3))
    sc = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    op = Op(box, tol, xi)
    op.set_positions(pos)
    op.set_self_coef(sc)
    ref_op = O.Operator(box, tol, xi)
    ref_op.set_positions(pos)
    for kb in ((0.15, -0.22), (-0.05, 0.31)):  # the second call re-phases the pair table
        op.set_bloch(*kb)
        ref_op.set_bloch(*kb)
        got1 = op.apply(p[:1])
        got6 = op.apply(p)
        for b in range(6):
            ref = ref_op.apply(p[b], sc)
            assert rel(got6[b], ref) < 1e-11, (kb, b)
        assert rel(got1[0], ref_op.apply(p[0], sc)) < 1e-11


@pytest.mark.parametrize("method", [C.GMRES, C.BICGSTAB])
def test_solve_matches_oracle(method):
    box, tol, xi, pos, grid, P = _geom("random3d")
    _case(f"solve method={method}")
    N = len(pos)
    eps = np.array([-2.5 + 0.6j, -1.2 + 0.3j, 3.0 + 0.1j])
    rows = np.stack([O.self_coefficient(np.full(N, e)) for e in eps])
    op = Op(box, tol, xi)
    op.set_positions(pos)
    op.set_self_coef(rows)
    b = np.zeros((3, N, 3), complex)
    for k in range(3):
        b[k, :, k] = 1.0
    stol = 1e-8 if method == C.GMRES else 1e-7
    x, reps = op.solve(b, tol=stol, method=method, restart=20)
    ref_op = O.Operator(box, tol, xi)
    ref_op.set_positions(pos)
    for k in range(3):
        xr, rr = ref_op.solve(b[k], rows[k], tol=stol, method="gmres" if method == C.GMRES else "bicgstab",
                              restart=20)
        assert reps[k]["converged"], reps[k]
        assert rel(x[k].ravel(), xr) < 1e-6, (k, rel(x[k].ravel(), xr), reps[k], rr)
        if method == C.GMRES:
            # same algorithm, same stopping rule: Arnoldi step counts agree to +-1
            assert abs(reps[k]["iters"] - rr["iters"]) <= 1, (reps[k], rr)


def test_batched_equals_one_at_a_time():
    box, tol, xi, pos, grid, P = _geom("hex_slab")
    _case("batched vs sequential")
    N = len(pos)
    eps = np.linspace(-3.0, -1.0, 5) + 0.3j
    rows = np.stack([O.self_coefficient(np.full(N, e)) for e in eps])
    b = np.zeros((5, N, 3), complex)
    for q in range(5):
        b[q, :, q % 3] = 1.0
    op = Op(box, tol, xi)
    op.set_positions(pos)
    op.set_self_coef(rows)
    xb, rb = op.solve(b, tol=1e-9, restart=30)
    for q in range(5):
        op1 = Op(box, tol, xi)
        op1.set_positions(pos)
        op1.set_self_coef(rows[q])
        x1, r1 = op1.solve(b[q:q + 1], tol=1e-9, restart=30)
        assert rel(xb[q], x1[0]) < 1e-7
        assert abs(rb[q]["iters"] - r1[0]["iters"]) <= 1


@pytest.mark.parametrize("precision", [C.SOLVE_DOUBLE, C.SOLVE_REFINED])
def test_stream_of_more_systems_than_slots(precision):
    """70 systems (more than the 32 slots) in one GMRES call: retired systems' slots are
    refilled, every system follows its own restart cycles (restart 12, so most need several),
    and each result equals its solve alone; some systems hit maxiter while others continue."""
    box, tol, xi, pos, grid, P = _geom("hex_slab")
    _case(f"stream of 70 systems precision={precision}")
    N = len(pos)
    rng = np.random.default_rng(9)
    B = 70
    eps = rng.uniform(-4.0, -1.2, B) + 1j * rng.uniform(0.3, 1.0, B)
    rows = np.stack([O.self_coefficient(np.full(N, e)) for e in eps])
    b = rng.standard_normal((B, N, 3)) + 1j * rng.standard_normal((B, N, 3))
    b[5] = 0.0  # a zero right-hand side retires at once
    op = Op(box, tol, xi)
    op.set_positions(pos)
    op.set_self_coef(rows)
    xb, rb = op.solve(b, tol=1e-9, restart=12, maxiter=600, precision=precision)
    assert np.all(xb[5] == 0) and rb[5]["converged"] and rb[5]["iters"] == 0
    for q in (0, 1, 17, 33, 40, 63, 69):
        op1 = Op(box, tol, xi)
        op1.set_positions(pos)
        op1.set_self_coef(rows[q])
        x1, r1 = op1.solve(b[q:q + 1], tol=1e-9, restart=12, maxiter=600, precision=precision)
        assert rb[q]["converged"] and r1[0]["converged"]
        assert rel(xb[q], x1[0]) < 1e-7, (q, rel(xb[q], x1[0]))
        assert abs(rb[q]["iters"] - r1[0]["iters"]) <= 1
    assert max(r["iters"] for r in rb) > 12  # several restart cycles happened
    # a maxiter budget that some systems exhaust while the others converge
    x2, r2 = op.solve(b, tol=1e-9, restart=12, maxiter=30, precision=precision)
    it = np.array([r["iters"] for r in r2])
    assert it.max() <= 30 and any(not r["converged"] for r in r2) and any(r["converged"] for r in r2)
    for q, r in enumerate(r2):
        if r["converged"] and q != 5:
            assert r["relres_pre"] < 1e-9


def test_refined_precision_reaches_fp64_accuracy():
    box, tol, xi, pos, grid, P = _geom("hex_slab")
    _case("refined")
    N = len(pos)
    eps = np.array([-2.2 + 0.2j, -1.5 + 0.4j, 2.0 + 0.1j, -4.0 + 1.0j])
    rows = np.stack([O.self_coefficient(np.full(N, e)) for e in eps])
    b = np.zeros((4, N, 3), complex)
    for q in range(4):
        b[q, :, q % 3] = 1.0
    op = Op(box, tol, xi)
    op.set_positions(pos)
    op.set_self_coef(rows)
    x64, r64 = op.solve(b, tol=1e-8, restart=30, precision=C.SOLVE_DOUBLE)
    xr, rr = op.solve(b, tol=1e-8, restart=30, precision=C.SOLVE_REFINED)
    for q in range(4):
        assert rr[q]["converged"] and rr[q]["relres_pre"] < 1e-8, rr[q]
        assert rel(xr[q], x64[q]) < 1e-6
        assert rr[q]["matvecs_fp64"] < rr[q]["matvecs"]
    # non-zero initial guesses: a perturbed solution, and the FP64 solution itself
    # (its FP32 initial residual is not below tol, its FP64 end-of-cycle one is)
    rng = np.random.default_rng(11)
    for x0 in (x64 * (1 + 1e-3 * rng.standard_normal(x64.shape)), x64):
        xg, rg = op.solve(b, x0=x0, tol=1e-8, restart=30, precision=C.SOLVE_REFINED)
        for q in range(4):
            assert rg[q]["converged"] and rg[q]["relres_pre"] < 1e-8, rg[q]
            assert rel(xg[q], x64[q]) < 1e-6


@pytest.mark.parametrize("bloch", [(0.0, 0.0), (0.21, -0.13)])
def test_adjoint_solve_satisfies_adjoint_identity(bloch):
    """mpm_solve_adjoint: A^H y = g, checked through <A x, y> = <x, g> with the
    forward operator, for random x (with and without a Bloch wave vector, where A
    is no longer complex symmetric)."""
    box, tol, xi, pos, grid, P = _geom("hex_slab")
    _case(f"adjoint bloch={bloch}")
    N = len(pos)
    rng = np.random.default_rng(21)
    eps = np.array([-2.2 + 0.3j, 3.0 + 0.2j])
    rows = np.stack([O.self_coefficient(np.full(N, e)) for e in eps])
    op = Op(box, tol, xi)
    op.set_positions(pos)
    op.set_self_coef(rows)
    op.set_bloch(*bloch)
    g = rng.standard_normal((2, N, 3)) + 1j * rng.standard_normal((2, N, 3))
    y, reps = op.solve_adjoint(g, tol=1e-11, restart=40)
    assert all(r["converged"] for r in reps), reps
    x = rng.standard_normal((2, N, 3)) + 1j * rng.standard_normal((2, N, 3))
    ax = op.apply(x)
    for q in range(2):
        lhs = np.vdot(ax[q], y[q])          # <A x, y>
        rhs = np.vdot(x[q], g[q])           # <x, A^H y> = <x, g>
        assert abs(lhs - rhs) < 1e-8 * abs(rhs), (q, lhs, rhs)
    # the self coefficients are restored: a forward solve still matches its own operator
    xs, rs = op.solve(g, tol=1e-10, restart=40)
    assert rel(op.apply(xs), g) < 1e-8


def test_zero_rhs_returns_zero():
    box, tol, xi, pos, grid, P = _geom("random3d")
    _case("zero rhs")
    N = len(pos)
    op = Op(box, tol, xi)
    op.set_positions(pos)
    op.set_self_coef(O.self_coefficient(np.full(N, -2.0 + 0.5j)))
    x, reps = op.solve(np.zeros((1, N, 3)), x0=np.ones((1, N, 3)))
    assert np.all(x == 0) and reps[0]["converged"]


def test_maxiter_is_respected():
    box, tol, xi, pos, grid, P = _geom("random3d")
    _case("maxiter")
    N = len(pos)
    op = Op(box, tol, xi)
    op.set_positions(pos)
    op.set_self_coef(O.self_coefficient(np.full(N, -2.0 + 0.01j)))
    b = np.zeros((1, N, 3), complex)
    b[0, :, 0] = 1
    x, reps = op.solve(b, tol=1e-14, maxiter=7, restart=5)
    assert reps[0]["iters"] == 7 and not reps[0]["converged"]


def test_dipole_solver_golden_fixture():
    """Reference tests/test_dipole_solvers.py:8-60 through the product API."""
    import cumpm_amd
    _case("golden fixture")
    ref = np.load(os.path.join(GOLDEN, "ref_dipole_monodisperse.npz"))
    d, a, eps_m = 12.0, 4.0, 2.13
    omega = np.linspace(2000, 6000, 3)
    eps = 4.0 - 12000 ** 2 / (omega ** 2 + 1j * omega * 700)
    pos = np.array([[0, 0, 0], [d, 0, 0], [0, d, 0], [d, d, 0]], float)
    epsm = np.repeat(eps[:, None], 4, axis=1)
    for solver_type in ("bicgstab", "gmres"):
        s = cumpm_amd.dipole_solver(np.array([5 * d] * 3), epsm, radius=a, eps_m=eps_m, tol=1e-4,
                                    field_type="monodisperse", solver_type=solver_type, quiet=True,
                                    precision="double")
        s.compute(pos)
        np.testing.assert_allclose(s.get_eff_polarizability(physical=False), ref["alpha_py"], rtol=1e-3, atol=1e-3)
        np.testing.assert_allclose(s.get_dipoles(physical=False), ref["dips_py"], rtol=1e-3, atol=1e-3)


@pytest.mark.parametrize("batched", [False, True])
def test_dipole_solver_matches_oracle_driver(batched):
    import cumpm_amd
    _case(f"driver batched={batched}")
    pos, box = hex_monolayer(10, 12, 11.0, jitter=0.2, seed=2)
    box = np.array([box[0], box[1], 100.0])
    omega = np.linspace(0.2, 0.8, 4)
    eps = 4.0 - 1.7 ** 2 / (omega ** 2 + 1j * omega * 0.08)
    s = cumpm_amd.dipole_solver(box, eps, radius=5.0, eps_m=1.0, tol=1e-8, quiet=True, precision="double",
                                batched=batched)
    s.compute(pos)
    od = O.DipoleSolver(box, eps.reshape(-1, 1), 5.0, 1.0, tol=1e-8)
    od.compute(pos)
    assert rel(s.get_dipoles(), od.get_dipoles()) < 1e-6
    assert rel(s.get_eff_polarizability(), od.get_eff_polarizability()) < 1e-6
    reps = s.get_reports()
    assert all(r["converged"] for w in reps[0] for r in w)


def test_dipole_solver_input_errors_need_gpu():
    import cumpm_amd
    _case("input errors")
    with pytest.raises(RuntimeError, match="half the box"):
        s = cumpm_amd.dipole_solver([8.0, 8.0, 8.0], -2.0 + 0.1j, tol=1e-6, quiet=True)
        s.compute(np.array([[1.0, 1.0, 1.0], [4.0, 4.0, 4.0]]))
    s = cumpm_amd.dipole_solver([40.0] * 3, -2.0 + 0.1j, radius=[1.0, 1.2], field_type="monodisperse", quiet=True)
    with pytest.raises(RuntimeError, match="not identical"):
        s.compute(np.array([[1.0, 1.0, 1.0], [6.0, 6.0, 6.0]]))
    s = cumpm_amd.dipole_solver([40.0] * 3, -2.0 + 0.1j, radius=[1.0, 1.2, 1.0], quiet=True)
    with pytest.raises(RuntimeError, match="inconsistent"):
        s.compute(np.array([[1.0, 1.0, 1.0], [6.0, 6.0, 6.0]]))
    s = cumpm_amd.dipole_solver([40.0] * 3, -2.0 + 0.1j, quiet=True)
    s.compute(np.array([[1.0, 1.0, 1.0], [6.0, 6.0, 6.0]]))
    with pytest.raises(RuntimeError, match="changed"):
        s.compute(np.array([[1.0, 1.0, 1.0]]))


def test_profile_copy_moves_exact_bytes():
    """mpm_profile_copy (the counter-calibration kernel) copies exactly `bytes`."""
    _case("profile copy")
    src = np.random.default_rng(3).standard_normal(4099 * 2)
    a, b = C.DeviceBuffer(src.nbytes), C.DeviceBuffer(src.nbytes)
    a.upload(src)
    b.upload(np.zeros_like(src))
    n = (src.nbytes // 16 - 1) * 16
    C.check(C.load().mpm_profile_copy(b.ptr, a.ptr, n, None))
    out = np.zeros_like(src)
    b.download(out)
    assert np.array_equal(out[: n // 8], src[: n // 8]) and not out[n // 8:].any()
    with pytest.raises(ValueError):
        C.check(C.load().mpm_profile_copy(b.ptr, a.ptr, 24, None))


POLY_GEOMS = [
    # name, box, errortol, xi, positions, radii, grid override, bloch
    ("poly_random", [22.0, 23.0, 24.0], 1e-6, 0.5,
     lambda: np.random.default_rng(31).uniform(0, 22, (24, 3)),
     lambda: np.random.default_rng(32).choice([1.0, 1.4, 2.0], 24), None, None),
    ("poly_overlap_bloch", [20.0, 21.0, 22.0], 1e-6, 0.5,
     lambda: np.array([[5.0, 5.0, 5.0], [5.0, 5.0, 6.5], [5.0, 5.9, 5.0], [12.0, 3.0, 17.0], [12.4, 3.1, 17.2]]),
     lambda: np.array([1.0, 1.6, 1.2, 2.5, 1.0]), None, (0.13, -0.21)),
    ("poly_two_classes_grid", [24.0, 24.0, 24.0], 1e-5, 0.55,
     lambda: np.random.default_rng(33).uniform(0, 24, (30, 3)),
     lambda: np.where(np.arange(30) % 3 == 0, 1.8, 1.0), (36, 30, 40), None),
]


def _poly(name):
    for g in POLY_GEOMS:
        if g[0] == name:
            _, box, tol, xi, pf, rf, grid, kb = g
            return np.asarray(box, float), tol, xi, pf(), rf(), grid, kb
    raise KeyError(name)


@pytest.mark.parametrize("name", [g[0] for g in POLY_GEOMS])
def test_poly_apply_fp64_matches_oracle(name):
    box, tol, xi, pos, rad, grid, kb = _poly(name)
    _case(f"poly apply fp64 {name} N={len(pos)}")
    N = len(pos)
    rng = np.random.default_rng(41)
    p = rng.standard_normal((5, N, 3)) + 1j * rng.standard_normal((5, N, 3))
    rows = rng.standard_normal((5, N)) + 1j * rng.standard_normal((5, N))
    op = Op(box, tol, xi, grid, radii=rad)
    ref_op = O.PolyOperator(box, tol, xi, rad, grid)
    if kb:
        op.set_bloch(*kb)
        ref_op.set_bloch(*kb)
    op.set_positions(pos)
    op.set_self_coef(rows)
    ref_op.set_positions(pos)
    info = op.info()
    oi = ref_op.info()
    assert info.polydisperse == 1 and tuple(info.grid) == oi["grid"] and info.support == oi["P"]
    assert info.rc == pytest.approx(oi["rc"]) and info.num_pairs == oi["num_pairs"]
    for parts in (C.PART_REAL, C.PART_RECIP, C.PART_ALL):
        got = op.apply(p, C.PREC_DOUBLE, parts)
        for b in range(5):
            ref = ref_op.apply(p[b], rows[b], parts)
            assert rel(got[b], ref) < 1e-11, (parts, b)


@pytest.mark.parametrize("name", ["poly_random", "poly_overlap_bloch"])
def test_poly_apply_fp32_within_reference_tolerance(name):
    box, tol, xi, pos, rad, grid, kb = _poly(name)
    _case(f"poly apply fp32 {name}")
    N = len(pos)
    rng = np.random.default_rng(42)
    p = rng.standard_normal((3, N, 3)) + 1j * rng.standard_normal((3, N, 3))
    op = Op(box, tol, xi, grid, radii=rad)
    if kb:
        op.set_bloch(*kb)
    op.set_positions(pos)
    op.set_self_coef(np.ones(N))
    ref = op.apply(p, C.PREC_DOUBLE)
    got = op.apply(p, C.PREC_MIXED)
    assert rel(got, ref) < 1e-4


def test_poly_solve_matches_oracle():
    box, tol, xi, pos, rad, grid, kb = _poly("poly_random")
    _case("poly gmres")
    N = len(pos)
    eps = -2.0 + 0.4j
    sc = O.self_coefficient(np.full(N, eps), rad)
    b = np.zeros((2, N, 3), complex)
    b[0, :, 0] = 1.0
    b[1, :, 2] = 1.0
    op = Op(box, tol, xi, grid, radii=rad)
    op.set_positions(pos)
    op.set_self_coef(np.stack([sc, sc]))
    x, reps = op.solve(b, tol=1e-8)
    ref_op = O.PolyOperator(box, tol, xi, rad, grid)
    ref_op.set_positions(pos)
    for q in range(2):
        xr, rr = ref_op.solve(b[q], sc, tol=1e-8)
        assert reps[q]["converged"] and rr["converged"]
        assert rel(x[q], xr.reshape(N, 3)) < 1e-6


def test_poly_identical_radii_matches_monodisperse():
    """Reference tests/test_electric_fields.py:30-82 on the device (1e-5)."""
    _case("poly vs mono")
    pos = np.array([[0.0, 0.0, 0.0], [5.0, 0.0, 0.0], [10.0, 0.0, 0.0], [15.0, 0.0, 0.0]])
    p = np.zeros((1, 4, 3), complex)
    p[0, :, 0] = 1.0
    mono = Op([20.0] * 3, 1e-6, 0.5, layout=1)
    poly = Op([20.0] * 3, 1e-6, 0.5, radii=[1.0] * 4)
    for op in (mono, poly):
        op.set_positions(pos)
        op.set_self_coef(np.ones(4))
    np.testing.assert_allclose(poly.apply(p), mono.apply(p), rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("batched,precision", [(False, "double"), (True, "refined")])
def test_dipole_solver_polydisperse_matches_oracle_driver(batched, precision):
    """field_type="auto" with two radius classes selects the polydisperse operator
    (dipole_solver.cu:527-536); physical units, nondimensionalised by the smallest radius."""
    import cumpm_amd
    _case(f"poly driver batched={batched} {precision}")
    rng = np.random.default_rng(7)
    cols, rows, sp = 6, 5, 11.0
    pos, box = hex_monolayer(cols, rows, sp, jitter=0.2, seed=3)
    box = np.array([box[0] + 40.0, box[1] + 40.0, 100.0])
    radii = np.where(rng.random(len(pos)) < 0.5, 4.0, 5.0)
    omega = np.linspace(0.3, 0.7, 3)
    eps = 4.0 - 1.7 ** 2 / (omega ** 2 + 1j * omega * 0.08)
    s = cumpm_amd.dipole_solver(box, eps, radius=radii, eps_m=1.0, tol=1e-8, quiet=True, precision=precision,
                                batched=batched)
    s.compute(pos)
    od = O.DipoleSolver(box, eps.reshape(-1, 1), radii, 1.0, tol=1e-8)
    od.compute(pos)
    assert rel(s.get_dipoles(), od.get_dipoles()) < 1e-6
    assert rel(s.get_eff_polarizability(), od.get_eff_polarizability()) < 1e-6
    reps = s.get_reports()
    assert all(r["converged"] for w in reps[0] for r in w)


def test_dev_polydisperse_field_matches_monodisperse_for_identical_radii():
    """Reference tests/test_electric_fields.py:30-82 through the dev API (1e-5; splits 1e-12)."""
    from cumpm_amd.dev import MonodisperseEwaldElectricField, PolydisperseEwaldElectricField
    _case("dev poly vs mono")
    pos = np.array([[0.0, 0.0, 0.0], [5.0, 0.0, 0.0], [10.0, 0.0, 0.0], [15.0, 0.0, 0.0]])
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    fields = []
    for ef in (MonodisperseEwaldElectricField(20.0, 20.0, 20.0, 1e-6, 0.5, True, precision="double"),
               PolydisperseEwaldElectricField(20.0, 20.0, 20.0, 1e-6, 0.5, True, [1.0] * 4, precision="double")):
        ef.update_particle_coordinates(x, y, z)
        ef.update_dipoles(np.ones(4), np.zeros(4), np.zeros(4))
        ef.set_self_coef(np.ones(4), np.zeros(4))
        ef.calculate()
        tot = ef.get_epoint_host()
        np.testing.assert_allclose(tot, ef.compute_real_field() + ef.compute_inverse_field(), rtol=1e-12, atol=1e-12)
        fields.append(tot)
    np.testing.assert_allclose(fields[1], fields[0], rtol=1e-5, atol=1e-5)


def _quad_geom():
    L = np.array([22.0, 23.0, 24.0])
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, 20, (10, 3))
    pos[1] = pos[0] + [1.5, 0.4, 0.3]   # overlapping
    pos[2] = pos[0] + [0.0, 2.6, 0.2]   # close
    return L, pos


@pytest.mark.parametrize("quads,bloch", [(True, None), ([0, 3, 7], (0.12, -0.2))])
def test_quad_apply_fp64_matches_oracle(quads, bloch):
    L, pos = _quad_geom()
    _case(f"quad apply fp64 quads={quads} bloch={bloch}")
    op = Op(L, 1e-8, 0.5, quads=quads)
    ref = O.Operator(L, 1e-8, 0.5)
    if bloch:
        op.set_bloch(*bloch)
        ref.set_bloch(*bloch)
    op.set_positions(pos)
    ref.set_positions(pos)
    ref.set_quads(None if quads is True else quads)
    assert op.n == ref.n and op.info().num_quads == ref.Q
    rng = np.random.default_rng(51)
    B = 3
    p = rng.standard_normal((B, op.n)) + 1j * rng.standard_normal((B, op.n))
    rows = rng.standard_normal((B, 10)) + 1j * rng.standard_normal((B, 10))
    op.set_self_coef(rows)
    for parts in (C.PART_REAL, C.PART_RECIP, C.PART_ALL):
        got = op.apply(p, C.PREC_DOUBLE, parts)
        for b in range(B):
            assert rel(got[b], ref.apply(p[b], rows[b], parts)) < 1e-11, (parts, b)
    got32 = op.apply(p, C.PREC_MIXED)
    assert rel(got32, op.apply(p, C.PREC_DOUBLE)) < 1e-4


def test_quad_golden_fixture_through_solve():
    """Reference tests/test_dipole_solvers.py:63-148: GMRES on the device against the
    fixture's dense solve (its bar 1e-5)."""
    _case("quad golden fixture")
    ref = np.load(os.path.join(GOLDEN, "ref_quadrupoles_monodisperse.npz"))
    op = Op([100.0] * 3, 1e-6, 0.4, quads=True)
    op.set_positions(np.array([[10.0, 10.0, 10.0], [20.0, 20.0, 20.0]]))
    assert op.n == 16
    sc = O.self_coefficient(np.full(2, 4.5 + 0.2j))
    op.set_self_coef(np.stack([sc] * 3))
    b = np.zeros((3, 16), complex)
    for dim in range(3):
        b[dim, dim:6:3] = 1.0
    x, reps = op.solve(b, tol=1e-12)
    for dim in range(3):
        assert reps[dim]["converged"]
        np.testing.assert_allclose(x[dim, :6].reshape(2, 3), ref["dips_py_all"][dim], rtol=1e-7, atol=1e-10)
        np.testing.assert_allclose(x[dim, 6:].reshape(2, 5), ref["quads_py_all"][dim], rtol=1e-7, atol=1e-10)


@pytest.mark.parametrize("precision", [C.SOLVE_DOUBLE, C.SOLVE_REFINED])
def test_quad_solve_matches_oracle(precision):
    L, pos = _quad_geom()
    _case(f"quad solve precision={precision}")
    op = Op(L, 1e-8, 0.5, quads=True)
    op.set_positions(pos)
    sc = O.self_coefficient(np.full(10, -2.0 + 0.5j))
    op.set_self_coef(sc)
    b = np.zeros((2, op.n), complex)
    b[0, 0:30:3] = 1.0
    b[1, 2:30:3] = 1.0
    x, reps = op.solve(b, tol=1e-10, precision=precision)
    ref = O.Operator(L, 1e-8, 0.5)
    ref.set_positions(pos)
    ref.set_quads()
    for q in range(2):
        xr, rr = ref.solve(b[q], sc, tol=1e-10)
        assert reps[q]["converged"] and rr["converged"]
        assert rel(x[q], xr) < 1e-6


@pytest.mark.parametrize("precision,batched", [("double", False), ("refined", True)])
def test_dipole_solver_quadrupole_fixture(precision, batched):
    """Reference tests/test_dipole_solvers.py:63-148 through the product API (its bar 1e-5)."""
    import cumpm_amd
    _case(f"driver quad fixture {precision} batched={batched}")
    ref = np.load(os.path.join(GOLDEN, "ref_quadrupoles_monodisperse.npz"))
    s = cumpm_amd.dipole_solver(np.array([100.0] * 3), 4.5 + 0.2j, radius=1.0, eps_m=1.0, xi=0.4, tol=1e-8,
                                field_type="monodisperse", quiet=True, quadrupoles=True, precision=precision,
                                batched=batched)
    s.compute(np.array([[10.0, 10.0, 10.0], [20.0, 20.0, 20.0]]))
    dips, quads = s.get_dipoles(), s.get_quadrupoles()
    assert dips.shape == (2, 3, 3) and quads.shape == (2, 3, 5)
    for dim in range(3):
        np.testing.assert_allclose(dips[:, dim], ref["dips_py_all"][dim], rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(quads[:, dim], ref["quads_py_all"][dim], rtol=1e-5, atol=1e-5)


def _quad_driver_case(a=1.0, eps_m=1.0):
    L, pos = _quad_geom()
    omega = np.linspace(0.3, 0.7, 4)
    eps = (4.0 - 1.7 ** 2 / (omega ** 2 + 1j * omega * 0.08)) * eps_m
    pos2 = pos + np.random.default_rng(11).normal(0.0, 0.05, pos.shape)
    return L * a, np.stack([pos, pos2]) * a, eps


@pytest.mark.parametrize("quads,batched,precision", [(True, False, "double"), ([0, 3, 7], False, "double"),
                                                     ([0, 3, 7], True, "refined")])
def test_dipole_solver_quadrupoles_match_oracle_driver(quads, batched, precision):
    """Two frames, four frequencies, derivative guesses with quadrupoles (dipole_solver.cu:233-345,
    :578-601) against the oracle's driver; overlapping and close pairs included."""
    import cumpm_amd
    _case(f"driver quads={quads} batched={batched} {precision}")
    L, frames, eps = _quad_driver_case()
    s = cumpm_amd.dipole_solver(L, eps, radius=1.0, tol=1e-8, quiet=True, quadrupoles=quads, precision=precision,
                                batched=batched)
    s.compute(frames)
    od = O.DipoleSolver(L, eps.reshape(-1, 1), 1.0, tol=1e-8, quadrupoles=quads)
    for f in frames:
        od.compute(f)
    Q = 10 if quads is True else 3
    assert s.get_num_quads() == Q
    q = s.get_quadrupoles()
    assert q.shape == (2, 4, Q, 3, 5)
    assert rel(s.get_dipoles(), od.get_dipoles()) < 1e-6
    assert rel(q, od.get_quadrupoles()) < 1e-6
    assert rel(s.get_eff_polarizability(), od.get_eff_polarizability()) < 1e-6
    assert all(r["converged"] for fr in s.get_reports() for w in fr for r in w)


def test_dipole_solver_quadrupole_units_and_full_rhs():
    """Physical units: lengths x a and eps x eps_m leave the nondimensional problem unchanged,
    so dipoles scale by eps_m a^3 and quadrupoles by eps_m a^4 (dipole_solver.cu:627-647).
    A full-length E0 (3N + 5Q, quadrupole rows nonzero) matches the oracle's driver."""
    import cumpm_amd
    _case("driver quad units / full rhs")
    out = []
    for a, eps_m in ((1.0, 1.0), (2.0, 1.5)):
        L, frames, eps = _quad_driver_case(a, eps_m)
        s = cumpm_amd.dipole_solver(L, eps[:2], radius=a, eps_m=eps_m, tol=1e-8, quiet=True, quadrupoles=[2, 5],
                                    precision="double")
        s.compute(frames[0])
        out.append((s.get_dipoles(), s.get_quadrupoles()))
    assert rel(out[1][0], out[0][0] * 1.5 * 2.0 ** 3) < 1e-9
    assert rel(out[1][1], out[0][1] * 1.5 * 2.0 ** 4) < 1e-9
    L, frames, eps = _quad_driver_case()
    rng = np.random.default_rng(4)
    E0 = rng.standard_normal((2, 45)) + 1j * rng.standard_normal((2, 45))
    s = cumpm_amd.dipole_solver(L, eps[:2], radius=1.0, tol=1e-8, quiet=True, quadrupoles=[2, 5, 8], E0=E0,
                                precision="double")
    s.compute(frames[0])
    od = O.DipoleSolver(L, eps[:2].reshape(-1, 1), 1.0, tol=1e-8, quadrupoles=[2, 5, 8], E0=E0)
    od.compute(frames[0])
    assert rel(s.get_dipoles(), od.get_dipoles()) < 1e-6
    assert rel(s.get_quadrupoles(), od.get_quadrupoles()) < 1e-6


def test_dev_field_quadrupoles_match_oracle():
    """Dev API with quadrupoles: get_epoint_host gives (field (N, 3), gradient (Q, 5)) as the
    reference's dev/electric_field.py:816-842; real + inverse = total."""
    from cumpm_amd.dev import MonodisperseEwaldElectricField
    _case("dev quads")
    L, pos = _quad_geom()
    ef = MonodisperseEwaldElectricField(*L, 1e-8, 0.5, True, solve_quadrupoles=True, quad_idxs=[0, 3, 7],
                                        precision="double")
    ef.update_particle_coordinates(pos[:, 0], pos[:, 1], pos[:, 2])
    assert ef.solve_quadrupoles and ef.num_quads == 3
    rng = np.random.default_rng(2)
    p = rng.standard_normal((10, 3)) + 1j * rng.standard_normal((10, 3))
    q = rng.standard_normal((3, 5)) + 1j * rng.standard_normal((3, 5))
    sc = rng.standard_normal(10) + 1j * rng.standard_normal(10)
    ef.set_dipoles(p)
    ef.update_quadrupoles_complex(*[v for c in range(5) for v in (q[:, c].real, q[:, c].imag)])
    ef.set_self_coef(sc.real, sc.imag)
    ef.calculate()
    E, G = ef.get_epoint_host()
    ref = O.Operator(L, 1e-8, 0.5)
    ref.set_positions(pos)
    ref.set_quads([0, 3, 7])
    y = ref.apply(np.concatenate([p.reshape(-1), q.reshape(-1)]), sc)
    assert rel(E, y[:30].reshape(10, 3)) < 1e-10
    assert rel(G, y[30:].reshape(3, 5)) < 1e-10
    Er, Gr = ef.compute_real_field()
    Ei, Gi = ef.compute_inverse_field()
    np.testing.assert_allclose(Er + Ei, E, rtol=1e-12, atol=1e-12 * np.abs(E).max())
    np.testing.assert_allclose(Gr + Gi, G, rtol=1e-12, atol=1e-12 * np.abs(G).max())


def test_dev_host_getters_tables_and_neighbor_list():
    """getRealSpaceTablesHost / getNeighborListHost / getRealSpacePrecalcsHost through the
    dev API: the tables equal the oracle's (independent closed forms, 1e-13), the neighbour
    rows hold exactly the pairs within rc (brute-force minimum image), and each pair's values
    are the tables interpolated at d_eff = max(d, 1) (monodisperse...cu:877-958)."""
    from cumpm_amd.dev import MonodisperseEwaldElectricField
    _case("dev host getters")
    box, tol, xi, pos, grid, P = _geom("random3d")
    pos = pos.copy()
    pos[1] = pos[0] + [0.6, 0.2, 0.1]  # d < 1: d_eff = 1
    ef = MonodisperseEwaldElectricField(*box, tol, xi, True, precision="double")
    ef.update_particle_coordinates(pos[:, 0], pos[:, 1], pos[:, 2])
    r, d1, d2 = ef.get_real_space_tables_host()
    rr, pe, pa, sp = O.tables(xi)
    np.testing.assert_array_equal(r, rr)
    np.testing.assert_allclose(d1, pe, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(d2, pa, rtol=1e-13, atol=1e-16)
    lst, counts = ef.get_neighbor_list_host()
    N = len(pos)
    width = lst.size // N
    assert width == ef.max_neighbors == counts.max()
    rc = ef.rc
    self_perp, perp, para = ef.get_real_space_precalcs_host()
    assert self_perp == pytest.approx(sp, rel=1e-13)
    k = 0
    for i in range(N):
        dv = pos[i] - pos
        dv -= box * np.round(dv / box)
        d = np.linalg.norm(dv, axis=1)
        expect = sorted(j for j in range(N) if j != i and d[j] < rc)
        row = lst[i * width:i * width + counts[i]]
        assert sorted(row.tolist()) == expect
        assert np.all(lst[i * width + counts[i]:(i + 1) * width] == -1)
        for j in row:
            de = max(d[j], 1.0)
            assert perp[k] == pytest.approx(np.interp(de, rr[1:], pe[1:]), rel=1e-12, abs=1e-16)
            assert para[k] == pytest.approx(np.interp(de, rr[1:], pa[1:]), rel=1e-12, abs=1e-16)
            k += 1
    assert k == perp.size


def test_near_field_golden_fixture_and_basic():
    """Reference tests/test_near_fields.py:5-80 through the product API: the pyMPM fixture (its
    bar 1e-4; here to 1e-8 of the oracle) and the basic shape / sign checks."""
    import cumpm_amd
    _case("near field fixture")
    ref = np.load(os.path.join(GOLDEN, "ref_near_field.npz"))["intensity_py"]
    box = np.array([50.0, 50.0, 50.0])
    dip_pos = np.array([[0.0, 0, 0], [10.0, 0, 0], [0, 10.0, 0], [0, 0, 10.0]])
    dip = np.array([[0.1 + 0.05j, 0, 0], [-0.1, 0.02j, 0], [0, 0.05, -0.01j], [0.01, 0, 0.1j]])
    fp = np.array([[5.0, 0, 0], [0, 5.0, 0], [2.5, 2.5, 2.5]])
    nf = cumpm_amd.Near_Field(box, [1.0, 0.0, 0.0], radius=1.0, dip=dip, dip_pos=dip_pos, field_points=fp,
                              precision="double")
    I = nf.calculate()
    np.testing.assert_allclose(I, ref, rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(I, O.near_field_intensity(box, [1.0, 0, 0], 1.0, dip, dip_pos, fp), rtol=1e-10)
    # the reference's default: an FP32 reciprocal grid
    I32 = cumpm_amd.Near_Field(box, [1.0, 0.0, 0.0], radius=1.0, dip=dip, dip_pos=dip_pos, field_points=fp).calculate()
    np.testing.assert_allclose(I32, ref, rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(I32, I, rtol=1e-5)
    nf2 = cumpm_amd.Near_Field(box, [1.0, 0.0, 0.0], radius=1.0, dip=[[0.1, 0, 0], [-0.1, 0, 0]],
                               dip_pos=[[0.0, 0, 0], [10.0, 0, 0]], field_points=[[5.0, 0, 0], [0, 5.0, 0]])
    I2 = nf2.calculate()
    assert I2.shape == (2,) and not np.any(np.isnan(I2)) and np.all(I2 >= 0.0)


def test_near_field_matches_oracle_with_overlap_and_radius():
    """|E_ind|^2 (E0 = 0) at field points inside, on and outside the spheres (d < 1, 1 < d < 2,
    d > 2), radius 2.5 (lengths scaled by it), errortol 1e-8: the oracle to 1e-10; the input
    errors of near_field.cu:105-140."""
    import cumpm_amd
    _case("near field vs oracle")
    rng = np.random.default_rng(12)
    box = np.array([60.0, 62.0, 64.0])
    dip_pos = rng.uniform(0, 55, (12, 3))
    dip = rng.standard_normal((12, 3)) + 1j * rng.standard_normal((12, 3))
    fp = np.concatenate([dip_pos[:3] + [[0.5, 0.3, 0.2], [2.0, 1.5, 0.0], [0.0, 0.0, 0.0]],
                         rng.uniform(0, 55, (6, 3))])
    nf = cumpm_amd.Near_Field(box, [0.0, 0.0, 0.0], radius=2.5, dip=dip, dip_pos=dip_pos, field_points=fp,
                              errortol=1e-8, precision="double")
    I = nf.calculate()
    ref = O.near_field_intensity(box, [0.0, 0, 0], 2.5, dip, dip_pos, fp, errortol=1e-8)
    assert rel(I, ref) < 1e-10
    with pytest.raises(RuntimeError, match="must set"):
        cumpm_amd.Near_Field(box, [1.0, 0, 0], dip=dip, dip_pos=dip_pos).calculate()
    with pytest.raises(RuntimeError, match="inconsistent"):
        cumpm_amd.Near_Field(box, [1.0, 0, 0], dip=dip, dip_pos=dip_pos[:5], field_points=fp).calculate()
    with pytest.raises(ValueError):
        cumpm_amd.Near_Field(box, [1.0, 0, 0], dip=dip, dip_pos=dip_pos, field_points=fp, precision="single")


@pytest.mark.parametrize("first_smallest", [True, False])
def test_near_field_polydisperse_matches_oracle(first_smallest):
    """Polydisperse Near_Field (near_field.cu:121-147; field points as sites of the reference's
    radii: particle f's radius for f < N, beyond it the first radius for the spread weight and the
    smallest for the pair functions) against the oracle at FP64 (1e-10) and the default FP32 grid
    (1e-5).  With the first radius not the smallest, the oracle covers the field points f < N."""
    import cumpm_amd
    _case(f"near field polydisperse first_smallest={first_smallest}")
    rng = np.random.default_rng(14)
    box = np.array([40.0, 41.0, 42.0])
    N = 10
    dip_pos = rng.uniform(0, 38, (N, 3))
    dip = rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))
    r = rng.uniform(1.0, 2.0, N)
    if first_smallest:
        r[0] = 0.9
    M = 14 if first_smallest else 8
    fp = rng.uniform(0, 38, (M, 3))
    fp[1] = dip_pos[4] + [0.6, 0.2, 0.1]   # inside a sphere
    fp[2] = dip_pos[5] + [2.2, 1.0, 0.0]   # near one
    nf = cumpm_amd.Near_Field(box, [0.3, 0.0, 0.1j], radius=r, dip=dip, dip_pos=dip_pos, field_points=fp,
                              errortol=1e-8, precision="double")
    I = nf.calculate()
    ref = O.near_field_intensity(box, [0.3, 0.0, 0.1j], r, dip, dip_pos, fp, errortol=1e-8)
    assert rel(I, ref) < 1e-10
    I32 = cumpm_amd.Near_Field(box, [0.3, 0.0, 0.1j], radius=r, dip=dip, dip_pos=dip_pos, field_points=fp,
                               errortol=1e-8).calculate()
    assert rel(I32, I) < 1e-5
    # forced polydisperse with equal radii: the monodisperse field to the Ewald tolerance at the
    # field points outside every sphere (inside one, the monodisperse tables clamp r < 1 to r = 1
    # where the polydisperse ones interpolate from the r -> 0 limit)
    Ip = cumpm_amd.Near_Field(box, [0.3, 0.0, 0.1j], radius=np.ones(N), dip=dip, dip_pos=dip_pos, field_points=fp,
                              errortol=1e-8, field_type="polydisperse", precision="double").calculate()
    Im = cumpm_amd.Near_Field(box, [0.3, 0.0, 0.1j], radius=1.0, dip=dip, dip_pos=dip_pos, field_points=fp,
                              errortol=1e-8, precision="double").calculate()
    dmin = np.array([np.min(np.linalg.norm(dip_pos - f, axis=1)) for f in fp])
    out = dmin > 2.0
    assert out.sum() >= M - 2
    assert rel(Ip[out], Im[out]) < 1e-5
    if not first_smallest:
        # two radii for the field points beyond the N-th: runs, finite
        more = cumpm_amd.Near_Field(box, [0.3, 0.0, 0.1j], radius=r, dip=dip, dip_pos=dip_pos,
                                    field_points=np.concatenate([fp, fp + 0.5]), errortol=1e-8).calculate()
        assert more.shape == (2 * M,) and np.all(np.isfinite(more))
        np.testing.assert_allclose(more[:M], I32, rtol=1e-5)


@pytest.mark.parametrize("bloch", [(0.0, 0.0), (0.12, -0.2)])
def test_adjoint_solve_with_quadrupoles(bloch):
    """With quadrupoles the field operator is not Hermitian in the (xx, xy, xz, yy, yz)
    components; M A is, for M = diag(I, W/2) with W the Frobenius product of traceless
    tensors, so mpm_solve_adjoint solves through M: checked by <A x, y> = <x, g> for random
    x, with and without a Bloch wave vector."""
    L, pos = _quad_geom()
    _case(f"adjoint quads bloch={bloch}")
    rng = np.random.default_rng(23)
    op = Op(L, 1e-8, 0.5, quads=[0, 3, 7])
    op.set_positions(pos)
    rows = np.stack([O.self_coefficient(np.full(10, e)) for e in (-2.2 + 0.3j, 3.0 + 0.2j)])
    op.set_self_coef(rows)
    op.set_bloch(*bloch)
    n = op.n
    g = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    y, reps = op.solve_adjoint(g, tol=1e-11, restart=60)
    assert all(r["converged"] for r in reps), reps
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    ax = op.apply(x)
    for q in range(2):
        lhs = np.vdot(ax[q], y[q])
        rhs = np.vdot(x[q], g[q])
        assert abs(lhs - rhs) < 1e-8 * abs(rhs), (q, lhs, rhs)


def test_memory_stats_track_operator_allocations():
    """mpm_memory_stats: an operator's grids and tables show up in the current bytes, the
    high-water mark covers a solve's workspace, and destroying the operator returns them."""
    _case("memory stats")
    base, _ = C.memory_stats(reset_peak=True)
    box, tol, xi, pos, grid, P = _geom("random3d")
    op = Op(box, tol, xi)
    op.set_positions(pos)
    op.set_self_coef(O.self_coefficient(np.full(len(pos), -2.0 + 0.5j)))
    held, _ = C.memory_stats()
    assert held > base
    b = np.zeros((2, len(pos), 3), complex)
    b[:, :, 0] = 1.0
    op.solve(b, tol=1e-8)
    cur, peak = C.memory_stats()
    assert peak >= cur >= held
    del op
    import gc
    gc.collect()
    assert C.memory_stats()[0] <= base + (cur - held)
