"""The autotuner (mpm_op_autotune, dipole_solver(autotune=...)) on the GPU: a tuned operator
applies the same operator as the untuned one (layout and FFT back-end are exact
re-arrangements), a grown grid equals an operator built on that grid (and the oracle on it),
the process cache answers the second call, and the driver's results do not depend on the
choice.  No counterpart in the reference (its pipeline is fixed)."""
import ctypes

import numpy as np
import pytest

import cumpm_amd
from cumpm_amd import _capi as C
from gpu_helpers import Op, hex_monolayer, rel
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _tune(op, B, precision, flags=0):
    info = C.TuneInfo()
    C.check(C.load().mpm_op_autotune(op.h, int(B), int(precision), int(flags), ctypes.byref(info)))
    return info.as_dict()


def _rhs(rng, B, n):
    return rng.standard_normal((B, n)) + 1j * rng.standard_normal((B, n))


def test_tuned_random_box_equals_untuned_and_caches():
    print("autotune random box", flush=True)
    rng = np.random.default_rng(1)
    L = 20.0
    pos = rng.uniform(0, L, (300, 3))
    op = Op([L, L, L], 1e-8, 0.5)
    op.set_positions(pos)
    op.set_self_coef(np.full(300, -0.3 + 0.05j))
    x = _rhs(rng, 4, op.n)
    y0 = op.apply(x)
    t = _tune(op, 4, C.PREC_DOUBLE)
    assert t["candidates"] >= 2 and not t["cached"]
    assert 0 < t["best_ms"] <= t["default_ms"]
    assert t["grid"] == list(op.info().grid)
    assert op.info().fft_backend == t["fft_backend"]
    assert rel(op.apply(x), y0) < 1e-12
    # the same shape again: the cache answers, same choice; FORCE times again
    op2 = Op([L, L, L], 1e-8, 0.5)
    op2.set_positions(pos + 0.01)
    op2.set_self_coef(np.full(300, -0.3 + 0.05j))
    t2 = _tune(op2, 3, C.PREC_DOUBLE)  # batch 3 falls in the same power-of-two bucket as 4
    assert t2["cached"] and (t2["layout"], t2["fft_backend"]) == (t["layout"], t["fft_backend"])
    t3 = _tune(op2, 4, C.PREC_DOUBLE, C.TUNE_FORCE)
    assert not t3["cached"]
    op3 = Op([L, L, L], 1e-8, 0.5, layout=1, fft_backend=0)
    op3.set_positions(pos + 0.01)
    op3.set_self_coef(np.full(300, -0.3 + 0.05j))
    assert rel(op2.apply(x), op3.apply(x)) < 1e-12


def test_tuned_slab_mixed_and_bloch_equal_untuned():
    # the monolayer has both layouts; the Bloch phase survives reconfiguration
    print("autotune slab mixed bloch", flush=True)
    pos, box = hex_monolayer(20, 16, 2.2, jitter=0.05, seed=3)
    pos[:, 2] += 12.0
    box = [box[0], box[1], 24.0]
    rng = np.random.default_rng(2)
    op = Op(box, 1e-6, 0.5)
    op.set_positions(pos)
    assert op.info().slab_planes > 0
    op.set_bloch(0.03, -0.02)
    N = pos.shape[0]
    op.set_self_coef(np.full(N, -0.3 + 0.05j))
    x = _rhs(rng, 8, op.n)
    y64 = op.apply(x)
    y32 = op.apply(x, precision=C.PREC_MIXED)
    t = _tune(op, 8, C.PREC_MIXED)
    assert t["candidates"] >= 4
    assert list(op.info().bloch) == [0.03, -0.02]
    assert rel(op.apply(x), y64) < 1e-12
    assert rel(op.apply(x, precision=C.PREC_MIXED), y64) < max(4 * rel(y32, y64), 1e-6)


def test_grid_growth_equals_operator_on_grown_grid_and_oracle():
    # grid 43 (prime) grows to 45 = 3^2 5 under MPM_TUNE_GRID
    print("autotune grid growth", flush=True)
    rng = np.random.default_rng(3)
    L = 20.0
    pos = rng.uniform(0, L, (120, 3))
    op = Op([L, L, L], 1e-8, 0.5, grid=(43, 43, 43))
    op.set_positions(pos)
    op.set_self_coef(np.full(120, -0.3 + 0.05j))
    P = op.info().support
    t = _tune(op, 2, C.PREC_DOUBLE, C.TUNE_GRID)
    assert t["candidates"] >= 8
    g = t["grid"]
    assert g in ([43] * 3, [45] * 3)
    x = _rhs(rng, 2, op.n)
    y = op.apply(x)
    ref = Op([L, L, L], 1e-8, 0.5, grid=g, support=P)
    ref.set_positions(pos)
    ref.set_self_coef(np.full(120, -0.3 + 0.05j))
    assert rel(y, ref.apply(x)) < 1e-12
    # the grown operator (whether or not chosen) against the oracle on 45^3
    grown = Op([L, L, L], 1e-8, 0.5, grid=(45, 45, 45), support=P)
    grown.set_positions(pos)
    grown.set_self_coef(np.full(120, -0.3 + 0.05j))
    o = O.Operator([L, L, L], 1e-8, 0.5, grid=[45, 45, 45], P=P)
    o.set_positions(pos)
    want = o.apply(x[0].reshape(120, 3), -0.3 + 0.05j)
    assert rel(grown.apply(x[:1])[0], want) < 1e-11


def test_quadrupole_and_poly_operators_have_one_configuration():
    print("autotune quads poly", flush=True)
    rng = np.random.default_rng(4)
    L = 20.0
    pos = rng.uniform(0, L, (40, 3))
    q = Op([L, L, L], 1e-6, 0.5, quads=[0, 5, 7])
    q.set_positions(pos)
    q.set_self_coef(np.full(40, -0.3 + 0.05j))
    x = _rhs(rng, 2, q.n)
    y0 = q.apply(x)
    t = _tune(q, 2, C.PREC_DOUBLE)
    assert t["candidates"] == 1 and t["best_ms"] == t["default_ms"]
    assert rel(q.apply(x), y0) < 1e-13
    p = Op([L, L, L], 1e-6, 0.5, radii=rng.uniform(1.0, 1.4, 40))
    p.set_positions(pos)
    p.set_self_coef(np.full(40, -0.3 + 0.05j))
    x = _rhs(rng, 2, p.n)
    y0 = p.apply(x)
    assert _tune(p, 2, C.PREC_DOUBLE)["candidates"] == 1
    assert rel(p.apply(x), y0) < 1e-13


def test_quadrupole_grid_growth_equals_operator_on_grown_grid_and_oracle():
    # quadrupoles: the grown grid is their only other candidate (full grid, built-in FFT)
    print("autotune quads grid growth", flush=True)
    rng = np.random.default_rng(6)
    L = 20.0
    pos = rng.uniform(0, L, (30, 3))
    q = Op([L, L, L], 1e-8, 0.5, grid=(43, 43, 43), quads=[0, 5, 7])
    q.set_positions(pos)
    rows = rng.standard_normal((2, 30)) + 1j * rng.standard_normal((2, 30))
    q.set_self_coef(rows)
    P = q.info().support
    t = _tune(q, 2, C.PREC_DOUBLE, C.TUNE_GRID)
    assert t["candidates"] == 2 and (t["layout"], t["fft_backend"]) == (1, 0)
    g = t["grid"]
    assert g in ([43] * 3, [45] * 3)
    x = _rhs(rng, 2, q.n)
    y = q.apply(x)
    ref = Op([L, L, L], 1e-8, 0.5, grid=g, support=P, quads=[0, 5, 7])
    ref.set_positions(pos)
    ref.set_self_coef(rows)
    assert rel(y, ref.apply(x)) < 1e-12
    o = O.Operator([L, L, L], 1e-8, 0.5, grid=g, P=P)
    o.set_positions(pos)
    o.set_quads([0, 5, 7])
    assert rel(y[0], o.apply(x[0], rows[0])) < 1e-11


def test_dipole_solver_autotune_same_results():
    print("autotune dipole_solver", flush=True)
    pos, box = hex_monolayer(16, 12, 2.2, jitter=0.05, seed=5)
    pos[:, 2] += 10.0
    box = [box[0], box[1], 20.0]
    eps = cumpm_amd.drude_dielectric(np.linspace(0.3, 0.7, 4), 0.08, 1.7, 4.0)
    out = []
    for tune in (False, True, "grid"):
        s = cumpm_amd.dipole_solver(box, eps, radius=1.0, tol=1e-8, quiet=True, field_type="monodisperse",
                                    batched=True, precision="double", autotune=tune)
        s.compute(pos)
        out.append(s.get_dipoles())
        info = s.get_tune_info()
        if tune:
            assert info["candidates"] >= 2 and info["best_ms"] > 0
            assert list(s.get_op_info().grid) == info["grid"]
        else:
            assert info["candidates"] == 0
    assert rel(out[1], out[0]) < 1e-7
    assert rel(out[2], out[0]) < 1e-5  # a grown grid is another discretisation at Ewald tol 1e-8
