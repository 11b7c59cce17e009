"""The operator-level dev API (cumpm_amd.dev) on the GPU: the reference's pipeline stages,
INTERACTION_FIELD mode, the direct field class and the device addresses.

Reference: src/cuMPM/dev/electric_field.py, src/cpp/electric_field_bindings.cpp:142-278,
tests/test_electric_fields.py (direct field :5-28, precision modes :287-327).
"""
import ctypes

import numpy as np
import pytest

from cumpm_amd import _capi as C
from cumpm_amd.dev import (DirectElectricField, FieldCalcMode, MonodisperseEwaldElectricField,
                           PolydisperseEwaldElectricField)
from cumpm_amd.solver import PrecisionMode
from gpu_helpers import hex_monolayer, rel
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _load(ef, pos, rng, sc=None):
    ef.update_particle_coordinates(pos[:, 0], pos[:, 1], pos[:, 2])
    N = pos.shape[0]
    p = rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))
    ef.update_dipoles_complex(p[:, 0].real, p[:, 0].imag, p[:, 1].real, p[:, 1].imag, p[:, 2].real, p[:, 2].imag)
    if sc is None:
        sc = -0.3 + 0.05j + 0.01 * rng.standard_normal(N)
    ef.set_self_coef(np.real(sc), np.imag(sc))
    return p


def _cases():
    rng = np.random.default_rng(2)
    L = 20.0
    rand = rng.uniform(0, L, (30, 3))
    slab, box = hex_monolayer(10, 8, 2.2, jitter=0.05, seed=3)
    slab[:, 2] += 10.0
    return [("random", MonodisperseEwaldElectricField, (L, L, L), rand, {}),
            ("slab", MonodisperseEwaldElectricField, (box[0], box[1], 20.0), slab, {}),
            ("quads", MonodisperseEwaldElectricField, (L, L, L), rand, dict(solve_quadrupoles=True, quad_idxs=[0, 3])),
            ("poly", PolydisperseEwaldElectricField, (L, L, L), rand, dict(particle_radii=rng.uniform(1.0, 1.4, 30)))]


@pytest.mark.parametrize("case", [0, 1, 2, 3])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_pipeline_stages_add_up_to_calculate(case, precision):
    name, cls, box, pos, kw = _cases()[case]
    print(f"dev pipeline {name} {precision}", flush=True)
    rng = np.random.default_rng(7)
    ef = cls(*box, 1e-6, 0.5, True, precision=precision, **kw)
    _load(ef, pos, rng)
    if ef.num_quads:
        q = rng.standard_normal((ef.num_quads, 5))
        ef.update_quadrupoles(*q.T)
    ef.calculate()
    total = ef.get_epoint_host(reshape=False)
    # the reference's sequence: clear, real space, spread, scale, contract (default buffers)
    ef.clear_epoint()
    ef.real_space()
    ef.spread()
    ef.scale()
    ef.contract()
    staged = ef.get_epoint_host(reshape=False)
    assert rel(staged, total) < (1e-12 if precision == "double" else 1e-6)
    # caller buffers: the grid and the field through device pointers
    nb = ctypes.c_size_t()
    C.check(C.load().mpm_op_grid_bytes(ef._h, int(ef._prec), ctypes.byref(nb)))
    grid = C.DeviceBuffer(nb.value)
    E = C.DeviceBuffer(total.nbytes)
    E.upload(np.zeros_like(total))
    ef.spread(grid.ptr.value)
    ef.scale(grid.ptr.value)
    ef.contract(E.ptr.value, grid.ptr.value)
    ef.real_space(E.ptr.value)
    out = np.zeros_like(total)
    E.download(out)
    assert rel(out, total) < (1e-12 if precision == "double" else 1e-6)
    # contract alone is the inverse field; real space alone the real field
    inv = ef.compute_inverse_field()
    real = ef.compute_real_field()
    if isinstance(inv, tuple):
        inv = np.concatenate([inv[0].reshape(-1), inv[1].reshape(-1)])
        real = np.concatenate([real[0].reshape(-1), real[1].reshape(-1)])
    E.upload(np.zeros_like(total))
    ef.spread(grid.ptr.value)
    ef.scale(grid.ptr.value)
    ef.contract(E.ptr.value, grid.ptr.value)
    E.download(out)
    assert rel(out, inv.reshape(-1)) < (1e-11 if precision == "double" else 1e-5)


def test_device_addresses_hold_the_operator_data():
    rng = np.random.default_rng(3)
    L = 20.0
    pos = rng.uniform(0, L, (25, 3))
    ef = MonodisperseEwaldElectricField(L, L, L, 1e-6, 0.5, True, precision="double")
    p = _load(ef, pos, rng)
    ef.calculate()
    x = np.zeros(25)
    for d, addr in enumerate((ef.dev_x_part, ef.dev_y_part, ef.dev_z_part)):
        assert addr
        C.check(C.load().mpm_buffer_download(x.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(addr), x.nbytes))
        np.testing.assert_array_equal(x, pos[:, d])
    dip = np.zeros((25, 3), np.complex128)
    C.check(C.load().mpm_buffer_download(dip.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ef.dev_dipoles),
                                         dip.nbytes))
    np.testing.assert_array_equal(dip, p)
    E = np.zeros((25, 3), np.complex128)
    C.check(C.load().mpm_buffer_download(E.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ef.dev_epoint),
                                         E.nbytes))
    np.testing.assert_array_equal(E, ef.get_epoint_host())
    sc = np.zeros(25, np.complex128)
    C.check(C.load().mpm_buffer_download(sc.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ef.dev_self_coef_real),
                                         sc.nbytes))
    np.testing.assert_array_equal(sc, ef._sc)
    assert ef.dev_self_coef_imag == ef.dev_self_coef_real + 8
    start, j, perp, para = ef._neighbors()
    M = j.size
    pp = np.zeros((M, 2))
    C.check(C.load().mpm_buffer_download(pp.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ef.dev_perp), pp.nbytes))
    nb = np.zeros(M, np.int32)
    C.check(C.load().mpm_buffer_download(nb.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ef.dev_neighbor_list),
                                         nb.nbytes))
    np.testing.assert_array_equal(nb, j)
    np.testing.assert_array_equal(pp[:, 0], perp)
    np.testing.assert_array_equal(pp[:, 1], para)
    for a in (ef.dev_spread_coef, ef.dev_spread_idxs, ef.dev_particle_index, ef.dev_neighbor_counts,
              ef.dev_scale_coef):
        assert a
    with pytest.raises(ValueError):
        ef._dev("no_such_buffer")


def test_interaction_field_mode_matches_oracle():
    # field points join the operator as sites without dipoles: E = -(rows of A [p; 0])
    rng = np.random.default_rng(4)
    L = 20.0
    pos = rng.uniform(0, L, (20, 3))
    fpts = rng.uniform(0, L, (7, 3))
    fpts[0] = pos[3] + [0.4, 0.2, 0.0]  # inside a sphere
    ef = MonodisperseEwaldElectricField(L, L, L, 1e-8, 0.5, False, precision="double")
    assert ef.mode == FieldCalcMode.INTERACTION_FIELD and not ef.calc_inter_dipole
    ef.update_particle_coordinates(pos[:, 0], pos[:, 1], pos[:, 2])
    ef.update_field_coordinates(fpts[:, 0], fpts[:, 1], fpts[:, 2])
    p = rng.standard_normal((20, 3)) + 1j * rng.standard_normal((20, 3))
    ef.set_dipoles(p)
    ef.calculate()
    E = ef.get_epoint_host()
    assert E.shape == (7, 3) and ef.num_field_points == 7 and ef.num_particles == 20
    op = O.Operator([L, L, L], 1e-8, 0.5)
    op.set_positions(np.concatenate([pos, fpts]))
    want = -op.apply(np.concatenate([p, np.zeros((7, 3))]), 0.0)[20:]
    assert rel(E, want) < 1e-11
    # the direct field in the same mode (reference tests/test_electric_fields.py:5-28)
    radius = np.array([1.0, 1.2, 1.5])
    d = DirectElectricField(radius, mode=False, precision="double")
    d.update_particle_coordinates(np.array([0.0, 2.0, 4.0]), np.zeros(3), np.zeros(3))
    d.update_field_coordinates([1.0, 3.0], [0.0, 0.0], [0.0, 0.0])
    d.update_dipoles_complex(np.ones(3), np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3))
    d.set_self_coef(np.ones(3), np.zeros(3))
    assert d.num_particles == 3 and d.num_field_points == 2
    d.calculate()
    ep = d.get_epoint_host()
    assert ep.shape == (2, 3) and not np.any(np.isnan(ep))
    D = O.DirectOperator(np.r_[radius, 1.0, 1.0])
    D.set_positions(np.array([[0, 0, 0], [2, 0, 0], [4, 0, 0], [1, 0, 0], [3, 0, 0]], float))
    x = np.zeros(15, complex)
    x[0:9:3] = 1.0
    np.testing.assert_allclose(ep, -D.apply(x, np.zeros(5))[9:].reshape(2, 3), rtol=1e-13, atol=1e-15)


def test_direct_field_class_solver_mode_and_precisions():
    # reference tests/test_electric_fields.py:287-327 and the solver mode against the oracle
    radius = np.array([1.0, 1.2, 1.5])
    e64 = DirectElectricField(radius, mode=False, precision=PrecisionMode.FP64)
    assert not e64.is_using_fp32 and e64.precision_mode == PrecisionMode.FP64
    e32 = DirectElectricField(radius, mode=False, precision=PrecisionMode.FP32)
    assert e32.is_using_fp32 and e32.precision_mode == PrecisionMode.FP32
    out = []
    for ef in (e64, e32):
        ef.update_particle_coordinates(np.array([0.0, 2.0, 4.0]), np.zeros(3), np.zeros(3))
        ef.update_field_coordinates([1.0, 3.0], [0.0, 0.0], [0.0, 0.0])
        ef.update_dipoles_complex(np.ones(3), np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3))
        ef.set_self_coef(np.ones(3), np.zeros(3))
        ef.calculate()
        out.append(ef.get_epoint_host())
    np.testing.assert_allclose(out[1], out[0], rtol=1e-5, atol=1e-5)
    rng = np.random.default_rng(8)
    pos = rng.uniform(0, 8, (16, 3))
    rad = rng.uniform(0.8, 1.2, 16)
    s = DirectElectricField(rad, mode=True, solve_quadrupoles=True, quad_idxs=[1, 5], precision="double")
    p = _load(s, pos, rng, sc=np.full(16, -0.2 + 0.1j))
    q = rng.standard_normal((2, 5)) + 1j * rng.standard_normal((2, 5))
    s.update_quadrupoles_complex(*[v for c in q.T for v in (c.real, c.imag)])
    s.calculate()
    E, G = s.get_epoint_host()
    D = O.DirectOperator(rad, [1, 5])
    D.set_positions(pos)
    y = D.apply(np.concatenate([p.reshape(-1), q.reshape(-1)]), np.full(16, -0.2 + 0.1j))
    assert rel(E, y[:48].reshape(16, 3)) < 1e-12
    assert rel(G, y[48:].reshape(2, 5)) < 1e-12
    real = s.compute_real_field()
    assert rel(real[0], E) < 1e-15
