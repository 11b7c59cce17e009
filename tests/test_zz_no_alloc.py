"""mpm_op_apply_noalloc (grids in the library's static scratch array) against mpm_op_apply.

Both entry points run the same kernels; the allocation-free one takes its reciprocal grids
from a file-scope device array instead of operator buffers.  Two calls in a row, the second
on a smaller operator with fewer right-hand sides, must each reproduce mpm_op_apply.
"""
import ctypes

import numpy as np
import pytest

from cumpm_amd import _capi as C
from gpu_helpers import Op, hex_monolayer, rel

pytestmark = pytest.mark.gpu


def _apply(op, p, precision, noalloc):
    p = np.ascontiguousarray(p, np.complex128)
    B = p.shape[0]
    xb = C.DeviceBuffer(p.nbytes)
    yb = C.DeviceBuffer(p.nbytes)
    xb.upload(p)
    yb.upload(np.full_like(p, np.nan))  # every output element must be written
    fn = C.load().mpm_op_apply_noalloc if noalloc else C.load().mpm_op_apply
    C.check(fn(op.h, xb.ptr, yb.ptr, B, int(precision), C.PART_ALL, None))
    out = np.empty_like(p)
    yb.download(out)
    return out


def _case(op, pos, B, seed):
    rng = np.random.default_rng(seed)
    op.set_positions(pos)
    N = pos.shape[0]
    eps = -2.0 + 0.3j + 0.1 * rng.standard_normal(B)
    rows = (-3.0 / (4.0 * np.pi * (1.0 - eps)))[:, None] * np.ones((1, N))
    op.set_self_coef(rows)
    return rng.standard_normal((B, N, 3)) + 1j * rng.standard_normal((B, N, 3))


@pytest.mark.parametrize("precision", [C.PREC_MIXED, C.PREC_DOUBLE])
def test_noalloc_apply_matches_apply_twice(precision):
    # first call: a jittered hex monolayer (slab layout, 4 right-hand sides)
    pos, box = hex_monolayer(16, 12, 2.2, jitter=0.04, seed=1)
    box[2] = 20.0
    big = Op(box, 1e-6, 0.5)
    p = _case(big, pos, 4, 2)
    print(f"noalloc: hex 16x12 slab, B=4, precision={precision}", flush=True)
    ref = _apply(big, p, precision, noalloc=False)
    got = _apply(big, p, precision, noalloc=True)
    assert np.all(np.isfinite(got))
    assert rel(got, ref) < 1e-12, rel(got, ref)
    # second call: a smaller 3-D cluster (full grid) with one right-hand side; the scratch
    # still holds the first call's grids
    rng = np.random.default_rng(3)
    L = 18.0
    pos2 = rng.uniform(0, L, (40, 3))
    small = Op([L, L, L], 1e-6, 0.5, layout=1)
    p2 = _case(small, pos2, 1, 4)
    print(f"noalloc: random 40 full grid, B=1, precision={precision}", flush=True)
    ref2 = _apply(small, p2, precision, noalloc=False)
    got2 = _apply(small, p2, precision, noalloc=True)
    assert np.all(np.isfinite(got2))
    assert rel(got2, ref2) < 1e-12, rel(got2, ref2)
    # and once more on the first operator after the smaller call
    got3 = _apply(big, p, precision, noalloc=True)
    assert rel(got3, ref) < 1e-12


def test_noalloc_apply_rejects_oversized_grid():
    rng = np.random.default_rng(5)
    L = 40.0
    op = Op([L, L, L], 1e-6, 0.5, grid=[128, 128, 128], layout=1)
    p = _case(op, rng.uniform(0, L, (20, 3)), 1, 6)
    xb = C.DeviceBuffer(p.nbytes)
    yb = C.DeviceBuffer(p.nbytes)
    xb.upload(p)
    # 3 x 128^3 complex doubles = 96 MiB of grid: more than the 64 MiB scratch
    st = C.load().mpm_op_apply_noalloc(op.h, xb.ptr, yb.ptr, 1, C.PREC_DOUBLE, C.PART_ALL, None)
    assert st == C.MPM_ERR_INVALID_ARGUMENT
    assert b"scratch" in C.load().mpm_last_error()
    with pytest.raises(ValueError):
        C.check(st)
