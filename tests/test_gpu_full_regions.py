"""The full grid's matrix-core path on the GPU: the region spread and contraction
(spread_mfma_kernel, contract_full_mfma_kernel with even z lengths, contract_mfma_kernel on
the full-grid regions with odd ones) and the fused z pass (zline_scale_kernel, FP32)
against the oracle's operator on the same grid and support (FP64 to 1e-11, the FP32
reciprocal part to the reference's 1e-4), with the launched kernels checked through the
profiling counters; and the per-particle path (taken where the regions would give too few
waves) against the same oracle.  The operator is the reference's monodisperse one
(monodisperse_ewald_electric_field.cu)."""
import numpy as np
import pytest

from cumpm_amd import _capi as C
from oracle import oracle as O
from gpu_helpers import Op, rel

pytestmark = pytest.mark.gpu


def _launched(op, p, precision, parts):
    C.profile_reset()
    C.profile_enable(True)
    y = op.apply(p, precision, parts)
    C.profile_enable(False)
    return y, C.profile_names()


@pytest.mark.parametrize("grid", [(48, 40, 48), (40, 36, 45)])
def test_full_grid_regions_match_oracle(grid):
    # 300 particles on 48 x 40 x 48: 12 x 5 x 3 regions of 4 x 8 x 16 voxels, enough waves
    # for the region path at 4 right-hand sides; nz = 45 (odd) takes the 4-byte form
    print(f"full-grid regions grid={grid}", flush=True)
    rng = np.random.default_rng(17)
    L = [20.0, 18.0, 21.0]
    N = 300
    pos = rng.uniform(0, 1, (N, 3)) * L
    op = Op(L, 1e-8, 0.5, grid=grid, support=8, layout=1)
    op.set_positions(pos)
    info = op.info()
    assert list(info.grid) == list(grid) and info.slab_planes == 0
    sc = -0.3 + 0.05j
    op.set_self_coef(np.full(N, sc))
    B = 4
    p = rng.standard_normal((B, N, 3)) + 1j * rng.standard_normal((B, N, 3))
    y64, names = _launched(op, p, C.PREC_DOUBLE, C.PART_ALL)
    contract = "contract_full_mfma_kernel<double" if grid[2] % 2 == 0 else "contract_mfma_kernel<double"
    assert any(n.startswith(contract) for n in names), names
    assert any(n.startswith("spread_mfma_kernel<double") for n in names), names
    ref = O.Operator(L, 1e-8, 0.5, grid=list(grid), P=8)
    ref.set_positions(pos)
    for b in range(B):
        assert rel(y64[b], ref.apply(p[b], np.full(N, sc))) < 1e-11, b
    # FP32 reciprocal part: regions in FP32 and the fused z transforms + projection
    r32, names = _launched(op, p, C.PREC_MIXED, C.PART_RECIP)
    if grid[2] == 48:  # 48 = 16 x 3: two register stages, the fused z pass applies
        assert any(n.startswith("zline_scale_kernel<float") for n in names), names
    assert any(n.startswith(contract.replace("double", "float")) for n in names), names
    r64 = op.apply(p, C.PREC_DOUBLE, C.PART_RECIP)
    assert rel(r32, r64) < 1e-4
    assert rel(r64[0], ref.apply(p[0], np.full(N, sc), parts=2)) < 1e-11


def test_small_full_grid_takes_per_particle_path_and_matches_oracle():
    # 600 particles on 32^3 with one right-hand side: 8 x 4 x 2 regions give 64 waves,
    # fewer than a quarter of the 600 per-particle waves, so the spread and gather run
    # per particle
    print("full-grid per-particle path", flush=True)
    rng = np.random.default_rng(19)
    L = [24.0, 24.0, 24.0]
    N = 600
    pos = rng.uniform(0, 24.0, (N, 3))
    op = Op(L, 1e-6, 0.5, grid=(32, 32, 32), support=8, layout=1)
    op.set_positions(pos)
    op.set_self_coef(np.full(N, 0.2 - 0.1j))
    p = rng.standard_normal((1, N, 3)) + 1j * rng.standard_normal((1, N, 3))
    y, names = _launched(op, p, C.PREC_DOUBLE, C.PART_ALL)
    assert any(n.startswith("gather_kernel<double") for n in names), names
    assert not any(n.startswith("contract_full_mfma_kernel") for n in names), names
    ref = O.Operator(L, 1e-6, 0.5, grid=[32, 32, 32], P=8)
    ref.set_positions(pos)
    assert rel(y[0], ref.apply(p[0], np.full(N, 0.2 - 0.1j))) < 1e-11
