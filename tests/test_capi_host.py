"""CPU-side checks of the product boundary: the library loads, exports every
symbol include/cumpm_amd.h declares, the ctypes structs match the header, and
the Python API validates its arguments like the reference (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

import cumpm_amd
from cumpm_amd import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cumpm_amd.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mpm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.load()
    names = _declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(C.EXPORTED)


def test_status_strings():
    lib = C.load()
    for s in range(7):
        assert lib.mpm_status_string(s)


def _c_struct_sizes():
    """Compile a tiny C program against the header and print sizeof of each struct."""
    import subprocess
    import tempfile
    src = r'''
#include <stdio.h>
#include "cumpm_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(mpm_op_desc), sizeof(mpm_op_info), sizeof(mpm_solve_desc),
         sizeof(mpm_solve_report), sizeof(mpm_dipole_desc), sizeof(mpm_eels_desc), sizeof(mpm_tune_info));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    return [int(v) for v in out]


def test_ctypes_structs_match_header():
    sizes = _c_struct_sizes()
    py = [ctypes.sizeof(C.OpDesc), ctypes.sizeof(C.OpInfo), ctypes.sizeof(C.SolveDesc),
          ctypes.sizeof(C.SolveReport), ctypes.sizeof(C.DipoleDesc), ctypes.sizeof(C.EelsDesc),
          ctypes.sizeof(C.TuneInfo)]
    assert sizes == py


def test_drude_dielectric_matches_formula():
    omega = np.linspace(0.2, 0.8, 10)
    eps = cumpm_amd.drude_dielectric(omega, 0.08, 1.7, 4.0)
    np.testing.assert_allclose(eps, 4.0 - 1.7 ** 2 / (omega ** 2 + 1j * omega * 0.08), rtol=1e-15)
    eps2 = cumpm_amd.drude_dielectric(omega, [0.08, 0.1], [1.7, 1.8], [4.0, 3.0])
    assert eps2.shape == (10, 2)
    with pytest.raises(ValueError):
        cumpm_amd.drude_dielectric(omega, [0.08, 0.1], [1.7], [4.0, 3.0])


def test_argument_validation_like_reference():
    # reference tests/test_inputs.py and src/cuMPM/solver.py:93-183
    with pytest.raises(ValueError):
        cumpm_amd.dipole_solver([10.0, 10.0], -2.0)
    with pytest.raises(ValueError):
        cumpm_amd.dipole_solver([10.0] * 3, -2.0, field_type="spectral")
    with pytest.raises(ValueError):
        cumpm_amd.dipole_solver([10.0] * 3, -2.0, precision="half")
    with pytest.raises(ValueError):
        cumpm_amd.dipole_solver([10.0] * 3, -2.0, E0=[1.0, 0.0])
    with pytest.raises(TypeError):
        cumpm_amd.dipole_solver([10.0] * 3, -2.0, quadrupoles=3.5)
    with pytest.raises(ValueError, match="length 3"):
        cumpm_amd.Near_Field([10.0, 10.0], [1.0, 0, 0])
    with pytest.raises(ValueError, match="E0"):
        cumpm_amd.Near_Field([10.0] * 3, [1.0, 0])
    with pytest.raises(RuntimeError, match="field_type"):
        cumpm_amd.Near_Field([10.0] * 3, [1.0, 0, 0], field_type="spectral")
    with pytest.raises(ValueError, match="multiple of 3"):
        cumpm_amd.Near_Field([10.0] * 3, [1.0, 0, 0], dip=[1.0, 0.0])
    # reference tests/test_eels_solver.py:179-181 and src/cuMPM/eels.py:85-104
    with pytest.raises(ValueError, match="Both split_dist and N_split must be specified"):
        cumpm_amd.EELS(box=[50, 50, 50], eps_p=[4.0], omega=[0.1], v=0.3, split_dist=2.0)
    with pytest.raises(ValueError, match="split_method"):
        cumpm_amd.EELS([50.0] * 3, [4.0], [0.1], 0.3, split_method="hex")
    with pytest.raises(ValueError, match="field_type"):
        cumpm_amd.EELS([50.0] * 3, [4.0], [0.1], 0.3, field_type="spectral")
    with pytest.raises(ValueError, match="v must lie"):
        cumpm_amd.EELS([50.0] * 3, [4.0], [0.1], 1.2)
    e = cumpm_amd.EELS([50.0] * 3, [4.0 + 0.1j], [0.1], 0.3, quiet=True)
    assert e.use_jacobi_precond is True
    e.use_jacobi_precond = False
    assert e.use_jacobi_precond is False
    cumpm_amd.dipole_solver([10.0] * 3, -2.0, quadrupoles=True, quiet=True)
    cumpm_amd.dipole_solver([10.0] * 3, -2.0, quadrupoles=(0, 2), quiet=True)
    s = cumpm_amd.dipole_solver([10.0] * 3, [-2.0 + 0.1j, -1.0 + 0.1j], E0=[[1, 0, 0], [0, 0, 1]], quiet=True)
    assert s.use_jacobi_precond is True
    s.use_jacobi_precond = False
    assert s.use_jacobi_precond is False


def test_failed_setup_leaves_solver_retryable():
    # a box too small for the real-space cutoff fails in the operator constructor (before any
    # device work); the solver must stay unconfigured, so a retry fails the same way instead of
    # using a half-built state (radius 2: a second nondimensionalisation would change the box)
    s = cumpm_amd.dipole_solver([10.0] * 3, -2.0 + 0.1j, radius=2.0, tol=1e-6, quiet=True)
    pos = np.array([[1.0, 1.0, 1.0], [6.0, 6.0, 6.0]])
    with pytest.raises(RuntimeError, match="Real space cutoff") as e1:
        s.compute(pos)
    assert s._dims()[0] == 0 and s._dims()[2] == 0
    with pytest.raises(RuntimeError, match="Real space cutoff") as e2:
        s.compute(pos)
    assert str(e1.value) == str(e2.value)
    assert s._dims()[0] == 0 and s._dims()[2] == 0
