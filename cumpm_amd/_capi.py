"""ctypes bindings of libcumpm_amd.so (declarations in include/cumpm_amd.h).

The library is loaded from ``cumpm_amd/lib/`` (built in-tree by
``cumpm_amd/build.py``).  There is no CPU fallback: if the library is missing
or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcumpm_amd.so")

MPM_OK = 0
MPM_ERR_INVALID_ARGUMENT = 1
MPM_ERR_RUNTIME = 2
MPM_ERR_HIP = 3
MPM_ERR_FFT = 4
MPM_ERR_UNSUPPORTED = 5
MPM_ERR_OUT_OF_MEMORY = 6

PREC_DOUBLE, PREC_MIXED, PREC_AUTO = 0, 1, 2
SOLVE_DOUBLE, SOLVE_MIXED, SOLVE_REFINED = 0, 1, 2
GMRES, BICGSTAB = 0, 1
PART_REAL, PART_RECIP, PART_ALL = 1, 2, 3
STAGE_SPREAD, STAGE_SCALE, STAGE_CONTRACT, STAGE_REAL_SPACE = 1, 2, 3, 4
TUNE_GRID, TUNE_FORCE = 1, 2
GUESS_DERIVATIVE, GUESS_MEAN_FIELD, GUESS_PREVIOUS = 0, 1, 2
FIELD_AUTO, FIELD_MONODISPERSE, FIELD_POLYDISPERSE, FIELD_DIRECT = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64


class OpDesc(ctypes.Structure):
    _fields_ = [("box", ctypes.c_double * 3), ("errortol", ctypes.c_double), ("xi", ctypes.c_double),
                ("grid", ctypes.c_int * 3), ("support", ctypes.c_int), ("layout", ctypes.c_int),
                ("fft_backend", ctypes.c_int)]


class OpInfo(ctypes.Structure):
    _fields_ = [("num_particles", _i64), ("num_pairs", _i64), ("grid", ctypes.c_int * 3),
                ("support", ctypes.c_int), ("stencil_width", ctypes.c_int),
                ("grid_spacing", ctypes.c_double * 3), ("spectral_split", ctypes.c_double * 3),
                ("rc", ctypes.c_double), ("self_perp", ctypes.c_double),
                ("self_coef_ewald", ctypes.c_double), ("bloch", ctypes.c_double * 2),
                ("slab_planes", ctypes.c_int), ("fft_backend", ctypes.c_int),
                ("polydisperse", ctypes.c_int), ("num_unique_radii", ctypes.c_int), ("num_quads", _i64)]


class SolveDesc(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("restart", ctypes.c_int), ("maxiter", ctypes.c_int),
                ("tol", ctypes.c_double), ("jacobi", ctypes.c_int), ("precision", ctypes.c_int)]


class SolveReport(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("matvecs", ctypes.c_int), ("matvecs_fp64", ctypes.c_int),
                ("converged", ctypes.c_int), ("relres_pre", ctypes.c_double),
                ("relres_true", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DipoleDesc(ctypes.Structure):
    _fields_ = [("box", ctypes.c_double * 3), ("eps_p", _dp), ("num_wavevectors", _i64),
                ("eps_cols", _i64), ("radius", _dp), ("num_radii", _i64), ("eps_m", ctypes.c_double),
                ("xi", ctypes.c_double), ("tol", ctypes.c_double), ("quiet", ctypes.c_int),
                ("guess", ctypes.c_int), ("method", ctypes.c_int), ("field_type", ctypes.c_int),
                ("E0", _dp), ("num_E0", _i64), ("E0_len", _i64), ("solve_quadrupoles", ctypes.c_int),
                ("quad_idxs", ctypes.POINTER(ctypes.c_int)), ("num_quad_idxs", _i64),
                ("precision", ctypes.c_int), ("refined", ctypes.c_int), ("batched", ctypes.c_int),
                ("restart", ctypes.c_int), ("maxiter", ctypes.c_int), ("grid", ctypes.c_int * 3),
                ("support", ctypes.c_int), ("ewald_tol", ctypes.c_double), ("autotune", ctypes.c_int)]


class TuneInfo(ctypes.Structure):
    _fields_ = [("layout", ctypes.c_int), ("fft_backend", ctypes.c_int), ("grid", ctypes.c_int * 3),
                ("best_ms", ctypes.c_double), ("default_ms", ctypes.c_double), ("candidates", ctypes.c_int),
                ("cached", ctypes.c_int)]

    def as_dict(self):
        return {"layout": self.layout, "fft_backend": self.fft_backend, "grid": list(self.grid),
                "best_ms": self.best_ms, "default_ms": self.default_ms, "candidates": self.candidates,
                "cached": bool(self.cached)}


class EelsDesc(ctypes.Structure):
    _fields_ = [("box", ctypes.c_double * 3), ("eps_p", _dp), ("num_wavevectors", _i64), ("eps_cols", _i64),
                ("omega", _dp), ("num_omega", _i64), ("v", ctypes.c_double), ("radius", _dp),
                ("num_radii", _i64), ("eps_m", ctypes.c_double), ("xi", ctypes.c_double),
                ("tol", ctypes.c_double), ("quiet", ctypes.c_int), ("guess", ctypes.c_int),
                ("method", ctypes.c_int), ("field_type", ctypes.c_int), ("integration_step", ctypes.c_double),
                ("solve_quadrupoles", ctypes.c_int), ("quad_idxs", ctypes.POINTER(ctypes.c_int)),
                ("num_quad_idxs", _i64), ("asm_flag", ctypes.c_int), ("asm_nx", ctypes.c_int),
                ("asm_ny", ctypes.c_int), ("precision", ctypes.c_int), ("refined", ctypes.c_int),
                ("batched", ctypes.c_int), ("restart", ctypes.c_int), ("maxiter", ctypes.c_int)]


# name -> (restype, argtypes)
_SIGS = {
    "mpm_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "mpm_last_error": (ctypes.c_char_p, []),
    "mpm_op_create": (ctypes.c_int, [ctypes.POINTER(OpDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "mpm_op_set_quads": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), _i64]),
    "mpm_op_create_poly": (ctypes.c_int, [ctypes.POINTER(OpDesc), _dp, _i64, ctypes.POINTER(ctypes.c_void_p)]),
    "mpm_op_create_direct": (ctypes.c_int, [_dp, _i64, ctypes.POINTER(ctypes.c_void_p)]),
    "mpm_op_destroy": (None, [ctypes.c_void_p]),
    "mpm_op_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OpInfo)]),
    "mpm_op_get_pair_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _i64, ctypes.POINTER(_i64)]),
    "mpm_op_get_neighbors": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i64), ctypes.POINTER(ctypes.c_int),
                                            _dp, _dp, _i64, ctypes.POINTER(_i64)]),
    "mpm_op_set_positions": (ctypes.c_int, [ctypes.c_void_p, _dp, _i64]),
    "mpm_op_set_bloch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "mpm_op_set_self_coef": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int]),
    "mpm_op_apply": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "mpm_op_grid_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]),
    "mpm_op_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_void_p]),
    "mpm_op_device_pointer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p),
                                             ctypes.POINTER(ctypes.c_size_t)]),
    "mpm_op_apply_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "mpm_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolveDesc), ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(SolveReport),
                                 ctypes.c_void_p]),
    "mpm_solve_adjoint": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolveDesc), ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(SolveReport),
                                         ctypes.c_void_p]),
    "mpm_near_field_create": (ctypes.c_int, [_dp, _dp, _dp, _i64, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "mpm_near_field_destroy": (None, [ctypes.c_void_p]),
    "mpm_near_field_set_dipoles": (ctypes.c_int, [ctypes.c_void_p, _dp, _i64]),
    "mpm_near_field_set_dipole_positions": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _i64]),
    "mpm_near_field_set_field_points": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _i64]),
    "mpm_near_field_calculate": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "mpm_near_field_set_precision": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "mpm_memory_stats": (ctypes.c_int, [ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.c_int]),
    "mpm_dipole_solver_create": (ctypes.c_int, [ctypes.POINTER(DipoleDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "mpm_dipole_solver_destroy": (None, [ctypes.c_void_p]),
    "mpm_dipole_solver_compute": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _i64]),
    "mpm_dipole_solver_get_dims": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i64)]),
    "mpm_dipole_solver_get_eff_polarizability": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "mpm_dipole_solver_get_dipoles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "mpm_dipole_solver_get_num_quads": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i64)]),
    "mpm_dipole_solver_get_quadrupoles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "mpm_dipole_solver_get_reports": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolveReport)]),
    "mpm_dipole_solver_get_op_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OpInfo)]),
    "mpm_dipole_solver_set_use_jacobi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "mpm_dipole_solver_get_use_jacobi": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "mpm_dipole_solver_get_timing": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "mpm_dipole_solver_get_tune_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TuneInfo)]),
    "mpm_op_radius_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                              ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "mpm_op_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(TuneInfo)]),
    "mpm_dipole_solver_reset": (ctypes.c_int, [ctypes.c_void_p]),
    "mpm_ebeam_field": (ctypes.c_int, [_dp, _i64, ctypes.POINTER(ctypes.c_int), _i64, _dp, _i64, ctypes.c_double,
                                       _dp, ctypes.c_double, _dp]),
    "mpm_op_set_self_coef_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "mpm_eels_solver_create": (ctypes.c_int, [ctypes.POINTER(EelsDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "mpm_eels_solver_destroy": (None, [ctypes.c_void_p]),
    "mpm_eels_solver_compute": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp, _i64]),
    "mpm_eels_solver_get_dims": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(_i64)]),
    "mpm_eels_solver_get_eels": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "mpm_eels_solver_get_dipoles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "mpm_eels_solver_get_quadrupoles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp]),
    "mpm_eels_solver_get_incident_field": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "mpm_eels_solver_get_reports": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _i64, ctypes.POINTER(_i64)]),
    "mpm_eels_solver_set_use_jacobi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "mpm_eels_solver_get_use_jacobi": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "mpm_eels_solver_get_timing": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp]),
    "mpm_buffer_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "mpm_buffer_free": (ctypes.c_int, [ctypes.c_void_p]),
    "mpm_buffer_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "mpm_buffer_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "mpm_device_synchronize": (ctypes.c_int, []),
    "mpm_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "mpm_profile_reset": (ctypes.c_int, []),
    "mpm_profile_get": (ctypes.c_int, [ctypes.c_char_p, _dp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "mpm_profile_names": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "mpm_profile_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
}

EXPORTED = tuple(_SIGS)

_lib = None


class MPMError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `python -m cumpm_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    """Map an mpm_status to the exception the reference raises for it."""
    if status == MPM_OK:
        return
    L = load()
    msg = (L.mpm_last_error() or b"").decode() or L.mpm_status_string(status).decode()
    if status == MPM_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == MPM_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if status == MPM_ERR_RUNTIME:
        raise RuntimeError(msg)
    raise MPMError(status, f"{L.mpm_status_string(status).decode()}: {msg}")


def dptr(a):
    return a.ctypes.data_as(_dp)


class DeviceBuffer:
    """Device allocation owned by the library (plumbing for numpy callers)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(load().mpm_buffer_alloc(self.nbytes, ctypes.byref(self.ptr)))

    def upload(self, arr):
        import numpy as np
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        check(load().mpm_buffer_upload(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))

    def download(self, out):
        assert out.nbytes <= self.nbytes and out.flags.c_contiguous
        check(load().mpm_buffer_download(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr and _lib is not None:
            _lib.mpm_buffer_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        self.free()


def profile_enable(on=True):
    check(load().mpm_profile_enable(int(bool(on))))


def profile_reset():
    check(load().mpm_profile_reset())


def profile_get(name):
    """(total ms, launches, right-hand sides summed over the launches) of a class or kernel."""
    ms = ctypes.c_double()
    cnt = _i64()
    units = _i64()
    check(load().mpm_profile_get(name.encode(), ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(units)))
    return ms.value, cnt.value, units.value


def profile_names():
    """Names (classes and kernels) with recorded launches."""
    need = ctypes.c_size_t()
    check(load().mpm_profile_names(None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(need.value + 64)
    check(load().mpm_profile_names(buf, len(buf), ctypes.byref(need)))
    return [n for n in buf.value.decode().split("\n") if n]


def memory_stats(reset_peak=False):
    """(current, peak) device bytes held by the library's own allocations."""
    cur, peak = _i64(), _i64()
    check(load().mpm_memory_stats(ctypes.byref(cur), ctypes.byref(peak), int(bool(reset_peak))))
    return int(cur.value), int(peak.value)
