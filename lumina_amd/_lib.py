"""ctypes binding of liblumina_hip.so (the C ABI declared in include/lumina_hip.h).

The library is loaded from this package directory only. There is no
fallback: if the library is missing or cannot be loaded, every entry point
raises, so a caller can never silently run something other than the HIP
path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblumina_hip.so")

LH_OK = 0
LH_UNAVAILABLE = 1
LH_DEVICE_ERROR = 2
LH_OUT_OF_MEMORY = 3
LH_INVALID_ARGUMENT = 4
LH_CONVERGENCE_FAILURE = 5

LH_PREC_FP64 = 0
LH_PREC_MIXED = 1

# Every symbol include/lumina_hip.h declares.
EXPORTS = [
    "lh_last_error", "lh_backend_create", "lh_backend_destroy", "lh_backend_device_info",
    "lh_session_create", "lh_session_upload_matrix", "lh_session_matvec", "lh_session_assemble",
    "lh_session_destroy",
    "lh_pairsum_create", "lh_pairsum_set_wavelength", "lh_pairsum_matvec", "lh_pairsum_solve",
    "lh_pairsum_incident", "lh_pairsum_cross_sections", "lh_pairsum_compute_cross_sections",
    "lh_pairsum_adjoint_grad", "lh_pairsum_field_at_points",
    "lh_pairsum_field_at_points_noalloc", "lh_pairsum_far_field",
    "lh_pairsum_destroy", "lh_batch_solve", "lh_batch_create", "lh_batch_run",
    "lh_batch_results", "lh_batch_destroy", "lh_batch_set_timing", "lh_batch_kernel_times",
    "lh_batch_set_graph_cycles",
    "lh_sweep_solve",
]


class c64(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("backend_type", C.c_int),
                ("memory_bytes", C.c_uint64), ("compute_units", C.c_uint32)]


class PairsumSpec(C.Structure):
    _fields_ = [
        ("n_dipoles", C.c_size_t),
        ("positions", C.POINTER(C.c_double)),
        ("polarisability", C.c_void_p),
        ("polarisability_uniform", C.c_int),
        ("k", C.c_double),
        ("use_fcd", C.c_int),
        ("cell_size", C.c_double),
        ("has_lattice", C.c_int),
        ("a1", C.c_double * 3),
        ("a2", C.c_double * 3),
        ("ewald_eta", C.c_double),
        ("truncation_real", C.c_int),
        ("truncation_recip", C.c_int),
        ("k_bloch", C.c_double * 3),
        ("has_substrate", C.c_int),
        ("z_interface", C.c_double),
        ("delta_eps", c64),
        ("precision", C.c_int), ("classify_positions", C.POINTER(C.c_double)),
    ]


class Incident(C.Structure):
    _fields_ = [("direction", C.c_double * 3), ("polarisation", c64 * 3),
                ("amplitude", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_uint64), ("cycles", C.c_uint64), ("matvecs", C.c_uint64),
                ("matvecs_fp32", C.c_uint64), ("residual", C.c_double),
                ("pair_evals", C.c_double), ("status", C.c_int), ("solve_group", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchProblem(C.Structure):
    _fields_ = [("spec", PairsumSpec), ("incident", Incident)]


class Material(C.Structure):
    _fields_ = [("n_points", C.c_size_t), ("wavelength_nm", C.POINTER(C.c_double)),
                ("epsilon", C.c_void_p)]


class SweepSpec(C.Structure):
    _fields_ = [("geometry", PairsumSpec), ("material", C.POINTER(C.c_int32)),
                ("spacing", C.POINTER(C.c_double)), ("uniform_spacing", C.c_double),
                ("hint", C.POINTER(C.c_int32)), ("n_hints", C.c_size_t),
                ("hint_depol", C.POINTER(C.c_double)), ("hint_rotation", C.POINTER(C.c_double)),
                ("n_env", C.c_double), ("substrate_material", C.c_int),
                ("substrate_retarded", C.c_int)]


class KernelTimes(C.Structure):
    _fields_ = [("far64_ms", C.c_double), ("far32_ms", C.c_double),
                ("far64_launches", C.c_uint64), ("far32_launches", C.c_uint64),
                ("far64_pairs", C.c_double), ("far32_pairs", C.c_double),
                ("run_ms", C.c_double), ("far64_block", C.c_int32),
                ("far32_block", C.c_int32), ("tune_ms", C.c_double),
                ("host_round_trips", C.c_int32), ("solve_groups", C.c_int32),
                ("cycles_issued", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LuminaError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class ConvergenceFailure(LuminaError):
    pass


_lib = None


def load():
    """Load liblumina_hip.so from the package directory (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LuminaError(LH_UNAVAILABLE,
                          f"{LIB_PATH} is not built; run lumina_amd/build.py (hipcc, gfx950)")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    PP = C.POINTER(C.c_void_p)
    D = C.POINTER(C.c_double)
    lib.lh_last_error.restype = C.c_char_p
    lib.lh_backend_create.argtypes = [C.c_int, PP]
    lib.lh_backend_destroy.argtypes = [P]
    lib.lh_backend_destroy.restype = None
    lib.lh_backend_device_info.argtypes = [P, C.POINTER(DeviceInfo)]
    lib.lh_session_create.argtypes = [P, C.c_size_t, PP]
    lib.lh_session_upload_matrix.argtypes = [P, P, C.c_size_t]
    lib.lh_session_matvec.argtypes = [P, P, P, C.c_size_t]
    lib.lh_session_assemble.argtypes = [P, D, C.c_size_t, C.c_double, P, PP]
    lib.lh_session_destroy.argtypes = [P]
    lib.lh_session_destroy.restype = None
    lib.lh_pairsum_create.argtypes = [P, C.POINTER(PairsumSpec), PP]
    lib.lh_pairsum_set_wavelength.argtypes = [P, C.c_double, P, C.c_int, c64]
    lib.lh_pairsum_matvec.argtypes = [P, P, P, C.c_int]
    lib.lh_pairsum_solve.argtypes = [P, P, C.c_double, C.c_size_t, P, C.POINTER(SolveInfo)]
    lib.lh_pairsum_incident.argtypes = [P, C.POINTER(Incident), P]
    lib.lh_pairsum_cross_sections.argtypes = [P, P, P, C.c_double, D]
    lib.lh_pairsum_compute_cross_sections.argtypes = [P, C.POINTER(Incident), C.c_double,
                                                      C.c_size_t, D, P, C.POINTER(SolveInfo)]
    lib.lh_pairsum_adjoint_grad.argtypes = [P, C.POINTER(Incident), C.c_double, C.c_size_t, P,
                                            C.c_int, D, D, D, C.POINTER(SolveInfo),
                                            C.POINTER(SolveInfo)]
    lib.lh_pairsum_field_at_points.argtypes = [P, P, C.POINTER(Incident), C.c_size_t, D, P, D]
    # device pointers for moments, points, field and intensity
    lib.lh_pairsum_field_at_points_noalloc.argtypes = [P, P, C.POINTER(Incident), C.c_size_t, P,
                                                       P, P]
    lib.lh_pairsum_far_field.argtypes = [P, P, C.c_size_t, C.c_size_t, D, D, D]
    lib.lh_pairsum_destroy.argtypes = [P]
    lib.lh_pairsum_destroy.restype = None
    lib.lh_batch_solve.argtypes = [P, C.POINTER(BatchProblem), C.c_size_t, C.c_double,
                                   C.c_size_t, C.c_int, D, C.POINTER(C.c_void_p),
                                   C.POINTER(SolveInfo)]
    lib.lh_batch_create.argtypes = [P, C.POINTER(BatchProblem), C.c_size_t, PP]
    lib.lh_batch_run.argtypes = [P, C.c_double, C.c_size_t, C.c_int]
    lib.lh_batch_results.argtypes = [P, D, C.POINTER(C.c_void_p), C.POINTER(SolveInfo)]
    lib.lh_batch_destroy.argtypes = [P]
    lib.lh_batch_destroy.restype = None
    lib.lh_batch_set_timing.argtypes = [P, C.c_int]
    lib.lh_batch_set_graph_cycles.argtypes = [P, C.c_int]
    lib.lh_batch_kernel_times.argtypes = [P, C.POINTER(KernelTimes)]
    lib.lh_sweep_solve.argtypes = [P, C.POINTER(SweepSpec), C.POINTER(Material), C.c_size_t, D,
                                   C.c_size_t, C.POINTER(Incident), C.c_double, C.c_size_t, D,
                                   C.POINTER(SolveInfo)]
    _lib = lib
    return lib


def check(status: int, lib=None):
    if status == LH_OK:
        return
    lib = lib or load()
    msg = lib.lh_last_error().decode(errors="replace")
    if status == LH_CONVERGENCE_FAILURE:
        raise ConvergenceFailure(status, msg)
    raise LuminaError(status, msg)


def cptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))
