"""Argument and state errors of the C ABI on a live backend.

The status codes mirror the reference's ComputeError / SolverError
(backend.rs:13-22, solver/mod.rs:18-31): a dimension mismatch of the
stored-matrix session is a DeviceError (gpu.rs:411-416, 429-435); invalid
inputs that the reference rejects with an error or a panic (no dipoles,
cda/mod.rs; a singular polarisability, assembly.rs:344-348) are
LH_INVALID_ARGUMENT here. Every call below fails on the host before any
kernel is launched, and the backend stays usable afterwards.
"""
import ctypes as C

import numpy as np
import pytest

from lumina_amd import _lib
from lumina_amd import inputs as I
from lumina_amd.solver import FP64, Batch, IncidentField, LatticeSpec, LuminaError, MatvecSession

pytestmark = pytest.mark.gpu


def sphere():
    pos = I.discretise(I.Sphere(4.0), 2.0)
    return pos, I.isotropic_block(40.0 + 10.0j), 2 * np.pi / 500.0


def status_of(fn, *a, **kw):
    with pytest.raises(LuminaError) as ei:
        fn(*a, **kw)
    return ei.value.status, str(ei.value)


def test_pairsum_spec_errors(backend):
    pos, alpha, k = sphere()
    mk = backend.create_pairsum_session
    st, msg = status_of(mk, np.zeros((0, 3)), alpha, k, use_fcd=True, cell_size=2.0)
    assert st == _lib.LH_INVALID_ARGUMENT and "No dipoles" in msg
    st, msg = status_of(mk, pos, np.zeros(9, complex), k, use_fcd=True, cell_size=2.0)
    assert st == _lib.LH_INVALID_ARGUMENT and "Singular" in msg
    for bad_k in (0.0, -1.0, float("nan"), float("inf")):
        st, _ = status_of(mk, pos, alpha, bad_k, use_fcd=True, cell_size=2.0)
        assert st == _lib.LH_INVALID_ARGUMENT
    st, _ = status_of(mk, pos, alpha, k, use_fcd=True, cell_size=0.0)
    assert st == _lib.LH_INVALID_ARGUMENT
    bad = pos.copy()
    bad[3, 1] = np.nan
    st, msg = status_of(mk, bad, alpha, k, use_fcd=True, cell_size=2.0)
    assert st == _lib.LH_INVALID_ARGUMENT and "non-finite" in msg
    lat = LatticeSpec((300.0, 0.0, 0.0), (0.0, 300.0, 0.0), None, 60, 5)
    st, msg = status_of(mk, pos, alpha, k, use_fcd=True, cell_size=2.0, lattice=lat)
    assert st == _lib.LH_INVALID_ARGUMENT and "truncation" in msg
    # the backend is still usable
    s = mk(pos, alpha, k, use_fcd=True, cell_size=2.0, precision=FP64)
    cs, _, info = s.compute_cross_sections(IncidentField(), 1e-10, 1000)
    assert info["status"] == 0 and cs[0] > 0.0
    s.close()


def test_pairsum_call_errors(backend):
    pos, alpha, k = sphere()
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=2.0)
    lib, n3 = backend._lib, 3 * len(pos)
    rhs = np.ones(n3, complex)
    x = np.zeros(n3, complex)
    info = _lib.SolveInfo()
    st = lib.lh_pairsum_solve(s._h, rhs.ctypes.data_as(C.c_void_p), -1.0, 100,
                              x.ctypes.data_as(C.c_void_p), C.byref(info))
    assert st == _lib.LH_INVALID_ARGUMENT
    st = lib.lh_pairsum_solve(s._h, None, 1e-8, 100, x.ctypes.data_as(C.c_void_p),
                              C.byref(info))
    assert st == _lib.LH_INVALID_ARGUMENT
    st, _ = status_of(s.set_wavelength, -2.0, alpha)
    assert st == _lib.LH_INVALID_ARGUMENT
    # an unknown precision code, a negative or NaN tolerance
    y = np.zeros(n3, complex)
    st = lib.lh_pairsum_matvec(s._h, rhs.ctypes.data_as(C.c_void_p),
                               y.ctypes.data_as(C.c_void_p), 7)
    assert st == _lib.LH_INVALID_ARGUMENT
    inc = IncidentField()._c()
    out = np.zeros(3)
    for bad_tol in (-1.0, float("nan")):
        st = lib.lh_pairsum_compute_cross_sections(s._h, C.byref(inc), bad_tol, 100,
                                                   out.ctypes.data_as(C.POINTER(C.c_double)), None,
                                                   None)
        assert st == _lib.LH_INVALID_ARGUMENT
    # the session is still usable
    cs, _, info = s.compute_cross_sections(IncidentField(), 1e-10, 1000)
    assert info["status"] == 0 and cs[0] > 0.0
    s.close()


def test_stored_matrix_session_errors(backend):
    n = 6
    ms = MatvecSession(backend, n)
    lib = backend._lib
    x = np.ones(n, complex)
    y = np.zeros(n, complex)
    # matvec before any upload
    st = lib.lh_session_matvec(ms._h, x.ctypes.data_as(C.c_void_p),
                               y.ctypes.data_as(C.c_void_p), n)
    assert st == _lib.LH_DEVICE_ERROR
    st, msg = status_of(ms.upload_matrix, np.eye(n + 1, dtype=complex))
    assert st == _lib.LH_DEVICE_ERROR and "dimension mismatch" in msg
    ms.upload_matrix(2.0 * np.eye(n, dtype=complex))
    st, msg = status_of(ms.matvec, np.ones(n + 3, complex))
    assert st == _lib.LH_DEVICE_ERROR and "dimension mismatch" in msg
    assert np.array_equal(ms.matvec(x), 2.0 * x)
    ms.close()


def test_backend_trait_methods(backend):
    """The rest of ComputeBackend (backend.rs:62-134) as the reference's GPU
    backend gives it: the one-off matvec through a temporary session
    (gpu.rs:634-655) against numpy, with its dimension-mismatch DeviceError;
    the host-side matrix fill (gpu.rs:611-632); dense_solve Unavailable."""
    r = np.random.default_rng(5)
    n = 37
    A = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    x = r.standard_normal(n) + 1j * r.standard_normal(n)
    y = backend.matvec(A, x)
    assert np.max(np.abs(y - A @ x)) < 1e-13 * np.max(np.abs(A @ x))
    st, msg = status_of(backend.matvec, A, x[:-1])
    assert st == _lib.LH_DEVICE_ERROR and "Dimension mismatch" in msg
    st, msg = status_of(backend.matvec, A[:, :-1], x[:-1])
    assert st == _lib.LH_DEVICE_ERROR
    F = backend.parallel_matrix_fill(3, 4, lambda i, j: complex(i, 10 * j))
    assert F.shape == (3, 4) and F[2, 3] == 2 + 30j and F[1, 0] == 1
    st, msg = status_of(backend.dense_solve, A, x)
    assert st == _lib.LH_UNAVAILABLE
    assert np.allclose(backend.matvec(np.eye(n), x), x, rtol=0, atol=0)


def test_batch_errors(backend):
    st, msg = status_of(Batch, backend, [])
    assert st == _lib.LH_INVALID_ARGUMENT
