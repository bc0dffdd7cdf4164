"""The ctypes wrappers refuse wrong-length arrays before any C call: the C
ABI copies fixed lengths (3N per vector, 9 or 9N per polarisability), so a
short array would be read or written past its end. CPU only: the session
object is built without a device and its library raises if it is called."""
import numpy as np
import pytest

from lumina_amd.solver import PairSumSession


class _NoCalls:
    def __getattr__(self, name):
        raise AssertionError(f"C entry point {name} reached with a wrong-length array")


def _session(n=4):
    s = object.__new__(PairSumSession)
    s._lib = _NoCalls()
    s._h = None
    s.n = n
    s.k = 0.01
    s.precision = 0
    return s


def test_solve_gmres_rejects_wrong_rhs_length():
    s = _session()
    with pytest.raises(ValueError, match="rhs"):
        s.solve_gmres(np.zeros(11, complex), 1e-8, 10)
    with pytest.raises(ValueError, match="rhs"):
        s.solve_gmres(np.zeros(13, complex), 1e-8, 10)


def test_cross_sections_rejects_wrong_lengths():
    s = _session()
    with pytest.raises(ValueError, match="p has"):
        s.cross_sections(np.zeros(6, complex), np.zeros(12, complex))
    with pytest.raises(ValueError, match="e_inc"):
        s.cross_sections(np.zeros(12, complex), np.zeros(3, complex))


def test_set_wavelength_rejects_wrong_polarisability_length():
    s = _session()
    for bad in (np.zeros(8, complex), np.zeros(18, complex), np.zeros(9 * 4 + 9, complex)):
        with pytest.raises(ValueError, match="polarisability"):
            s.set_wavelength(0.02, bad)


def test_matvec_rejects_wrong_length():
    from lumina_amd._lib import LuminaError
    s = _session()
    with pytest.raises(LuminaError):
        s.matvec(np.zeros(5, complex))
