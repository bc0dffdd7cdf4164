"""lh_pairsum_field_at_points_noalloc (device buffers, fixed scratch, no
allocation) against lh_pairsum_field_at_points on the same session: the
same kernels on the same point batches, so the results are bit-identical.
Two calls in a row, the second smaller and with explicit moments and
another incident field, then the first again (the scratch keeps the last
call's contents; nothing may depend on them). Device buffers are made with
the HIP runtime the library itself links (libamdhip64.so.7, through
ctypes: plumbing only); the calls go through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from lumina_amd import _lib
from lumina_amd import inputs as I
from lumina_amd.fields import field_at_points
from lumina_amd.solver import FP64, IncidentField, PairSumSession

pytestmark = pytest.mark.gpu

LCP = (1 / np.sqrt(2), 1j / np.sqrt(2), 0.0)


class DevBuf:
    """A device buffer of the runtime the library links (hipMalloc through
    ctypes), filled from or read into numpy arrays."""
    hip = None

    def __init__(self, nbytes, host=None):
        if DevBuf.hip is None:
            DevBuf.hip = C.CDLL("libamdhip64.so.7")
        self.ptr = C.c_void_p()
        assert DevBuf.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(max(nbytes, 16))) == 0
        self.nbytes = nbytes
        if host is not None:
            h = np.ascontiguousarray(host)
            assert h.nbytes == nbytes
            assert DevBuf.hip.hipMemcpy(self.ptr, C.c_void_p(h.ctypes.data),
                                        C.c_size_t(nbytes), 1) == 0   # host to device

    def read(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes == self.nbytes
        assert DevBuf.hip.hipMemcpy(C.c_void_p(out.ctypes.data), self.ptr,
                                    C.c_size_t(self.nbytes), 2) == 0   # device to host
        return out

    def free(self):
        if self.ptr:
            DevBuf.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def device_free_bytes():
    free, total = C.c_size_t(), C.c_size_t()
    if DevBuf.hip is None:
        DevBuf.hip = C.CDLL("libamdhip64.so.7")
    assert DevBuf.hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def make_bufs(points, moments=None, want_field=True):
    m = len(points)
    return (DevBuf(points.nbytes, points),
            DevBuf(16 * 3 * m) if want_field else None,
            DevBuf(8 * m),
            DevBuf(16 * len(moments), moments.astype(np.complex128)) if moments is not None
            else None)


def noalloc(s, points, inc, moments=None, want_field=True, bufs=None):
    m = len(points)
    own = bufs is None
    if own:
        bufs = make_bufs(points, moments, want_field)
    pts, field, inten, mom = bufs
    ci = inc._c()
    st = s._lib.lh_pairsum_field_at_points_noalloc(
        s._h, mom.ptr if mom else None, C.byref(ci), m, pts.ptr, field.ptr if field else None,
        inten.ptr)
    _lib.check(st)
    f = field.read(np.complex128, (m, 3)) if field is not None else None
    i = inten.read(np.float64, (m,))
    if own:
        for b in bufs:
            if b is not None:
                b.free()
    return f, i


def points_around(n, half, seed):
    return np.random.default_rng(seed).uniform(-half, half, size=(n, 3))


@pytest.fixture(scope="module")
def sphere(backend):
    pos = I.au_sphere_r10_d2()
    lam, n_env = 520.0, 1.33
    alpha = I.polarisability_block("Au", lam, 2.0, n_env)
    k = I.wavenumber(lam, n_env)
    s = PairSumSession(backend, pos, alpha, k, True, 2.0, precision=FP64)
    _, p, _ = s.compute_cross_sections(IncidentField(), 1e-10, 1000)
    yield s, p
    s.close()


def test_noalloc_matches_field_at_points_twice(sphere):
    s, p = sphere
    inc1 = IncidentField()
    pts1 = points_around(100_000, 40.0, 1)
    # first call on this session: no field-map buffer exists yet, and the
    # call must not make one (the session's own path would need ~36 MB)
    bufs = make_bufs(pts1)
    free0 = device_free_bytes()
    f1, i1 = noalloc(s, pts1, inc1, bufs=bufs)
    free1 = device_free_bytes()
    for b in bufs:
        if b is not None:
            b.free()
    # the same call again: once the runtime has set up the kernels' first
    # launch, a call changes the device's free memory by nothing at all
    bufs = make_bufs(pts1)
    free2 = device_free_bytes()
    f1b, i1b = noalloc(s, pts1, inc1, bufs=bufs)
    free3 = device_free_bytes()
    for b in bufs:
        if b is not None:
            b.free()
    print(f"device free memory change over the first call: {free0 - free1} bytes, "
          f"over the second: {free2 - free3} bytes")
    assert free0 - free1 < 8 << 20
    assert free2 - free3 == 0
    assert np.array_equal(f1b, f1) and np.array_equal(i1b, i1)
    r1, ri1 = field_at_points(s, pts1, inc1)
    assert np.array_equal(f1, r1) and np.array_equal(i1, ri1)
    # second call: fewer points, explicit moments, circular light along -x
    inc2 = IncidentField(direction=(-1.0, 0.0, 0.0), polarisation=(0.0, LCP[0], LCP[1]),
                         amplitude=2.0)
    pts2 = points_around(700, 25.0, 2)
    mom = 0.5 * p + 1j * np.roll(p, 3)
    f2, i2 = noalloc(s, pts2, inc2, moments=mom)
    r2, ri2 = field_at_points(s, pts2, inc2, moments=mom)
    assert np.array_equal(f2, r2) and np.array_equal(i2, ri2)
    # intensity only
    _, i3 = noalloc(s, pts2, inc2, moments=mom, want_field=False)
    assert np.array_equal(i3, ri2)
    # a third call like the first, after the scratch held the second's
    # contents (5,000 points: the field map splits the dipoles into other
    # chunks than for 100,000, so it is compared with its own reference)
    f4, i4 = noalloc(s, pts1[:5000], inc1)
    r4, ri4 = field_at_points(s, pts1[:5000], inc1)
    assert np.array_equal(f4, r4) and np.array_equal(i4, ri4)
    assert np.max(np.abs(f4 - r1[:5000])) <= 1e-13 * np.max(np.abs(r1))


def test_noalloc_argument_errors(sphere):
    s, _ = sphere
    ci = IncidentField()._c()
    st = s._lib.lh_pairsum_field_at_points_noalloc(s._h, None, C.byref(ci), 10, None, None, None)
    assert st == _lib.LH_INVALID_ARGUMENT
    # m = 0 is a no-op
    st = s._lib.lh_pairsum_field_at_points_noalloc(s._h, None, C.byref(ci), 0, None, None, None)
    assert st == _lib.LH_OK
