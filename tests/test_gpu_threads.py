"""Sessions and batches used from several host threads at once.

The reference's backend and sessions are Send + Sync and its wavelength loop
holds one session per rayon thread (gpu.rs:24-30); its
test_parallel_wavelengths_gpu (gpu.rs:1036-1072) gives each of 4 threads its
own stored-matrix session of one 12 x 12 matrix and a thread-dependent
vector and checks every product against the CPU. Here every session and
batch owns a non-blocking HIP stream and the library holds no global lock;
ctypes releases the GIL around each library call, so the threads' calls
(uploads, launches, graph captures, allocations and frees) overlap inside
the library.

- the reference's test, restated: FP64 stored-matrix products, bar 1e-12
  (the reference's 1e-4 is its FP32 bar);
- one on-the-fly session per wavelength on 4 threads, each solving the
  515-dipole sphere in FP64 and in mixed precision (device GMRES, one graph
  captured per restart cycle while the other threads work): the cross
  sections, iteration counts and moments equal bit for bit those of the same
  solves run one after another, since no result depends on timing once the
  far-field form is fixed (the autotuner's timing-based choice is pinned by
  LUMINA_FAR_BLOCK);
- two batches solved on two threads at once equal the same batches solved
  one after another, bit for bit;
- a failing call on one thread leaves its message on that thread only
  (lh_last_error is thread-local).
"""
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from lumina_amd import inputs as I
from lumina_amd import _lib
from lumina_amd.solver import FP64, MIXED, Batch, IncidentField, LuminaError

pytestmark = pytest.mark.gpu

THREADS = 4


def run_threads(fn, args):
    """fn(*a) for every a in args on its own thread, started together."""
    gate = threading.Barrier(len(args))

    def body(a):
        gate.wait()
        return fn(*a)

    with ThreadPoolExecutor(len(args)) as ex:
        return list(ex.map(body, args))


def test_parallel_stored_sessions(backend):
    """gpu.rs:1036-1072: 4 threads, one session each, the same matrix."""
    dim = 12
    i, j = np.meshgrid(np.arange(dim), np.arange(dim), indexing="ij")
    A = np.where(i == j, 5.0 + 1.0j, -0.1 + 0.05j * (i - j))

    def work(tid):
        s = backend.create_session(dim)
        try:
            s.upload_matrix(A)
            v = np.arange(dim) + 1j * (0.1 * tid)
            return s.matvec(v), A @ v
        finally:
            s.close()

    for tid, (got, want) in enumerate(run_threads(work, [(t,) for t in range(THREADS)])):
        err = np.max(np.abs(got - want))
        print(f"thread {tid}: max err {err:.2e}")
        assert err < 1e-12 * np.max(np.abs(want))


WAVELENGTHS = (400.0, 440.0, 480.0, 520.0)


def sphere_solve(backend, lam, precision):
    pos = I.au_sphere_r10_d2()
    s = backend.create_pairsum_session(pos, I.polarisability_block("Au", lam, 2.0, 1.33),
                                       I.wavenumber(lam, 1.33), use_fcd=True, cell_size=2.0,
                                       precision=precision)
    try:
        cs, p, info = s.compute_cross_sections(IncidentField(), 1e-10, 1000)
        return cs, p, info["iterations"]
    finally:
        s.close()


@pytest.mark.parametrize("precision", [FP64, MIXED], ids=["fp64", "mixed"])
def test_parallel_wavelength_sessions(backend, monkeypatch, precision):
    monkeypatch.setenv("LUMINA_FAR_BLOCK", "256")
    monkeypatch.setenv("LUMINA_FFT", "0")
    seq = [sphere_solve(backend, lam, precision) for lam in WAVELENGTHS]
    par = run_threads(lambda lam: sphere_solve(backend, lam, precision),
                      [(lam,) for lam in WAVELENGTHS])
    for lam, (cs0, p0, it0), (cs1, p1, it1) in zip(WAVELENGTHS, seq, par):
        print(f"{lam} nm: C_ext {cs1[0]:.12g}, iterations {it1} (sequential {it0})")
        assert it1 == it0
        assert np.array_equal(cs1, cs0)
        assert np.array_equal(p1, p0)


def batch_problems(lams):
    out = []
    for shape in ("sphere", "cuboid"):
        pos = (I.au_sphere_r10_d2() if shape == "sphere"
               else I.discretise(I.Cuboid((16.0, 12.0, 10.0)), 2.0))
        for lam in lams:
            out.append(dict(positions=pos, polarisability=I.polarisability_block("Au", lam, 2.0, 1.33),
                            k=I.wavenumber(lam, 1.33), use_fcd=True, cell_size=2.0))
    return out


def batch_solve(backend, problems):
    b = Batch(backend, problems)
    try:
        b.run(1e-8, 1000, MIXED)
        return b.results(want_moments=True)
    finally:
        b.close()


def test_parallel_batches(backend, monkeypatch):
    monkeypatch.setenv("LUMINA_FAR_BLOCK", "256")
    monkeypatch.setenv("LUMINA_FFT", "0")
    sets = [batch_problems((410.0, 450.0, 490.0)), batch_problems((430.0, 470.0, 510.0))]
    seq = [batch_solve(backend, p) for p in sets]
    par = run_threads(lambda p: batch_solve(backend, p), [(p,) for p in sets])
    for (cs0, inf0, m0), (cs1, inf1, m1) in zip(seq, par):
        assert [i["iterations"] for i in inf1] == [i["iterations"] for i in inf0]
        assert np.array_equal(cs1, cs0)
        for a, b in zip(m1, m0):
            assert np.array_equal(a, b)
    print("batch C_ext", seq[0][0][:, 0], seq[1][0][:, 0])


def test_last_error_is_per_thread(backend):
    """One thread's failure (k <= 0 rejected at session creation) sets its own
    lh_last_error; a thread whose calls succeed reads an empty message."""
    pos = I.au_sphere_r10_d2()
    alpha = I.polarisability_block("Au", 500.0, 2.0, 1.33)
    lib = backend._lib

    def work(fail):
        if fail:
            with pytest.raises(LuminaError) as ei:
                backend.create_pairsum_session(pos, alpha, -1.0, use_fcd=True, cell_size=2.0)
            return ei.value.status, lib.lh_last_error().decode()
        s = backend.create_pairsum_session(pos, alpha, I.wavenumber(500.0, 1.33), use_fcd=True,
                                           cell_size=2.0)
        s.matvec(np.ones(3 * len(pos), complex))
        s.close()
        return _lib.LH_OK, lib.lh_last_error().decode()

    (st_bad, msg_bad), (st_ok, msg_ok) = run_threads(work, [(True,), (False,)])
    print("failing thread:", st_bad, repr(msg_bad), "| other thread:", repr(msg_ok))
    assert st_bad == _lib.LH_INVALID_ARGUMENT and msg_bad
    assert st_ok == _lib.LH_OK and msg_ok == ""
