"""Device assembly of the stored interaction matrix
(assemble_and_create_session, backend.rs:109-117; the wgpu form gpu.rs:508-607
with shaders/matrix_assemble.wgsl), which CdaSolver uses with FCD off, no
lattice or substrate and dipoles off a regular grid (cda/mod.rs:298-311).

The reference's GPU tests compare its FP32 assembly with the CPU assembly
elementwise (< 5e-5 absolute, gpu.rs:878-953) and check the diagonal patch.
Here the assembly is FP64:

- every column (A e_k through the session's matvec, which reproduces a
  column exactly) against the oracle's dense assembly (assembly.rs
  restated) of the same dipoles with FCD off: off-diagonal blocks to 1e-13
  relative per block, diagonal blocks bit for bit equal to the alpha^-1
  blocks passed in (anisotropic, a different block per dipole);
- one dipole (gpu.rs's N = 1 case): A = alpha^-1;
- a coincident pair gets a zero block, as in matrix_assemble.wgsl;
- at the largest size the reference would assemble (its 2 GiB budget,
  cda/mod.rs:81: 3,800 dipoles, a 2.08 GB matrix), the stored product
  equals the on-the-fly FP64 product to 1e-12;
- invalid arguments are rejected before any allocation.
"""
import time

import numpy as np
import pytest

from lumina_amd import _lib
from lumina_amd import inputs as I
from lumina_amd.solver import FP64, LuminaError
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

LAM, N_ENV = 600.0, 1.33


def off_grid(n, seed, box=20.0):
    r = np.random.default_rng(seed)
    pos = r.uniform(-box, box, size=(n, 3))
    return pos


def alpha_blocks(n, seed):
    """Anisotropic per-dipole polarisabilities: the J&C Au block rotated
    and scaled per dipole."""
    base = I.polarisability_block("Au", LAM, 2.0, N_ENV).reshape(3, 3)
    r = np.random.default_rng(seed)
    out = np.empty((n, 3, 3), complex)
    for i in range(n):
        q, _ = np.linalg.qr(r.standard_normal((3, 3)))
        out[i] = q @ (base * np.diag(r.uniform(0.5, 1.5, 3))) @ q.T
    return out


def columns(sess, dim):
    A = np.empty((dim, dim), complex)
    e = np.zeros(dim, complex)
    for c in range(dim):
        e[c] = 1.0
        A[:, c] = sess.matvec(e)
        e[c] = 0.0
    return A


def test_assembly_matches_oracle_and_diagonal_is_exact(backend):
    n = 40
    pos = off_grid(n, 1)
    alpha = alpha_blocks(n, 2)
    ainv = np.stack([O.invert_3x3(a.reshape(9)).reshape(3, 3) for a in alpha])
    k = I.wavenumber(LAM, N_ENV)
    s = backend.assemble_and_create_session(pos, k, ainv)
    A = columns(s, 3 * n)
    s.close()
    ref = O.Problem(pos, alpha.reshape(n, 9), k, use_fcd=False, cell_size=2.0).assemble()
    worst = 0.0
    for i in range(n):
        for j in range(n):
            blk = A[3 * i:3 * i + 3, 3 * j:3 * j + 3]
            if i == j:
                assert np.array_equal(blk, ainv[i])
                continue
            rb = ref[3 * i:3 * i + 3, 3 * j:3 * j + 3]
            worst = max(worst, np.max(np.abs(blk - rb)) / np.max(np.abs(rb)))
    print(f"off-diagonal blocks vs oracle: max rel err {worst:.2e}")
    assert worst < 1e-13


def test_single_dipole_and_coincident_pair(backend):
    k = I.wavenumber(LAM, N_ENV)
    ainv = alpha_blocks(3, 5)
    s = backend.assemble_and_create_session(np.array([[1.0, 2.0, 3.0]]), k, ainv[:1])
    x = np.array([1.0 + 2.0j, -0.5j, 3.0])
    want = ainv[0] @ x
    assert np.max(np.abs(s.matvec(x) - want)) <= 1e-15 * np.max(np.abs(want))
    s.close()
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [5.0, 0.0, 0.0]])
    s = backend.assemble_and_create_session(pos, k, ainv)
    A = columns(s, 9)
    s.close()
    assert np.all(A[0:3, 3:6] == 0) and np.all(A[3:6, 0:3] == 0)
    assert A[0, 6] != 0 and A[1, 7] != 0
    assert np.array_equal(A[0:3, 6:9], A[3:6, 6:9])


def test_budget_size_matches_pair_sum(backend):
    """3,800 off-grid dipoles: 3 x 3,800 = 11,400 unknowns, a 2.08 GB FP64
    matrix (the reference assembles up to 2 GiB)."""
    n = 3800
    pos = off_grid(n, 3, box=40.0)
    alpha = I.polarisability_block("Au", LAM, 2.0, N_ENV)
    k = I.wavenumber(LAM, N_ENV)
    ainv = np.broadcast_to(O.invert_3x3(alpha).reshape(3, 3), (n, 3, 3))
    t0 = time.perf_counter()
    s = backend.assemble_and_create_session(pos, k, ainv)
    t_asm = time.perf_counter() - t0
    r = np.random.default_rng(4)
    x = r.standard_normal(3 * n) + 1j * r.standard_normal(3 * n)
    y = s.matvec(x)
    s.close()
    p = backend.create_pairsum_session(pos, alpha, k, use_fcd=False, cell_size=2.0)
    y_otf = p.matvec(x, FP64)
    p.close()
    err = np.max(np.abs(y - y_otf)) / np.max(np.abs(y_otf))
    print(f"N={n}: assembly + upload of the session {t_asm * 1e3:.1f} ms; "
          f"stored vs on-the-fly {err:.2e}")
    assert err < 1e-12


def test_assembly_errors(backend):
    k = I.wavenumber(LAM, N_ENV)
    ainv = alpha_blocks(2, 6)
    pos = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    bad = [
        (np.zeros((0, 3)), k, ainv[:0], "No dipoles"),
        (pos, 0.0, ainv, "k must be"),
        (pos, float("nan"), ainv, "k must be"),
        (np.array([[0.0, np.inf, 0.0], [3.0, 0.0, 0.0]]), k, ainv, "non-finite dipole"),
        (pos, k, np.where(np.arange(9).reshape(3, 3) == 4, np.nan, ainv), "non-finite diagonal"),
    ]
    for p, kk, blocks, msg in bad:
        with pytest.raises(LuminaError) as ei:
            backend.assemble_and_create_session(p, kk, blocks)
        assert ei.value.status == _lib.LH_INVALID_ARGUMENT and msg in str(ei.value)
    with pytest.raises(ValueError):
        backend.assemble_and_create_session(pos, k, ainv[:1])
    # the backend is still usable
    s = backend.assemble_and_create_session(pos, k, ainv)
    assert np.all(np.isfinite(s.matvec(np.ones(6, complex))))
    s.close()
