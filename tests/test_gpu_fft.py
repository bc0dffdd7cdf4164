"""FFT block-Toeplitz matvec (fft_matvec.rs) on the GPU.

The reference's own tests (fft_matvec.rs:429-527: 5x5x5 and 10x10x10 grids
at 3 nm spacing, 600 nm, alpha = 1000 + 200i, FCD off, the deterministic
vector ((7i+1) mod 17)/17 + i ((11i+3) mod 13)/13, FFT against the on-the-fly
pair sum to 1e-10) are restated against the oracle's pair sum, with the
tighter bar 1e-13. Beyond the reference, the device also takes the FFT path
with FCD on (the near list then carries the FCD correction to the point
tensor); that path is checked against the oracle's FCD matvec, full solves
against the device pair sum (C_ext to 1e-10), mixed precision (FP32 FFT)
against FP64, graph replay against eager issue (bitwise), and the adjoint
gradients against the pair-sum session's.

Which path a solve took shows in its statistics: pair_evals is 0 for the
FFT matvec.
"""
import numpy as np
import pytest

from lumina_amd import inputs as I
from lumina_amd.solver import (FP64, MIXED, ConvergenceFailure, IncidentField, LatticeSpec,
                               PairSumSession, SubstrateRuntime)
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def relnorm(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def make_grid(n, spacing):
    """fft_matvec.rs:394-409: i outer, k inner."""
    return np.array([[i * spacing, j * spacing, k * spacing]
                     for i in range(n) for j in range(n) for k in range(n)], dtype=float)


def box(nx, ny, nz, h=(1.5, 1.5, 1.5), origin=(-7.0, 3.0, 1.0), seed=None):
    pts = np.array([[origin[0] + i * h[0], origin[1] + j * h[1], origin[2] + k * h[2]]
                    for i in range(nx) for j in range(ny) for k in range(nz)])
    if seed is not None:   # dipoles in any order
        pts = pts[np.random.default_rng(seed).permutation(len(pts))]
    return pts


def ref_vector(n):
    i = np.arange(3 * n)
    return ((i * 7 + 1) % 17) / 17.0 + 1j * (((i * 11 + 3) % 13) / 13.0)


def session(monkeypatch, mode, *a, graph=True, **kw):
    if mode is None:
        monkeypatch.delenv("LUMINA_FFT", raising=False)
    else:
        monkeypatch.setenv("LUMINA_FFT", mode)
    if graph:
        monkeypatch.delenv("LUMINA_GRAPH", raising=False)
    else:
        monkeypatch.setenv("LUMINA_GRAPH", "0")
    s = PairSumSession(*a, **kw)
    monkeypatch.delenv("LUMINA_FFT", raising=False)
    monkeypatch.delenv("LUMINA_GRAPH", raising=False)
    return s


def uses_fft(s):
    """A one-iteration solve: pair_evals is 0 on the FFT path."""
    try:
        _, info = s.solve_gmres(np.ones(3 * s.n, dtype=complex), 1e-30, 1)
    except ConvergenceFailure as e:   # carries the statistics
        info = e.info
    return info["pair_evals"] == 0


@pytest.mark.parametrize("n", [5, 10])
def test_reference_grids_match_pair_sum(backend, monkeypatch, n):
    k = 2 * np.pi / 600.0
    pos = make_grid(n, 3.0)
    alpha = I.isotropic_block(1000 + 200j)
    s = session(monkeypatch, "force", backend, pos, alpha, k, False, 3.0)
    assert uses_fft(s)
    x = ref_vector(len(pos))
    y = s.matvec(x)
    yo = O.Problem(pos, alpha, k, False, 3.0).matvec(x)
    print("reference grid", n, relnorm(y, yo))
    assert relnorm(y, yo) < 1e-10          # the reference's bar
    assert relnorm(y, yo) < 1e-13
    # the device pair sum agrees too
    sp = session(monkeypatch, "0", backend, pos, alpha, k, False, 3.0)
    assert not uses_fft(sp)
    assert relnorm(y, sp.matvec(x)) < 1e-13
    s.close()
    sp.close()


@pytest.mark.parametrize("case", ["cube_fcd", "aniso_fcd", "aniso_nofcd", "plane", "line"])
def test_fft_matvec_matches_oracle(backend, monkeypatch, case):
    k = I.wavenumber(520.0, 1.33)
    alpha = I.polarisability_block("Au", 520.0, 1.5, 1.33)
    fcd, cell = True, 1.5
    if case == "cube_fcd":
        pos = box(20, 16, 14, seed=1)        # 4,480 dipoles: the FFT path by default
        mode = None
    elif case == "aniso_fcd":
        pos = box(12, 10, 8, h=(1.5, 2.0, 2.5), seed=2)
        mode = "force"
    elif case == "aniso_nofcd":
        pos = box(9, 13, 7, h=(2.5, 1.0, 1.75))
        mode, fcd = "force", False
    elif case == "plane":
        pos = box(30, 25, 1, seed=3)
        mode = "force"
    else:
        pos = box(1, 1, 40)
        mode = "force"
    s = session(monkeypatch, mode, backend, pos, alpha, k, fcd, cell)
    assert uses_fft(s)
    x = np.random.default_rng(5).standard_normal(3 * len(pos)) * (1 + 0.5j)
    y = s.matvec(x)
    yo = O.Problem(pos, alpha, k, fcd, cell).matvec(x)
    print(case, len(pos), relmax(y, yo), relnorm(y, yo))
    assert relmax(y, yo) < 1e-12
    # FP32 FFT: single-precision agreement
    y32 = s.matvec(x, precision=MIXED)
    print(case, "fp32", relmax(y32, yo))
    assert 0 < relmax(y32, yo) < 2e-5
    s.close()


def test_threshold_and_eligibility(backend, monkeypatch):
    k = I.wavenumber(520.0, 1.33)
    alpha = I.polarisability_block("Au", 520.0, 2.0, 1.33)
    # a full grid below kFftMinDipoles keeps the pair sum by default
    s = session(monkeypatch, None, backend, box(10, 10, 10, h=(2, 2, 2)), alpha, k, True, 2.0)
    assert not uses_fft(s)
    s.close()
    # a sphere of more than 4,096 dipoles is not a full grid
    sph = I.discretise(I.Sphere(20.5), 2.0)
    assert len(sph) > 4096
    s = session(monkeypatch, None, backend, sph, alpha, k, True, 2.0)
    assert not uses_fft(s)
    s.close()
    # a missing node breaks the Toeplitz structure
    full = box(20, 16, 14)
    s = session(monkeypatch, "force", backend, np.delete(full, 100, axis=0), alpha, k, True, 2.0)
    assert not uses_fft(s)
    s.close()
    # substrate and lattice solves keep their own kernels
    s = session(monkeypatch, "force", backend, full, alpha, k, True, 1.5,
                substrate=SubstrateRuntime(z_interface=-1.0, delta_eps=0.3 + 0.01j))
    assert not uses_fft(s)
    s.close()
    lat = LatticeSpec((60.0, 0.0, 0.0), (0.0, 60.0, 0.0), truncation_real=1, truncation_recip=1)
    s = session(monkeypatch, "force", backend, box(6, 6, 3), alpha, k, True, 1.5, lattice=lat)
    assert not uses_fft(s)
    s.close()


def test_fft_solves_match_pair_sum(backend, monkeypatch):
    lam, n_env, d = 520.0, 1.33, 1.5
    k = I.wavenumber(lam, n_env)
    alpha = I.polarisability_block("Au", lam, d, n_env)
    pos = box(20, 16, 14, seed=4)
    inc = IncidentField(direction=(0.0, 0.6, 0.8), polarisation=(1.0, 0.0, 0.0))
    res = {}
    for name, mode, prec, graph in (("pairs", "0", FP64, True), ("fft", None, FP64, True),
                                    ("fft_eager", None, FP64, False),
                                    ("fft_mixed", None, MIXED, True)):
        s = session(monkeypatch, mode, backend, pos, alpha, k, True, d, precision=prec,
                    graph=graph)
        out, p, info = s.compute_cross_sections(inc, 1e-10, 2000)
        res[name] = (out, p, info)
        print(name, out, info)
        s.close()
    cs0, p0, i0 = res["pairs"]
    cs1, p1, i1 = res["fft"]
    assert i0["pair_evals"] > 0 and i1["pair_evals"] == 0
    assert np.all(np.abs(cs1 - cs0) <= 1e-10 * np.abs(cs0))
    assert relmax(p1, p0) < 1e-8
    assert abs(i1["iterations"] - i0["iterations"]) <= 2
    # a replayed cycle graph issues the same kernels as eager issue
    cs2, p2, i2 = res["fft_eager"]
    assert np.array_equal(p2, p1) and np.array_equal(cs2, cs1)
    # mixed: FP32 FFT inner iterations, FP64 residuals
    cs3, p3, i3 = res["fft_mixed"]
    assert i3["matvecs_fp32"] > 0
    assert abs(cs3[0] - cs0[0]) <= 1e-8 * abs(cs0[0])


def test_fft_after_wavelength_change(backend, monkeypatch):
    """set_wavelength recomputes the tensor spectrum."""
    d = 1.5
    pos = box(18, 16, 15)
    k1, k2 = I.wavenumber(520.0, 1.33), I.wavenumber(640.0, 1.33)
    a1 = I.polarisability_block("Au", 520.0, d, 1.33)
    a2 = I.polarisability_block("Au", 640.0, d, 1.33)
    s = session(monkeypatch, None, backend, pos, a1, k1, True, d)
    assert uses_fft(s)
    s.set_wavelength(k2, a2)
    x = np.random.default_rng(8).standard_normal(3 * len(pos)) + 0j
    y = s.matvec(x)
    yo = O.Problem(pos, a2, k2, True, d).matvec(x)
    assert relmax(y, yo) < 1e-12
    s.close()


def test_fft_adjoint_gradients(backend, monkeypatch):
    """Forward and adjoint solves on the FFT path; gradients as on the
    pair-sum path."""
    lam, n_env, d = 560.0, 1.33, 2.0
    k = I.wavenumber(lam, n_env)
    eps = I.material("Au").epsilon(lam)
    alpha = I.polarisability_block("Au", lam, d, n_env)
    D = I.inverse_polarisability_derivative(eps, d, n_env)
    pos = box(17, 16, 16, h=(d, d, d))
    out = {}
    for mode in ("0", None):
        s = session(monkeypatch, mode, backend, pos, alpha, k, True, d)
        cs, ge, gp, fi, ai = s.adjoint_grad(IncidentField(), 1e-12, 2000, D)
        out[mode] = (cs, ge, gp, fi)
        s.close()
    assert out[None][3]["pair_evals"] == 0 and out["0"][3]["pair_evals"] > 0
    print("adjoint fft vs pairs", relmax(out[None][1], out["0"][1]),
          relmax(out[None][2], out["0"][2]))
    assert relmax(out[None][1], out["0"][1]) < 1e-9
    assert relmax(out[None][2], out["0"][2]) < 1e-9


def test_full_size_cuboid_fft_vs_pair_sum(backend, monkeypatch):
    """The 80 x 40 x 40 nm cuboid at 1.5 nm (38,637 dipoles, a full grid):
    C_ext of the FFT solve against the pair-sum solve, FP64, tol 1e-10."""
    pos = I.au_cuboid_80x40x40_d1p5()
    lam, n_env, d = 520.0, 1.33, 1.5
    k = I.wavenumber(lam, n_env)
    alpha = I.polarisability_block("Au", lam, d, n_env)
    cs = {}
    for mode in ("0", None):
        s = session(monkeypatch, mode, backend, pos, alpha, k, True, d)
        out, p, info = s.compute_cross_sections(IncidentField(), 1e-10, 1000)
        cs[mode] = (out, info)
        print(mode, out, info)
        s.close()
    assert cs[None][1]["pair_evals"] == 0
    assert np.all(np.abs(cs[None][0] - cs["0"][0]) <= 1e-10 * np.abs(cs["0"][0]))


def test_fft_with_offset_dependent_fcd_flags(backend, monkeypatch):
    """Classification positions that move one pair across 2 cell_size make
    the FCD flags depend on more than the grid offset: the near list then
    carries the per-pair FCD correction on top of the FFT point sum."""
    h = 1.5
    pos = box(20, 16, 14, h=(h, h, h), origin=(0.0, 0.0, 0.0))
    cls = pos.copy()
    cls[(1 * 16 + 3) * 14 + 4, 0] -= 1e-9      # ix = 1: its pair with ix = 3 leaves FCD
    k = I.wavenumber(520.0, 1.33)
    alpha = I.polarisability_block("Au", 520.0, h, 1.33)
    s = session(monkeypatch, None, backend, pos, alpha, k, True, h, classify_positions=cls)
    assert uses_fft(s)
    x = np.random.default_rng(9).standard_normal(3 * len(pos)) * (0.3 - 1j)
    y = s.matvec(x)
    yo = O.Problem(pos, alpha, k, True, h, classify_positions=cls).matvec(x)
    y_plain = O.Problem(pos, alpha, k, True, h).matvec(x)
    print("offset-dependent flags", relmax(y, yo), relmax(y_plain, yo))
    assert relmax(y_plain, yo) > 1e-6           # the flipped pair matters
    assert relmax(y, yo) < 1e-12
    s.close()


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_fft_solve_behind_other_solves_in_a_batch(backend, monkeypatch, precision):
    """An FFT solve whose partial slices do not start the buffer (a pair-sum
    solve before it in the batch) gives the same answer, bit for bit, as the
    same solve alone in a batch: in mixed precision its FP32 slice is written
    and read at the same element offset (a mismatch there once read unwritten
    slices as NaN)."""
    from lumina_amd.solver import Batch
    monkeypatch.setenv("LUMINA_FFT", "force")
    monkeypatch.setenv("LUMINA_FAR_BLOCK", "256")
    h = 1.5
    cub = box(6, 5, 4, h=(h, h, h), origin=(-4.0, 2.0, 1.0))
    k = I.wavenumber(600.0, 1.33)
    fft = dict(positions=cub, k=k, use_fcd=True, cell_size=h,
               polarisability=I.polarisability_block("Au", 600.0, h, 1.33))
    sph = I.au_sphere_r10_d2()
    other = dict(positions=sph, k=k, use_fcd=True, cell_size=2.0,
                 polarisability=I.polarisability_block("Au", 600.0, 2.0, 1.33))
    out = {}
    for name, probs in (("alone", [fft]), ("behind", [other, fft])):
        b = Batch(backend, probs)
        b.run(1e-10, 150, precision)
        try:
            cs, infos, p = b.results(want_moments=True)
        except ConvergenceFailure as e:
            cs, infos, p = e.cross_sections, e.infos, e.moments
        b.close()
        out[name] = (cs[-1], infos[-1], p[-1])
        print(name, cs[-1], infos[-1])
    assert out["alone"][1]["pair_evals"] == 0
    assert np.all(np.isfinite(out["behind"][0]))
    assert out["behind"][1]["iterations"] == out["alone"][1]["iterations"]
    assert np.array_equal(out["behind"][0], out["alone"][0])
    assert np.array_equal(out["behind"][2], out["alone"][2])
