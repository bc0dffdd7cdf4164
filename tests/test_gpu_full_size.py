"""Size-independent properties of the matvec at the full workload sizes,
where the oracle is too slow to run: the 38,637-dipole cuboid and the
35,013-dipole helix.

- Complex symmetry: the free-space operator A = alpha^-1 - G is complex
  symmetric (G(r_i, r_j) = G(r_j, r_i)^T, greens.rs:175-190, with a uniform
  symmetric alpha block), so y^T (A x) = x^T (A y) for any x, y. The
  one-tensor-per-pair kernels write every pair's column sum through a
  separate path (lane moves, the transposed partial slice), so a misplaced
  column sum breaks this identity.
- Linearity: A (a x + b y) = a A x + b A y.
- Agreement of the far-field forms: row kernel, one-tensor-per-pair kernel
  and (for the full-grid cuboid) the FFT matvec.

Bars: FP64 1e-12 relative; FP32 far field (mixed) 2e-6 relative, the FP32
matvec bar of the parity tests.
"""
import numpy as np
import pytest

from lumina_amd import inputs as I
from lumina_amd.solver import FP64, MIXED

pytestmark = pytest.mark.gpu


def rand_c(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def geometry(name):
    if name == "cuboid":
        return I.au_cuboid_80x40x40_d1p5(), 1.5, 520.0
    return I.au_helix_r25_p40_t4_w8_d1p5(), 1.5, 600.0


@pytest.fixture(scope="module")
def problems():
    out = {}
    for name in ("cuboid", "helix"):
        pos, d, lam = geometry(name)
        out[name] = (pos, I.polarisability_block("Au", lam, d, 1.33), I.wavenumber(lam, 1.33), d)
    return out


@pytest.mark.parametrize("name", ["cuboid", "helix"])
@pytest.mark.parametrize("form", ["512", "sym", "1024", "256"])
def test_complex_symmetry_and_linearity(backend, problems, monkeypatch, name, form):
    pos, alpha, k, d = problems[name]
    monkeypatch.setenv("LUMINA_FFT", "0")
    monkeypatch.setenv("LUMINA_FAR_BLOCK", form)
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=d)
    n3 = 3 * len(pos)
    x, y = rand_c(n3, 1), rand_c(n3, 2)
    for prec, bar in ((FP64, 1e-12), (MIXED, 2e-6)):
        ax, ay = s.matvec(x, prec), s.matvec(y, prec)
        lhs, rhs = y @ ax, x @ ay
        scale = np.linalg.norm(y) * np.linalg.norm(ax)
        sym_err = abs(lhs - rhs) / scale
        a, b = 0.7 - 1.3j, 2.1 + 0.4j
        lin = s.matvec(a * x + b * y, prec)
        lin_err = np.max(np.abs(lin - (a * ax + b * ay))) / np.max(np.abs(lin))
        print(f"{name} form {form} prec {prec}: symmetry {sym_err:.2e}, linearity {lin_err:.2e}")
        assert sym_err < bar
        assert lin_err < bar
    s.close()


@pytest.mark.parametrize("name", ["cuboid", "helix"])
def test_far_field_forms_agree(backend, problems, monkeypatch, name):
    pos, alpha, k, d = problems[name]
    x = rand_c(3 * len(pos), 3)
    out = {}
    for form in ("256", "sym", "1024"):
        monkeypatch.setenv("LUMINA_FFT", "0")
        monkeypatch.setenv("LUMINA_FAR_BLOCK", form)
        s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=d)
        out[form] = (s.matvec(x, FP64), s.matvec(x, MIXED))
        s.close()
    ref = out["256"][0]
    e64 = max(np.max(np.abs(out[f][0] - ref)) / np.max(np.abs(ref)) for f in ("sym", "1024"))
    e32 = max(np.max(np.abs(out[f][1] - ref)) / np.max(np.abs(ref)) for f in out)
    print(f"{name}: FP64 unordered-pair forms vs row {e64:.2e}; FP32 vs FP64 row {e32:.2e}")
    assert e64 < 1e-12
    assert e32 < 2e-6
    if name == "cuboid":
        monkeypatch.setenv("LUMINA_FFT", "force")
        s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=d)
        ef = np.max(np.abs(s.matvec(x, FP64) - ref)) / np.max(np.abs(ref))
        s.close()
        print(f"cuboid: FFT vs row {ef:.2e}")
        assert ef < 1e-12
