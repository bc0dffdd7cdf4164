"""Host-side geometry preparation (lumina_amd/csrc/lh_host.cpp), compiled
with g++ and checked on the CPU: the near list against a brute-force FP32
exclusion test, the distinct-displacement table and its FCD flags against the
reference's test (greens.rs:102-106), grid detection, and the Ewald records
against the oracle's lattice helpers (lattice.rs:75-233)."""
import os
import subprocess

import numpy as np
import pytest

from lumina_amd import inputs as I
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lumina_amd", "csrc")

HARNESS = r"""
#include <cstdio>
#include <vector>
#include "lh_host.h"
using namespace lh;
int main() {
  int mode;
  if (std::scanf("%d", &mode) != 1) return 1;
  if (mode == 0 || mode == 1 || mode == 3) {
    // near list (0) or grid (1): n fcd cell rskip2f has_cls, then positions, then cls
    Geometry g;
    int n, fcd, has_cls;
    double cell, rskip;
    std::scanf("%d %d %lf %lf %d", &n, &fcd, &cell, &rskip, &has_cls);
    g.n = n; g.use_fcd = fcd; g.cell = cell; g.rskip2f = (float)rskip;
    g.x.resize(n); g.y.resize(n); g.z.resize(n);
    for (int i = 0; i < n; ++i) std::scanf("%lf %lf %lf", &g.x[i], &g.y[i], &g.z[i]);
    if (has_cls) {
      g.cls.resize(3 * n);
      for (int i = 0; i < 3 * n; ++i) std::scanf("%lf", &g.cls[i]);
    }
    if (mode == 3) {
      build_near_list(g);
      detect_grid(g);
      std::vector<int> idx, offs;
      const bool full = full_grid_indices(g, idx);
      const bool ok = full && fcd_offsets(g, idx, offs);
      std::printf("%d %d %d\n", full ? 1 : 0, ok ? 1 : 0, (int)offs.size() / 3);
      for (size_t q = 0; q < offs.size(); q += 3)
        std::printf("%d %d %d\n", offs[q], offs[q + 1], offs[q + 2]);
    } else if (mode == 0) {
      build_near_list(g);
      std::printf("%d %lld %d\n", g.kw, g.nnz, g.ntab);
      for (int kk = 0; kk < g.kw; ++kk)
        for (int i = 0; i < n; ++i) {
          const size_t e = (size_t)kk * n + i;
          std::printf("%d %d %d\n", i, g.ell_col[e], g.ell_tid[e]);
        }
      for (int u = 0; u < g.ntab; ++u)
        std::printf("%.17g %.17g %.17g %g\n", g.disp[4 * u], g.disp[4 * u + 1], g.disp[4 * u + 2],
                    g.disp[4 * u + 3]);
    } else {
      detect_grid(g);
      std::vector<int> idx;
      const bool full = full_grid_indices(g, idx);
      std::printf("%d %d %d %d %.17g %.17g %.17g %d\n", g.grid, g.gn[0], g.gn[1], g.gn[2],
                  g.gh[0], g.gh[1], g.gh[2], full ? 1 : 0);
      if (g.grid)
        for (int i = 0; i < n; ++i)
          std::printf("%d %d %d %d %d\n", g.gpack[i], g.gizj2[i], full ? idx[3 * i] : -1,
                      full ? idx[3 * i + 1] : -1, full ? idx[3 * i + 2] : -1);
    }
  } else if (mode == 4) {
    // natural spline second derivatives: n, then n (x, y) pairs
    int n;
    std::scanf("%d", &n);
    std::vector<double> xs(n), ys(n), y2(n);
    for (int i = 0; i < n; ++i) std::scanf("%lf %lf", &xs[i], &ys[i]);
    natural_spline(xs.data(), ys.data(), n, y2.data());
    for (int i = 0; i < n; ++i) std::printf("%.17g\n", y2[i]);
  } else if (mode == 6) {
    // far-field phase range: n k has_sub z_interface, then positions
    Geometry g;
    Problem p;
    int n;
    std::scanf("%d %lf %d %lf", &n, &p.k, &p.has_sub, &p.z_interface);
    g.n = n;
    g.x.resize(n); g.y.resize(n); g.z.resize(n);
    for (int i = 0; i < n; ++i) std::scanf("%lf %lf %lf", &g.x[i], &g.y[i], &g.z[i]);
    const double rev = max_phase_revolutions(g, p);
    std::printf("%.17g %d\n", rev, rev > kMaxHwRevolutions ? 1 : 0);
  } else {
    // Ewald records: a1 a2 eta nr nq kb k
    Problem p;
    for (int c = 0; c < 3; ++c) std::scanf("%lf", &p.a1[c]);
    for (int c = 0; c < 3; ++c) std::scanf("%lf", &p.a2[c]);
    std::scanf("%lf %d %d", &p.eta, &p.n_real, &p.n_recip);
    for (int c = 0; c < 3; ++c) std::scanf("%lf", &p.kb[c]);
    std::scanf("%lf", &p.k);
    std::vector<double> rec;
    int nr, nq;
    double eta;
    ewald_records(p, rec, nr, nq, eta);
    std::printf("%d %d %.17g\n", nr, nq, eta);
    for (size_t q = 0; q < rec.size(); q += 6)
      std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", rec[q], rec[q + 1], rec[q + 2],
                  rec[q + 3], rec[q + 4], rec[q + 5]);
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("host")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas",
                    "-I", CSRC, str(src), os.path.join(CSRC, "lh_host.cpp"), "-o", str(exe)],
                   check=True)
    return exe


def run(exe, text):
    return subprocess.run([str(exe)], input=text, capture_output=True, text=True,
                          check=True).stdout.split("\n")


def near_list(exe, pos, cell, fcd=True, cls=None):
    r = 2.0 * cell * 1.05
    rskip = np.float32(r * r)
    lines = [f"0 {len(pos)} {1 if fcd else 0} {cell!r} {float(rskip)!r} {0 if cls is None else 1}"]
    lines += [" ".join(repr(float(v)) for v in p) for p in pos]
    if cls is not None:
        lines += [" ".join(repr(float(v)) for v in p) for p in cls]
    out = run(exe, "\n".join(lines) + "\n")
    kw, nnz, ntab = map(int, out[0].split())
    n = len(pos)
    ell = np.array([list(map(int, l.split())) for l in out[1:1 + kw * n]]).reshape(-1, 3)
    disp = np.array([list(map(float, l.split())) for l in out[1 + kw * n:1 + kw * n + ntab]])
    return kw, nnz, ntab, ell, disp, rskip


def brute_near(pos, rskip):
    p32 = pos.astype(np.float32)
    d = p32[:, None, :] - p32[None, :, :]
    sq = d * d                                   # float32, no contraction
    r2 = (sq[..., 0] + sq[..., 1]) + sq[..., 2]
    m = r2 <= rskip
    np.fill_diagonal(m, False)
    return m


@pytest.mark.parametrize("case", ["sphere", "random", "cli_cell"])
def test_near_list_matches_brute_force(harness, case):
    if case == "sphere":
        pos, cell = I.au_sphere_r10_d2(), 2.0
    elif case == "cli_cell":
        pos, cell = I.au_sphere_r10_d2(), 3.0
    else:
        pos, cell = np.random.default_rng(4).uniform(-9, 9, size=(400, 3)), 1.7
    kw, nnz, ntab, ell, disp, rskip = near_list(harness, pos, cell)
    want = brute_near(pos, rskip)
    got = np.zeros_like(want)
    for i, j, u in ell:
        if u < 0:
            continue
        assert not got[i, j]
        got[i, j] = True
        # the table entry is this pair's FP64 displacement, flagged by the
        # reference's FCD test
        dd = pos[i] - pos[j]
        assert np.array_equal(disp[u, :3], dd)
        dist = np.sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2])
        assert disp[u, 3] == (0.0 if dist > 2.0 * cell else 1.0)
    assert np.array_equal(got, want)
    assert nnz == want.sum()
    # rows are padded at the end only, and the ELL width is the longest row
    per_row = want.sum(axis=1)
    assert kw == per_row.max()
    tid = ell[:, 2].reshape(kw, len(pos))
    for i in range(len(pos)):
        assert np.all(tid[:per_row[i], i] >= 0) and np.all(tid[per_row[i]:, i] == -1)
    # a lattice-built particle has few distinct displacements
    if case != "random":
        assert ntab < 200


def test_near_list_classification_positions(harness):
    """A dipole moved by 1e-6 keeps the FCD flags of its unperturbed pairs
    when the classification positions are given."""
    pos = I.au_sphere_r10_d2()
    moved = pos.copy()
    moved[100, 0] += 1e-6
    _, _, _, ell, disp, _ = near_list(harness, moved, 2.0, cls=pos)
    for i, j, u in ell:
        if u < 0:
            continue
        d0 = pos[i] - pos[j]
        dist0 = np.sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2])
        assert disp[u, 3] == (0.0 if dist0 > 2.0 * 2.0 else 1.0)
        assert np.array_equal(disp[u, :3], moved[i] - moved[j])


def grid(exe, pos, full=False):
    lines = [f"1 {len(pos)} 0 1.0 0.0 0"] + [" ".join(repr(float(v)) for v in p) for p in pos]
    out = run(exe, "\n".join(lines) + "\n")
    head = out[0].split()
    ok, gn, gh = int(head[0]), list(map(int, head[1:4])), list(map(float, head[4:7]))
    packs = np.array([list(map(int, l.split())) for l in out[1:1 + len(pos)]]) if ok else None
    if full:
        return int(head[7]), packs[:, 2:] if ok else None
    return ok, gn, gh, packs[:, :2] if ok else None


def test_grid_detection(harness):
    pos = I.discretise(I.Cylinder(10.0, 8.0), 2.0)
    ok, gn, gh, packs = grid(harness, pos)
    assert ok == 1 and np.allclose(gh, 2.0)
    idx = np.rint((pos - pos.min(axis=0)) / 2.0).astype(int)
    assert np.array_equal(np.array(gn), idx.max(axis=0) + 1)
    nyd, nzd = 2 * gn[1] - 1, 2 * gn[2] - 1
    assert np.array_equal(packs[:, 0], idx[:, 0] * nyd * nzd + idx[:, 1] * nzd + idx[:, 2])
    assert np.array_equal(packs[:, 1], 2 * idx[:, 2] - (gn[2] - 1))
    # the table index pack_i - pack_j + centre covers [0, NXd NYd NZd)
    centre = (gn[0] - 1) * nyd * nzd + (gn[1] - 1) * nzd + (gn[2] - 1)
    e = packs[:, 0][:, None] - packs[:, 0][None, :] + centre
    assert e.min() >= 0 and e.max() < (2 * gn[0] - 1) * nyd * nzd
    # off-grid positions are refused; a single plane is a grid
    off = pos.copy()
    off[7, 1] += 1e-7
    assert grid(harness, off)[0] == 0
    assert grid(harness, np.random.default_rng(0).uniform(-5, 5, (50, 3)))[0] == 0
    plane = pos[np.abs(pos[:, 2] - pos[:, 2].min()) < 1e-12]
    okp, gnp, _, _ = grid(harness, plane)
    assert okp == 1 and gnp[2] == 1


def test_ewald_records_match_lattice_helpers(harness):
    a1, a2 = (300.0, 0.0, 0.0), (40.0, 250.0, 0.0)
    k = 2 * np.pi / 600.0
    kb = (0.002, -0.001, 0.0)
    line = " ".join(repr(float(v)) for v in (*a1, *a2)) + f" 0.0 4 3 " + \
        " ".join(repr(float(v)) for v in kb) + f" {k!r}"
    out = run(harness, "2 " + line + "\n")
    nr, nq, eta = out[0].split()
    nr, nq, eta = int(nr), int(nq), float(eta)
    rec = np.array([list(map(float, l.split())) for l in out[1:1 + nr + nq]])
    assert nr == 81 and abs(eta - O.lattice_eta(a1, a2)) < 1e-15
    # real-space images: m1 outer, m2 inner, zero flag and Bloch phase
    m = [(m1, m2) for m1 in range(-4, 5) for m2 in range(-4, 5)]
    R = np.array([[m1 * a1[c] + m2 * a2[c] for c in range(3)] for m1, m2 in m])
    assert np.allclose(rec[:nr, :3], R, rtol=0, atol=1e-12)
    assert np.array_equal(rec[:nr, 3], np.all(np.abs(R) < 1e-12, axis=1).astype(float))
    ph = R @ np.array(kb)
    assert np.allclose(rec[:nr, 4], np.cos(ph), atol=1e-15)
    assert np.allclose(rec[:nr, 5], np.sin(ph), atol=1e-15)
    # spectral terms: Q = k_par + G, q = csqrt+(|Q|^2 - k^2), c = (i / 2A) / q
    b1, b2 = O.recip_vectors(a1, a2)
    area = abs(a1[0] * a2[1] - a1[1] * a2[0])
    want = []
    for n1 in range(-3, 4):
        for n2 in range(-3, 4):
            qx = kb[0] + (n1 * b1[0] + n2 * b2[0])
            qy = kb[1] + (n1 * b1[1] + n2 * b2[1])
            q = O.csqrt_positive_imag(complex(qx * qx + qy * qy - k * k))
            if abs(q) < 1e-10:
                continue
            c = (1j / (2 * area)) / q
            want.append([qx, qy, q.real, q.imag, c.real, c.imag])
    want = np.array(want)
    assert nq == len(want)
    assert np.allclose(rec[nr:], want, rtol=1e-13, atol=1e-18)


def test_full_grid_indices(harness):
    """A box with every node occupied once (fft_matvec.rs:20-75, the
    reference's make_grid and its displaced-dipole case) gives each dipole's
    (ix, iy, iz); a missing, displaced or doubled node does not."""
    n = 4
    pos = np.array([[i * 1.0, j * 1.0, k * 1.0] for i in range(n) for j in range(n)
                    for k in range(n)])
    full, idx = grid(harness, pos, full=True)
    assert full == 1
    assert np.array_equal(idx, np.rint(pos).astype(int))
    perm = np.random.default_rng(1).permutation(len(pos))
    aniso = pos[perm] * np.array([1.5, 2.0, 0.5]) + np.array([-3.0, 7.0, 0.25])
    full, idx = grid(harness, aniso, full=True)
    assert full == 1 and np.array_equal(idx, np.rint(pos[perm]).astype(int))
    moved = pos.copy()
    moved[7, 0] += 0.1
    assert grid(harness, moved, full=True)[0] == 0
    assert grid(harness, pos[1:], full=True)[0] == 0
    doubled = pos.copy()
    doubled[5] = doubled[6]
    assert grid(harness, doubled, full=True)[0] == 0
    # one plane and one line are full grids too
    assert grid(harness, pos[pos[:, 2] == 0.0], full=True)[0] == 1
    assert grid(harness, pos[(pos[:, 1] == 0.0) & (pos[:, 2] == 0.0)], full=True)[0] == 1


def fcd_offsets(exe, pos, cell, cls=None):
    r = 2.0 * cell * 1.05
    rskip = np.float32(r * r)
    lines = [f"3 {len(pos)} 1 {cell!r} {float(rskip)!r} {0 if cls is None else 1}"]
    lines += [" ".join(repr(float(v)) for v in p) for p in pos]
    if cls is not None:
        lines += [" ".join(repr(float(v)) for v in p) for p in cls]
    out = run(exe, "\n".join(lines) + "\n")
    full, ok, m = map(int, out[0].split())
    offs = np.array([list(map(int, l.split())) for l in out[1:1 + m]]).reshape(-1, 3)
    return full, ok, offs


def test_fcd_offsets_of_a_full_grid(harness):
    """The FCD flag as a function of the grid offset: on a 6 x 5 x 4 box at
    the cell spacing every offset with |o| h <= 2 cell is FCD (the reference's
    dist <= 2 cell_size, greens.rs:102-106); a classification position that
    moves one pair across 2 cell makes the flags offset-dependent."""
    h = 1.5
    pos = np.array([[i * h, j * h, k * h] for i in range(6) for j in range(5)
                    for k in range(4)])
    full, ok, offs = fcd_offsets(harness, pos, h)
    assert full == 1 and ok == 1
    want = sorted((a, b, c) for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3)
                  if (a, b, c) != (0, 0, 0) and a * a + b * b + c * c <= 4)
    assert [tuple(o) for o in offs] == want
    # a cell of 1.2 h: offsets up to |o| h <= 2.4 h, i.e. |o|^2 <= 5.76
    full, ok, offs = fcd_offsets(harness, pos, 1.2 * h)
    assert ok == 1 and len(offs) == sum(1 for a in range(-2, 3) for b in range(-2, 3)
                                        for c in range(-2, 3)
                                        if (a, b, c) != (0, 0, 0) and a * a + b * b + c * c <= 5)
    cls = pos.copy()
    cls[37, 0] -= 1e-9        # (ix = 1): the pair with ix = 3 now sits just beyond 2 cell
    full, ok, _ = fcd_offsets(harness, pos, h, cls=cls)
    assert full == 1 and ok == 0


@pytest.mark.parametrize("name", ["Au", "Ag"])
def test_natural_spline_matches_restatement(harness, name):
    """The spline second derivatives of the sweep set-up (spline.rs:32-66)
    against the restated NaturalSpline of the host inputs, on both
    Johnson-Christy tables (Re and Im eps), and the two-point case (a
    straight line: y2 = 0)."""
    m = I.material(name)
    for spl in (m.re, m.im):
        text = f"4 {len(spl.xs)}\n" + "\n".join(f"{x!r} {y!r}" for x, y in zip(spl.xs, spl.ys))
        y2 = np.array([float(v) for v in run(harness, text + "\n")[:len(spl.xs)]])
        assert np.array_equal(y2, np.array(spl.y2))
    y2 = [float(v) for v in run(harness, "4 2\n300 2.25\n1200 2.5\n")[:2]]
    assert y2 == [0.0, 0.0]


def phase_range(exe, pos, k, sub=False, z0=0.0):
    lines = [f"6 {len(pos)} {k!r} {1 if sub else 0} {z0!r}"]
    lines += [" ".join(repr(float(v)) for v in p) for p in pos]
    out = run(exe, "\n".join(lines) + "\n")
    rev, off = out[0].split()
    return float(rev), int(off) == 1


def test_hardware_sincos_range_decision(harness):
    """The FP32 far field leaves the hardware sin/cos (valid to 256
    revolutions) once k times the bounding-box diagonal exceeds 200
    revolutions; the engine takes this decision at build and again whenever
    a solve is re-targeted to a new k (set_problem)."""
    pos = np.array([[0.0, 0.0, 0.0], [3000.0, 4000.0, 0.0], [1000.0, 0.0, 0.0]])
    diag = 5000.0
    for lam in (500.0, 26.0, 24.0):
        k = 2 * np.pi / lam
        rev, off = phase_range(harness, pos, k)
        assert abs(rev - diag / lam) < 1e-9 * rev
        assert off == (diag / lam > 200.0)
    # with a substrate the image of a source lies up to 2 diag + 2|z0| + 2 max|z| away
    pos2 = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 5.0]])
    k = 2 * np.pi / 10.0
    rev, _ = phase_range(harness, pos2, k, sub=True, z0=-2.0)
    assert abs(rev - (2 * 4.0 + 4.0 + 10.0) / 10.0) < 1e-12
    # a single dipole has no phase range
    assert phase_range(harness, pos[:1], k)[0] == 0.0
