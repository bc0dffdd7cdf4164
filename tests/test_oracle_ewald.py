"""Pins the oracle's 2D Ewald spectral sum (EwaldGreens::recip_space,
ewald.rs:186-273) beyond the invariants the reference's own tests hold.

The reference documents the spectral term as

    G_spec = (i / 2A) sum_G e^{i Q.rho} e^{-q|z|} / q  (k^2 delta - kt kt^T),
    Q = k_par + G,  q = csqrt+(|Q|^2 - k^2),  kt = (Qx, Qy, i q sgn z),

with sgn 0 = +1 and terms with |q| < 1e-10 left out. Three kinds of check:

* closed forms the formula reduces to: far above the lattice (|z| >> a,
  every G != 0 evanescent and the erfc-damped real-space sum zero) only the
  (0,0) order survives, G_per = (k/2A) e^{-ik|z|} diag(1, 1, 0) at normal
  incidence (the textbook sheet field (ik/2A) e^{+ik|z|} is the conjugate of
  this times -i: the reference's convention, kept for parity), and at an
  oblique k_par = (kx, 0, 0) e^{i kx x - i kappa|z|} / (2 A kappa) times
  [[kappa^2, 0, s kx kappa], [0, k^2, 0], [s kx kappa, 0, kx^2]] with
  kappa = sqrt(k^2 - kx^2); one step closer, the first evanescent shell
  (+-b1, +-b2) adds (i/2A) e^{-q_a |z|}/q_a diag(2k^2 - 2b1^2, 2k^2, 2b1^2)
  from (+-b1, 0) and the like from (0, +-b2), which fixes the reciprocal
  vectors, the 1/q factor and the M tensor;
* a numpy evaluation of the documented sum written here from the formula
  (reciprocal vectors from b_i . a_j = 2 pi delta_ij, not from the oracle's
  code), compared term for term with the oracle for z < 0, z = 0, z > 0,
  oblique Bloch vectors and a skewed lattice;
* planted mistakes in that numpy evaluation (prefactor 1/2A, exponent
  e^{+q|z|}, kt_z = q sgn z, phase e^{-iQ.rho}, a swapped reciprocal basis),
  each of which fails one of the closed forms.
"""
import numpy as np
import pytest

from oracle import pyoracle as O


def recip_basis(a1, a2):
    """b1, b2 in the xy plane with b_i . a_j = 2 pi delta_ij."""
    A = np.array([[a1[0], a1[1]], [a2[0], a2[1]]], dtype=float)
    B = 2 * np.pi * np.linalg.inv(A).T   # rows b1, b2
    return B[0], B[1]


def spectral_numpy(a1, a2, r, kb, k, n=5, mistake=None):
    """The documented spectral sum, evaluated directly (see module doc)."""
    b1, b2 = recip_basis(a1, a2)
    if mistake == "swap_basis":
        b1, b2 = np.array([b1[1], b1[0]]), np.array([b2[1], b2[0]])
    area = abs(a1[0] * a2[1] - a1[1] * a2[0])
    pre = (1.0 if mistake == "real_prefactor" else 1j) / (2.0 * area)
    z = float(r[2])
    s = 1.0 if z >= 0 else -1.0
    g = np.zeros((3, 3), complex)
    for n1 in range(-n, n + 1):
        for n2 in range(-n, n + 1):
            Q = np.array([kb[0], kb[1]]) + n1 * b1 + n2 * b2
            q = np.sqrt(complex(Q @ Q - k * k))
            if q.imag < 0:
                q = -q
            if abs(q) < 1e-10:
                continue
            ph = np.exp((-1j if mistake == "phase_sign" else 1j) * (Q[0] * r[0] + Q[1] * r[1]))
            ez = np.exp((q if mistake == "growing" else -q) * abs(z))
            ktz = (1.0 if mistake == "real_ktz" else 1j) * q * s
            kt = np.array([Q[0], Q[1], ktz])
            M = k * k * np.eye(3) - np.outer(kt, kt)
            g += pre * ph * ez / q * M
    return g


def far_zone_normal(k, area, z):
    return (k / (2 * area)) * np.exp(-1j * k * abs(z)) * np.diag([1.0, 1.0, 0.0])


def far_zone_oblique(k, kx, area, x, z):
    kap = np.sqrt(k * k - kx * kx)
    s = 1.0 if z >= 0 else -1.0
    M = np.array([[kap * kap, 0, s * kx * kap], [0, k * k, 0], [s * kx * kap, 0, kx * kx]])
    return np.exp(1j * kx * x - 1j * kap * abs(z)) / (2 * area * kap) * M


def first_shell(k, b1, b2, area, z):
    """The (+-b1, 0) and (0, +-b2) terms at rho = 0 for a rectangular lattice."""
    out = np.zeros((3, 3), complex)
    for b, ax in ((b1, 0), (b2, 1)):
        q = np.sqrt(b * b - k * k)
        d = np.full(3, 2 * k * k)
        d[ax] = 2 * k * k - 2 * b * b
        d[2] = 2 * b * b
        out += (1j / (2 * area)) * np.exp(-q * abs(z)) / q * np.diag(d)
    return out


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


A1, A2 = [300.0, 0.0, 0.0], [0.0, 300.0, 0.0]
K500 = 2 * np.pi / 500.0   # vacuum 500 nm: lambda > a, every G != 0 evanescent


@pytest.mark.parametrize("z", [6000.0, -6000.0, 9000.0])
@pytest.mark.parametrize("rho", [(0.0, 0.0), (70.0, -130.0)])
def test_far_zone_normal_incidence(z, rho):
    want = far_zone_normal(K500, 9e4, z)
    got = O.ewald(A1, A2, [rho[0], rho[1], z], [0, 0, 0], K500)   # real + spectral
    assert rel(got, want) < 1e-12
    assert rel(spectral_numpy(A1, A2, [rho[0], rho[1], z], [0, 0, 0], K500), want) < 1e-12


@pytest.mark.parametrize("z", [8000.0, -8000.0])
def test_far_zone_oblique_bloch(z):
    kx = 0.45 * K500
    x = 55.0
    want = far_zone_oblique(K500, kx, 9e4, x, z)
    got = O.ewald(A1, A2, [x, 20.0, z], [kx, 0, 0], K500)
    assert rel(got, want) < 1e-11


def test_first_evanescent_shell_rectangular():
    a1, a2 = [300.0, 0.0, 0.0], [0.0, 200.0, 0.0]
    area = 6e4
    b1, b2 = 2 * np.pi / 300.0, 2 * np.pi / 200.0
    z = 700.0
    spec = O.ewald(a1, a2, [0.0, 0.0, z], [0, 0, 0], K500, 2)   # spectral part only
    rest = spec - far_zone_normal(K500, area, z)
    want = first_shell(K500, b1, b2, area, z)
    # next shells: (+-b1, +-b2) and (+-2 b1, 0), e^{-q|z|} at least 1e3 times smaller
    assert rel(rest, want) < 2e-3
    # the x and y entries differ (a1 != a2): a transposed basis would swap them
    assert abs(want[0, 0] - want[1, 1]) > 0.5 * abs(want[0, 0])


@pytest.mark.parametrize("case", [
    (A1, A2, [10.0, 5.0, 2.0], [0, 0, 0]),
    (A1, A2, [-40.0, 25.0, -3.5], [0, 0, 0]),
    (A1, A2, [12.0, -7.0, 0.0], [0, 0, 0]),
    (A1, A2, [0.0, 0.0, 0.0], [0, 0, 0]),
    (A1, A2, [33.0, 8.0, -12.0], [0.004, -0.002, 0.0]),
    ([80.0, 0.0, 0.0], [0.0, 80.0, 0.0], [10.0, 5.0, 2.0], [0, 0, 0]),
    ([250.0, 0.0, 0.0], [120.0, 210.0, 0.0], [30.0, -15.0, 6.0], [0.003, 0.001, 0.0]),
])
@pytest.mark.parametrize("lam", [500.0, 900.0])
def test_spectral_sum_matches_documented_formula(case, lam):
    a1, a2, r, kb = case
    k = 2 * np.pi / lam
    got = O.ewald(a1, a2, r, kb, k, 2)
    want = spectral_numpy(a1, a2, r, kb, k)
    assert rel(got, want) < 1e-12


def test_propagating_orders_are_kept():
    # a > lambda: the (+-1, 0) orders propagate (q imaginary); far above the
    # lattice they survive next to (0,0) and the oracle still follows the
    # documented sum
    a1, a2 = [700.0, 0.0, 0.0], [0.0, 700.0, 0.0]
    k = 2 * np.pi / 500.0
    r = [0.0, 0.0, 5000.0]
    got = O.ewald(a1, a2, r, [0, 0, 0], k)
    assert rel(got, spectral_numpy(a1, a2, r, [0, 0, 0], k)) < 1e-11
    assert rel(got, far_zone_normal(k, 4.9e5, 5000.0)) > 0.1


@pytest.mark.parametrize("mistake", ["real_prefactor", "growing", "real_ktz", "phase_sign",
                                     "swap_basis"])
def test_planted_mistakes_fail_a_closed_form(mistake):
    """Each plausible slip in the spectral sum fails at least one of the
    closed forms the oracle passes."""
    # a non-finite result (the growing exponential overflows) counts as failing
    fails = []
    z = 6000.0
    with np.errstate(all="ignore"):
        fails.append(not rel(spectral_numpy(A1, A2, [0.0, 0.0, z], [0, 0, 0], K500,
                                            mistake=mistake),
                             far_zone_normal(K500, 9e4, z)) <= 1e-6)
        kx, x = 0.45 * K500, 55.0
        fails.append(not rel(spectral_numpy(A1, A2, [x, 20.0, 8000.0], [kx, 0, 0], K500,
                                            mistake=mistake),
                             far_zone_oblique(K500, kx, 9e4, x, 8000.0)) <= 1e-6)
        a1, a2 = [300.0, 0.0, 0.0], [0.0, 200.0, 0.0]
        zz = 700.0
        rest = (spectral_numpy(a1, a2, [0.0, 0.0, zz], [0, 0, 0], K500, mistake=mistake)
                - far_zone_normal(K500, 6e4, zz))
        fails.append(not rel(rest, first_shell(K500, 2 * np.pi / 300.0, 2 * np.pi / 200.0, 6e4,
                                               zz)) <= 2e-3)
    print(mistake, fails)
    assert any(fails), mistake
