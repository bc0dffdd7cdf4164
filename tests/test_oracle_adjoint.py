"""The oracle's adjoint C_ext gradients against central finite differences
of C_ext computed from exact (LAPACK) solves of the oracle's assembled
operator, with the FCD near/far classification frozen at the unperturbed
positions. The gradient pair sums (dG/dr) and the d(alpha^-1)/d eps model are
pinned here by finite differences, which share no formula with them."""
import numpy as np
import pytest

from lumina_amd import inputs as I
from oracle import pyoracle as O


def test_point_gradient_matches_finite_differences():
    k = I.wavenumber(600.0, 1.33)
    rs = np.array([0.3, -0.2, 0.1])
    for ro in ([3.0, 1.0, -2.0], [0.0, 0.0, 2.0], [40.0, -25.0, 10.0]):
        ro = np.array(ro)
        g = O.greens_point_grad(ro, rs, k)
        r = np.linalg.norm(ro - rs)
        h = 1e-5 * r
        for c in range(3):
            e = np.zeros(3)
            e[c] = h
            fd = (O.greens_point(ro + e, rs, k) - O.greens_point(ro - e, rs, k)) / (2 * h)
            assert np.max(np.abs(g[c] - fd)) < 1e-7 * np.max(np.abs(g)), (ro, c)


def _cext(pos, alphas, k, lam_nm, pol, cls=None, cell=2.0):
    P = O.Problem(pos, alphas, k, True, cell, classify_positions=cls)
    e = P.incident(polarisation=pol)
    p = np.linalg.solve(P.assemble(), e)
    return P.cross_sections(p, e)[0], P, p, e


@pytest.mark.parametrize("pol", [(1, 0, 0), (1 / np.sqrt(2), 1j / np.sqrt(2), 0)])
def test_adjoint_matches_finite_differences(pol):
    lam, n_env, d = 560.0, 1.33, 2.0
    pos = I.discretise(I.Sphere(4.5), d)
    n = len(pos)
    k = I.wavenumber(lam, n_env)
    eps0 = I.material("Au").epsilon(lam)
    rng = np.random.default_rng(11)
    eps = eps0 * (1.0 + 0.05 * rng.standard_normal(n))       # per-dipole permittivity
    alphas = np.array([I.polarisability_from_epsilon(e, lam, d, n_env) for e in eps])
    c0, P, p, e = _cext(pos, alphas, k, lam, pol)
    A = P.assemble()
    assert np.max(np.abs(A - A.T)) < 1e-12 * np.max(np.abs(A))   # complex symmetric
    lam_v = np.linalg.solve(A, np.conj(e))
    dainv = np.array([I.inverse_polarisability_derivative(x, d, n_env) for x in eps])
    ge, gr = P.adjoint_grad(p, lam_v, dainv, polarisation=pol)
    assert gr is not None
    scale_e = np.max(np.abs(ge))
    scale_r = np.max(np.abs(gr))
    for i in rng.choice(n, 5, replace=False):
        for part, de in ((0, 1e-6), (1, 1e-6j)):
            h = abs(de) * abs(eps[i])
            cs = []
            for sgn in (1, -1):
                e2 = eps.copy()
                e2[i] = eps[i] + sgn * h * (de / abs(de))
                a2 = alphas.copy()
                a2[i] = I.polarisability_from_epsilon(e2[i], lam, d, n_env)
                cs.append(_cext(pos, a2, k, lam, pol)[0])
            fd = (cs[0] - cs[1]) / (2 * h)
            assert abs(fd - ge[i, part]) < 1e-6 * scale_e, (i, part, fd, ge[i, part])
        for c in range(3):
            h = 1e-5
            cs = []
            for sgn in (1, -1):
                p2 = pos.copy()
                p2[i, c] += sgn * h
                cs.append(_cext(p2, alphas, k, lam, pol, cls=pos)[0])
            fd = (cs[0] - cs[1]) / (2 * h)
            assert abs(fd - gr[i, c]) < 1e-6 * scale_r, (i, c, fd, gr[i, c])


def test_adjoint_with_substrate_gives_epsilon_gradient_only():
    lam, n_env, d = 560.0, 1.0, 2.0
    pos = I.discretise(I.Sphere(3.0), d)
    k = I.wavenumber(lam, n_env)
    a = I.polarisability_block("Au", lam, d, n_env)
    P = O.Problem(pos, a, k, True, d, substrate=(-3.0, 0.2 + 0.0j))
    e = P.incident()
    A = P.assemble()
    # the Fresnel image operator keeps A complex symmetric
    assert np.max(np.abs(A - A.T)) < 1e-12 * np.max(np.abs(A))
    p = np.linalg.solve(A, e)
    lam_v = np.linalg.solve(A, np.conj(e))
    ge, gr = P.adjoint_grad(p, lam_v, I.inverse_polarisability_derivative(
        I.material("Au").epsilon(lam), d, n_env))
    assert gr is None and np.all(np.isfinite(ge))


def test_lattice_operator_is_not_symmetric():
    """With a 2D lattice the operator is not complex symmetric, even at
    k_par = 0: for two dipoles at the same height the spectral sum's xz and yz
    entries take sgn(0) = +1 for r and for -r (ewald.rs:186-273), so
    G_per(-r) differs from G_per(r) there. The adjoint therefore reuses A only
    without a lattice."""
    lam, n_env, d = 600.0, 1.0, 2.0
    pos = I.discretise(I.Cylinder(4.0, 4.0), d)
    k = I.wavenumber(lam, n_env)
    a = I.polarisability_block("Au", lam, d, n_env)
    lat = dict(a1=(40.0, 0.0, 0.0), a2=(0.0, 40.0, 0.0), eta=0.0, n_real=3, n_recip=3)
    A = O.Problem(pos, a, k, True, d, lattice=lat).assemble()
    assert np.max(np.abs(A - A.T)) > 1e-3 * np.max(np.abs(A))
    # the asymmetric entries are exactly the in-plane pairs' xz / yz entries
    n = len(pos)
    D = np.abs(A - A.T).reshape(n, 3, n, 3).transpose(0, 2, 1, 3)   # [i, j, a, b]
    same_z = np.abs(pos[:, None, 2] - pos[None, :, 2]) < 1e-12
    assert np.max(D[~same_z]) < 1e-12 * np.max(np.abs(A))
