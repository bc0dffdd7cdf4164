"""Input preparation (geometry, materials, polarisability) against closed
forms, scipy and the dipole counts of the workloads."""
import numpy as np
import pytest
import scipy.interpolate

from lumina_amd import inputs as I


def test_workload_dipole_counts():
    assert len(I.au_sphere_r10_d2()) == 515
    assert len(I.au_cuboid_80x40x40_d1p5()) == 38637
    assert len(I.au_cylinder_d100_h40_d2()) == 41181


def test_helix_count():
    assert len(I.au_helix_r25_p40_t4_w8_d1p5()) == 35013


def test_sphere_lattice_centred_and_symmetric():
    p = I.au_sphere_r10_d2()
    assert np.allclose(p.mean(0), 0.0, atol=1e-12)
    # x-major, z-minor ordering of discretise.rs:43-55
    assert np.all(np.diff(p[:, 0]) >= 0)
    assert set(np.unique(p[:, 0])) == set(np.arange(-10, 11, 2.0))


def test_cuboid_grid_full():
    p = I.au_cuboid_80x40x40_d1p5()
    assert [len(np.unique(p[:, c])) for c in range(3)] == [53, 27, 27]


def test_spline_matches_scipy_natural():
    mat = I.material("Au")
    cs = scipy.interpolate.CubicSpline(mat.wl, [mat.re(w) for w in mat.wl], bc_type="natural")
    for w in np.linspace(190, 890, 57):
        assert abs(mat.re(w) - cs(w)) < 1e-9 * max(1.0, abs(cs(w)))


def test_spline_passes_through_table():
    mat = I.material("Ag")
    for w, (lum, n, k) in zip(mat.wl, I._JC_AG):
        e = mat.epsilon(w)
        assert abs(e.real - (n * n - k * k)) < 1e-12
        assert abs(e.imag - 2 * n * k) < 1e-12


def test_material_range():
    with pytest.raises(ValueError):
        I.material("Au").epsilon(900.0)
    I.material("Au").epsilon(892.0)


def test_polarisability_closed_forms():
    eps, em, V, k = complex(-3.9, 2.6), 1.33 ** 2, 8.0, 0.016
    a_cm = I.clausius_mossotti(V, eps, em)
    assert abs(a_cm - 3 * V * (eps - em) / (eps + 2 * em)) < 1e-12 * abs(a_cm)
    a = I.radiative_correction(a_cm, k)
    assert abs(a - a_cm / (1 - 1j * k ** 3 / (6 * np.pi) * a_cm)) < 1e-12 * abs(a)
    # ellipsoid tensor with L = 1/3 and R = I equals the isotropic CM + RRCM
    t = I.ellipsoid_tensor(V, (1 / 3, 1 / 3, 1 / 3), eps, em, k)
    assert np.allclose(t.reshape(3, 3), a * np.eye(3), rtol=1e-12, atol=1e-14)


def test_depolarisation_factors():
    L = I.depol_factors((5.0, 5.0, 5.0))
    assert np.allclose(L, 1 / 3, atol=1e-12)
    for ax in [(10.0, 4.0, 3.0), (30.0, 12.0, 20.0), (7.0, 7.0, 2.8)]:
        assert abs(sum(I.depol_factors(ax)) - 1.0) < 1e-6
    # prolate spheroid closed form, a > b = c
    a, b = 10.0, 4.0
    e = np.sqrt(1 - (b / a) ** 2)
    Lx = (1 - e * e) / e ** 3 * (np.arctanh(e) - e)
    assert abs(I.depol_factors((a, b, b))[0] - Lx) < 1e-6


def test_batch_particles_fixed_seed():
    ps = I.batch_particles()
    assert len(ps) == 64
    ns = [len(p.positions) for p in ps]
    assert min(ns) > 0 and max(ns) < 15000
    assert ns == [len(p.positions) for p in I.batch_particles()]
