"""The wavelength sweep with its per-wavelength set-up on the device
(lh_sweep_solve; runner.rs:145-218).

The device computes eps(lambda) from the material splines, alpha, alpha^-1,
the substrate factor and the incident field; everything else is the batched
solve. Checks: against the same solves set up on the host (inputs.py, the
restatement of spline.rs / types.rs / substrate.rs): C_ext, C_abs, C_sca
to 1e-10 relative (the solves stop at a relative residual of 1e-12 and the
operators differ by rounding);
single-dipole closed forms C_ext = k Im(e . alpha . e) (isotropic and
ellipsoid-hinted) that pin the device polarisability; the reference's
NaN row for a wavelength that does not converge; MaterialError::OutOfRange
as an argument error.
"""
import numpy as np
import pytest

from lumina_amd import inputs as I
from lumina_amd import sweep as W
from lumina_amd.solver import (FP64, MIXED, IncidentField, LatticeSpec, LuminaError,
                               SubstrateRuntime, batch_compute, substrate_reflection_factor)

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.abs(np.asarray(b))))


def rel_ext(a, b):
    """Largest difference of (C_ext, C_abs, C_sca) rows relative to each
    row's C_ext: C_sca = C_ext - C_abs of a small particle is a difference
    of nearly equal numbers."""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.abs(b[:, :1])))


def test_sphere_sweep_matches_host_setup(backend):
    """The 515-dipole Au sphere in water, 31 wavelengths 400..700 nm, FP64
    at tol 1e-12: the device set-up against the host set-up of the same
    batch."""
    pos = I.au_sphere_r10_d2()
    wls = [400.0 + 10.0 * t for t in range(31)]
    cs, infos = W.run_sweep(backend, pos, wls, [W.johnson_christy("Au")], spacing=2.0, n_env=1.33,
                            tolerance=1e-12, max_iterations=5000)
    probs = [dict(positions=pos, polarisability=I.polarisability_block("Au", lam, 2.0, 1.33),
                  k=I.wavenumber(lam, 1.33), use_fcd=True, cell_size=2.0) for lam in wls]
    csh, ih, _ = batch_compute(backend, probs, 1e-12, 5000, precision=FP64)
    print("sweep vs host set-up", rel(cs, csh), [i["iterations"] for i in infos])
    assert all(i["status"] == 0 for i in infos)
    # alpha agrees to rounding; GMRES stops at a relative residual of 1e-12,
    # and near 700 nm (1,300+ iterations) the operator's condition turns
    # rounding-level differences of alpha into ~1e-12 of C_ext
    assert rel(cs, csh) < 1e-10
    assert rel(cs[:9], csh[:9]) < 1e-13
    for a, b in zip(infos, ih):
        assert abs(a["iterations"] - b["iterations"]) <= max(2, b["iterations"] // 100)


@pytest.mark.parametrize("name", ["Au", "Ag"])
def test_single_dipole_closed_form(backend, name):
    """One isotropic dipole: p = alpha E, C_ext = k Im(alpha) (cda/mod.rs:
    187-235 with a single dipole), alpha from the host restatement of
    types.rs:243-259 at the spline's eps."""
    wls = np.linspace(I.material(name).wl[0] + 1.0, I.material(name).wl[-1] - 1.0, 23)
    n_env, d = 1.33, 3.0
    cs, infos = W.run_sweep(backend, np.zeros((1, 3)), wls, [W.johnson_christy(name)], spacing=d,
                            n_env=n_env, tolerance=1e-14, max_iterations=10)
    want = []
    for lam in wls:
        k = I.wavenumber(lam, n_env)
        a = I.polarisability_block(name, lam, d, n_env)[0]
        want.append(k * a.imag)
    print(name, rel(cs[:, 0], want))
    assert rel(cs[:, 0], want) < 1e-13


def test_ellipsoid_hint_closed_form(backend):
    """One dipole with an ellipsoid hint (depolarisation factors and a
    rotation about z): C_ext = k Im(alpha_xx) for x-polarised light along z
    (types.rs:286-317 through the host restatement)."""
    depol = I.depol_factors((12.0, 7.0, 4.0))
    th = 0.7
    rot = (np.cos(th), -np.sin(th), 0.0, np.sin(th), np.cos(th), 0.0, 0.0, 0.0, 1.0)
    wls = np.linspace(450.0, 850.0, 9)
    cs, _ = W.run_sweep(backend, np.zeros((1, 3)), wls, [W.johnson_christy("Au")], spacing=2.5,
                        hints=[W.EllipsoidHint(depol, rot)], hint_index=[0], n_env=1.0,
                        tolerance=1e-14, max_iterations=10)
    want = []
    for lam in wls:
        k = I.wavenumber(lam, 1.0)
        t = I.ellipsoid_tensor(2.5 ** 3, depol, I.material("Au").epsilon(lam), 1.0, k, rot)
        want.append(k * t[0].imag)
    print("ellipsoid", rel(cs[:, 0], want))
    assert rel(cs[:, 0], want) < 1e-13


def test_mixed_materials_spacings_hints(backend):
    """A particle of two materials, two spacings and hinted dipoles: the
    per-dipole alpha^-1 of the device against per-dipole host blocks."""
    rng = np.random.default_rng(3)
    pos = I.discretise(I.Sphere(9.0), 2.0)
    n = len(pos)
    mat = (pos[:, 0] > 0).astype(np.int32)          # Au for x <= 0, Ag for x > 0
    spacing = np.where(pos[:, 2] > 1.0, 2.0, 2.0 * 0.999)
    hint = np.where(rng.uniform(size=n) < 0.3, 0, -1).astype(np.int32)
    depol = I.depol_factors((8.0, 5.0, 5.0))
    hints = [W.EllipsoidHint(depol, (0.0, 1.0, 0.0, -1.0, 0.0, 0.0, 0.0, 0.0, 1.0))]
    wls = [505.0, 610.0, 777.0]
    names = ["Au", "Ag"]
    cs, infos = W.run_sweep(backend, pos, wls, [W.johnson_christy(m) for m in names],
                            material_index=mat, spacing=spacing, hints=hints, hint_index=hint,
                            n_env=1.2, cell_size=2.0, tolerance=1e-10, max_iterations=6000)
    probs = []
    for lam in wls:
        k = I.wavenumber(lam, 1.2)
        blocks = []
        for i in range(n):
            eps = I.material(names[mat[i]]).epsilon(lam)
            vol = spacing[i] * spacing[i] * spacing[i]
            if hint[i] >= 0:
                blocks.append(I.ellipsoid_tensor(vol, depol, eps, 1.44, k, hints[0].rotation))
            else:
                a = I.radiative_correction(I.clausius_mossotti(vol, eps, 1.44), k)
                blocks.append(I.isotropic_block(a))
        probs.append(dict(positions=pos, polarisability=np.array(blocks), k=k, use_fcd=True,
                          cell_size=2.0))
    csh, ih, _ = batch_compute(backend, probs, 1e-10, 6000, precision=FP64)
    print("mixed particle", rel_ext(cs, csh), [i["iterations"] for i in infos])
    assert all(i["status"] == 0 for i in infos)
    assert rel_ext(cs, csh) < 1e-9


@pytest.mark.parametrize("retarded", [True, False])
def test_substrate_factor_on_device(backend, retarded):
    """A sphere above a substrate whose table varies with wavelength (the Ag
    table as the substrate): the device reflection factor against
    substrate_reflection_factor on the host."""
    pos = I.discretise(I.Sphere(6.0), 2.0)
    z0 = float(pos[:, 2].min()) - 1.0
    wls = [430.0, 560.0, 690.0]
    mats = [W.johnson_christy("Au"), W.johnson_christy("Ag")]
    cs, _ = W.run_sweep(backend, pos, wls, mats, spacing=2.0, n_env=1.0,
                        substrate=W.SubstrateSpec(z0, 1, retarded), tolerance=1e-12,
                        max_iterations=3000)
    probs = []
    for lam in wls:
        delta = substrate_reflection_factor(I.material("Ag").epsilon(lam), 1.0, retarded)
        probs.append(dict(positions=pos, polarisability=I.polarisability_block("Au", lam, 2.0, 1.0),
                          k=I.wavenumber(lam, 1.0), use_fcd=True, cell_size=2.0,
                          substrate=SubstrateRuntime(z0, delta)))
    csh, _, _ = batch_compute(backend, probs, 1e-12, 3000, precision=FP64)
    print("substrate", retarded, rel_ext(cs, csh))
    # the two set-ups agree to rounding; the solves (tol 1e-12) to 1e-10
    assert rel_ext(cs, csh) < 1e-10


def test_lattice_sweep_mixed(backend):
    """A small periodic array on a substrate, mixed precision: device set-up
    against host set-up of the same batch."""
    pos = I.discretise(I.Cylinder(8.0, 6.0), 2.0)
    z0 = float(pos[:, 2].min()) - 1.0
    lat = LatticeSpec((60.0, 0.0, 0.0), (0.0, 60.0, 0.0), truncation_real=2, truncation_recip=2)
    glass = W.TabulatedMaterial("glass", np.array([300.0, 1200.0]), np.array([2.25, 2.25]) + 0j)
    wls = [520.0, 640.0]
    cs, infos = W.run_sweep(backend, pos, wls, [W.johnson_christy("Au"), glass], spacing=2.0,
                            n_env=1.0, substrate=W.SubstrateSpec(z0, 1, True), lattice=lat,
                            tolerance=1e-10, max_iterations=2000, precision=MIXED)
    probs = [dict(positions=pos, polarisability=I.polarisability_block("Au", lam, 2.0, 1.0),
                  k=I.wavenumber(lam, 1.0), use_fcd=True, cell_size=2.0, lattice=lat,
                  substrate=SubstrateRuntime(z0, substrate_reflection_factor(2.25 + 0j, 1.0, True)))
             for lam in wls]
    csh, _, _ = batch_compute(backend, probs, 1e-10, 2000, precision=MIXED)
    print("lattice sweep", rel_ext(cs, csh), infos)
    assert rel_ext(cs, csh) < 1e-9


def test_nan_row_and_out_of_range(backend):
    pos = I.discretise(I.Sphere(12.0), 2.0)
    # runner.rs:189-203: a wavelength that does not converge is a NaN row
    cs, infos = W.run_sweep(backend, pos, [500.0, 700.0], [W.johnson_christy("Au")], spacing=2.0,
                            n_env=1.33, tolerance=1e-14, max_iterations=3)
    assert all(i["status"] != 0 for i in infos)
    assert np.all(np.isnan(cs))
    # MaterialError::OutOfRange
    with pytest.raises(LuminaError, match="outside"):
        W.run_sweep(backend, pos, [500.0, 1500.0], [W.johnson_christy("Au")], spacing=2.0)
    with pytest.raises(LuminaError):
        W.run_sweep(backend, pos, [500.0], [W.johnson_christy("Au")], material_index=[1] * len(pos),
                    spacing=2.0)
