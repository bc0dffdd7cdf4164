"""Near-field and far-field maps on the GPU (fields.rs) against the oracle.

Parity: the device field at observation points and the far-field pattern
agree with the oracle's FP64 restatement (fields.rs:18-48, 156-221) to 1e-12
relative (max-norm) on the same solved moments; theta and phi of the far
field are bit-identical. The reference's own field tests
(tests/v012_v013_features.rs:129-197, 361-470) are restated with the device
path. At full size (the 80 x 40 x 40 nm cuboid at 1.5 nm, 38,637 dipoles)
the two entry points are checked against each other: |E_sca(r rhat)|^2
(4 pi r / k^2)^2 tends to the far-field intensity as r grows.
"""
import numpy as np
import pytest

from lumina_amd import fields as F
from lumina_amd import inputs as I
from lumina_amd.solver import FP64, MIXED, IncidentField, PairSumSession
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

LAM = 520.0


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def sphere(backend):
    """Au sphere r = 10 nm at 3 nm spacing (the reference's
    gold_sphere_dipoles(10, 3, 520)), solved for x-polarised light."""
    pos = I.discretise(I.Sphere(10.0), 3.0)
    k = I.wavenumber(LAM, 1.0)
    alpha = I.polarisability_block("Au", LAM, 3.0, 1.0)
    s = PairSumSession(backend, pos, alpha, k, True, 3.0)
    _, p, _ = s.compute_cross_sections(IncidentField(), 1e-12, 1000)
    yield s, pos, k, p
    s.close()


def test_field_at_points_matches_oracle(sphere):
    s, pos, k, p = sphere
    rng = np.random.default_rng(11)
    pts = np.concatenate([rng.uniform(-25, 25, size=(300, 3)),
                          pos[[0, 7, len(pos) - 1]],               # on a dipole: skipped
                          pos[[3]] + 1e-11,                        # |d|^2 < 1e-20: skipped
                          [[0.0, 0.0, 4000.0], [1e4, -2e4, 3e3]]])
    for inc in (IncidentField(), IncidentField((0.0, 0.6, 0.8), (1.0, 0.0, 0.0), 2.5),
                IncidentField(polarisation=(1 / np.sqrt(2), 1j / np.sqrt(2), 0.0))):
        e, inten = F.field_at_points(s, pts, inc, p)
        eo, io = O.field_at_points(pos, p, k, pts, inc.direction, inc.polarisation,
                                   inc.amplitude)
        print("field vs oracle", relmax(e, eo), relmax(inten, io))
        assert relmax(e, eo) < 1e-12
        assert relmax(inten, io) < 1e-12
        # pointwise: every point to 1e-11 of its own magnitude
        assert np.all(np.abs(e - eo).max(axis=1) <= 1e-11 * np.abs(eo).max(axis=1))
    # the device-resident solution (moments=None) is the same data
    e1, i1 = F.field_at_points(s, pts, IncidentField(), None)
    e2, i2 = F.field_at_points(s, pts, IncidentField(), p)
    assert np.array_equal(e1, e2) and np.array_equal(i1, i2)


def test_field_incident_only_and_zero_moments(sphere):
    s, pos, k, p = sphere
    pts = np.random.default_rng(2).uniform(-40, 40, size=(257, 3))
    inc = IncidentField((0.0, 0.0, 1.0), (0.3, 0.4j, 0.0), 1.7)
    e, inten = F.field_at_points(s, pts, inc, np.zeros(3 * len(pos), dtype=complex))
    ph = np.exp(1j * k * pts[:, 2])
    want = 1.7 * np.array([0.3, 0.4j, 0.0])[None, :] * ph[:, None]
    assert np.max(np.abs(e - want)) < 1e-15
    assert np.max(np.abs(inten - 1.7 ** 2 * 0.25)) < 1e-14


def test_near_field_map_dimensions_and_grid(sphere):
    """test_near_field_map_dimensions, ..._intensity_is_nonnegative,
    ..._nonzero_outside_particle (tests/v012_v013_features.rs:129-197)."""
    s, pos, k, p = sphere
    plane = F.NearFieldPlane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 20.0, 20.0, 15, 12)
    nf = F.compute_near_field_map(s, plane, IncidentField())
    assert nf.nx == 15 and nf.ny == 12
    assert nf.positions.shape == (180, 3) and nf.field_intensity.shape == (180,)
    # x fastest, u = x_min + ix dx, the z = 0 plane
    assert np.array_equal(nf.positions[:15, 0], -20.0 + np.arange(15) * (40.0 / 14))
    assert np.all(nf.positions[:, 2] == 0.0)
    assert nf.extent == (-20.0, 20.0, -20.0, 20.0)
    _, io = O.field_at_points(pos, p, k, nf.positions)
    assert relmax(nf.field_intensity, io) < 1e-12
    nf2 = F.compute_near_field_map(s, F.NearFieldPlane((0, 0, 0), (0, 0, 1), 30.0, 30.0, 10, 10))
    assert np.all(nf2.field_intensity >= 0.0)
    nf3 = F.compute_near_field_map(s, F.NearFieldPlane((0, 0, 0), (0, 0, 1), 100.0, 100.0, 5, 5))
    assert np.count_nonzero(nf3.field_intensity > 0.5) > 0


def test_near_field_tilted_plane(sphere):
    s, pos, k, p = sphere
    plane = F.NearFieldPlane((1.0, -2.0, 3.0), (0.95, 0.1, 0.2), 15.0, 9.0, 21, 13)
    nf = F.compute_near_field_map(s, plane)
    u, v = F.plane_basis(plane.normal)
    assert abs(np.dot(u, v)) < 1e-15
    _, io = O.field_at_points(pos, p, k, nf.positions)
    assert relmax(nf.field_intensity, io) < 1e-12


def test_far_field_matches_oracle(sphere):
    s, pos, k, p = sphere
    for nt, nph in ((18, 36), (3, 4), (1, 5), (37, 72)):
        ff = F.compute_far_field(s, nt, nph, p, wavelength_nm=LAM)
        th, ph, io = O.far_field(pos, p, k, nt, nph)
        assert ff.n_theta == nt and ff.n_phi == nph
        assert ff.intensity.shape == (nt * nph,)
        assert np.array_equal(ff.theta, th) and np.array_equal(ff.phi, ph)
        print("far field vs oracle", nt, nph, relmax(ff.intensity, io))
        assert relmax(ff.intensity, io) < 1e-12
        assert np.all(ff.intensity >= 0.0)
    ff0 = F.compute_far_field(s, 0, 7)
    assert ff0.intensity.size == 0


def test_far_field_single_dipole(backend):
    """test_far_field_x_dipole_has_node_along_x and
    test_far_field_z_direction_has_maximum_for_x_dipole
    (tests/v012_v013_features.rs:399-470): one x dipole at the origin."""
    k = 2 * np.pi / 500.0
    s = PairSumSession(backend, np.zeros((1, 3)), I.isotropic_block(500 + 50j), k, True, 3.0)
    p = np.array([1.0, 0.0, 0.0], dtype=complex)
    ff = F.compute_far_field(s, 3, 4, p)
    assert ff.intensity[1 * 4 + 0] < 1e-20
    assert np.all(np.abs(ff.intensity[:4] - 1.0) < 1e-10)
    # closed form |p|^2 - |rhat . p|^2 everywhere
    ff = F.compute_far_field(s, 19, 24, p)
    want = 1.0 - (np.sin(ff.theta) * np.cos(ff.phi)) ** 2
    assert np.max(np.abs(ff.intensity - want)) < 1e-15
    s.close()


def test_circular_dichroism(backend):
    """test_cd_isotropic_sphere_is_zero (tests/v012_v013_features.rs:476-500):
    an isotropic sphere under x- and y-polarised light, then a chiral helix
    whose CD is non-zero and odd under mirroring."""
    lam = 500.0
    k = 2 * np.pi / lam
    pos = I.discretise(I.Sphere(8.0), 2.0)
    alpha = I.polarisability_block("Au", lam, 2.0, 1.0)
    s = PairSumSession(backend, pos, alpha, k, True, 2.0)
    _, px, _ = s.compute_cross_sections(IncidentField(polarisation=(1, 0, 0)), 1e-12, 1000)
    _, py, _ = s.compute_cross_sections(IncidentField(polarisation=(0, 1, 0)), 1e-12, 1000)
    cd = F.compute_circular_dichroism(px, py, k)
    assert abs(cd) < 1e-6
    s.close()
    hel = I.discretise(I.Helix(12.0, 20.0, 1.5, 4.0), 2.0)
    cds = []
    for mirror in (1.0, -1.0):
        q = hel * np.array([mirror, 1.0, 1.0])
        s = PairSumSession(backend, q, alpha, k, True, 2.0)
        _, px, _ = s.compute_cross_sections(IncidentField(polarisation=(1, 0, 0)), 1e-12, 2000)
        _, py, _ = s.compute_cross_sections(IncidentField(polarisation=(0, 1, 0)), 1e-12, 2000)
        cds.append(F.compute_circular_dichroism(px, py, k))
        s.close()
        if mirror > 0:
            scale = k * np.sum(np.abs(px))
    print("helix CD", cds, scale)
    # chiral: non-zero; odd under the mirror x -> -x
    assert abs(cds[0]) > 1e-6 * scale
    assert abs(cds[0] + cds[1]) < 1e-9 * scale


def test_full_size_near_field_tends_to_far_field(backend):
    """38,637 dipoles: the scattered field at r = 1e7 nm along each grid
    direction (incident amplitude 0) against the far-field pattern,
    |E|^2 (4 pi r / k^2)^2 -> |(I - rhat rhat) F|^2. The Fraunhofer and
    1/(kr) corrections are below 1e-5 here."""
    lam, n_env, d = 600.0, 1.33, 1.5
    pos = I.au_cuboid_80x40x40_d1p5()
    k = I.wavenumber(lam, n_env)
    alpha = I.polarisability_block("Au", lam, d, n_env)
    s = PairSumSession(backend, pos, alpha, k, True, d, precision=MIXED)
    _, p, info = s.compute_cross_sections(IncidentField(), 1e-8, 3000)
    nt, nph = 9, 16
    ff = F.compute_far_field(s, nt, nph)
    r = 1e7
    rh = np.stack([np.sin(ff.theta) * np.cos(ff.phi), np.sin(ff.theta) * np.sin(ff.phi),
                   np.cos(ff.theta)], axis=1)
    _, inten = F.field_at_points(s, r * rh, IncidentField(amplitude=0.0))
    scaled = inten * (4 * np.pi * r / k ** 2) ** 2
    print("near->far", relmax(scaled, ff.intensity), info)
    assert relmax(scaled, ff.intensity) < 1e-4
    # and the far field of the full solve against the oracle on a sample of
    # directions (the oracle sums all 38,637 dipoles per direction)
    th, ph, io = O.far_field(pos, p, k, 3, 4)
    ff2 = F.compute_far_field(s, 3, 4, p)
    assert relmax(ff2.intensity, io) < 1e-12
    s.close()


def test_cda_solver_response_methods(backend):
    """The rest of the reference's OpticalSolver surface on CdaSolver
    (cda/mod.rs:123-166, 421-445): cross sections from an already solved
    response equal those of the solve and the oracle's formula on the same
    moments; the near-field map and the far-field pattern of the response
    equal the oracle's fields.rs restatement; the method name."""
    from lumina_amd.solver import CdaSolver, SimulationParams
    lam, n_env, d = 540.0, 1.33, 2.0
    pos = I.au_sphere_r10_d2()
    alpha = I.polarisability_block("Au", lam, d, n_env)
    inc = IncidentField(direction=(0.0, 0.0, 1.0), polarisation=(1.0, 0.0, 0.0), amplitude=1.5)
    solver = CdaSolver(backend, incident_field=inc, use_fcd=True, cell_size=d)
    params = SimulationParams(environment_n=n_env, solver_tolerance=1e-12, max_iterations=2000)
    assert solver.method_name() == "Coupled Dipole Approximation (CDA)"
    resp = solver.solve_dipoles(pos, alpha, lam, params)
    cs = solver.compute_cross_sections(pos, alpha, lam, params)
    got = solver.cross_sections_from_response(pos, alpha, resp, params)
    a = np.array([got.extinction, got.absorption, got.scattering])
    b = np.array([cs.extinction, cs.absorption, cs.scattering])
    k = I.wavenumber(lam, n_env)
    P = O.Problem(pos, alpha, k, True, d)
    p = resp.moments.reshape(-1)
    want = P.cross_sections(p, P.incident(amplitude=1.5), 1.5)
    print("from response", a, "solve", b, "oracle", want)
    assert got.wavelength_nm == lam
    assert np.all(np.abs(a - b) <= 1e-10 * abs(b[0]))
    assert np.all(np.abs(a - want) <= 1e-12 * abs(want[0]))
    plane = F.NearFieldPlane(centre=(0.0, 0.0, 0.0), normal=(0.0, 1.0, 1.0), half_width=25.0,
                             half_height=15.0, nx=13, ny=7)
    nf = solver.compute_near_field(pos, alpha, resp, plane)
    pts, _ = F.plane_points(plane)
    _, io = O.field_at_points(pos, p, k, pts, amplitude=1.5)
    assert nf.nx == 13 and nf.ny == 7 and np.array_equal(nf.positions, pts)
    assert relmax(nf.field_intensity, io) < 1e-12
    ff = solver.compute_far_field(pos, alpha, resp, 7, 8)
    th, ph, fo = O.far_field(pos, p, k, 7, 8)
    assert ff.wavelength_nm == lam and ff.n_theta == 7 and ff.n_phi == 8
    assert np.array_equal(ff.theta, th) and np.array_equal(ff.phi, ph)
    assert relmax(ff.intensity, fo) < 1e-12
