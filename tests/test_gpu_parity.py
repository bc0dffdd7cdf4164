"""Parity of the HIP path (through the C ABI) with the oracle.

Bars: FP64 matvec <= 1e-12 relative (max-norm) against the oracle's
on-the-fly matvec; FP64 solve C_ext/C_abs/C_sca within 1e-10 relative of the
oracle's GMRES at tol 1e-12; FP32 far-field matvec within 2e-6 relative;
mixed-precision solve C_ext within 1e-6 of the FP64 oracle.
"""
import numpy as np
import pytest

from lumina_amd import inputs as I
from lumina_amd.solver import (FP64, MIXED, ConvergenceFailure, IncidentField, SubstrateRuntime,
                               batch_compute)
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def sphere_problem(lam=520.0, n_env=1.33, mat="Au", cell=2.0):
    pos = I.au_sphere_r10_d2()
    alpha = I.polarisability_block(mat, lam, 2.0, n_env)
    k = I.wavenumber(lam, n_env)
    return pos, alpha, k, cell


def rand_x(n, seed=0):
    r = np.random.default_rng(seed)
    return r.standard_normal(3 * n) + 1j * r.standard_normal(3 * n)


def test_device_info(backend):
    info = backend.device_info()
    print(info)
    assert info["backend_type"] == "Gpu"
    assert "gfx950" in info["arch"]
    assert info["name"]
    assert info["compute_units"] == 256


@pytest.mark.parametrize("use_fcd", [True, False])
def test_matvec_fp64_sphere(backend, use_fcd):
    pos, alpha, k, cell = sphere_problem()
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=use_fcd, cell_size=cell)
    x = rand_x(len(pos))
    y = s.matvec(x, FP64)
    yo = O.Problem(pos, alpha, k, use_fcd=use_fcd, cell_size=cell).matvec(x)
    assert rel(y, yo) < 1e-12


def test_matvec_fp64_cli_cell_size(backend):
    # CLI default cell_size = 3.0 with spacing 2 (cda/mod.rs:89): 122 near pairs per dipole
    pos, alpha, k, _ = sphere_problem()
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=3.0)
    x = rand_x(len(pos), 1)
    assert rel(s.matvec(x, FP64), O.Problem(pos, alpha, k, True, 3.0).matvec(x)) < 1e-12


def test_matvec_substrate_and_anisotropic(backend):
    pos = I.discretise(I.Ellipsoid((12.0, 8.0, 6.0)), 2.0)
    pos = pos.copy()
    pos[:, 2] += 10.0
    lam, n_env = 600.0, 1.0
    k = I.wavenumber(lam, n_env)
    depol = I.depol_factors((12.0, 8.0, 6.0))
    a = I.polarisability_block("Ag", lam, 2.0, n_env, depol=depol)
    # per-dipole anisotropic blocks with a small variation
    blocks = np.tile(a, (len(pos), 1)) * (1.0 + 0.01 * np.arange(len(pos)))[:, None]
    sub = SubstrateRuntime(z_interface=-1.0, delta_eps=I_delta())
    s = backend.create_pairsum_session(pos, blocks, k, use_fcd=True, cell_size=2.0, substrate=sub)
    x = rand_x(len(pos), 2)
    yo = O.Problem(pos, blocks, k, True, 2.0, substrate=(-1.0, sub.delta_eps)).matvec(x)
    assert rel(s.matvec(x, FP64), yo) < 1e-12
    # the FP32 far field with the image term (k_far_row<float, true, ...>)
    e32 = rel(s.matvec(x, MIXED), yo)
    print(f"substrate + anisotropic fp32 matvec rel err {e32:.2e}")
    assert e32 < 2e-6


def I_delta():
    from lumina_amd.solver import substrate_reflection_factor
    return substrate_reflection_factor(2.25 + 0.0j, 1.0, True)


@pytest.mark.parametrize("n", [1, 2, 65, 257, 515])
@pytest.mark.parametrize("form", ["sym", "256"])
def test_matvec_substrate_pair_forms(backend, n, form, monkeypatch):
    """The Fresnel image term (assembly.rs:278-295) in the one-tensor-per-
    unordered-pair kernels (form sym: k_far_sym64/32<., true>, where each
    image tensor serves both directions and the diagonal adds the self-image
    j == i one way) and in the row kernels (256), FP64 and FP32, against the
    oracle: a single dipole (the self-image alone), ragged tiles and the
    515-dipole sphere on glass."""
    monkeypatch.setenv("LUMINA_FAR_BLOCK", form)
    if n == 515:
        pos, alpha, k, cell = sphere_problem()
    else:
        r = np.random.default_rng(200 + n)
        pos = np.unique(np.round(r.uniform(-12, 12, size=(n, 3)) / 2.0) * 2.0, axis=0)
        alpha = I.isotropic_block(40.0 + 10.0j)
        k = 2 * np.pi / 480.0
        cell = 2.0
    z0 = float(pos[:, 2].min()) - 1.5
    sub = SubstrateRuntime(z_interface=z0, delta_eps=I_delta())
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=cell, substrate=sub)
    x = rand_x(len(pos), 31)
    yo = O.Problem(pos, alpha, k, True, cell, substrate=(z0, sub.delta_eps)).matvec(x)
    e64, e32 = rel(s.matvec(x, FP64), yo), rel(s.matvec(x, MIXED), yo)
    print(f"substrate form {form} n={len(pos)}: fp64 {e64:.2e}, fp32 {e32:.2e}")
    assert e64 < 1e-12
    assert e32 < 2e-6


@pytest.mark.parametrize("runs", ["1", "2", "5"])
def test_substrate_pair_form_clean_tiles_and_solve(backend, runs, monkeypatch):
    """A rod of 97 x 4 x 4 dipoles on glass, ordered along its axis: tile
    pairs beyond the exclusion radius take the image kernels without the
    exclusion test, in runs of 1, 2 and 5 source tiles (FP32). Matvecs in
    both precisions against the oracle, and solves in the pair form against
    the row form (tol 1e-10)."""
    g = np.stack(np.meshgrid(np.arange(97), np.arange(4), np.arange(4), indexing="ij"), -1)
    pos = 1.5 * g.reshape(-1, 3).astype(float)
    alpha = I.isotropic_block(40.0 + 10.0j)
    k = 2 * np.pi / 480.0
    z0 = -1.0
    sub = SubstrateRuntime(z_interface=z0, delta_eps=I_delta())
    x = rand_x(len(pos), 32)
    yo = O.Problem(pos, alpha, k, True, 1.5, substrate=(z0, sub.delta_eps)).matvec(x)
    monkeypatch.setenv("LUMINA_RUN_TILES", runs)
    out = {}
    for form in ("sym", "256"):
        monkeypatch.setenv("LUMINA_FAR_BLOCK", form)
        s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=1.5,
                                           substrate=sub)
        e64, e32 = rel(s.matvec(x, FP64), yo), rel(s.matvec(x, MIXED), yo)
        cs, _, info = s.compute_cross_sections(IncidentField(), 1e-10, 2000)
        s.close()
        print(f"rod on glass, form {form}, runs {runs}: fp64 {e64:.2e}, fp32 {e32:.2e}, "
              f"{info['iterations']} iterations, C_ext {cs[0]:.10e}")
        assert e64 < 1e-12
        assert e32 < 2e-6
        out[form] = (cs, info["iterations"])
    assert abs(out["sym"][1] - out["256"][1]) <= 1
    assert np.all(np.abs(out["sym"][0] - out["256"][0]) <= 1e-9 * abs(out["256"][0][0]))


def test_matvec_random_positions(backend):
    # off-grid positions, no coincident pairs, FCD radius crossing many pairs
    r = np.random.default_rng(3)
    pos = r.uniform(-8, 8, size=(300, 3))
    alpha = I.isotropic_block(50.0 + 20.0j)
    k = 2 * np.pi / 500.0
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=1.7)
    x = rand_x(300, 4)
    yo = O.Problem(pos, alpha, k, True, 1.7).matvec(x)
    assert rel(s.matvec(x, FP64), yo) < 1e-12
    # FP32 far field: the exclusion test on FP32-rounded positions gives the
    # same pair split as the FP64 near list
    e32 = rel(s.matvec(x, MIXED), yo)
    print(f"random positions fp32 matvec rel err {e32:.2e}")
    assert e32 < 2e-6


@pytest.mark.parametrize("n", [1, 2, 3, 257])
def test_matvec_small_and_ragged(backend, n):
    r = np.random.default_rng(n)
    pos = r.uniform(-20, 20, size=(n, 3))
    alpha = I.isotropic_block(30.0 + 5.0j)
    k = 2 * np.pi / 450.0
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=2.0)
    x = rand_x(n, 5)
    assert rel(s.matvec(x, FP64), O.Problem(pos, alpha, k, True, 2.0).matvec(x)) < 1e-12


@pytest.mark.parametrize("sincos", ["hw", "sw"])
def test_matvec_fp32_far_field(backend, sincos, monkeypatch):
    """FP32 far field with the hardware sin/cos (default) and with the
    software sincos_f32."""
    monkeypatch.setenv("LUMINA_FP32_SINCOS", sincos)
    for lam in (400.0, 520.0, 800.0):
        pos, alpha, k, cell = sphere_problem(lam=lam)
        s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=cell)
        x = rand_x(len(pos), 6)
        yo = O.Problem(pos, alpha, k, True, cell).matvec(x)
        e = rel(s.matvec(x, MIXED), yo)
        print(f"fp32 far-field matvec rel err ({sincos}, {lam} nm)", e)
        assert e < 2e-6
    # a large object: phases up to k * 95 nm
    pos = I.au_cuboid_80x40x40_d1p5()
    alpha = I.polarisability_block("Au", 400.0, 1.5, 1.33)
    k = I.wavenumber(400.0, 1.33)
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=1.5)
    x = rand_x(len(pos), 9)
    y32 = s.matvec(x, MIXED)
    y64 = s.matvec(x, FP64)
    e = rel(y32, y64)
    print(f"fp32 vs fp64 matvec rel err ({sincos}, cuboid 400 nm)", e)
    assert e < 2e-6


@pytest.mark.parametrize("form", ["256", "512"])
def test_set_wavelength_beyond_hardware_sincos_range(backend, form, monkeypatch):
    """A session built at a k whose far-field phases lie within the hardware
    sin/cos range (v_sin_f32: 256 revolutions) and re-targeted to a k beyond
    it switches the FP32 far field to the software sincos_f32 (and off the
    unordered-pair form, block 512, which uses the hardware functions)."""
    monkeypatch.setenv("LUMINA_FAR_BLOCK", form)
    r = np.random.default_rng(11)
    n = 200
    pos = np.stack([np.sort(r.uniform(0.0, 6000.0, n)), r.uniform(-5, 5, n),
                    r.uniform(-5, 5, n)], axis=1)
    alpha = I.isotropic_block(40.0 + 8.0j)
    k1, k2 = 2 * np.pi / 1000.0, 2 * np.pi / 20.0   # 6 and 300 revolutions over 6 um
    s = backend.create_pairsum_session(pos, alpha, k1, use_fcd=False, cell_size=2.0,
                                       precision=MIXED)
    x = rand_x(n, 12)
    e1 = rel(s.matvec(x, MIXED), O.Problem(pos, alpha, k1, False, 2.0).matvec(x))
    s.set_wavelength(k2, alpha)
    yo = O.Problem(pos, alpha, k2, False, 2.0).matvec(x)
    e2 = rel(s.matvec(x, MIXED), yo)
    e64 = rel(s.matvec(x, FP64), yo)
    print(f"form {form}: fp32 rel err at 6 rev {e1:.2e}, at 300 rev {e2:.2e}; fp64 {e64:.2e}")
    # FP32 positions of up to 6 um carry 5e-4 nm rounding: ~1e-4 relative
    # on the 30 nm neighbour pairs; FP32 phases of up to 1,900 rad ~1e-4 more
    assert e1 < 5e-4
    assert e2 < 2e-3
    assert e64 < 1e-11


def test_solve_fp64_parity_sphere(backend):
    pos, alpha, k, cell = sphere_problem()
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=cell)
    out, p, info = s.compute_cross_sections(IncidentField(), 1e-12, 1000)
    P = O.Problem(pos, alpha, k, True, cell)
    cso, po, infoo = P.compute_cross_sections(1e-12, 1000)
    print("gpu", out, info, "oracle", cso, infoo)
    assert np.max(np.abs(out - cso) / np.abs(cso)) < 1e-10
    assert np.linalg.norm(p - po) / np.linalg.norm(po) < 1e-9


def test_solve_fp64_tol_1e10(backend):
    pos, alpha, k, cell = sphere_problem(lam=450.0)
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=cell)
    out, p, info = s.compute_cross_sections(IncidentField(), 1e-10, 1000)
    cso, po, infoo = O.Problem(pos, alpha, k, True, cell).compute_cross_sections(1e-10, 1000)
    print("tol 1e-10", out, cso, info, infoo)
    assert np.max(np.abs(out - cso) / np.abs(cso)) < 1e-8
    assert info["residual"] < 1e-10


def test_solve_mixed_error(backend):
    pos, alpha, k, cell = sphere_problem()
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=cell,
                                       precision=MIXED)
    out, p, info = s.compute_cross_sections(IncidentField(), 1e-10, 1000)
    cso, _, _ = O.Problem(pos, alpha, k, True, cell).compute_cross_sections(1e-12, 1000)
    err = abs(out[0] - cso[0]) / abs(cso[0])
    print("mixed C_ext err", err, info)
    assert err < 1e-6
    assert info["matvecs_fp32"] > 0


def test_convergence_failure(backend):
    pos, alpha, k, cell = sphere_problem()
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=cell)
    with pytest.raises(ConvergenceFailure) as ei:
        s.compute_cross_sections(IncidentField(), 1e-12, 5)
    assert ei.value.info["iterations"] == 5


def _batch_probs():
    pos = I.au_sphere_r10_d2()
    pos2 = I.discretise(I.Ellipsoid((8.0, 5.0, 4.0)), 2.0)
    d = I.depol_factors((8.0, 5.0, 4.0))
    probs = []
    for lam in (400.0, 450.0, 500.0):
        k = I.wavenumber(lam, 1.33)
        probs.append(dict(positions=pos, polarisability=I.polarisability_block("Au", lam, 2.0, 1.33),
                          k=k, use_fcd=True, cell_size=2.0))
        probs.append(dict(positions=pos2,
                          polarisability=I.polarisability_block("Ag", lam, 2.0, 1.33, depol=d),
                          k=k, use_fcd=True, cell_size=2.0))
    return probs


def test_batch_matches_single_including_failures(backend):
    """Six solves of two geometries in one batch: every solve that the
    reference's GMRES(30) converges (tol 1e-12, 1000 iterations) matches the
    oracle to 1e-10; the one it does not (Ag ellipsoid at 450 nm) is reported
    as a convergence failure after the same 1000 iterations."""
    probs = _batch_probs()
    with pytest.raises(ConvergenceFailure) as ei:
        batch_compute(backend, probs, 1e-12, 1000, precision=FP64)
    cs, infos = ei.value.cross_sections, ei.value.infos
    for q, pr in enumerate(probs):
        P = O.Problem(pr["positions"], pr["polarisability"], pr["k"], True, 2.0)
        cso, po, io = P.compute_cross_sections(1e-12, 1000)
        print(q, infos[q]["iterations"], io["iterations"], infos[q]["status"], io["status"])
        assert (infos[q]["status"] == 0) == (io["status"] == 0), q
        if io["status"] == 0:
            assert np.max(np.abs(cs[q] - cso) / np.abs(cso)) < 1e-10, (q, cs[q], cso)
            assert abs(infos[q]["iterations"] - io["iterations"]) <= max(2, io["iterations"] // 20)
        else:
            assert infos[q]["iterations"] == 1000


def test_batch_mixed(backend):
    probs = [p for p in _batch_probs() if len(p["positions"]) == 515]
    cs, infos, _ = batch_compute(backend, probs, 1e-10, 1000, precision=MIXED)
    for q, pr in enumerate(probs):
        cso, _, _ = O.Problem(pr["positions"], pr["polarisability"], pr["k"], True,
                              2.0).compute_cross_sections(1e-12, 1000)
        err = abs(cs[q][0] - cso[0]) / abs(cso[0])
        print("batch mixed C_ext err", err, infos[q])
        assert err < 1e-6
        assert infos[q]["matvecs_fp32"] > 0


def test_batch_resident_rerun(backend):
    """A batch stays on the device: two runs (FP64, then mixed) of the same
    batch give the FP64 answer and the mixed answer within 1e-6."""
    from lumina_amd.solver import Batch
    probs = [p for p in _batch_probs() if len(p["positions"]) == 515]
    b = Batch(backend, probs)
    b.run(1e-10, 1000, FP64)
    cs64, i64, _ = b.results()
    b.set_timing(True)
    b.run(1e-10, 1000, MIXED)
    cs32, i32, _ = b.results()
    kt = b.kernel_times()
    b.close()
    assert np.max(np.abs(cs32[:, 0] - cs64[:, 0]) / np.abs(cs64[:, 0])) < 1e-6
    assert kt["far32_launches"] > 0 and kt["far32_ms"] > 0
    assert kt["far32_pairs"] == sum(i["matvecs_fp32"] * 515 * 514 for i in i32)


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_batch_launch_options_bitwise(backend, precision, monkeypatch):
    """Launch options change scheduling only: the captured-graph cycle and the
    host-issued cycle, and every row-kernel block size, give bit-identical
    moments and iteration counts, because each target's sum over a
    256-source tile is formed in the same order. The one-tensor-per-pair
    kernels (block 512; FP64 with rotating sources, block 1024) sum in
    another order, and two Gram-Schmidt passes in the mixed form's FP32
    steps (LUMINA_CGS2=1, one by default) orthogonalise differently, as does
    modified Gram-Schmidt in FP64 steps (LUMINA_ORTHO=mgs), and FP64 partial
    slices of FP32 matvecs (LUMINA_PART64=1) round differently: their results
    agree to the solve tolerance."""
    from lumina_amd.solver import Batch
    probs = _batch_probs()
    runs = {}
    opts = [("auto", {}), ("nograph", {"LUMINA_GRAPH": "0", "LUMINA_FAR_BLOCK": "256"}),
            ("b64", {"LUMINA_FAR_BLOCK": "64"}), ("b128", {"LUMINA_FAR_BLOCK": "128"}),
            ("b256", {"LUMINA_FAR_BLOCK": "256"}), ("sym", {"LUMINA_FAR_BLOCK": "512"}),
            ("rot", {"LUMINA_FAR_BLOCK": "1024"}),
            ("cgs2", {"LUMINA_CGS2": "1", "LUMINA_FAR_BLOCK": "256"}),
            ("k3", {"LUMINA_GRAPH_CYCLES": "3", "LUMINA_FAR_BLOCK": "256"}),
            ("kall", {"LUMINA_GRAPH_CYCLES": "all", "LUMINA_FAR_BLOCK": "256"}),
            ("timed", {"LUMINA_FAR_BLOCK": "256"}),
            ("mgs", {"LUMINA_ORTHO": "mgs", "LUMINA_FAR_BLOCK": "256"}),
            ("p64", {"LUMINA_PART64": "1", "LUMINA_FAR_BLOCK": "256"})]
    for name, env in opts:
        for key in ("LUMINA_GRAPH", "LUMINA_FAR_BLOCK", "LUMINA_CGS2", "LUMINA_GRAPH_CYCLES",
                    "LUMINA_ORTHO", "LUMINA_PART64"):
            monkeypatch.delenv(key, raising=False)
        for key, val in env.items():
            monkeypatch.setenv(key, val)
        b = Batch(backend, probs)
        if name == "timed":
            # far-field events recorded as nodes of the captured graphs
            b.set_timing(True)
        b.run(1e-10, 200, precision)
        try:
            cs, infos, p = b.results(want_moments=True)
        except ConvergenceFailure as e:
            cs, infos, p = e.cross_sections, e.infos, e.moments
        kt = b.kernel_times()
        b.close()
        if name.startswith("b"):
            blk = int(name[1:])
            assert kt["far64_block"] == blk
            if precision == MIXED:
                assert kt["far32_block"] == blk
        elif name == "sym":
            assert kt["far64_block"] == 256
            if precision == MIXED:
                assert kt["far32_block"] == 512
        elif name == "rot":
            assert kt["far64_block"] == 1024
            if precision == MIXED:
                assert kt["far32_block"] == 256
        elif name == "cgs2":
            assert kt["far64_block"] == 256
        elif name in ("mgs", "p64"):
            assert kt["far64_block"] == 256
        elif name in ("k3", "kall", "timed"):
            assert kt["far64_block"] == 256
            cyc = max(i["cycles"] for i in infos)
            if name == "kall":
                # 200 iterations: 7 cycles per launch
                assert kt["host_round_trips"] == -(-cyc // 7)
                assert kt["cycles_issued"] == 7 * kt["host_round_trips"]
            elif name == "k3":
                assert kt["host_round_trips"] == -(-cyc // 3)
            else:
                # 1, 2, 4, 8, 8, ... cycles per launch
                assert kt["cycles_issued"] >= cyc
                assert kt["host_round_trips"] <= 3 + -(-max(cyc - 7, 0) // 8)
                n_fp = kt["far32_launches"] if precision == MIXED else kt["far64_launches"]
                assert n_fp > 0 and kt["far64_launches"] > 0
                assert (kt["far32_ms"] if precision == MIXED else kt["far64_ms"]) > 0
        else:
            assert kt["far64_block"] in (64, 128, 256, 512, 1024)
        runs[name] = (cs, [i["iterations"] for i in infos], p,
                      (kt["far64_block"], kt["far32_block"]))
        print(name, runs[name][1], kt["far64_block"], kt["far32_block"], kt["tune_ms"])
    cs0, it0, p0, _ = runs["b256"]
    for name, (cs, it, p, (f64, f32)) in runs.items():
        # the one-tensor-per-pair kernels (block 512) sum in another order
        # modified Gram-Schmidt (FP64 steps only) orthogonalises in another
        # order; FP64 partial slices of the FP32 matvecs (p64) round them
        # differently
        exact = (f64 not in (512, 1024) and not (precision == MIXED and f32 == 512)
                 and not (precision == MIXED and name in ("cgs2", "p64"))
                 and not (precision == FP64 and name == "mgs"))
        if exact:
            assert it == it0, name
            assert np.array_equal(cs, cs0), name
            for a, b_ in zip(p, p0):
                assert np.array_equal(a, b_), name
        else:
            for q in range(len(it)):
                if it0[q] < 200:
                    assert abs(it[q] - it0[q]) <= 2, (name, q)
                    # C_sca = C_ext - C_abs carries their absolute error: the
                    # bar scales with C_ext for all three
                    assert np.all(np.abs(cs[q] - cs0[q]) <= 1e-9 * np.abs(cs0[q][0])), (name, q)


@pytest.mark.parametrize("form", ["sym", "1024"])
@pytest.mark.parametrize("n", [1, 2, 65, 257, 515, 700])
def test_matvec_fp64_symmetric_kernel(backend, n, form, monkeypatch):
    """The FP64 one-tensor-per-unordered-pair kernels (LDS-staged sources,
    "sym"; sources rotating around the wave, "1024") against the oracle (FP64
    bar), and a solve through each against the row kernel's."""
    monkeypatch.setenv("LUMINA_FAR_BLOCK", form)
    if n == 515:
        pos, alpha, k, cell = sphere_problem()
    else:
        r = np.random.default_rng(200 + n)
        pos = np.unique(np.round(r.uniform(-15, 15, size=(n, 3)) / 1.5) * 1.5, axis=0)
        alpha = I.isotropic_block(40.0 + 10.0j)
        k = 2 * np.pi / 480.0
        cell = 1.5
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=cell)
    x = rand_x(len(pos), 13)
    # a solve first, so that the session's tiling is in force for the matvec
    out, _, info = s.compute_cross_sections(IncidentField(), 1e-10, 1000)
    yo = O.Problem(pos, alpha, k, True, cell).matvec(x)
    e = rel(s.matvec(x, FP64), yo)
    print(f"sym fp64 matvec n={len(pos)} rel err {e:.2e}")
    assert e < 1e-12
    monkeypatch.setenv("LUMINA_FAR_BLOCK", "256")
    s2 = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=cell)
    out2, _, info2 = s2.compute_cross_sections(IncidentField(), 1e-10, 1000)
    assert abs(info["iterations"] - info2["iterations"]) <= 1
    assert np.all(np.abs(out - out2) <= 1e-9 * np.abs(out2))


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 255, 256, 257, 515, 700])
def test_matvec_fp32_symmetric_kernel(backend, n, monkeypatch):
    """The FP32 one-tensor-per-unordered-pair kernel against the oracle: ragged
    tile edges, the diagonal tile, FCD near pairs excluded."""
    monkeypatch.setenv("LUMINA_FAR_BLOCK", "512")
    if n == 515:
        pos, alpha, k, cell = sphere_problem()
    else:
        r = np.random.default_rng(100 + n)
        pos = np.round(r.uniform(-15, 15, size=(n, 3)) / 1.5) * 1.5
        pos = np.unique(pos, axis=0)
        alpha = I.isotropic_block(40.0 + 10.0j)
        k = 2 * np.pi / 480.0
        cell = 1.5
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=cell)
    x = rand_x(len(pos), 11)
    yo = O.Problem(pos, alpha, k, True, cell).matvec(x)
    e = rel(s.matvec(x, MIXED), yo)
    print(f"sym fp32 matvec n={len(pos)} rel err {e:.2e}")
    assert e < 2e-6


@pytest.mark.parametrize("nx", [40, 97])
def test_matvec_fp32_symmetric_kernel_far_tiles(backend, nx, monkeypatch):
    """A rod of nx x 4 x 4 dipoles ordered along its axis: 256-dipole tiles
    two or more apart lie beyond the exclusion radius, so the FP32
    one-tensor-per-pair kernel runs in both forms (with and without the
    exclusion test), on a ragged last tile for nx = 97. Against the oracle
    and against the row kernel."""
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(4), np.arange(4), indexing="ij"), -1)
    pos = 1.5 * g.reshape(-1, 3).astype(float)
    alpha = I.isotropic_block(40.0 + 10.0j)
    k = 2 * np.pi / 480.0
    x = rand_x(len(pos), 12)
    yo = O.Problem(pos, alpha, k, True, 1.5).matvec(x)
    out = {}
    for blk in ("512", "256"):
        monkeypatch.setenv("LUMINA_FAR_BLOCK", blk)
        s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=1.5)
        out[blk] = s.matvec(x, MIXED)
        s.close()
    e, e_row = rel(out["512"], yo), rel(out["512"], out["256"])
    print(f"sym fp32 rod n={len(pos)} rel err {e:.2e} vs row kernel {e_row:.2e}")
    assert e < 2e-6
    assert e_row < 2e-6


@pytest.mark.parametrize("nx", [97, 203])
def test_fp32_symmetric_runs_of_source_tiles(backend, nx, monkeypatch):
    """The FP32 one-tensor-per-pair kernel with 1, 2, 3, 5 and 64 source tiles
    per work item on rods of 7 and 13 tiles (near and clean runs, a ragged last
    tile, runs longer than a tile row): every run length against the oracle and
    against one tile per item, and a mixed solve (which reduces through the
    per-tile slot counts) against the oracle's FP64 GMRES."""
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(4), np.arange(4), indexing="ij"), -1)
    pos = 1.5 * g.reshape(-1, 3).astype(float)
    alpha = I.isotropic_block(40.0 + 10.0j)
    k = 2 * np.pi / 480.0
    x = rand_x(len(pos), 13)
    P = O.Problem(pos, alpha, k, True, 1.5)
    yo = P.matvec(x)
    monkeypatch.setenv("LUMINA_FAR_BLOCK", "512")
    out = {}
    for kt in ("1", "2", "3", "5", "64"):
        monkeypatch.setenv("LUMINA_RUN_TILES", kt)
        s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=1.5,
                                           precision=MIXED)
        out[kt] = s.matvec(x, MIXED)
        if kt == "3":
            cs, _, info = s.compute_cross_sections(IncidentField(), 1e-8, 1000)
        s.close()
    for kt, y in out.items():
        e, e1 = rel(y, yo), rel(y, out["1"])
        print(f"rod n={len(pos)} run_tiles {kt}: rel err {e:.2e}, vs one tile {e1:.2e}")
        assert e < 2e-6
        assert e1 < 2e-6
    ref, _, _ = P.compute_cross_sections(1e-10, 1000)
    print(f"mixed solve with runs of 3: {info['iterations']} iterations, C_ext {cs[0]:.6e} "
          f"vs {ref[0]:.6e}")
    assert np.all(np.abs(cs - ref) <= 1e-6 * np.abs(ref[0]))


def test_stored_matrix_session(backend):
    r = np.random.default_rng(7)
    n = 300
    A = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    x = r.standard_normal(n) + 1j * r.standard_normal(n)
    s = backend.create_session(n)
    s.upload_matrix(A)
    assert rel(s.matvec(x), A @ x) < 1e-13


def test_large_cuboid_rows(backend):
    """38,637-dipole cuboid: FP64 and FP32 matvec rows against the oracle on a
    sample of rows (the oracle's cost is N per row)."""
    pos = I.au_cuboid_80x40x40_d1p5()
    lam, n_env = 520.0, 1.33
    alpha = I.polarisability_block("Au", lam, 1.5, n_env)
    k = I.wavenumber(lam, n_env)
    s = backend.create_pairsum_session(pos, alpha, k, use_fcd=True, cell_size=1.5)
    n = len(pos)
    x = rand_x(n, 8)
    y64 = s.matvec(x, FP64)
    y32 = s.matvec(x, MIXED)
    P = O.Problem(pos, alpha, k, True, 1.5)
    for lo in (0, n // 2, n - 64):
        yo = P.matvec_rows(x, lo, lo + 64)
        assert rel(y64[3 * lo:3 * lo + 192], yo) < 1e-12
        assert rel(y32[3 * lo:3 * lo + 192], yo) < 5e-6


def test_sphere_spectrum_31_wavelengths(backend):
    """The 515-dipole Au sphere in water over 400..700 nm (31 wavelengths),
    one FP64 batch at tol 1e-12: every C_ext, C_abs, C_sca within 1e-10 of
    the oracle's GMRES (the reference algorithm) at the same tolerance, and
    within 1e-9 of the exact (LAPACK) solution of the assembled operator
    (GMRES stops at a relative residual of 1e-12; the operator's condition
    near 700 nm turns that into a few 1e-10 of C_ext)."""
    pos = I.au_sphere_r10_d2()
    wls = [400.0 + 10.0 * t for t in range(31)]
    probs = [dict(positions=pos, polarisability=I.polarisability_block("Au", lam, 2.0, 1.33),
                  k=I.wavenumber(lam, 1.33), use_fcd=True, cell_size=2.0) for lam in wls]
    cs, infos, _ = batch_compute(backend, probs, 1e-12, 5000, precision=FP64)
    worst_g, worst_x = 0.0, 0.0
    for q, pr in enumerate(probs):
        P = O.Problem(pos, pr["polarisability"], pr["k"], True, 2.0)
        cso, _, io = P.compute_cross_sections(1e-12, 5000)
        assert io["status"] == 0 and infos[q]["status"] == 0
        worst_g = max(worst_g, float(np.max(np.abs(cs[q] - cso) / np.abs(cso))))
        e = P.incident()
        csx = P.cross_sections(np.linalg.solve(P.assemble(), e), e)
        worst_x = max(worst_x, float(np.max(np.abs(cs[q] - csx) / np.abs(csx))))
    print("spectrum worst rel err vs oracle GMRES", worst_g, "vs exact", worst_x,
          [i["iterations"] for i in infos])
    assert worst_g < 1e-10
    assert worst_x < 1e-9


def test_sphere_spectrum_tol_1e10_mgs(backend, monkeypatch):
    """The 31-wavelength sphere spectrum at its own tolerance (tol 1e-10)
    with modified Gram-Schmidt in the reference's order (LUMINA_ORTHO=mgs,
    iterative.rs:92-96): the same Arnoldi trajectory as the oracle's GMRES,
    up to the last bits of the dot products (parallel sums on the device):
    the same iteration count on 29 of the 31 wavelengths and one apart on
    the two slowest, and C_ext, C_abs, C_sca within 1e-10 of it (measured
    3.3e-11; CGS2, the default, stops along another trajectory: within
    1.4e-9 at this tolerance)."""
    monkeypatch.setenv("LUMINA_ORTHO", "mgs")
    pos = I.au_sphere_r10_d2()
    wls = [400.0 + 10.0 * t for t in range(31)]
    probs = [dict(positions=pos, polarisability=I.polarisability_block("Au", lam, 2.0, 1.33),
                  k=I.wavenumber(lam, 1.33), use_fcd=True, cell_size=2.0) for lam in wls]
    cs, infos, _ = batch_compute(backend, probs, 1e-10, 5000, precision=FP64)
    worst, same = 0.0, 0
    for q, pr in enumerate(probs):
        P = O.Problem(pos, pr["polarisability"], pr["k"], True, 2.0)
        cso, _, io = P.compute_cross_sections(1e-10, 5000)
        assert io["status"] == 0 and infos[q]["status"] == 0
        e = float(np.max(np.abs(cs[q] - cso) / np.abs(cso[0])))
        worst = max(worst, e)
        same += int(infos[q]["iterations"] == io["iterations"])
        assert abs(infos[q]["iterations"] - io["iterations"]) <= 1
        print(f"{wls[q]:.0f} nm: iterations {infos[q]['iterations']} (oracle {io['iterations']}), "
              f"rel {e:.1e}")
    print("mgs spectrum at tol 1e-10: worst", worst, "same iteration count", same, "of 31")
    assert same >= 28
    assert worst < 1e-10
