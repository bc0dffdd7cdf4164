"""Solve groups: a batch whose solves are split into two groups, each group's
restart cycles issued on its own stream with the far fields of the groups
taking turns, gives bit-identical moments, cross sections and iteration
counts to the one-group engine: every solve runs the same kernels on the
same data in the same order, only interleaved differently with the other
group's. The batch mixes every far-field path (free-space pair sums in the
row and one-tensor-per-pair forms, an FFT block-Toeplitz solve, a Fresnel
image solve and a 2D-lattice solve on the Ewald grid table), in FP64 and in
mixed precision, with captured graphs (adaptive and whole-solve), host-issued
cycles and timed graphs."""
import numpy as np
import pytest

from lumina_amd import inputs as I
from lumina_amd.solver import (FP64, MIXED, Batch, ConvergenceFailure, LatticeSpec,
                               SubstrateRuntime, substrate_reflection_factor)

pytestmark = pytest.mark.gpu

KEYS = ("LUMINA_GROUPS", "LUMINA_GRAPH", "LUMINA_FAR_BLOCK", "LUMINA_FFT", "LUMINA_GRAPH_CYCLES")


def mixed_batch():
    sphere = I.au_sphere_r10_d2()
    ell = I.discretise(I.Ellipsoid((8.0, 5.0, 4.0)), 2.0)
    d = I.depol_factors((8.0, 5.0, 4.0))
    cub = np.array([[-4.0 + 1.5 * i, 2.0 + 1.5 * j, 1.0 + 1.5 * k]
                    for i in range(6) for j in range(5) for k in range(4)])
    cyl = I.discretise(I.Cylinder(8.0, 6.0), 2.0)
    n_env = 1.33
    probs = []
    for lam in (400.0, 520.0):
        k = I.wavenumber(lam, n_env)
        probs.append(dict(positions=sphere, k=k, use_fcd=True, cell_size=2.0,
                          polarisability=I.polarisability_block("Au", lam, 2.0, n_env)))
    k = I.wavenumber(450.0, n_env)
    probs.append(dict(positions=ell, k=k, use_fcd=True, cell_size=2.0,
                      polarisability=I.polarisability_block("Ag", 450.0, 2.0, n_env, depol=d)))
    k = I.wavenumber(600.0, n_env)
    probs.append(dict(positions=cub, k=k, use_fcd=True, cell_size=1.5,
                      polarisability=I.polarisability_block("Au", 600.0, 1.5, n_env)))
    glass = substrate_reflection_factor(2.25 + 0j, n_env, True)
    probs.append(dict(positions=sphere, k=k, use_fcd=True, cell_size=2.0,
                      polarisability=I.polarisability_block("Au", 600.0, 2.0, n_env),
                      substrate=SubstrateRuntime(z_interface=float(sphere[:, 2].min()) - 1.0,
                                                 delta_eps=glass)))
    k1 = I.wavenumber(600.0, 1.0)
    probs.append(dict(positions=cyl, k=k1, use_fcd=True, cell_size=2.0,
                      polarisability=I.polarisability_block("Au", 600.0, 2.0, 1.0),
                      lattice=LatticeSpec((300.0, 0.0, 0.0), (0.0, 300.0, 0.0), None, 5, 5),
                      substrate=SubstrateRuntime(
                          z_interface=float(cyl[:, 2].min()) - 1.0,
                          delta_eps=substrate_reflection_factor(2.25 + 0j, 1.0, True))))
    return probs


def run(backend, probs, precision, monkeypatch, env, timing=False):
    for key in KEYS:
        monkeypatch.delenv(key, raising=False)
    monkeypatch.setenv("LUMINA_FFT", "force")
    for key, val in env.items():
        monkeypatch.setenv(key, val)
    b = Batch(backend, probs)
    if timing:
        b.set_timing(True)
    b.run(1e-10, 300, precision)
    try:
        cs, infos, p = b.results(want_moments=True)
    except ConvergenceFailure as e:
        cs, infos, p = e.cross_sections, e.infos, e.moments
    kt = b.kernel_times()
    b.close()
    return cs, infos, p, kt


@pytest.mark.parametrize("precision", [FP64, MIXED])
@pytest.mark.parametrize("form", ["256", "sym"])
def test_groups_bitwise(backend, precision, form, monkeypatch):
    probs = mixed_batch()
    base = run(backend, probs, precision, monkeypatch,
               {"LUMINA_GROUPS": "1", "LUMINA_FAR_BLOCK": form})
    assert base[3]["solve_groups"] == 1
    assert all(i["solve_group"] == 0 for i in base[1])
    variants = {
        "graph": ({"LUMINA_GROUPS": "2", "LUMINA_FAR_BLOCK": form}, False),
        "host": ({"LUMINA_GROUPS": "2", "LUMINA_FAR_BLOCK": form, "LUMINA_GRAPH": "0"}, False),
        "whole": ({"LUMINA_GROUPS": "2", "LUMINA_FAR_BLOCK": form,
                   "LUMINA_GRAPH_CYCLES": "all"}, False),
        "timed": ({"LUMINA_GROUPS": "2", "LUMINA_FAR_BLOCK": form}, True),
    }
    it0 = [i["iterations"] for i in base[1]]
    print("one group", it0, [i["status"] for i in base[1]])
    for name, (env, timing) in variants.items():
        cs, infos, p, kt = run(backend, probs, precision, monkeypatch, env, timing)
        print(name, [i["iterations"] for i in infos], [i["solve_group"] for i in infos], kt)
        assert kt["solve_groups"] == 2, name
        assert sorted({i["solve_group"] for i in infos}) == [0, 1], name
        assert [i["iterations"] for i in infos] == it0, name
        assert [i["status"] for i in infos] == [i["status"] for i in base[1]], name
        assert np.array_equal(cs, base[0]), name
        for a, b_ in zip(p, base[2]):
            assert np.array_equal(a, b_), name
        if timing:
            n = kt["far32_launches"] if precision == MIXED else kt["far64_launches"]
            assert n > 0 and kt["far64_launches"] > 0, name
            assert (kt["far32_ms"] if precision == MIXED else kt["far64_ms"]) > 0, name


def test_groups_default_threshold(backend, monkeypatch):
    """Small batches keep one group by default; LUMINA_GROUPS=2 splits even
    two solves, one per group."""
    probs = mixed_batch()[:2]
    for key in KEYS:
        monkeypatch.delenv(key, raising=False)
    cs1, infos1, _, kt1 = run(backend, probs, FP64, monkeypatch, {})
    assert kt1["solve_groups"] == 1
    cs2, infos2, _, kt2 = run(backend, probs, FP64, monkeypatch, {"LUMINA_GROUPS": "2"})
    assert kt2["solve_groups"] == 2
    assert sorted(i["solve_group"] for i in infos2) == [0, 1]


def test_graph_cycles_entry_point(backend, monkeypatch):
    """lh_batch_set_graph_cycles: -1 solves the whole batch in one graph
    launch, N fixes the cycles per launch; the results equal the adaptive
    default's bit for bit. Values below -1 are refused."""
    from lumina_amd.solver import LuminaError
    probs = mixed_batch()
    for key in KEYS:
        monkeypatch.delenv(key, raising=False)
    monkeypatch.setenv("LUMINA_FAR_BLOCK", "256")
    monkeypatch.setenv("LUMINA_GROUPS", "2")
    b = Batch(backend, probs)
    out = {}
    for cyc in (0, -1, 3):
        b.set_graph_cycles(cyc)
        b.run(1e-10, 300, MIXED)
        try:
            cs, infos, _ = b.results()
        except ConvergenceFailure as e:
            cs, infos = e.cross_sections, e.infos
        out[cyc] = (cs, [i["iterations"] for i in infos], b.kernel_times())
    with pytest.raises(LuminaError):
        b.set_graph_cycles(-2)
    b.close()
    ncyc = max(i for i in out[0][1]) // 30 + 1
    assert out[-1][2]["host_round_trips"] == 1
    assert out[3][2]["host_round_trips"] == -(-out[3][2]["cycles_issued"] // 3)
    assert out[0][2]["host_round_trips"] > 1 and ncyc > 1
    for cyc in (-1, 3):
        assert out[cyc][1] == out[0][1]
        assert np.array_equal(out[cyc][0], out[0][0])


def test_sweep_groups_bitwise(backend, monkeypatch):
    """The device wavelength sweep (lh_sweep_solve) solves its wavelengths as
    one batch: with two solve groups its rows equal the one-group rows bit
    for bit."""
    from lumina_amd import sweep as W
    pos = I.au_sphere_r10_d2()
    wls = [400.0 + 25.0 * t for t in range(9)]
    out = {}
    for g in ("1", "2"):
        for key in KEYS:
            monkeypatch.delenv(key, raising=False)
        monkeypatch.setenv("LUMINA_FAR_BLOCK", "256")
        monkeypatch.setenv("LUMINA_GROUPS", g)
        cs, infos = W.run_sweep(backend, pos, wls, [W.johnson_christy("Au")], spacing=2.0,
                                n_env=1.33, tolerance=1e-10, max_iterations=2000)
        out[g] = (cs, [i["iterations"] for i in infos], [i["solve_group"] for i in infos])
        print(g, out[g][1], out[g][2])
    assert sorted(set(out["2"][2])) == [0, 1] and set(out["1"][2]) == {0}
    assert out["2"][1] == out["1"][1]
    assert np.array_equal(out["2"][0], out["1"][0])


@pytest.mark.parametrize("precision", [FP64, MIXED])
@pytest.mark.parametrize("form", ["256", "sym"])
def test_solve_independent_of_batch_position(backend, precision, form, monkeypatch):
    """A solve's result does not depend on the other solves of its batch or
    on its place among them: every problem of the mixed batch (free-space
    pair sums, FFT, Fresnel image, 2D lattice) solved alone in a batch of one
    equals its row of the whole batch bit for bit. Offsets into the shared
    device buffers (vectors, source records, partial slices, tables) differ
    between the two runs, so a mix-up of any of them shows here."""
    probs = mixed_batch()
    env = {"LUMINA_GROUPS": "2", "LUMINA_FAR_BLOCK": form}
    cs, infos, p, _ = run(backend, probs, precision, monkeypatch, env)
    for q, pr in enumerate(probs):
        cs1, inf1, p1, _ = run(backend, [pr], precision, monkeypatch, env)
        print(q, infos[q]["iterations"], inf1[0]["iterations"], cs[q], cs1[0])
        assert np.all(np.isfinite(cs1[0])), q
        assert inf1[0]["iterations"] == infos[q]["iterations"], q
        assert np.array_equal(cs1[0], cs[q]), q
        assert np.array_equal(p1[0], p[q]), q
