"""Benchmark of the CDA solve hot path on one MI355X.

Default workload: 64 Au/Ag spheres and ellipsoids (numpy PCG64 seed 0,
radius U(5, 30) nm, semi-axis ratios U(1, 2.5), dipole spacing 2 nm, FCD with
cell_size = spacing) in water (n = 1.33), x-polarised plane wave along +z,
each at 16 wavelengths 400..800 nm: 1024 GMRES(30) solves of uneven size,
solved together by the batched device engine in mixed precision (FP32
far-field pair sum, FP64 Krylov/Givens/residuals/near field).

A step is one batched solve of all 1024 problems (inputs resident in HBM),
issued as one hipGraph launch that holds every restart cycle up to max_iter.
Printed: one JSON line with Green pair-evals/s (the value), solves/s and
converged solves/s, the same batch with adaptive graph launches, the plain
FP64 one-solve-per-session rate on a subset of the same problems, the C_ext
error of the mixed results against an FP64 batch of all 1024 problems (same
tolerance), the far-field kernel roofline from HIP events, and the CPU
baseline (the oracle's FP64 restatement of the reference CPU path on this
host's cores).

Other workloads: --workload cuboid (80x40x40 nm Au cuboid, 1.5 nm spacing,
38,637 dipoles, 520 nm, one solve FP64 and mixed); --workload substrate (the
same cuboid on glass with the retarded Fresnel image term, one solve FP64 and
mixed, the image-pair kernel timed by HIP events); --workload metasurface
(Au cylinder D=100 nm H=40 nm at 2 nm spacing, 41,181 dipoles, on glass with
the retarded Fresnel image, 300 nm square lattice with 2D Ewald sums, Bloch
k = 0, a batch of wavelengths in 500..890 nm, FP64 and mixed); --workload
sphere (Au sphere r=10 nm, d=2 nm, 515 dipoles, water, 31 wavelengths
400..700 nm, FP64 GMRES tol 1e-10, one batched solve, compared with the
oracle); --workload helix (Au helix, coil 25 nm, pitch 40 nm, 4 turns, wire
8 nm, d=1.5 nm, 35,013 dipoles, water, 600 nm, left and right circular
light: forward + adjoint solves and the dC_ext/d(eps_i, r_i) gradients,
checked against the oracle's CPU adjoint on sampled dipoles and against
central differences); --workload fields (the solved cuboid's 400x400
near-field map and 181x360 far-field pattern, FP64, checked against the
oracle on sampled points); --workload sweep (the sphere's 31-wavelength CLI
run end to end with the per-wavelength set-up on the device, beside the
same run set up on the host).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X published vector FP32
FP64_PEAK_TFLOPS = 78.6    # MI355X published vector FP64
FLOPS_PER_FAR_PAIR = 90    # point-G far pair: y += A x + B d (d . x) (FMA = 2 flops)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def batch_problems(n_particles=64, n_wl=16, n_env=1.33, spacing=2.0):
    from lumina_amd import inputs as I
    return I.batch_workload(n_particles, n_wl, n_env, spacing)


def pairs(n):
    return n * (n - 1.0)


def cpu_model():
    """The host CPU as /proc/cpuinfo names it: model name, family, model,
    stepping."""
    keys = {"model name": None, "cpu family": None, "model": None, "stepping": None}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k in keys and keys[k] is None:
                    keys[k] = v.strip()
                if line.strip() == "":
                    break
    except OSError:
        pass
    name = keys["model name"] or "unknown"
    return (f"{name} (family {keys['cpu family']}, model {keys['model']}, "
            f"stepping {keys['stepping']})")


def cpu_threads():
    """Threads of the CPU baseline: every CPU this process may run on (its
    affinity mask), at most 16 -- one GPU's share of the host's cores."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def host_cpus():
    out = {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "threads_used": cpu_threads()}
    try:
        out["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        out["affinity"] = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            out["cgroup_cpu_max"] = f.read().strip()
    except OSError:
        out["cgroup_cpu_max"] = None
    return out


def time_rows(P, x, n, rows, nthreads):
    """Seconds for `rows` target rows of the oracle's on-the-fly matvec, as
    nthreads blocks spread evenly over the dipoles, one block per thread
    (the ctypes call releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    per = max(1, -(-rows // nthreads))
    nb = max(1, min(nthreads, n // per)) if per <= n else 1
    per = min(per, n)
    starts = np.linspace(0, n - per, nb).astype(int)
    with ThreadPoolExecutor(nb) as ex:
        t0 = time.perf_counter()
        list(ex.map(lambda a: P.matvec_rows(x, int(a), int(a) + per), starts))
        dt = time.perf_counter() - t0
    return dt, nb * per


def cpu_baseline(probs, infos64, budget_s=24.0):
    """The oracle (C++ FP64 restatement of the reference CPU path:
    matvec_on_the_fly's rows over a thread pool, assembly.rs:224-226) on the
    batch's own size mix. For every particle (at its middle wavelength) a
    sample of target rows spread over its dipoles is timed with all threads
    and with one, giving its row rate; the CPU time of the whole workload is
    then sum over solves of matvecs x N(N-1) / rate, with each solve's FP64
    matvec count (iterations + 2 cycles, as the reference's GMRES performs
    them). value = the workload's pair evaluations / that time. Extrapolated
    from the sampled rows; GMRES vector work (serial in the reference) is not
    timed."""
    from oracle import pyoracle as O
    threads = cpu_threads()
    parts = {}
    for q, p in enumerate(probs):
        parts.setdefault(p["particle"], []).append(q)
    per_part = budget_s / (1.5 * len(parts))
    rate_all, rate_one = {}, {}
    t_start = time.perf_counter()
    rng = np.random.default_rng(0)
    for ip, qs in parts.items():
        p = probs[qs[len(qs) // 2]]
        n = p["n"]
        P = O.Problem(p["positions"], p["polarisability"], p["k"], True, p["cell_size"])
        x = rng.standard_normal(3 * n) + 1j * rng.standard_normal(3 * n)
        for nt, budget, store in ((threads, per_part, rate_all), (1, 0.5 * per_part, rate_one)):
            # a warm-up pass sizes the sample to the budget; then the median of 3
            rows = min(n, 4 * nt)
            dt, done = time_rows(P, x, n, rows, nt)
            rows = int(min(n, max(rows, rows * budget / max(dt, 1e-6))))
            rates = []
            for _ in range(3):
                dt, done = time_rows(P, x, n, rows, nt)
                rates.append(done * (n - 1) / dt)
            store[ip] = float(np.median(rates))
    work = t_all = t_one = 0.0
    for q, p in enumerate(probs):
        i = infos64[q]
        w = (i["iterations"] + 2 * i["cycles"]) * pairs(p["n"])
        work += w
        t_all += w / rate_all[p["particle"]]
        t_one += w / rate_one[p["particle"]]
    hc = host_cpus()
    return {"value": work / t_all, "unit": "pair-evals/s", "cores": threads, "kind": "port",
            "extrapolated": True, **hc,
            "one_thread_value": work / t_one, "workload_cpu_s_all_threads": t_all,
            "workload_cpu_s_one_thread": t_one, "workload_pair_evals": work,
            "sample": f"oracle matvec rows of each of the {len(parts)} particles (middle "
                      f"wavelength) spread over its dipoles, on {threads} threads and on 1 "
                      f"(median of 3 after a warm-up); "
                      f"the whole {len(probs)}-solve workload's FP64 matvecs (iterations + 2 "
                      f"cycles each) priced at those row rates (extrapolated); "
                      f"{time.perf_counter() - t_start:.1f} s of sampling"}


def cpu_matvec_rate(P, n, pairs_per_matvec, rows=None, reps=3):
    """pair-evals/s of the oracle's on-the-fly matvec on this host: full
    matvecs (median of reps after a warm-up), or, with rows, the first
    ``rows`` target rows spread over a thread pool (sampled, scaled by
    rows / n)."""
    from concurrent.futures import ThreadPoolExecutor
    threads = cpu_threads()
    x = np.random.default_rng(0).standard_normal(3 * n) + 0j
    if rows is None:
        P.matvec(x, nthreads=threads)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            P.matvec(x, nthreads=threads)
            ts.append(time.perf_counter() - t0)
        dt = float(np.median(ts))
        return {"value": pairs_per_matvec / dt, "unit": "pair-evals/s", "cores": threads,
                "kind": "port", **host_cpus(),
                "sample": f"median of {reps} full oracle matvecs (N={n}) after a warm-up, "
                          f"{dt:.2f} s each"}
    # `threads` blocks of rows spread evenly over the dipoles
    per = max(1, rows // threads)
    starts = np.linspace(0, n - per, threads).astype(int)
    chunks = [(int(a), int(a) + per) for a in starts]
    rows = per * threads
    def work(c):
        P.matvec_rows(x, c[0], c[1])
    with ThreadPoolExecutor(threads) as ex:
        t0 = time.perf_counter()
        list(ex.map(work, chunks))
        dt = time.perf_counter() - t0
    return {"value": pairs_per_matvec * rows / n / dt, "unit": "pair-evals/s", "cores": threads,
            "kind": "port", "extrapolated": True, **host_cpus(),
            "sample": f"oracle matvec, {threads} blocks of {per} rows spread over N={n}, on "
                      f"{threads} threads in {dt:.2f} s, scaled to the full matvec (extrapolated)"}


def results_allow_failures(batch):
    """Batch results; solves that hit max_iter are reported like the
    reference's ConvergenceFailure rows (status 5) instead of aborting."""
    from lumina_amd.solver import ConvergenceFailure
    try:
        cs, infos, _ = batch.results()
    except ConvergenceFailure as e:
        cs, infos = e.cross_sections, e.infos
    return cs, infos


def far_kernel_name(prec_mixed, blk):
    # block 512: the one-tensor-per-unordered-pair kernels with LDS-staged
    # sources (with and without the exclusion test); 1024: the FP64 one whose
    # sources rotate around the wave; otherwise the row kernels
    hw = os.environ.get("LUMINA_FP32_SINCOS", "hw") != "sw"
    if blk == 512:
        return "k_far_sym32<true|false>" if prec_mixed else "k_far_sym64<true|false>"
    if blk == 1024:
        return "k_far_sym<double>"
    return (f"k_far_row<float, false, {str(hw).lower()}, true|false>" if prec_mixed
            else "k_far_row<double, false, false, true|false>")


def run_batch(args):
    from lumina_amd.solver import (FP64, MIXED, Batch, ConvergenceFailure, HipBackend,
                                   IncidentField, PairSumSession)
    probs = batch_problems(args.particles, args.wavelengths)
    S = len(probs)
    nw = args.wavelengths
    log(f"[bench] batch: {S} solves, N in [{min(p['n'] for p in probs)}, "
        f"{max(p['n'] for p in probs)}], total dipoles {sum(p['n'] for p in probs)}")
    be = HipBackend(0)
    t0 = time.perf_counter()
    batch = Batch(be, probs)
    log(f"[bench] batch resident in HBM ({time.perf_counter() - t0:.1f} s to build)")
    # far-field HIP events, recorded as nodes of the captured cycle graphs:
    # the timed path is the graph path that ships
    batch.set_timing(True)
    # the whole batch solve as one graph launch (every restart cycle up to
    # max_iter captured once) and one host read
    batch.set_graph_cycles(-1)

    prec = MIXED if args.precision == "mixed" else FP64
    for w in range(args.warmup):
        batch.run(args.tol, args.max_iter, prec)
        log(f"[bench] warmup {w} done")
    times, evals = [], 0.0
    kt_tot = {"far32_ms": 0.0, "far32_pairs": 0.0, "far32_launches": 0,
              "far64_ms": 0.0, "far64_pairs": 0.0, "far64_launches": 0}
    trips = []
    groups = None
    rounds = 0   # far-field launches that had work: lock-step rounds up to the longest solve
    for s_ in range(args.steps):
        t1 = time.perf_counter()
        batch.run(args.tol, args.max_iter, prec)
        dt = time.perf_counter() - t1
        times.append(dt)
        cs_mixed, infos = results_allow_failures(batch)
        evals += sum(i["pair_evals"] for i in infos)
        # far-field launches with work: each solve group launches its own,
        # one per lock-step round up to the group's longest solve
        for g in sorted({i["solve_group"] for i in infos}):
            rounds += max((i["matvecs_fp32"] if prec == MIXED else i["matvecs"] - i["matvecs_fp32"])
                          for i in infos if i["solve_group"] == g)
        kt = batch.kernel_times()
        for key in kt_tot:
            kt_tot[key] += kt[key]
        trips.append((kt["host_round_trips"], kt["cycles_issued"]))
        groups = kt["solve_groups"]
        blocks = (kt["far64_block"], kt["far32_block"], kt["tune_ms"])
        log(f"[bench] step {s_}: {dt:.2f} s, iters mean "
            f"{np.mean([i['iterations'] for i in infos]):.1f}, max "
            f"{max(i['iterations'] for i in infos)}, {kt['host_round_trips']} host round trips "
            f"({kt['cycles_issued']} cycles)")
    total = sum(times)
    value = evals / total
    iters = [i["iterations"] for i in infos]
    conv = np.array([i["status"] == 0 for i in infos])

    # the same batch with adaptive graph launches (1, 2, 4, then 8 restart
    # cycles per launch, one host read after each): one warm-up run that
    # captures those graphs, one timed run
    adaptive = None
    if not args.no_adaptive:
        batch.set_timing(False)
        batch.set_graph_cycles(0)
        batch.run(args.tol, args.max_iter, prec)
        t1 = time.perf_counter()
        batch.run(args.tol, args.max_iter, prec)
        dt1 = time.perf_counter() - t1
        cs1, inf1 = results_allow_failures(batch)
        kt1 = batch.kernel_times()
        adaptive = {"pair_evals_per_s": sum(i["pair_evals"] for i in inf1) / dt1,
                    "s": dt1, "host_round_trips": kt1["host_round_trips"],
                    "cycles_issued": kt1["cycles_issued"],
                    "identical_to_timed_path": bool(np.array_equal(cs1, cs_mixed))}
        log(f"[bench] adaptive graph launches: {dt1:.2f} s, "
            f"{kt1['host_round_trips']} host round trips, same results: "
            f"{adaptive['identical_to_timed_path']}")
    batch.close()

    # FP64 answer of every solve, one FP64 batch at the same tolerance (the
    # FP64 far-field form the mixed batch's tuner chose for its residuals)
    err, fp64_batch, infos64 = None, None, None
    if not args.no_error:
        os.environ["LUMINA_FAR_BLOCK"] = {512: "sym"}.get(blocks[0], str(blocks[0]))
        b64 = Batch(be, probs)
        os.environ.pop("LUMINA_FAR_BLOCK")
        t1 = time.perf_counter()
        b64.run(args.tol, args.max_iter, FP64)
        dt64 = time.perf_counter() - t1
        cs_ref, infos64 = results_allow_failures(b64)
        b64.close()
        conv64 = np.array([i["status"] == 0 for i in infos64])
        ok = conv & conv64
        errs = np.abs(cs_mixed[:, 0] - cs_ref[:, 0]) / np.abs(cs_ref[:, 0])
        err = float(np.max(errs[ok])) if ok.any() else None
        fail = ~conv & ~conv64
        one = conv != conv64
        fp64_batch = {
            "pair_evals_per_s": sum(i["pair_evals"] for i in infos64) / dt64,
            "solves_per_s": S / dt64, "s": dt64, "solves": S, "converged": int(conv64.sum()),
            "note": "a single run: its time includes the graph capture",
            "cext_err_solves": int(ok.sum()),
            "cext_rel_err_median_vs_fp64": float(np.median(errs[ok])) if ok.any() else None,
            "max_iter_rows": {
                "both_precisions": int(fail.sum()),
                "cext_rel_spread_max": float(np.max(errs[fail])) if fail.any() else None,
                "cext_rel_spread_median": float(np.median(errs[fail])) if fail.any() else None},
            "converged_in_one_precision_only": [
                {"solve": int(q), "particle": probs[q]["particle"], "n": probs[q]["n"],
                 "wl": probs[q]["wl"], "mixed": [infos[q]["iterations"], infos[q]["residual"]],
                 "fp64": [infos64[q]["iterations"], infos64[q]["residual"]]}
                for q in np.flatnonzero(one)]}
        log(f"[bench] batched FP64 ({S} solves): {dt64:.1f} s; max C_ext err "
            f"(both converged, {int(ok.sum())} solves): {err}")

    # plain FP64: one session and solve per problem, the problems at
    # args.plain_wavelengths of the wavelengths (every particle)
    plain = None
    if not args.no_fp64_plain:
        npw = max(1, min(args.plain_wavelengths, nw))
        widx = sorted(set(np.linspace(0, nw - 1, npw).round().astype(int).tolist()))
        sub = [p for q, p in enumerate(probs) if (q % nw) in widx]
        t_pl, ev_pl, conv_pl = 0.0, 0.0, 0
        # each session as a caller would run it: its own tuner picks the
        # far-field form for its problem (the time includes the tuning)
        os.environ.pop("LUMINA_FAR_BLOCK", None)
        for p in sub:
            s_ = PairSumSession(be, p["positions"], p["polarisability"], p["k"], True,
                                p["cell_size"], precision=FP64)
            t1 = time.perf_counter()
            try:
                out, _, info = s_.compute_cross_sections(IncidentField(), args.tol, args.max_iter)
                conv_pl += 1
            except ConvergenceFailure as e:
                info = e.info
            t_pl += time.perf_counter() - t1
            ev_pl += info["pair_evals"]
            s_.close()
        plain = {"pair_evals_per_s": ev_pl / t_pl, "solves_per_s": len(sub) / t_pl,
                 "solves": len(sub), "converged": conv_pl,
                 "far64_form": "each session's own tuner (included in the time)",
                 "sample": f"{len(sub)} problems (every particle at wavelength indices {widx} "
                           f"of {nw}), one FP64 session and solve each, tol {args.tol:g}"}
        log(f"[bench] plain FP64: {plain['pair_evals_per_s']:.3e} pair-evals/s on {len(sub)} solves")

    far_ms = kt_tot["far32_ms"] if prec == MIXED else kt_tot["far64_ms"]
    far_pairs = kt_tot["far32_pairs"] if prec == MIXED else kt_tot["far64_pairs"]
    far_n = kt_tot["far32_launches"] if prec == MIXED else kt_tot["far64_launches"]
    achieved = FLOPS_PER_FAR_PAIR * far_pairs / (far_ms * 1e-3) / 1e12 if far_ms > 0 else 0.0
    peak = FP32_PEAK_TFLOPS if prec == MIXED else FP64_PEAK_TFLOPS
    blk = blocks[1] if prec == MIXED else blocks[0]
    kname = far_kernel_name(prec == MIXED, blk)
    # HBM bytes per launch of this kernel on this workload, from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a plain copy)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "far_field_traffic.json")
    if (prec == MIXED and args.particles == 64 and args.wavelengths == 16
            and os.path.exists(tf)):
        with open(tf) as f:
            tj = json.load(f)
        traffic = tj.get("bytes_per_launch_by_solve_groups", {}).get(str(groups),
                                                                     tj.get("bytes_per_launch"))
    roof = {"bound": "valu", "kernel": kname,
            "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": traffic, "avg_launch_ms": far_ms / max(rounds, 1), "launches": rounds,
            "event_pairs": far_n,
            "flops_per_pair": FLOPS_PER_FAR_PAIR, "targets_per_item": blk,
            "far_pairs_per_s": far_pairs / (far_ms * 1e-3) if far_ms > 0 else 0.0,
            "events": ("HIP events on each solve group's stream around every far-field launch "
                       "of the timed steps (a launch is one group's far field; the groups' far "
                       "fields take turns and never overlap), " + ("recorded by the host-issued cycles"
                                          if os.environ.get("LUMINA_GRAPH") == "0" else
                                          "event-record nodes of the captured cycle graphs") +
                       "; avg_launch_ms divides their total by the launches that had work (per "
                       "group, the lock-step rounds up to its longest solve's matvec count); "
                       "event_pairs "
                       "also counts the launches of cycles issued after every solve had "
                       "finished, whose kernels exit at once")}

    line = {
        "metric": ("Green pair-evals/s (batched mixed-precision CDA solves)" if prec == MIXED
                   else "Green pair-evals/s (batched FP64 CDA solves)"),
        "value": value, "unit": "pair-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * total / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 far field + fp64 Krylov" if prec == MIXED else "fp64",
        "data": "synthetic (seeded discretised particles, J&C Au/Ag)",
        "config": {"workload": f"{args.particles} Au/Ag spheres+ellipsoids x {nw} "
                               f"wavelengths 400-800 nm, d=2 nm, FCD, water, GMRES(30) "
                               f"tol {args.tol:g}, one batched solve per step",
                   "solves": S, "dipoles": int(sum(p["n"] for p in probs)),
                   "n_min": int(min(p["n"] for p in probs)),
                   "n_max": int(max(p["n"] for p in probs)), "parallelism": "single GPU"},
        "solves_per_s": S * args.steps / total,
        "converged_solves_per_s": int(conv.sum()) * args.steps / total,
        "converged": int(conv.sum()), "failed_max_iter": int((~conv).sum()),
        "iterations_mean": float(np.mean(iters)), "iterations_max": int(max(iters)),
        "max_iter": args.max_iter,
        "host_round_trips_per_step": trips[-1][0] if trips else None,
        "cycles_issued_per_step": trips[-1][1] if trips else None,
        "graph_cycles_per_launch": "every cycle up to max_iter (lh_batch_set_graph_cycles -1)",
        "solve_groups": groups,
        "cext_rel_err_max_vs_fp64": err,
        "cext_err_covers": (f"{fp64_batch['cext_err_solves']} of {S} solves (those that "
                            f"converge in both precisions)") if fp64_batch else None,
        "fp64_batched": fp64_batch, "fp64_plain": plain, "adaptive_graph_launches": adaptive,
        "speedup_vs_fp64_plain_pair_evals": (value / plain["pair_evals_per_s"]) if plain else None,
        # solves/s of the whole batch over solves/s of the plain sessions on
        # their sample (every particle at evenly spaced wavelengths)
        "speedup_vs_fp64_plain_solves": ((S * args.steps / total) / plain["solves_per_s"])
                                        if plain else None,
        "roofline": roof,
    }
    if not args.no_cpu_baseline and infos64 is not None:
        line["cpu_baseline"] = cpu_baseline(probs, infos64)
        log(f"[bench] cpu baseline {line['cpu_baseline']['value']:.3e} pair-evals/s "
            f"({line['cpu_baseline']['cores']} threads)")
    return line


def run_cuboid(args):
    from lumina_amd import inputs as I
    from lumina_amd.solver import FP64, MIXED, IncidentField, PairSumSession, HipBackend
    pos = I.au_cuboid_80x40x40_d1p5()
    lam, n_env, d = 520.0, 1.33, 1.5
    alpha = I.polarisability_block("Au", lam, d, n_env)
    k = I.wavenumber(lam, n_env)
    be = HipBackend(0)
    res = {}
    n = len(pos)
    # the on-the-fly pair sum (LUMINA_FFT=0), then the FFT block-Toeplitz
    # matvec the full grid qualifies for
    for name, prec, fft in (("fp64", FP64, "0"), ("mixed", MIXED, "0"), ("fft_fp64", FP64, "1"),
                            ("fft_mixed", MIXED, "1")):
        os.environ["LUMINA_FFT"] = fft
        s = PairSumSession(be, pos, alpha, k, True, d, precision=prec)
        os.environ.pop("LUMINA_FFT", None)
        for _ in range(args.warmup):
            s.compute_cross_sections(IncidentField(), args.tol, args.max_iter)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out, p, info = s.compute_cross_sections(IncidentField(), args.tol, args.max_iter)
        dt = (time.perf_counter() - t0) / args.steps
        res[name] = {"cs": out.tolist(), "s": dt, "pair_evals_per_s": info["pair_evals"] / dt,
                     "info": info}
        if fft == "1":
            # no pairs are evaluated: the rate of the pair sum it replaces
            res[name]["equivalent_pair_evals_per_s"] = info["matvecs"] * pairs(n) / dt
        log(f"[bench] cuboid {name}: {dt:.4f} s, {info}")
        s.close()
    err = abs(res["mixed"]["cs"][0] - res["fp64"]["cs"][0]) / abs(res["fp64"]["cs"][0])
    fft_err = max(abs(res[f]["cs"][0] - res["fp64"]["cs"][0]) / abs(res["fp64"]["cs"][0])
                  for f in ("fft_fp64", "fft_mixed"))
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as O
        cpu = cpu_matvec_rate(O.Problem(pos, alpha, k, True, d), n, pairs(n), rows=4096)
        mv = res["fp64"]["info"]["matvecs"]
        cpu["solve_s_extrapolated"] = mv * pairs(n) / cpu["value"]
        log(f"[bench] cuboid cpu baseline {cpu['value']:.3e} pair-evals/s")
    return {
        "cpu_baseline": cpu,
        "metric": "Green pair-evals/s (single large mixed-precision CDA solve)",
        "value": res["mixed"]["pair_evals_per_s"], "unit": "pair-evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * res["mixed"]["s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 far field + fp64 Krylov", "data": "synthetic",
        "config": {"workload": "Au cuboid 80x40x40 nm, d=1.5 nm (38,637 dipoles), 520 nm, water, "
                               f"FCD, GMRES(30) tol {args.tol:g}", "parallelism": "single GPU"},
        "fp64": res["fp64"], "mixed": res["mixed"], "cext_rel_err_mixed_vs_fp64": err,
        "fft": {"fp64": res["fft_fp64"], "mixed": res["fft_mixed"],
                "cext_rel_err_vs_pair_sum_fp64": fft_err,
                "note": "FFT block-Toeplitz matvec (fft_matvec.rs) with the FCD correction on "
                        "the near list; no pair evaluations, equivalent rate = matvecs x N(N-1) "
                        "/ time"},
    }


def run_substrate(args):
    """The free-space Fresnel image term (assembly.rs:278-295): the 80x40x40 nm
    Au cuboid (38,637 dipoles, d = 1.5 nm) at 520 nm in water on glass (eps
    2.25, retarded factor, interface half a spacing below the lowest dipole
    plane). Per matvec N(N-1) direct pairs plus N^2 image pairs. One FP64 and
    one mixed solve, each a batch of one whose far-field launches (direct and
    image tensor per pair, k_far_row<R, true, ...>) are timed by HIP events."""
    from lumina_amd import inputs as I
    from lumina_amd.solver import (FP64, MIXED, Batch, HipBackend, SubstrateRuntime,
                                   substrate_reflection_factor)
    pos = I.au_cuboid_80x40x40_d1p5()
    lam, n_env, d = 520.0, 1.33, 1.5
    k = I.wavenumber(lam, n_env)
    sub = SubstrateRuntime(z_interface=float(pos[:, 2].min()) - 0.5 * d,
                           delta_eps=substrate_reflection_factor(2.25 + 0j, n_env, True))
    prob = dict(positions=pos, k=k, use_fcd=True, cell_size=d, substrate=sub,
                polarisability=I.polarisability_block("Au", lam, d, n_env))
    n = len(pos)
    be = HipBackend(0)
    res = {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        b = Batch(be, [prob])
        b.set_timing(True)
        for _ in range(args.warmup):
            b.run(args.tol, args.max_iter, prec)
        t1 = time.perf_counter()
        for _ in range(args.steps):
            b.run(args.tol, args.max_iter, prec)
        dt = (time.perf_counter() - t1) / args.steps
        cs, infos = results_allow_failures(b)
        kt = b.kernel_times()
        b.close()
        info = infos[0]
        far_ms = kt["far32_ms"] if prec == MIXED else kt["far64_ms"]
        far_pairs = kt["far32_pairs"] if prec == MIXED else kt["far64_pairs"]
        far_n = kt["far32_launches"] if prec == MIXED else kt["far64_launches"]
        # direct pairs 90 flops, image pairs 91: 90.5 per pair of the count
        tflops = 90.5 * far_pairs / (far_ms * 1e-3) / 1e12 if far_ms > 0 else None
        peak = FP32_PEAK_TFLOPS if prec == MIXED else FP64_PEAK_TFLOPS
        res[name] = {"s": dt, "pair_evals_per_s": info["pair_evals"] / dt, "cs": cs[0].tolist(),
                     "info": info, "far_kernel_ms_per_launch": far_ms / far_n if far_n else None,
                     "far_kernel_pairs_per_s": far_pairs / (far_ms * 1e-3) if far_ms > 0 else None,
                     "far_kernel_tflops": tflops,
                     "far_kernel_frac": tflops / peak if tflops else None,
                     "far_block": kt["far32_block"] if prec == MIXED else kt["far64_block"]}
        log(f"[bench] substrate {name}: {dt:.4f} s, {info['iterations']} iterations, far kernel "
            f"{res[name]['far_kernel_pairs_per_s']:.3e} pairs/s ({tflops:.1f} TFLOP/s)")
    err = abs(res["mixed"]["cs"][0] - res["fp64"]["cs"][0]) / abs(res["fp64"]["cs"][0])
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as O
        P = O.Problem(pos, prob["polarisability"], k, True, d,
                      substrate=(sub.z_interface, sub.delta_eps))
        cpu = cpu_matvec_rate(P, n, pairs(n) + float(n) * n, rows=2048)
        log(f"[bench] substrate cpu baseline {cpu['value']:.3e} pair-evals/s")
    return {
        "cpu_baseline": cpu,
        "metric": "Green pair-evals/s (single mixed-precision CDA solve on a substrate)",
        "value": res["mixed"]["pair_evals_per_s"], "unit": "pair-evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * res["mixed"]["s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 far field + fp64 Krylov", "data": "synthetic",
        "config": {"workload": f"Au cuboid 80x40x40 nm, d=1.5 nm ({n} dipoles), 520 nm, water, "
                               f"on glass (retarded Fresnel image), FCD, GMRES(30) tol "
                               f"{args.tol:g}", "parallelism": "single GPU"},
        "fp64": res["fp64"], "mixed": res["mixed"], "cext_rel_err_mixed_vs_fp64": err,
    }


def metasurface_problems(n_wl, lam_lo=500.0, lam_hi=890.0, n_env=1.0):
    """Au cylinder D=100 H=40 at d=2 (41,181 dipoles) on glass (eps 2.25,
    retarded Fresnel factor, interface 1 nm below the lowest dipole plane:
    z0 = -21 nm), 300 nm square lattice, truncations 5/5, eta = sqrt(pi/A),
    Bloch k = 0, vacuum host (the reference's default n_env = 1). The J&C
    table ends at 892 nm, so the 500..1100 nm sweep stops at 890 nm."""
    from lumina_amd import inputs as I
    from lumina_amd.solver import LatticeSpec, SubstrateRuntime, substrate_reflection_factor
    pos = I.au_cylinder_d100_h40_d2()
    z0 = float(pos[:, 2].min()) - 1.0
    lat = LatticeSpec((300.0, 0.0, 0.0), (0.0, 300.0, 0.0), None, 5, 5)
    sub = SubstrateRuntime(z_interface=z0, delta_eps=substrate_reflection_factor(2.25 + 0j, n_env,
                                                                                 True))
    wls = np.linspace(lam_lo, lam_hi, n_wl) if n_wl > 1 else np.array([600.0])
    probs = [dict(positions=pos, polarisability=I.polarisability_block("Au", float(lam), 2.0, n_env),
                  k=I.wavenumber(float(lam), n_env), use_fcd=True, cell_size=2.0, substrate=sub,
                  lattice=lat, n=len(pos), wl=float(lam)) for lam in wls]
    return probs


def run_metasurface(args):
    from lumina_amd.solver import FP64, MIXED, Batch, HipBackend
    nw = args.wavelengths if args.wavelengths != 16 else 4
    probs = metasurface_problems(nw)
    n = probs[0]["n"]
    log(f"[bench] metasurface: N={n}, {len(probs)} wavelengths "
        f"{probs[0]['wl']:.0f}..{probs[-1]['wl']:.0f} nm")
    be = HipBackend(0)
    res = {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        t0 = time.perf_counter()
        b = Batch(be, probs)
        t_build = time.perf_counter() - t0
        b.set_timing(True)
        for _ in range(args.warmup):
            b.run(args.tol, args.max_iter, prec)
        t1 = time.perf_counter()
        for _ in range(args.steps):
            b.run(args.tol, args.max_iter, prec)
        dt = (time.perf_counter() - t1) / args.steps
        cs, infos = results_allow_failures(b)
        kt = b.kernel_times()
        b.close()
        ev = sum(i["pair_evals"] for i in infos)
        far_ms = kt["far32_ms"] + kt["far64_ms"]
        far_pairs = kt["far32_pairs"] + kt["far64_pairs"]
        res[name] = {"s": dt, "pair_evals_per_s": ev / dt, "solves_per_s": len(probs) / dt,
                     "build_s": t_build, "cs": cs.tolist(),
                     "iterations": [i["iterations"] for i in infos],
                     "converged": int(sum(i["status"] == 0 for i in infos)),
                     "table_pairs_per_s": far_pairs / (far_ms * 1e-3) if far_ms > 0 else None,
                     "kernel_times": kt}
        log(f"[bench] metasurface {name}: {dt:.2f} s, iterations {res[name]['iterations']}, "
            f"table kernel {res[name]['table_pairs_per_s']:.3e} pairs/s")
    c64 = np.array(res["fp64"]["cs"])[:, 0]
    c32 = np.array(res["mixed"]["cs"])[:, 0]
    ok = (np.array(res["fp64"]["iterations"]) < args.max_iter) & \
         (np.array(res["mixed"]["iterations"]) < args.max_iter)
    # only solves that converged in both precisions compare
    err = float(np.max(np.abs(c32[ok] - c64[ok]) / np.abs(c64[ok]))) if ok.any() else None
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as O
        p0 = probs[0]
        lat = p0["lattice"]
        P = O.Problem(p0["positions"], p0["polarisability"], p0["k"], True, 2.0,
                      substrate=(p0["substrate"].z_interface, p0["substrate"].delta_eps),
                      lattice=dict(a1=lat.a1, a2=lat.a2, eta=0.0, n_real=5, n_recip=5))
        cpu = cpu_matvec_rate(P, n, 2.0 * n * n, rows=32)
        log(f"[bench] metasurface cpu baseline {cpu['value']:.3e} pair-evals/s")
    return {
        "cpu_baseline": cpu,
        "metric": "Green pair-evals/s (periodic metasurface, batched mixed-precision solves)",
        "value": res["mixed"]["pair_evals_per_s"], "unit": "pair-evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * res["mixed"]["s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 Ewald tables + fp64 Krylov", "data": "synthetic",
        "config": {"workload": f"Au cylinder D=100 H=40 d=2 ({n} dipoles) on glass, 300 nm square "
                               f"lattice (2D Ewald 5/5), {len(probs)} wavelengths "
                               f"{probs[0]['wl']:.0f}..{probs[-1]['wl']:.0f} nm, GMRES(30) "
                               f"tol {args.tol:g}", "parallelism": "single GPU"},
        "fp64": res["fp64"], "mixed": res["mixed"], "cext_rel_err_mixed_vs_fp64": err,
    }


def run_sphere(args):
    """31 wavelengths of the 515-dipole sphere in one FP64 batch (tol 1e-10),
    every C_ext/C_abs/C_sca compared with the oracle's GMRES."""
    from lumina_amd import inputs as I
    from lumina_amd.solver import FP64, Batch, HipBackend
    from oracle import pyoracle as O
    pos = I.au_sphere_r10_d2()
    n_env, d = 1.33, 2.0
    wls = [400.0 + 10.0 * t for t in range(31)]
    probs = [dict(positions=pos, polarisability=I.polarisability_block("Au", lam, d, n_env),
                  k=I.wavenumber(lam, n_env), use_fcd=True, cell_size=d) for lam in wls]
    be = HipBackend(0)

    def spectrum(env):
        for key, val in env.items():
            os.environ[key] = val
        b = Batch(be, probs)
        for key in env:
            os.environ.pop(key)
        for _ in range(args.warmup):
            b.run(1e-10, args.max_iter, FP64)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.run(1e-10, args.max_iter, FP64)
        dt = (time.perf_counter() - t0) / args.steps
        cs, infos = results_allow_failures(b)
        b.close()
        return dt, cs, infos

    # CGS2 (the default) and modified Gram-Schmidt in the reference's order
    dt, cs, infos = spectrum({})
    dt_m, cs_m, infos_m = spectrum({"LUMINA_ORTHO": "mgs"})
    ev = sum(i["pair_evals"] for i in infos)
    t1 = time.perf_counter()
    errs, errs_m, it_diff, ev_cpu = [], [], [], 0.0
    threads = cpu_threads()
    for q, lam in enumerate(wls):
        P = O.Problem(pos, probs[q]["polarisability"], probs[q]["k"], True, d)
        cso, _, io = P.compute_cross_sections(1e-10, args.max_iter, nthreads=threads)
        ev_cpu += io["matvecs"] * pairs(len(pos))
        if io["status"] == 0 and infos[q]["status"] == 0:
            errs.append(float(np.max(np.abs(cs[q] - cso) / np.abs(cso))))
        if io["status"] == 0 and infos_m[q]["status"] == 0:
            errs_m.append(float(np.max(np.abs(cs_m[q] - cso) / np.abs(cso[0]))))
        it_diff.append(int(infos_m[q]["iterations"]) - int(io["iterations"]))
    t_cpu = time.perf_counter() - t1
    log(f"[bench] sphere spectrum: GPU {dt:.3f} s (CGS2), {dt_m:.3f} s (MGS), oracle {t_cpu:.1f} s, "
        f"max rel err {max(errs) if errs else None} (CGS2), {max(errs_m) if errs_m else None} (MGS)")
    return {
        "metric": "Green pair-evals/s (31-wavelength FP64 spectrum, one batched solve)",
        "value": ev / dt, "unit": "pair-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * dt, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
        "config": {"workload": "Au sphere r=10 nm d=2 nm (515 dipoles), water, 31 wavelengths "
                               f"400..700 nm, FCD, GMRES(30) tol 1e-10, max_iter {args.max_iter}",
                   "parallelism": "single GPU"},
        "solves_per_s": len(wls) / dt, "iterations": [i["iterations"] for i in infos],
        "spectrum_nm": wls, "c_ext": cs[:, 0].tolist(), "c_abs": cs[:, 1].tolist(),
        "c_sca": cs[:, 2].tolist(),
        "max_rel_err_vs_oracle": max(errs) if errs else None,
        "mgs": {"s": dt_m, "iterations": [i["iterations"] for i in infos_m],
                "iterations_minus_oracle": it_diff,
                "max_rel_err_vs_oracle": max(errs_m) if errs_m else None,
                "note": "LUMINA_ORTHO=mgs: modified Gram-Schmidt in the reference's order "
                        "(iterative.rs:92-96); error relative to C_ext"},
        "cpu_baseline": {"value": ev_cpu / t_cpu, "unit": "pair-evals/s", "cores": threads,
                         "kind": "port", "sample": "the same 31 solves by the oracle (FP64 "
                                                   "restatement, rows over a thread pool)"},
    }


def run_helix(args):
    """Forward and adjoint solves of the 35,013-dipole helix under left and
    right circular light, FP64 and mixed, with the gradient checks."""
    from lumina_amd import inputs as I
    from lumina_amd.solver import FP64, MIXED, HipBackend, IncidentField, PairSumSession
    from oracle import pyoracle as O
    lam, n_env, d = 600.0, 1.33, 1.5
    pos = I.au_helix_r25_p40_t4_w8_d1p5()
    n = len(pos)
    k = I.wavenumber(lam, n_env)
    eps = I.material("Au").epsilon(lam)
    a = I.polarisability_block("Au", lam, d, n_env)
    D = I.inverse_polarisability_derivative(eps, d, n_env)
    tol = 1e-10
    be = HipBackend(0)
    pols = {"LCP": (1 / np.sqrt(2), 1j / np.sqrt(2), 0.0), "RCP": (1 / np.sqrt(2), -1j / np.sqrt(2), 0.0)}
    res = {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        s = PairSumSession(be, pos, a, k, True, d, precision=prec)
        for _ in range(args.warmup):
            s.adjoint_grad(IncidentField(polarisation=pols["LCP"]), tol, args.max_iter, D)
        out = {}
        t0 = time.perf_counter()
        for hand, pol in pols.items():
            cs, ge, gp, fi, ai = s.adjoint_grad(IncidentField(polarisation=pol), tol,
                                                args.max_iter, D)
            out[hand] = dict(cs=cs, ge=ge, gp=gp, fi=fi, ai=ai)
        dt = time.perf_counter() - t0
        ev = sum(o["fi"]["pair_evals"] + o["ai"]["pair_evals"] for o in out.values())
        res[name] = dict(s=dt, ev=ev, out=out, session=s)
        log(f"[bench] helix {name}: {dt:.2f} s for 2 x (forward + adjoint + gradients), "
            f"iterations {[(o['fi']['iterations'], o['ai']['iterations']) for o in out.values()]}")
    # checks on the LCP gradients: oracle CPU adjoint on 32 seeded dipoles
    # (given the device's p and lam) and central differences of C_ext in Re eps
    s = res["fp64"]["session"]
    inc = IncidentField(polarisation=pols["LCP"])
    g = res["fp64"]["out"]["LCP"]
    _, p, _ = s.compute_cross_sections(inc, 1e-12, args.max_iter)
    lv, _ = s.solve_gmres(np.conj(s.incident(inc)), 1e-12, args.max_iter)
    cs12, ge12, gp12, _, _ = s.adjoint_grad(inc, 1e-12, args.max_iter, D)
    rows = np.random.default_rng(1).choice(n, 32, replace=False)
    t1 = time.perf_counter()
    geo, gro = O.Problem(pos, a, k, True, d).adjoint_grad(p, lv, D, polarisation=pols["LCP"],
                                                          rows=rows)
    t_or = time.perf_counter() - t1
    e_or = float(np.max(np.abs(gp12[rows] - gro[rows])) / np.max(np.abs(gro[rows])))
    e_or_eps = float(np.max(np.abs(ge12[rows] - geo[rows])) / np.max(np.abs(geo[rows])))
    # central differences on the same 32 dipoles, all perturbed problems in
    # one FP64 batch: Re eps of each (relative step 1e-6 of |eps|), and the
    # position of the first fd_dipoles along a seeded axis (step 1e-6 nm, FCD
    # classification held at the unperturbed positions)
    from lumina_amd.solver import Batch, ConvergenceFailure
    h, hr = 1e-6 * abs(eps), 1e-6
    alphas = np.tile(a, (n, 1))
    probs = []
    for i in rows:
        for sg in (1, -1):
            a2 = alphas.copy()
            a2[i] = I.polarisability_from_epsilon(eps + sg * h, lam, d, n_env)
            probs.append(dict(positions=pos, polarisability=a2, k=k, use_fcd=True, cell_size=d,
                              incident=inc))
    nfd = min(args.fd_dipoles, len(rows))
    axes = np.random.default_rng(4).integers(0, 3, nfd)
    for i, ax in zip(rows[:nfd], axes):
        for sg in (1, -1):
            p2 = pos.copy()
            p2[i, ax] += sg * hr
            probs.append(dict(positions=p2, polarisability=a, k=k, use_fcd=True, cell_size=d,
                              incident=inc, classify_positions=pos))
    t1 = time.perf_counter()
    b = Batch(be, probs)
    b.run(1e-12, args.max_iter, FP64)
    try:
        cfd, _, _ = b.results()
    except ConvergenceFailure as ex:
        cfd = ex.cross_sections
    b.close()
    t_fd = time.perf_counter() - t1
    ne = 2 * len(rows)
    fd_eps = (cfd[0:ne:2, 0] - cfd[1:ne:2, 0]) / (2 * h)
    fd_err = float(np.max(np.abs(fd_eps - ge12[rows, 0])) / np.max(np.abs(ge12[:, 0])))
    fd_r = (cfd[ne::2, 0] - cfd[ne + 1::2, 0]) / (2 * hr)
    fd_err_r = float(np.max(np.abs(fd_r - gp12[rows[:nfd], axes])) / np.max(np.abs(gp12)))
    log(f"[bench] helix finite differences: {len(probs)} FP64 solves in one batch, {t_fd:.1f} s")
    g64, g32 = res["fp64"]["out"], res["mixed"]["out"]
    mixed_err = max(float(np.max(np.abs(g32[h]["gp"] - g64[h]["gp"])) / np.max(np.abs(g64[h]["gp"])))
                    for h in pols)
    cext_err = max(abs(g32[h]["cs"][0] - g64[h]["cs"][0]) / abs(g64[h]["cs"][0]) for h in pols)
    for r in res.values():
        r["session"].close()
    log(f"[bench] helix checks: oracle rows {e_or:.2e}/{e_or_eps:.2e} ({t_or:.1f} s), FD eps "
        f"{fd_err:.2e}, FD r {fd_err_r:.2e}, mixed vs fp64 gradient {mixed_err:.2e}")
    m = res["mixed"]
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_matvec_rate(O.Problem(pos, a, k, True, d), n, pairs(n), rows=4096)
        log(f"[bench] helix cpu baseline {cpu['value']:.3e} pair-evals/s")
    return {
        "cpu_baseline": cpu,
        "metric": "Green pair-evals/s (helix forward + adjoint solves, mixed precision)",
        "value": m["ev"] / m["s"], "unit": "pair-evals/s", "n_gpus": 1, "steps": 1,
        "warmup": args.warmup, "ms_per_step": 1e3 * m["s"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 far field + fp64 Krylov",
        "data": "synthetic",
        "config": {"workload": f"Au helix coil 25 pitch 40 4 turns wire 8 d=1.5 ({n} dipoles), "
                               f"water, 600 nm, LCP and RCP, forward + adjoint GMRES(30) tol "
                               f"{tol:g} + gradients", "parallelism": "single GPU"},
        "c_ext": {h: float(g64[h]["cs"][0]) for h in pols},
        "circular_dichroism_ext": float(g64["LCP"]["cs"][0] - g64["RCP"]["cs"][0]),
        "fp64": {"s": res["fp64"]["s"], "pair_evals_per_s": res["fp64"]["ev"] / res["fp64"]["s"],
                 "iterations": {h: (o["fi"]["iterations"], o["ai"]["iterations"])
                                for h, o in g64.items()}},
        "mixed_iterations": {h: (o["fi"]["iterations"], o["ai"]["iterations"]) for h, o in g32.items()},
        "cext_rel_err_mixed_vs_fp64": cext_err, "grad_rel_err_mixed_vs_fp64": mixed_err,
        "grad_check": {"oracle_rows": 32, "oracle_rel_err_positions": e_or,
                       "oracle_rel_err_eps": e_or_eps, "fd_eps_dipoles": len(rows),
                       "fd_eps_rel_err": fd_err, "fd_eps_step": "1e-6 |eps|",
                       "fd_position_dipoles": nfd, "fd_position_rel_err": fd_err_r,
                       "fd_position_step_nm": hr, "fd_batch_s": t_fd, "tol": 1e-12,
                       "bar": "1e-5 of the largest gradient component (FD), 1e-8 (oracle)"},
    }


def run_fields(args):
    """Near-field map and far-field pattern of the solved 38,637-dipole
    cuboid: a 400 x 400 map on the z = 0 plane over +-100 nm (160,000 points,
    fields.rs:51-108) and a 181 x 360 far-field grid (fields.rs:156-221),
    both FP64 on the device from the resident solution."""
    from concurrent.futures import ThreadPoolExecutor
    from lumina_amd import fields as F
    from lumina_amd import inputs as I
    from lumina_amd.solver import MIXED, HipBackend, IncidentField, PairSumSession
    from oracle import pyoracle as O
    pos = I.au_cuboid_80x40x40_d1p5()
    n = len(pos)
    lam, n_env, d = 520.0, 1.33, 1.5
    alpha = I.polarisability_block("Au", lam, d, n_env)
    k = I.wavenumber(lam, n_env)
    be = HipBackend(0)
    s = PairSumSession(be, pos, alpha, k, True, d, precision=MIXED)
    _, p, info = s.compute_cross_sections(IncidentField(), args.tol, args.max_iter)
    plane = F.NearFieldPlane((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 100.0, 100.0, 400, 400)
    pts, _ = F.plane_points(plane)
    nt, nph = 181, 360
    for _ in range(args.warmup):
        F.field_at_points(s, pts, IncidentField())
        F.compute_far_field(s, nt, nph)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e, inten = F.field_at_points(s, pts, IncidentField())
    t_near = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ff = F.compute_far_field(s, nt, nph)
    t_far = (time.perf_counter() - t0) / args.steps
    near_pairs = float(len(pts)) * n
    far_pairs = float(nt * nph) * n
    log(f"[bench] fields: near map {len(pts)} points {1e3 * t_near:.2f} ms "
        f"({near_pairs / t_near:.3e} pairs/s), far field {nt}x{nph} {1e3 * t_far:.2f} ms "
        f"({far_pairs / t_far:.3e} dipole-directions/s)")
    # parity on a sample of map points and directions against the oracle
    rows = np.random.default_rng(0).choice(len(pts), 64, replace=False)
    eo, _ = O.field_at_points(pos, p, k, pts[rows])
    e_near = float(np.max(np.abs(e[rows] - eo)) / np.max(np.abs(eo)))
    cpu = None
    if not args.no_cpu_baseline:
        threads = cpu_threads()
        per = 64
        chunks = [pts[i * per:(i + 1) * per] for i in range(threads)]
        with ThreadPoolExecutor(threads) as ex:
            t1 = time.perf_counter()
            list(ex.map(lambda c: O.field_at_points(pos, p, k, c), chunks))
            dt = time.perf_counter() - t1
        cpu = {"value": per * threads * n / dt, "unit": "pair-evals/s", "cores": threads,
               "kind": "port", **host_cpus(),
               "sample": f"oracle field_at_points, {threads} x {per} map points x {n} dipoles on "
                         f"{threads} threads in {dt:.2f} s"}
        log(f"[bench] fields cpu baseline {cpu['value']:.3e} pair-evals/s")
    s.close()
    return {
        "cpu_baseline": cpu,
        "metric": "Green pair-evals/s (near-field map, FP64)",
        "value": near_pairs / t_near, "unit": "pair-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t_near, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Au cuboid 80x40x40 nm d=1.5 ({n} dipoles), 520 nm, water: "
                               "400x400 near-field map (z=0, +-100 nm) and 181x360 far field "
                               "from the mixed-precision solution", "parallelism": "single GPU"},
        "far_field": {"ms": 1e3 * t_far, "dipole_directions_per_s": far_pairs / t_far,
                      "directions": nt * nph},
        "near_rel_err_vs_oracle_64_points": e_near, "solve_info": info,
        "timing": "wall clock of the C-ABI call: upload of the points, kernels, download",
    }


def run_sweep(args):
    """The CLI spectrum run (runner.rs:145-218) of the 515-dipole sphere, 31
    wavelengths 400..700 nm, FP64 tol 1e-10, end to end from the material
    table: lh_sweep_solve (geometry, near list, per-wavelength set-up on the
    device, batched solve, cross sections) against the same run with the
    set-up on the host (spline, alpha, alpha^-1 per wavelength, then
    lh_batch_solve)."""
    from lumina_amd import inputs as I
    from lumina_amd import sweep as W
    from lumina_amd.solver import FP64, HipBackend, batch_compute
    pos = I.au_sphere_r10_d2()
    n_env, d = 1.33, 2.0
    wls = [400.0 + 10.0 * t for t in range(31)]
    be = HipBackend(0)
    au = W.johnson_christy("Au")
    max_iter = 5000   # near 700 nm the solve needs ~1,100 iterations at tol 1e-10

    def device():
        return W.run_sweep(be, pos, wls, [au], spacing=d, n_env=n_env, tolerance=1e-10,
                           max_iterations=max_iter)

    def host():
        probs = [dict(positions=pos, polarisability=I.polarisability_block("Au", lam, d, n_env),
                      k=I.wavenumber(lam, n_env), use_fcd=True, cell_size=d) for lam in wls]
        return batch_compute(be, probs, 1e-10, max_iter, precision=FP64)

    for _ in range(args.warmup):
        device()
        host()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cs, infos = device()
    t_dev = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        csh, _, _ = host()
    t_host = (time.perf_counter() - t0) / args.steps
    err = float(np.max(np.abs(cs - csh) / np.abs(csh[:, :1])))
    log(f"[bench] sweep: device set-up {1e3 * t_dev:.1f} ms, host set-up {1e3 * t_host:.1f} ms "
        f"for {len(wls)} wavelengths, max diff / C_ext {err:.2e}")
    ev = sum(i["pair_evals"] for i in infos)
    return {
        "metric": "wavelengths/s (CLI spectrum run end to end, device set-up)",
        "value": len(wls) / t_dev, "unit": "wavelengths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t_dev, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Au sphere r=10 nm d=2 ({len(pos)} dipoles), water, "
                               f"{len(wls)} wavelengths 400..700 nm, FCD, GMRES(30) tol 1e-10 "
                               "(max 5000 iterations), from the Johnson-Christy table",
                   "parallelism": "single GPU"},
        "host_setup_ms": 1e3 * t_host, "pair_evals_per_s": ev / t_dev,
        "max_diff_over_cext_vs_host_setup": err,
        "timing": "wall clock of the whole call: geometry, near list, set-up, solve, results",
    }


def run_copy(args):
    """A plain device copy of 256 MiB (torch copy_): the reference point that
    scales the HBM traffic counters (FETCH_SIZE, WRITE_SIZE) of a profile."""
    import torch
    n = 64 * 1024 * 1024
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    nbytes = 2 * 4 * n
    return {"metric": "device copy GB/s (read + write)", "value": nbytes / dt / 1e9, "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": {"workload": "copy of 256 MiB (64 Mi float32)"},
            "bytes_read_per_copy": 4 * n, "bytes_written_per_copy": 4 * n}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["batch", "cuboid", "substrate", "metasurface", "sphere",
                                           "helix", "fields", "sweep", "copy"],
                    default="batch")
    ap.add_argument("--fd-dipoles", type=int, default=8)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-iter", type=int, default=1000)
    ap.add_argument("--particles", type=int, default=64)
    ap.add_argument("--wavelengths", type=int, default=16)
    ap.add_argument("--plain-wavelengths", type=int, default=4)
    ap.add_argument("--no-adaptive", action="store_true",
                    help="skip the adaptive-graph-launch leg of the batch workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-plain", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); sharding this path "
              f"over several GPUs is designed but not built.", file=sys.stderr)
        sys.exit(2)
    run = {"batch": run_batch, "cuboid": run_cuboid, "substrate": run_substrate,
           "metasurface": run_metasurface,
           "sphere": run_sphere, "helix": run_helix, "fields": run_fields, "sweep": run_sweep, "copy": run_copy}
    line = run[args.workload](args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
