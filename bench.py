"""Frequency-solve throughput of the MI355X NEMOH solver core.

Workload: rectangular barge 100 x 20 x 10 m (draft 10 m), no symmetry,
4,000 quad panels (bottom 90 x 20, sides 10 high), finite depth 50 m,
6 radiation DOF + 3 headings (0/45/90 deg) = 9 right-hand sides,
40 frequencies uniform in 0.2-2.0 rad/s (2.0 rad/s takes the infinite-depth
path, k h >= 20).  One step = one pass over the 40 frequencies, each a
frequency solve: influence-matrix assembly, factorisation, all 9 solves,
phi = S sigma and the force integration.  `value` = 40 x steps / timed
seconds (freq-solves/s); ms_per_step x steps is the timed region.

`value` is the whole-job rate of the mixed-precision path (complex64 LU on
MFMA + FP64 refinement) with the 40 frequencies solved as one batch: one
handle of the mesh per frequency, each heavy step (assembly, every block-LU
step, every triangular solve, the refinement's steps, the gemvs, the
forces) one launch over all 40.  The plain FP64 path is measured the same
way, and both also through the native request loop (several frequencies in
flight on their own streams); the speed-up is taken against the faster
FP64 rate.  `roofline` is that of value's dominant kernel, the grouped
rank-128 LU update, timed over the value leg's timed passes.  The
single-stream rates, per-phase times and the same kernel's roofline one
matrix at a time come from a one-frequency-at-a-time leg.  Also reported: the mixed-vs-FP64 error, the 12k-panel
semi-submersible, the 256-body batch, serving latency under Poisson arrivals
and the reference Solver routines' CPU time on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

# Hardware queues of this process (read by the HIP runtime when it starts,
# before any device call), raised to 8 when the environment asks for fewer:
# the throughput legs run 8 frequencies on 8 streams, and with 4 queues (the
# runtime's default) two streams share each queue and their kernels run one
# after the other (8 queues: 4000-panel barge 338-342 against 310-311
# freq-solves/s, 256-body batch 917-930 against 875)
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "freq-solves/s at 4k panels, mixed vs FP64 (<=1e-6 rel err); serving p99 ms"
PEAK = {"fp64": 78.6, "fp32": 157.3}  # MI355X dense TFLOP/s (FP64 / FP32 matrix)
HBM_PEAK_GBS = 8000.0
# HBM bytes of lu_update128_kernel<float> per launch over its algorithmic
# bytes (16 m^2 + 2048 m for an m x m trailing block and two 64-wide panels),
# from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this workload, each
# scaled by the counter's reading for a 256 MiB copy: 58.36 MB fetched +
# 41.00 MB written against 85.93 MB algorithmic on average over 6386 launches
# (profiles/lu_update128_barge4000_hbm_traffic.md)
LU_UPDATE_TRAFFIC_RATIO = 1.156
# The same for group_update128_kernel<float>, the grouped update of the 40
# matrices of a batch pass (value's dominant kernel): 2270.9 MB fetched +
# 1640.0 MB written against 3437.4 MB algorithmic per launch on average over
# 62 launches (profiles/group_update128_barge4000_batch40_hbm_traffic.md)
BATCH_UPDATE_TRAFFIC_RATIO = 1.138


def workload():
    from nemoh_amd import bodycond as bcm
    from nemoh_amd import mesh as nmesh
    m = nmesh.box(100.0, 20.0, 10.0, 90, 20, 10)
    assert m.npanels == 4000
    omegas = np.linspace(0.2, 2.0, 40)
    modes = bcm.six_dof()
    om, kind, nvel = bcm.problems(m, omegas, [0.0, 45.0, 90.0], modes, 50.0, 9.81, 1025.0)
    nds = bcm.integration(m, modes)
    return m, omegas, om, kind, nvel, nds


def workload_c3():
    """Semi-submersible: 4 columns r = 6 m at 50 m spacing, pontoons 8 x 6 m,
    draft 20 m, 12,048 panels, infinite depth, 6 DOF + 8 headings (0-315 deg)
    = 14 RHS, 60 frequencies 0.1-2.5 rad/s (the problems are generated on
    the device)."""
    from nemoh_amd import mesh as nmesh
    return nmesh.semisub(), np.linspace(0.1, 2.5, 60), np.deg2rad(np.arange(8) * 45.0)


def run_c3(steps):
    """C3 on `steps` frequencies spread over the band, FP64 then mixed."""
    from nemoh_amd import bodycond as bcm
    from nemoh_amd import core as ncore
    m, omegas, betas = workload_c3()
    idx = np.linspace(0, omegas.size - 1, steps).round().astype(int)
    c = ncore.Core(m.vertices, m.panels, 0, 0.0, 1025.0, 9.81)
    c.set_dofs(bcm.six_dof())
    out = {}
    for name, prec in (("fp64", ncore.FP64), ("mixed", ncore.MIXED)):
        dt, forces, ph = run_precision(c, prec, omegas[idx], None, None, steps, 1, betas=betas)
        out[name] = (steps / dt, forces, ph)
    c.close()
    err = max(float((np.abs(out["mixed"][1][k] - out["fp64"][1][k]).max(0) /
                     np.abs(out["fp64"][1][k]).max(0)).max()) for k in out["fp64"][1])
    # the same frequencies as one batch (a handle per frequency, each heavy
    # step one launch over all of them), one untimed and two timed passes
    batched, same = {}, True
    tables = ncore.build_tables()
    for name, prec in (("mixed", ncore.MIXED), ("fp64", ncore.FP64)):
        cores = []
        for _ in idx:
            h = ncore.Core(m.vertices, m.panels, 0, 0.0, 1025.0, 9.81, tables=tables)
            h.set_dofs(bcm.six_dof())
            cores.append(h)
        b = ncore.Batch(cores, prec, max_rhs=betas.size + 6)
        b.solve_waves(omegas[idx], betas)
        t0 = time.perf_counter()
        for _ in range(2):
            got = b.solve_waves(omegas[idx], betas)
        batched[name] = round(2 * steps / (time.perf_counter() - t0), 3)
        # run_precision keys its forces by position in the frequency list
        same = same and all(np.array_equal(got[i][0], out[name][1][i]) for i in range(steps) if i in out[name][1])
        b.close()
        for h in cores:
            h.close()
    batched["identical_to_single_stream"] = bool(same)
    return {"workload": f"semi-submersible, {m.npanels} panels, infinite depth, 14 RHS (6 DOF + 8 headings), "
                        f"{steps} of 60 frequencies 0.1-2.5 rad/s: " + ", ".join(f"{w:.3f}" for w in omegas[idx]),
            "unit": "freq-solves/s", "mixed": round(out["mixed"][0], 3), "fp64": round(out["fp64"][0], 3),
            "speedup_mixed_vs_fp64": round(out["mixed"][0] / out["fp64"][0], 3), "max_rel_err_mixed_vs_fp64": err,
            "batched": batched,
            "refinement": "max|r| <= 1e-8 max|b| (normwise backward error <= 1e-8)",
            "phase_ms_median": {name: {k: round(float(np.median(v)), 2) for k, v in out[name][2].items()
                                       if k in ("assembly", "factor", "solve")} for name in out}}


BETAS = np.deg2rad([0.0, 45.0, 90.0])


def run_precision(core, prec, omegas, nvel, nds, steps, warmup, nrhs=9, host_inputs=False, betas=BETAS,
                  timing=True):
    """One frequency at a time on one handle.  By default the 9 problems of
    each frequency are generated on the device (nemoh_core_solve_waves: no
    per-frequency input crosses PCIe); host_inputs=True hands the host
    arrays over instead (nemoh_core_solve_frequency: NVEL and NDS uploaded,
    the drop-in's per-call data flow)."""
    core.set_precision(prec)
    # per-phase events launch every kernel one by one; without them each
    # frequency replays the handle's HIP graph
    core.enable_timing(timing)
    forces = {}
    phases = {"assembly": [], "factor": [], "solve": [], "update": [], "update_flops": [], "it": []}

    def one(k):
        w = omegas[k % omegas.size]
        if host_inputs:
            cols = slice((k % omegas.size) * nrhs, (k % omegas.size + 1) * nrhs)
            _, f = core.solve_frequency(2 * math.pi / w, nvel[:, cols], nds=nds, want_pressure=False)
        else:
            f, _ = core.solve_waves(w, betas)
        return k % omegas.size, f

    for k in range(warmup):
        one(k)
    t0 = time.perf_counter()
    for k in range(steps):
        idx, f = one(warmup + k)
        forces[idx] = f
        st = core.stats()
        phases["assembly"].append(st.ms_assembly)
        phases["factor"].append(st.ms_factor)
        phases["solve"].append(st.ms_solve)
        phases["it"].append(st.refine_iterations)
        phases["update"].append(st.ms_lu_update)
        phases["update_flops"].append(st.lu_update_flops)
        phases.setdefault("update_launches", []).append(st.lu_update_launches)
    dt = time.perf_counter() - t0
    return dt, forces, phases


def run_outputs(core, omega=0.9, ntheta=181, ngrid=50, reps=3):
    """The Solver's optional outputs after one mixed solve of the barge at
    omega (finite depth): the Kochin function at ntheta angles and the
    potential on an ngrid x ngrid free-surface grid (the sizes the reference's
    Postprocessing case asks for), each call timed end to end (sources
    resident, results copied back) and reported as the median of reps."""
    from nemoh_amd import core as ncore
    core.set_precision(ncore.MIXED)
    core.solve_waves(omega, BETAS)
    n = core.nfa
    th = np.linspace(0.0, 2.0 * math.pi, ntheta)
    xs, ys = np.linspace(-150.3, 149.7, ngrid), np.linspace(-75.3, 74.7, ngrid)
    pts = np.stack([g.ravel() for g in np.meshgrid(xs, ys)] + [np.zeros(ngrid * ngrid)], axis=1)
    tk, tp = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        h = core.kochin(th)
        tk.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        phi = core.potential(pts)
        tp.append(time.perf_counter() - t0)
    assert np.all(np.isfinite(h)) and np.all(np.isfinite(phi))
    mk, mp = float(np.median(tk)), float(np.median(tp))
    return {"workload": f"barge 4000 panels at w = {omega} rad/s (finite depth), {h.shape[1]} problems: Kochin at "
                        f"{ntheta} angles, potential at {ngrid}x{ngrid} free-surface points",
            "kochin_ms": round(1e3 * mk, 3), "kochin_angle_panels_per_s": round(ntheta * n / mk, 1),
            "potential_ms": round(1e3 * mp, 3), "potential_point_panels_per_s": round(pts.shape[0] * n / mp, 1)}


def run_concurrent(m, omegas, prec, passes, warmup, workers):
    """`passes` passes over every frequency with `workers` frequencies in
    flight, each on its own solver handle and HIP stream (the native request
    loop): the latency-bound phases of one frequency (the factorisation's
    dependent block steps, the triangular-solve chains) leave most compute
    units idle, and other frequencies fill them.  Results are bit for bit
    those of the single-stream path.  Returns (seconds for the timed passes,
    forces by frequency index of the last pass)."""
    from nemoh_amd import bodycond as bcm
    from nemoh_amd import serving
    nw = omegas.size

    def request(idx):  # problems generated on the device, as in the single-stream leg
        return serving.WaveRequest(m, 50.0, omegas[idx], BETAS, bcm.six_dof())

    with serving.Server(workers=workers) as srv:  # one frequency per handle and stream
        # warm-up passes (at least one frequency per worker binds every handle)
        srv.wait(srv.submit(request(np.arange(max(warmup * nw, workers)) % nw), prec))
        idx = np.arange(passes * nw) % nw
        t0 = time.perf_counter()
        f, _ = srv.wait(srv.submit(request(idx), prec))
        dt = time.perf_counter() - t0
    last = idx.size - nw
    return dt, {int(idx[last + i]): f[last + i] for i in range(nw)}


def run_batched_frequencies(m, omegas, prec, passes, warmup, ref):
    """The 40 frequencies as one batch (nemoh_batch_solve_waves_multi): one
    handle of the mesh per frequency, each heavy step one launch over all of
    them (the frequencies in both depth regimes take one assembly launch
    each).  Handles, matrices and job tables resident; set-up and `warmup`
    passes (at least one: it builds the job tables) untimed.  Returns
    (seconds for the timed passes, whether every force equals the
    single-stream one bit for bit, and the grouped rank-128 updates of the
    timed passes as (ms, algorithmic flops, algorithmic bytes, launches),
    timed with HIP events on the batch's stream)."""
    from nemoh_amd import bodycond as bcm
    from nemoh_amd import core as ncore
    tables = ncore.build_tables()
    cores = []
    for _ in omegas:
        h = ncore.Core(m.vertices, m.panels, 0, 50.0, 1025.0, 9.81, tables=tables)
        h.set_dofs(bcm.six_dof())
        cores.append(h)
    b = ncore.Batch(cores, prec, max_rhs=9)
    for _ in range(max(1, warmup)):
        got = b.solve_waves(omegas, BETAS)
    b.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(passes):
        got = b.solve_waves(omegas, BETAS)
    dt = time.perf_counter() - t0
    upd = b.update_time()
    same = all(np.array_equal(got[k][0], ref[k]) for k in range(omegas.size) if k in ref)
    b.close()
    for h in cores:
        h.close()
    return dt, same, upd


def run_batched_streams(m, omegas, prec, passes, warmup, streams=4):
    """The 40 frequencies as `streams` batches of 40 / streams handles, each
    batch driven by its own host thread on its own stream, so that one
    batch's assembly (FP64 VALU) runs beside another's factorisation (MFMA).
    Returns the seconds of the timed passes and the forces of the last pass
    by frequency index."""
    import threading
    from nemoh_amd import bodycond as bcm
    from nemoh_amd import core as ncore
    tables = ncore.build_tables()
    parts = np.array_split(np.arange(omegas.size), streams)
    batches, cores, got = [], [], [None] * streams
    for p in parts:
        cs = []
        for _ in p:
            h = ncore.Core(m.vertices, m.panels, 0, 50.0, 1025.0, 9.81, tables=tables)
            h.set_dofs(bcm.six_dof())
            cs.append(h)
        cores += cs
        batches.append(ncore.Batch(cs, prec, max_rhs=9))

    def run(i, n):
        for _ in range(n):
            got[i] = batches[i].solve_waves(omegas[parts[i]], BETAS)

    def go(n):
        th = [threading.Thread(target=run, args=(i, n)) for i in range(streams)]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        return time.perf_counter() - t0

    go(max(1, warmup))
    dt = go(passes)
    forces = {int(k): got[i][j][0] for i, p in enumerate(parts) for j, k in enumerate(p)}
    for b in batches:
        b.close()
    for h in cores:
        h.close()
    return dt, forces


REF_HARNESS = ROOT / "oracle" / "_ref" / "harness"


def cpu_baseline_reference(m, omegas, nrhs=9, procs=8, stride=4, m_small=800, m_large=1600):
    """The reference Solver's own routines (FP64 build of the unmodified
    sources, oracle/_ref) on the host cores: `procs` single-threaded processes,
    one per sampled frequency (the reference is single-threaded and
    frequencies are independent).  Each times VAVFD+VNSFD (or the INFD pair)
    for every `stride`-th field row against all panels, and GAUSSZ on an
    m_large x 2 m_large system; then the same processes time GAUSSZ on
    m_small x 2 m_small.  One frequency is N^2 (1 + NRHS) pairs (assembly +
    phi per problem, SOLVE_BEM_FD_DIRECT.f90:103-183) + GAUSSZ at N,
    extrapolated from m_large with the power of n that the two sizes show
    (GAUSSZ leaves the caches as n grows, so its time grows faster than n^3:
    the n^3 scaling of an in-cache size underestimates it)."""
    import subprocess
    import tempfile
    n = m.npanels
    head = f"mesh\n{m.npoints} {m.npanels} {m.isym} {50.0:.17e}\n"
    head += "".join(f"{x:.17e} {y:.17e} {z:.17e}\n" for x, y, z in m.vertices)
    head += "".join(f"{a} {b} {c} {d}\n" for a, b, c, d in m.panels + 1)
    rows = (n + stride - 1) // stride
    idx = np.linspace(0, omegas.size - 1, procs).round().astype(int)

    def batch(nrows, mg):
        # each process writes to a file of its own: the harness echoes more
        # than a pipe holds, and a full pipe would make the processes wait
        # for each other
        outs = [tempfile.TemporaryFile(mode="w+") for _ in idx]
        ps = [subprocess.Popen([str(REF_HARNESS)], stdin=subprocess.PIPE, stdout=f, text=True) for f in outs]
        for p, k in zip(ps, idx):
            p.stdin.write(head + f"bench\n{2 * math.pi / omegas[k]:.17e} {nrows} {stride} {mg}\nend\n")
            p.stdin.close()
        res = []
        for p, f in zip(ps, outs):
            if p.wait() != 0:
                raise RuntimeError("reference harness failed")
            f.seek(0)
            last = f.read().strip().split("\n")[-1].split()
            f.close()
            res.append((float(last[2]), float(last[4])))
        return res

    t0 = time.perf_counter()
    big = batch(rows, m_large)
    small = batch(1, m_small)
    wall = time.perf_counter() - t0
    g_large = float(np.mean([g for _, g in big]))
    g_small = float(np.mean([g for _, g in small]))
    expo = math.log(g_large / g_small) / math.log(m_large / m_small)
    t_freq = [tp / (rows * n) * n * n * (1 + nrhs) + tg * (n / m_large) ** expo for tp, tg in big]
    mean_t = float(np.mean(t_freq))
    return {"value": procs / mean_t, "unit": "freq-solves/s", "cores": procs, "kind": "reference",
            "sample": f"reference Solver routines (flang -O2 FP64 build of the unmodified sources), {procs} "
                      f"single-threaded processes at w = {', '.join(f'{omegas[k]:.3f}' for k in idx)} rad/s: "
                      f"each times {rows} field rows x {n} panels of VAVFD+VNSFD pairs and GAUSSZ on "
                      f"{m_large}x{2 * m_large} ({g_large:.1f} s) and {m_small}x{2 * m_small} ({g_small:.2f} s); "
                      f"one frequency = N^2(1+{nrhs}) pairs + GAUSSZ extrapolated as n^{expo:.2f} from "
                      f"{m_large}; mean {mean_t:.1f} s per frequency per core (sample wall {wall:.1f} s; the "
                      f"complete reference Solver took 302 s per frequency on one core of another host, "
                      f"profiles/cpu_reference_barge4000.txt)"}


def cpu_baseline(m, omegas, budget_s=20.0):
    """Reference algorithm (the oracle, FP64, one core) on a bounded sample:
    Green-function pairs of a block of matrix rows at one finite-depth
    frequency, and the Gauss-Jordan inversion of a smaller complex matrix;
    scaled to one full frequency the way the reference spends it:
    N^2 (1 + NRHS) pair evaluations (assembly, then phi re-evaluated per
    problem) + 10.7 N^3 flops of GAUSSZ."""
    from oracle import pyoracle as po
    c = po.Case(m.vertices, m.panels, 0, depth=50.0, rho=1025.0, g=9.81)
    period = 2 * math.pi / float(omegas[len(omegas) // 2])
    rows = 8
    t0 = time.perf_counter()
    c.pairs_rows(period, 0, rows)
    t_rows = time.perf_counter() - t0
    n = m.npanels
    per_pair = t_rows / (rows * n)
    ns = 600
    rng = np.random.default_rng(1)
    a = rng.standard_normal((ns, ns)) + 1j * rng.standard_normal((ns, ns)) + ns * np.eye(ns)
    t0 = time.perf_counter()
    po.gaussz_inverse(a)
    t_g = time.perf_counter() - t0
    nrhs = 9
    t_freq = per_pair * n * n * (1 + nrhs) + t_g * (n / ns) ** 3
    return {"value": 1.0 / t_freq, "unit": "freq-solves/s", "cores": 1, "kind": "port",
            "sample": f"oracle (C restatement, FP64, 1 core): {rows} matrix rows x {n} panels of finite-depth "
                      f"Green-function pairs at w={2 * math.pi / period:.3f} rad/s ({per_pair * 1e6:.2f} us/pair) "
                      f"and GAUSSZ on a {ns}x{ns} complex matrix ({t_g:.2f} s); scaled to one 4000-panel frequency "
                      f"as N^2(1+9) pairs + (N/{ns})^3 GAUSSZ = {t_freq:.1f} s"}


def run_serving(rates, nreq, workers, seed=2024, budget_gib=32.0):
    """Serving: Poisson arrivals (fixed seed) of two-cylinder requests
    (r ~ U(2,15) m, spacing U(3,10) r, N log-uniform in [200, 16000], 12 DOF
    + 1-4 headings, 1-64 frequencies, infinite depth), admitted by the
    native request loop under a budget of budget_gib of device memory
    (each frequency is admitted with its handle's resident bytes charged:
    FP64 D and S plus the complex64 copy and factors on the mixed path).
    Latency = submission to completion, stamped by the loop.  The first rate
    runs on the mixed and the plain FP64 paths, the others on the mixed path
    only; the loop saturates where req_per_s stops following the rate."""
    from nemoh_amd import core as ncore
    from nemoh_amd import serving
    rng = np.random.default_rng(seed)
    reqs = [serving.two_cylinder_waves(rng) for _ in range(nreq)]
    gaps = rng.exponential(1.0, nreq)
    nfreq = sum(r.nfreq for r in reqs)
    tables = ncore.build_tables()

    def stream(rate, prec):
        arrivals = np.cumsum(gaps) / rate
        with serving.Server(workers=workers, tables=tables, budget_bytes=budget_gib * 2 ** 30) as srv:
            t0 = time.perf_counter()
            ids = []
            for r, ta in zip(reqs, arrivals):
                dt = t0 + ta - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids.append(srv.submit(r, prec))
            tim = [srv.wait(i)[1] for i in ids]
            wall = time.perf_counter() - t0
            _, peak = srv.memory()
        lat = np.array([t["latency"] for t in tim])
        svc = np.array([t["service"] for t in tim])
        return {"p50_ms": round(float(np.percentile(lat, 50)), 1), "p99_ms": round(float(np.percentile(lat, 99)), 1),
                "mean_service_ms": round(float(svc.mean()), 1), "wall_s": round(wall, 2),
                "req_per_s": round(nreq / wall, 3), "freq_solves_per_s": round(nfreq / wall, 2),
                "peak_charged_gib": round(peak / 2 ** 30, 2)}

    out = {"workload": f"{nreq} requests, Poisson arrivals (seed {seed}): two cylinders r ~ U(2,15) m, spacing "
                       "U(3,10) r, N log-uniform 200-16000 panels, 12 DOF + 1-4 headings, 1-64 frequencies, "
                       "infinite depth",
           "panels": [r.npanels for r in reqs], "frequencies": nfreq, "workers": workers,
           "budget": f"{budget_gib:g} GiB of device memory: each in-flight frequency charged its handle's "
                     "resident bytes (about 40 N^2 B mixed, 32 N^2 B FP64, per symmetry block)", "rates": {}}
    for k, rate in enumerate(rates):
        res = {"mixed": stream(rate, ncore.MIXED)}
        if k == 0:
            res["fp64"] = stream(rate, ncore.FP64)
        out["rates"][str(rate)] = res
        print(f"serving at {rate} req/s: {res}", file=sys.stderr, flush=True)
    out["saturation_req_per_s"] = max(v["mixed"]["req_per_s"] for v in out["rates"].values())
    return out


def run_c4(nbodies, workers, seed=256):
    """Batched bodies: nbodies independent bodies (cylinder / box / prolate
    spheroid, dimensions U(2,50) m, draft U(0.2,0.8) of height, N
    log-uniform in [300, 6000], seeded PCG64), 7 RHS (6 DOF + heading 0), 8
    frequencies in 0.2-2.0 rad/s, infinite depth, all submitted at once to
    the native request loop; freq-solves/s = bodies x frequencies / wall."""
    from nemoh_amd import core as ncore
    from nemoh_amd import serving
    rng = np.random.default_rng(seed)
    reqs = [serving.random_body_waves(rng) for _ in range(nbodies)]
    ns = np.array([r.npanels for r in reqs], dtype=np.float64)
    nfreq = sum(r.nfreq for r in reqs)
    tables = ncore.build_tables()
    res, forces = {}, {}
    for name, prec in (("mixed", ncore.MIXED), ("fp64", ncore.FP64)):
        with serving.Server(workers=workers, tables=tables) as srv:
            t0 = time.perf_counter()
            ids = [srv.submit(r, prec) for r in reqs]
            forces[name] = [srv.wait(i)[0] for i in ids]
            wall = time.perf_counter() - t0
        res[name] = round(nfreq / wall, 2)
        print(f"c4 {name}: {nfreq} freq-solves in {wall:.2f} s", file=sys.stderr, flush=True)
    # normwise over each body's force matrix at each frequency: a cylinder's
    # yaw about its own axis moves no water, so that column is rounding
    # noise on both paths
    err = max(float((np.abs(a - b).max(axis=(1, 2)) / np.abs(b).max(axis=(1, 2))).max())
              for a, b in zip(forces["mixed"], forces["fp64"]))
    # the same bodies as one batch (nemoh_batch: each heavy step one launch
    # over all 256 bodies), frequency by frequency, matrices resident
    batched, same = {}, True
    for name, prec in (("mixed", ncore.MIXED), ("fp64", ncore.FP64)):
        t0 = time.perf_counter()
        cores = []
        for r in reqs:
            c = ncore.Core(r.mesh.vertices, r.mesh.panels, r.mesh.isym, 0.0, r.rho, r.g, tables=tables)
            c.set_dofs(r.radiation)
            cores.append(c)
        b = ncore.Batch(cores, prec, max_rhs=7)
        b.solve_waves(float(reqs[0].omegas[0]), reqs[0].betas)  # first use: the batch's job tables
        setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        got = [b.solve_waves(float(w), reqs[0].betas) for w in reqs[0].omegas]
        wall = time.perf_counter() - t0
        same = same and all(np.array_equal(got[fi][k][0], forces[name][k][fi])
                            for k in range(len(reqs)) for fi in range(len(reqs[0].omegas)))
        b.close()
        for c in cores:
            c.close()
        batched[name] = round(nfreq / wall, 2)
        batched[name + "_setup_s"] = round(setup, 2)
        print(f"c4 batched {name}: {nfreq} freq-solves in {wall:.2f} s (set-up {setup:.2f} s)", file=sys.stderr,
              flush=True)
    batched["identical_to_request_loop"] = bool(same)
    batched["how"] = ("nemoh_batch: one frequency of all bodies at a time, every step one launch over all bodies "
                      "(assembly, each block-LU step, each triangular solve, right-hand sides, refinement steps, "
                      "residual and phi gemvs, combination, forces), body conditions per body; handles, matrices "
                      "and job tables resident (set-up untimed, first frequency included in it)")
    return {"workload": f"{nbodies} bodies (seed {seed}): cylinder/box/prolate spheroid, U(2,50) m, draft "
                        "U(0.2,0.8) of height, N log-uniform 300-6000, 7 RHS, 8 frequencies 0.2-2.0 rad/s, "
                        "infinite depth",
            "unit": "freq-solves/s", "mixed": res["mixed"], "fp64": res["fp64"],
            "speedup_mixed_vs_fp64": round(res["mixed"] / res["fp64"], 3), "max_rel_err_mixed_vs_fp64": err,
            "E_N2": float((ns ** 2).mean()), "E_N3": float((ns ** 3).mean()), "workers": workers,
            "batched": batched}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # steps = passes over the 40 frequencies (one batch call each; the
    # request loop's legs run as many, where with 8 in flight four passes,
    # 20 waves, make the start and the drain of the pipeline a small share)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-outputs", dest="outputs", action="store_false",
                    help="skip the Kochin / free-surface potential timing")
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--single-passes", type=int, default=1,
                    help="passes over the 40 frequencies of the one-at-a-time legs")
    # 8 in flight: 316-320 freq-solves/s mixed against 294-304 with 4 and
    # 298-302 with 6 (160-step jobs on the 4000-panel barge)
    ap.add_argument("--workers", type=int, default=8, help="frequencies in flight for the throughput legs")
    ap.add_argument("--c3-steps", type=int, default=3, help="frequencies of the 12k-panel case (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=32, help="requests of the serving run (0: skip)")
    ap.add_argument("--serve-rates", default="5,10,20,50,100,200", help="comma-separated arrival rates (req/s)")
    # 8 workers: saturation 12.6 req/s and p50 61 ms at 5 req/s, against 11.4
    # req/s and 79 ms with 4 (the 32 GiB budget then binds)
    ap.add_argument("--serve-workers", type=int, default=8)
    ap.add_argument("--serve-budget-gib", type=float, default=32.0, help="device memory budget of the request loop")
    ap.add_argument("--c4-bodies", type=int, default=256, help="bodies of the batched run (0: skip)")
    ap.add_argument("--c4-workers", type=int, default=8)
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: each frequency solve is independent, so N GPUs "
                                   "would shard frequencies with no communication; run with --gpus 1"}))
        sys.exit(2)
    from nemoh_amd import core as ncore

    from nemoh_amd import bodycond as bcm

    m, omegas, om, kind, nvel, nds = workload()
    c = ncore.Core(m.vertices, m.panels, 0, 50.0, 1025.0, 9.81)
    c.set_dofs(bcm.six_dof())
    n = m.npanels
    nw = omegas.size
    fp64_steps = a.fp64_steps or a.steps
    # one frequency at a time: phases, per-kernel roofline, host-input rate
    ns = a.single_passes * nw
    t64, f64, ph64 = run_precision(c, ncore.FP64, omegas, nvel, nds, ns, 2)
    tmx, fmx, phmx = run_precision(c, ncore.MIXED, omegas, nvel, nds, ns, 2)
    # the drop-in's per-call data flow: NVEL and NDS handed over from host memory
    th, fh, _ = run_precision(c, ncore.MIXED, omegas, nvel, nds, ns, 2, host_inputs=True)
    # the same frequencies without per-phase timing: HIP-graph replays
    tg64, _, _ = run_precision(c, ncore.FP64, omegas, nvel, nds, ns, 2, timing=False)
    tgmx, _, _ = run_precision(c, ncore.MIXED, omegas, nvel, nds, ns, 2, timing=False)
    host_same = max(float((np.abs(fh[k] - fmx[k]).max(0) / np.abs(fmx[k]).max(0)).max()) for k in fmx if k in fh)
    # the mixed path with the optional FP32 assembly (FP32 far field, wave
    # terms and exponential images, complex64 D and S): its speed and its
    # error against the FP64 path on every force
    c.set_option(ncore.OPT_FP32_ASSEMBLY, 1)
    t32, f32, ph32 = run_precision(c, ncore.MIXED, omegas, nvel, nds, ns, 2)
    c.set_option(ncore.OPT_FP32_ASSEMBLY, 0)
    err32 = max(float((np.abs(f32[k] - f64[k]).max(0) / np.abs(f64[k]).max(0)).max()) for k in f64 if k in f32)
    outputs = run_outputs(c) if a.outputs else None
    # the autotuner's timings of the factorisation forms at this size (one
    # factorisation at a time; the throughput legs below keep the default
    # form, rank 128 with 16-tile bands: they do not opt in to the table)
    tuning = {"objective": "one factorisation at a time, fastest of 3 runs per form; the in-flight legs keep the "
                           "default form (rank 128, band 16)"}
    for name, prec in (("mixed", ncore.MIXED), ("fp64", ncore.FP64)):
        c.set_precision(prec)
        forms, k = c.autotune(reps=3)
        tuning[name] = {"forms_ms": [[r, b, round(ms, 3)] for r, b, ms in forms], "chosen": list(forms[k][:2])}
    c.close()
    ss_mixed = ns / tmx
    ss_fp64 = ns / t64
    # the timed job (value): a.steps passes over the 40 frequencies, each pass
    # one batch call over 40 handles of the mesh, one frequency each
    bmx, bsame_mx, bupd = run_batched_frequencies(m, omegas, ncore.MIXED, a.steps, a.warmup, fmx)
    b64, bsame_64, _ = run_batched_frequencies(m, omegas, ncore.FP64, fp64_steps, a.warmup, f64)
    rate_mixed = a.steps * nw / bmx
    rate_fp64_batch = fp64_steps * nw / b64
    # the same passes as 4 batches of 10 on 4 streams at once
    smx, fsmx = run_batched_streams(m, omegas, ncore.MIXED, a.steps, a.warmup)
    s64, fs64 = run_batched_streams(m, omegas, ncore.FP64, fp64_steps, a.warmup)
    streams_same = all(np.array_equal(fsmx[k], fmx[k]) for k in fmx if k in fsmx) and \
        all(np.array_equal(fs64[k], f64[k]) for k in f64 if k in fs64)
    # the same passes through the request loop, a.workers frequencies in flight
    cmx, gmx = run_concurrent(m, omegas, ncore.MIXED, a.steps, a.warmup, a.workers)
    c64, g64 = run_concurrent(m, omegas, ncore.FP64, fp64_steps, a.warmup, a.workers)
    rate_mixed_loop = a.steps * nw / cmx
    rate_fp64_loop = fp64_steps * nw / c64
    # the comparison is the fastest FP64 rate measured
    fp64_rates = {"batch": (rate_fp64_batch, b64), "4 batches on 4 streams": (fp64_steps * nw / s64, s64),
                  "request loop": (rate_fp64_loop, c64)}
    fp64_method = max(fp64_rates, key=lambda k: fp64_rates[k][0])
    rate_fp64, fp64_time = fp64_rates[fp64_method]
    same = all(np.array_equal(gmx[k], fmx[k]) for k in fmx if k in gmx) and \
        all(np.array_equal(g64[k], f64[k]) for k in f64 if k in g64)
    common = sorted(set(g64) & set(gmx))
    err = max(float((np.abs(gmx[k] - g64[k]).max(0) / np.abs(g64[k]).max(0)).max()) for k in common)
    # roofline of value's dominant kernel, the batch's grouped rank-128 LU
    # update (group_update128_kernel<float>): algorithmic flops of its
    # launches in the timed passes (8 * 128 * m^2 per matrix and launch, m =
    # trailing order) over their summed durations, HIP events on the batch's
    # stream; bytes likewise, scaled by the measured HBM traffic ratio
    b_ms, b_flops, b_bytes, b_launches = bupd
    b_achieved = b_flops / (b_ms * 1e-3) / 1e12 if b_ms > 0 else 0.0
    b_bytes_launch = b_bytes / max(b_launches, 1)
    # the same kernel body one frequency at a time (lu_update128_kernel<float>
    # on the single-stream leg, HIP events on the solver stream)
    upd_ms = float(np.sum(phmx["update"]))
    upd_flops = float(np.sum(phmx["update_flops"]))
    upd_launches = int(np.sum(phmx["update_launches"]))
    achieved = upd_flops / (upd_ms * 1e-3) / 1e12 if upd_ms > 0 else 0.0
    # trailing orders of the rank-128 update launches (block steps in pairs):
    # each reads and writes its m x m complex64 trailing matrix and reads the
    # m x 128 L and 128 x m A panels
    nblk = (n + 63) // 64
    trail = n - 64 * (np.arange(0, nblk, 2) + 2)
    trail = trail[trail > 0]
    upd_bytes = float(np.mean(16.0 * trail ** 2 + 2048.0 * trail))
    # LU + solve flops over the factor + solve phases (complex64 LU 8/3 N^3,
    # 2 triangular solves 8 N^2 NRHS per refinement sweep)
    lu_flops = 8.0 / 3.0 * n ** 3
    fac_ms = float(np.median(phmx["factor"]))
    out = {
        "metric": METRIC, "value": round(rate_mixed, 3), "unit": "freq-solves/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(1e3 * bmx / a.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "complex64 LU + fp64 refinement (fp64 assembly)", "data": "synthetic",
        "config": {"workload": "barge 100x20x10 m, 4000 panels, no symmetry, depth 50 m, 6 DOF + 3 headings "
                               "(9 RHS), 40 frequencies 0.2-2.0 rad/s", "panels": n, "rhs": 9,
                   "frequencies": int(omegas.size)},
        "step": f"one pass over the {nw} frequencies ({nw} freq-solves): one nemoh_batch_solve_waves_multi call "
                f"over {nw} handles of the mesh, one frequency each, every heavy step one launch over all of them",
        "identical_to_single_stream": bool(bsame_mx and bsame_64),
        "fp64_value": round(rate_fp64, 3), "fp64_ms_per_step": round(1e3 * fp64_time / fp64_steps, 3),
        "fp64_method": fp64_method,
        "speedup_mixed_vs_fp64": round(rate_mixed / rate_fp64, 3),
        "max_rel_err_mixed_vs_fp64": err,
        "max_rel_err_mixed_fp32_assembly_vs_fp64": err32,
        "batch": {"mixed": round(rate_mixed, 3), "fp64": round(rate_fp64_batch, 3)},
        "batch_4_streams": {"mixed": round(a.steps * nw / smx, 3), "fp64": round(fp64_steps * nw / s64, 3),
                            "identical_to_single_stream": bool(streams_same),
                            "how": "4 batches of 10 frequencies, each on its own host thread and stream"},
        "request_loop": {"mixed": round(rate_mixed_loop, 3), "fp64": round(rate_fp64_loop, 3),
                         "frequencies_in_flight": a.workers, "hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"]),
                         "identical_to_single_stream": bool(same)},
        "single_stream": {"passes": a.single_passes, "mixed": round(ss_mixed, 3), "fp64": round(ss_fp64, 3),
                          "mixed_graphs": round(ns / tgmx, 3), "fp64_graphs": round(ns / tg64, 3),
                          "mixed_host_inputs": round(ns / th, 3),
                          "mixed_fp32_assembly": round(ns / t32, 3),
                          "host_vs_device_inputs_max_rel_diff": host_same},
        "inputs": "problems generated in HBM by nemoh_core_solve_waves (no per-frequency PCIe transfer); "
                  "mixed_host_inputs = the same solves with NVEL/NDS handed over from host memory",
        "phase_ms_median": {
            "fp64": {k: round(float(np.median(v)), 3) for k, v in ph64.items() if k in ("assembly", "factor", "solve")},
            "mixed": {k: round(float(np.median(v)), 3) for k, v in phmx.items() if k in ("assembly", "factor", "solve")},
            "mixed_fp32_assembly": {k: round(float(np.median(v)), 3) for k, v in ph32.items()
                                    if k in ("assembly", "factor", "solve")},
            "mixed_refine_iterations_max": int(max(phmx["it"]) if phmx["it"] else 0)},
        "roofline": {"bound": "mfma", "kernel": "group_update128_kernel<float> (rank-128 LU trailing update of the "
                                                 "40 matrices of a batch pass, complex64 MFMA)",
                     "measured_on": "value's timed batch passes",
                     "achieved": round(b_achieved, 3), "peak": PEAK["fp32"], "unit": "TFLOP/s",
                     "frac": round(b_achieved / PEAK["fp32"], 4),
                     "traffic": round(BATCH_UPDATE_TRAFFIC_RATIO * b_bytes_launch),
                     "algorithmic_bytes_per_launch": round(b_bytes_launch),
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated, x algorithmic",
                     "launches": b_launches, "avg_launch_us": round(1e3 * b_ms / max(b_launches, 1), 3),
                     "flops_per_launch_avg": b_flops / max(b_launches, 1)},
        "roofline_single_stream": {
            "kernel": "lu_update128_kernel<float> (the same body, one matrix per launch)",
            "measured_on": "single-stream leg (one frequency in flight)",
            "achieved": round(achieved, 3), "peak": PEAK["fp32"], "unit": "TFLOP/s",
            "frac": round(achieved / PEAK["fp32"], 4), "traffic": round(LU_UPDATE_TRAFFIC_RATIO * upd_bytes),
            "algorithmic_bytes_per_launch": round(upd_bytes), "launches": upd_launches,
            "avg_launch_us": round(1e3 * upd_ms / max(upd_launches, 1), 3)},
        "autotune": tuning,
        "lu_factor_tflops": round(lu_flops / (fac_ms * 1e-3) / 1e12, 3),
        "assembly_gpairs_per_s": round(n * n / (float(np.median(phmx["assembly"])) * 1e-3) / 1e9, 3),
    }
    if outputs:
        out["outputs"] = outputs
    if a.c3_steps > 0:
        out["c3"] = run_c3(a.c3_steps)
    if a.c4_bodies > 0:
        out["c4"] = run_c4(a.c4_bodies, a.c4_workers)
    if a.serve_requests > 0:
        out["serving"] = run_serving([float(x) for x in a.serve_rates.split(",")], a.serve_requests,
                                     a.serve_workers, budget_gib=a.serve_budget_gib)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = (cpu_baseline_reference(m, omegas) if REF_HARNESS.exists()
                               else cpu_baseline(m, omegas))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
