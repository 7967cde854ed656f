#!/usr/bin/env python3
"""Benchmark of the batched FixBackbone path on one MI355X.

Workload: 256 independent chains, lengths ~ U[50, 400], each with its own
seeded sequence, self-avoiding C-alpha trace and 20-memory fragment library
(20 perturbed copies of the trace, sigma 1 A, one memory per copy at every
9-residue window).  All FixBackbone terms of the shipped coeff file are on
plus fragment memory, the excluded-volume pair style and the harmonic bonds
of the reference deck; FP64 Langevin dynamics at 300 K with a 2 fs step.

A "step" is one Langevin step of every replica (one batch-step).  `value`
is replica-steps/s of the batched path (FP32 pair arithmetic, FP64
accumulation), whole job, inputs resident in HBM.  Reported beside it:
  * the plain single-chain FP64 GPU drop-in on the same hardware (one chain
    at a time through awsem_run: the phase engine's step sequence replayed
    from a captured graph), replica-steps/s on a sample of the same chains;
  * the max relative force error of the batched path against that FP64
    path (max-norm over atoms);
  * p50 / p99 latency of one batch-step;
  * the HBM roofline of the resident kernel (algorithmic bytes / kernel time);
  * the CPU oracle (FP64 restatement, 1 core) on a bounded sample;
  * the other workloads, bounded: one N=1000 chain through the FP64 drop-in;
    64 replicas of an 8 x 300-residue
    complex; the serving mix at offered arrival rates of 200-1600 jobs/s
    (below and past saturation); fragment-memory table mode; each
    frustratometer mode (ms per call, decoy energies/s beside the oracle).

    python bench.py [--gpus 1] [--steps 640] [--warmup 20]
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MVV2E = 48.88821291 ** 2
BOLTZ = 0.0019872067


def _gen(args):
    L, seed = args
    from awsemmd_amd import workload as W
    return W.make_system([L], seed=seed)


def make_workload(n_rep: int, seed: int, lmin=50, lmax=400, workers=None):
    rng = np.random.default_rng(seed)
    lengths = rng.integers(lmin, lmax + 1, size=n_rep)
    jobs = [(int(L), seed * 100003 + k) for k, L in enumerate(lengths)]
    workers = workers or min(16, os.cpu_count() or 1)
    if workers > 1:
        with mp.get_context("fork").Pool(workers) as pool:
            systems = pool.map(_gen, jobs, chunksize=4)
    else:
        systems = [_gen(j) for j in jobs]
    return systems


def velocities(sys_, T, seed):
    """Maxwell-Boltzmann velocities (A/fs) at T for CA 27, CB 60 (H-beta 2), O 28."""
    rng = np.random.default_rng(seed)
    n = sys_.n
    m = np.zeros((3, n))
    m[0] = 27.0
    m[1] = np.array([2.0 if c == "G" else 60.0 for c in sys_.seq])
    m[2] = 28.0
    sd = np.sqrt(BOLTZ * T / (m * MVV2E))
    v = rng.normal(size=(3, n, 3)) * sd[:, :, None]
    v -= (v * m[:, :, None]).sum(axis=(0, 1)) / m.sum()
    return v


def pair_counts(s, P_cut=13.184136, dssp_cut=7.0, pap_cut=9.9736, ev_cut=4.5):
    """Contact-list entries per evaluation (for the algorithmic byte count)."""
    x = s.coords()
    site = np.where(np.array([c == "G" for c in s.seq])[:, None], x[0], x[1])
    def cnt(a, b, cut, minsep):
        d2 = ((a[:, None, :] - b[None, :, :]) ** 2).sum(-1)
        i, j = np.triu_indices(len(a), k=minsep)
        return int((d2[i, j] < cut * cut).sum())
    p_dens = cnt(site, site, P_cut, 2)
    p_pap = cnt(x[0], x[0], pap_cut, 1)
    p_dssp = 2 * cnt(x[2], x[0], dssp_cut, 3)
    allx = x.reshape(-1, 3)
    p_ev = cnt(allx, allx, ev_cut, 1)
    return p_dens, p_dssp, p_pap, p_ev


def algorithmic_bytes(systems):
    """Per batch-step: 360 L + 336 M + 64 L + 4 (3 P_dens + P_dssp + P_pap + P_ev)."""
    tot = 0
    for s in systems:
        L = s.n
        M = len(s.memories)
        pd, pdssp, ppap, pev = pair_counts(s)
        tot += 360 * L + 336 * M + 64 * L + 4 * (3 * pd + pdssp + ppap + pev)
    return tot


def profiled_traffic(chunk):
    """HBM bytes per launch of the resident kernel from the committed rocprofv3
    summary of this workload (FETCH_SIZE and WRITE_SIZE, copy-calibrated),
    rescaled to `chunk` steps per launch; None without the summary."""
    path = os.path.join(ROOT, "profiles", "k_resident_256chains.md")
    try:
        kv = {}
        for line in open(path):
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        return float(kv["traffic_bytes_per_launch"]) * chunk / float(kv["steps_per_launch"]), float(kv["valu_busy_pct"])
    except (OSError, KeyError, ValueError):
        return None, None


def rel_force_err(f, fref):
    d = np.linalg.norm((f - fref).reshape(-1, 3), axis=1).max()
    m = np.linalg.norm(fref.reshape(-1, 3), axis=1).max()
    return float(d / m) if m > 0 else float(d)


def _cpu_chain(args):
    """One host core: the oracle's FP64 Langevin steps on one chain for
    `seconds`; returns (residues, steps, seconds)."""
    s, seconds = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from awsemmd_amd import workload as W
    d = tempfile.mkdtemp(prefix="awsem_cpu_")
    p = W.write_system(s, d)
    o = oracle.Oracle(p["coeff"], p["seq"], d)
    x = s.coords()
    f, _ = o.compute(x)
    v = velocities(s, 300.0, 7)
    x, v, f, _ = o.langevin(x, v, f, 1)  # warm-up
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < seconds:
        x, v, f, e = o.langevin(x, v, f, 2, step0=k)
        k += 2
    return s.n, k, time.perf_counter() - t0


def cpu_baseline(systems, budget_s=30.0, max_cores=16):
    """The FP64 oracle (a plain C restatement of the path: FixBackbone terms,
    pair style, bonds, Langevin) on the host cores, one chain per core at
    once, like one LAMMPS rank per replica.  The sample is `cores` chains
    spread evenly over the workload's lengths; the batch rate on these cores
    is cores / (mean CPU seconds per replica-step)."""
    cores = max(1, min(max_cores, os.cpu_count() or 1))
    order = np.argsort([s.n for s in systems])
    pick = [systems[order[int((k + 0.5) * len(order) / cores)]] for k in range(cores)]
    per = max(1.5, budget_s / cores)
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_chain, [(s, per) for s in pick], chunksize=1)
    sec_per_step = np.array([t / k for _, k, t in res])
    mean_t = float(sec_per_step.mean())
    lengths = [n for n, _, _ in res]
    return {"value": cores / mean_t, "unit": "replica-steps/s", "cores": cores, "kind": "port",
            "per_core_value": 1.0 / mean_t,
            "sample": f"oracle FP64 FixBackbone+pair style+bonds+Langevin (all-pairs sweeps, no neighbour list) on "
                      f"{cores} host cores at once, one chain each (L {min(lengths)}-{max(lengths)}, evenly spaced "
                      f"over the workload's lengths), {per:.1f} s each; rate = cores / mean seconds per replica-step"}


def serving_pool(A, ctx, n_pool, seed, sigma=0.6):
    """The serving mix's job pool: lengths lognormal around a median of 150
    residues (log-sd `sigma`), capped at 1500, each system generated like the
    single-chain workload with its own seed (sequence, trace, 20-memory
    library)."""
    rng = np.random.default_rng(seed)
    lengths = np.clip(np.round(rng.lognormal(np.log(150.0), sigma, n_pool)), 20, 1500).astype(int)
    jobs = [(int(L), seed * 7919 + k) for k, L in enumerate(lengths)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(min(16, os.cpu_count() or 1)) as pool:
        systems = pool.map(_gen, jobs, chunksize=2)
    fixes = [ctx(s, A.PREC_FP32) for s in systems]
    log(f"serving: pool of {n_pool} jobs generated in {time.perf_counter() - t0:.1f}s "
        f"(L median {int(np.median(lengths))}, max {lengths.max()}, {int((lengths > 520).sum())} above the LDS cap)")
    return systems, fixes, lengths


def serving_run(A, systems, fixes, lengths, rate_hz, duration_s, seed, steps=500, round_steps=50, sigma=0.6):
    """One arrival rate of the serving mix: a Poisson stream of annealing jobs
    (600 -> 300 K over `steps` steps, then the final per-term energies) drawn
    from the pool, submitted at their arrival times for duration_s to a
    background-scheduled server that admits them by slots and device memory
    while others run.  Latency runs from a job's arrival to its result."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, int(rate_hz * duration_s * 1.5) + 16))
    arrivals = arrivals[arrivals < duration_s]
    n_jobs = len(arrivals)
    pick = rng.integers(0, len(systems), n_jobs)
    srv = A.Server(fixes[0], A.PREC_FP32, max_jobs=1024, steps_per_round=round_steps)
    # warm-up job (kernel load, first allocations), not counted
    w = srv.submit(fixes[0], systems[0].coords(), None, round_steps, seed=1)
    srv.pump()
    srv.result(w, systems[0].n)
    srv.release(w)
    srv.start()
    # requests arrive on their own threads (the library packs a job's device
    # image on the submitting thread); latency runs from the arrival time
    from concurrent.futures import ThreadPoolExecutor

    def submit(k):
        t_arr = t_start + arrivals[k]
        q = pick[k]
        jid = srv.submit(fixes[q], systems[q].coords(), None, steps, t_start=600.0, t_stop=300.0, seed=1000 + k)
        return jid, time.perf_counter() - t_arr

    t_start = time.perf_counter()
    futs = []
    with ThreadPoolExecutor(max_workers=8) as pool:
        for k in range(n_jobs):
            dt = t_start + arrivals[k] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            futs.append(pool.submit(submit, k))
        subs = [f.result() for f in futs]
    t_sub = time.perf_counter() - t_start
    lat, energies_ok = [], True
    for (jid, t_pack), q in zip(subs, pick):
        srv.wait(jid, 600.0)
        x, e, t = srv.result(jid, systems[q].n)
        energies_ok &= bool(np.all(np.isfinite(e)))
        lat.append(1e3 * t_pack + t["total_ms"])
        srv.release(jid)
    t_all = time.perf_counter() - t_start
    st = srv.stats()
    srv.stop()
    srv.close()
    lat = np.array(lat)
    Ls = lengths[pick]
    return {"offered_rate_hz": rate_hz, "jobs": n_jobs, "achieved_arrival_rate_hz": n_jobs / t_sub,
            "jobs_per_s": n_jobs / t_all, "replica_steps_per_s": n_jobs * steps / t_all,
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "max": float(lat.max())},
            "lengths": {"median": int(np.median(Ls)), "max": int(Ls.max()), "above_lds_cap": int((Ls > 520).sum())},
            "peak_active_jobs": st["peak_active"], "rounds": st["rounds"], "device_ms": st["device_ms"],
            "final_energies_finite": energies_ok, "wall_s": t_all}


def serving_sweep(A, ctx, rates, duration_s, seed, n_pool=600, steps=500, round_steps=50):
    """The serving mix at several offered arrival rates, below and at
    saturation, each for duration_s, from one job pool."""
    systems, fixes, lengths = serving_pool(A, ctx, n_pool, seed)
    out = []
    for k, r in enumerate(rates):
        res = serving_run(A, systems, fixes, lengths, r, duration_s, seed + 17 * k, steps, round_steps)
        log(f"serving at {r:.0f} jobs/s offered: {res['jobs_per_s']:.1f} jobs/s done, latency p50 "
            f"{res['latency_ms']['p50']:.1f} ms p99 {res['latency_ms']['p99']:.1f} ms, peak {res['peak_active_jobs']:.0f} active")
        out.append(res)
    return {"steps_per_job": steps, "steps_per_round": round_steps, "duration_s_per_rate": duration_s,
            "lengths": {"median": 150, "lognormal_sigma": 0.6, "cap": 1500}, "pool": n_pool, "rates": out}


def calib_copy(nbytes):
    """Device-to-device copy of nbytes: the reference kernel for scaling the
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    hip.hipMalloc(C.byref(a), C.c_size_t(nbytes))
    hip.hipMalloc(C.byref(b), C.c_size_t(nbytes))
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    for _ in range(3):
        hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"calibration copy of {nbytes} bytes x3 done")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def force_field(A, W, system):
    """Shipped coeff with fragment memory on, parameter tables in a temp dir;
    returns ctx(system, precision) -> FixBackbone."""
    coeff_dir = tempfile.mkdtemp(prefix="awsem_bench_")
    W.write_system(system, coeff_dir, fm=False)  # parameter tables
    coeff = os.path.join(coeff_dir, "fix_backbone_coeff.data")
    with open(coeff, "w") as f:
        f.write(W.coeff_text(fm=True, mem_file="unused.mem", gamma_file="uniform.gamma"))

    def ctx(s, prec):
        return A.FixBackbone.from_arrays(coeff, coeff_dir, precision=prec, **W.arrays(s))
    return ctx


def run_single_chain(a, A, W):
    """Plain FP64 GPU drop-in on one N=1000 chain: `steps` Langevin steps
    through awsem_run (the phase engine's step sequence replayed from a
    captured graph; one untimed call of `check_every` steps first captures
    it).  Every `check_every` steps the per-term energies that awsem_run
    returns (those of its final coordinates) are compared, untimed, with
    awsem_compute_res at the same coordinates (the two entry points share no
    launch sequence; the forces awsem_run returns also hold the last step's
    Langevin drag and random force, as LAMMPS's do); the comparison with the
    FP64 oracle at this size is a GPU test (tests/test_gpu_configs.py, N=1000
    checked every 100 steps)."""
    s = W.make_system([a.n_single], seed=a.seed)
    d = tempfile.mkdtemp(prefix="awsem_c1_")
    p = W.write_system(s, d)
    fix = A.FixBackbone(p["coeff"], p["seq"], d, A.PREC_FP64)
    x = s.coords()
    v = velocities(s, 300.0, 3)
    worst_e, t_gpu, done, checks = 0.0, 0.0, 0, 0
    fix.run(x.copy(), v.copy(), min(a.check_every, a.steps), dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0,
            seed=a.seed, step0=0)  # warm-up: kernel load, graph capture
    finite = True
    while done < a.steps:
        k = min(a.check_every, a.steps - done)
        t0 = time.perf_counter()
        x, v, f, e = fix.run(x, v, k, dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0, seed=a.seed, step0=done)
        t_gpu += time.perf_counter() - t0
        done += k
        _, eg = fix.compute_res(x)
        finite = finite and bool(np.all(np.isfinite(e)) and np.all(np.isfinite(x)))
        worst_e = max(worst_e, float(np.max(np.abs(e - eg) / np.maximum(1.0, np.abs(eg)))))
        checks += 1
        if checks % 10 == 0:
            log(f"c1: {done} steps, run vs compute_res: per-term energy {worst_e:.2e}")
    return {"metric": "plain FP64 GPU drop-in, one chain, steps/s", "value": done / t_gpu,
            "unit": "steps/s", "n_gpus": 1, "steps": done, "warmup": 0, "ms_per_step": 1e3 * t_gpu / done,
            "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"single chain N={a.n_single}, FP64 end-to-end, Langevin 300 K, dt 2 fs",
                       "check_every": a.check_every, "checks": checks},
            "ns_per_day": done / t_gpu * 2.0 * 86400 / 1e6,
            "max_rel_energy_diff_run_vs_compute_res": worst_e,
            "final_energies_finite": finite}


def run_multimer(a, A, W):
    """64 replicas of one 8 x 300-residue compact multimer (centroids 10-15 A
    apart, seeded), each with its own random stream, FP32 pair arithmetic,
    annealing 600 -> 300 K; a bounded number of steps is timed, with the
    annealing ramp laid out over the full run (a.anneal_steps)."""
    t0 = time.perf_counter()
    complex_ = _gen_multimer(a.seed * 31)
    systems = [complex_] * a.replicas
    log(f"c3: multimer of {complex_.n} residues generated in {time.perf_counter() - t0:.1f}s, {a.replicas} replicas")
    ctx = force_field(A, W, systems[0])
    batch = A.Batch([ctx(s, A.PREC_FP32) for s in systems], A.PREC_FP32)
    x0 = A.pack_states([s.coords() for s in systems])
    batch.set_engine(a.engine)
    batch.set_state(x0, np.zeros_like(x0))
    lg = dict(dt=2.0, damp=1000.0, seed=a.seed)
    ramp = lambda step: 600.0 - 300.0 * step / a.anneal_steps
    batch.run(a.warmup, t_start=ramp(0), t_stop=ramp(a.warmup), step0=0, **lg)
    tuned = None if a.no_autotune else batch.autotune(probe_steps=20, t_start=ramp(a.warmup), t_stop=ramp(a.warmup),
                                                       step0=a.warmup, **lg)
    chunk = a.chunk if a.chunk > 0 else (tuned["steps_per_launch"] if tuned else 10)
    step, ms = a.warmup, []
    t0 = time.perf_counter()
    while step < a.warmup + a.steps:
        k = min(chunk, a.warmup + a.steps - step)
        c0 = time.perf_counter()
        batch.run(k, t_start=ramp(step), t_stop=ramp(step + k), step0=step, **lg)
        ms.append(1e3 * (time.perf_counter() - c0) / k)
        step += k
    t = time.perf_counter() - t0
    _, _, _, e = batch.get_state()
    value = a.replicas * a.steps / t
    return {"metric": "replica-steps/s, multimer annealing batch", "value": value, "unit": "replica-steps/s",
            "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * t / a.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 pair math, f64 accumulation/integration", "data": "synthetic",
            "config": {"workload": "64 replicas of an 8-chain x 300-residue compact complex, centroids 10-15 A "
                                   "apart, annealing 600->300 K", "replicas": a.replicas, "residues_per_replica": systems[0].n,
                       "anneal_steps": a.anneal_steps, "steps_per_launch": chunk,
                       "verlet_skin_A": tuned["skin"] if tuned else 2.0,
                       "engine": tuned["engine"] if tuned else a.engine},
            "ns_per_day_per_replica": value / a.replicas * 2.0 * 86400 / 1e6,
            "latency_ms_per_batch_step": {"p50": float(np.percentile(ms, 50)), "p99": float(np.percentile(ms, 99))},
            "final_energies_finite": bool(np.all(np.isfinite(e)))}


def _gen_multimer(seed):
    from awsemmd_amd import workload as W
    return W.make_system([300] * 8, seed=seed, compact=True)


# phase-clock slots of the resident engine (AWSEM_PHASE_PROF=1); slot 4 is
# unused (the xi scaling is part of F); in table mode the table pass is
# clocked in B_fm
PHASE_NAMES = ["A", "B", "C", "D", None, "F", "int1", "int2", "B_fm", "B_local", "B_pair", "list",
               "pair_scan", "pair_dssp", "pair_pap", "pair_ev"]


def run_fm_table(a, A, W):
    """Fragment-memory table mode on replicas of one deck-sized protein: a
    63-residue chain with 20 memories per 9-residue window, the shipped
    decks' grid (r from 0 to 50 A in 0.01 A steps: 5002 points per curve),
    well width 0.1, sigma exponent 0.15.  One context is given once per
    replica, so the batch stores its curves once; each replica has its own
    velocities and random stream (Langevin 300 K, dt 2 fs), FP32 pair
    arithmetic.  The same batch with the fragment memory evaluated from the
    memories is timed beside it (a different potential with the decks' well
    width: the table's Gaussians are 10 times narrower)."""
    s = W.make_system([a.table_len], seed=a.seed * 7)
    d = tempfile.mkdtemp(prefix="awsem_tb_")
    grid = (0.0, 50.0, 0.01, 0.1, 0, 0.15)
    res = {}
    x0 = A.pack_states([s.coords()] * a.table_replicas)
    v0 = A.pack_states([velocities(s, 300.0, 500 + k) for k in range(a.table_replicas)])
    lg = dict(dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0, seed=a.seed)
    for mode in ("table", "memory"):
        p = W.write_system(s, os.path.join(d, mode), table=grid if mode == "table" else None)
        t0 = time.perf_counter()
        fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP32)
        t_ctx = time.perf_counter() - t0
        batch = A.Batch([fix] * a.table_replicas, A.PREC_FP32)
        batch.set_state(x0, v0)
        clocked = bool(os.environ.get("AWSEM_PHASE_PROF"))
        batch.run(a.warmup, step0=0, **lg)
        if clocked or a.no_autotune:  # fixed settings (a clocked run counts every evaluation)
            batch.set_engine("resident")
            tuned = {"engine": "resident", "steps_per_launch": 160, "skin": 2.0}
        else:
            tuned = batch.autotune(probe_steps=160, step0=a.warmup, **lg)  # restores the state
        chunk = tuned["steps_per_launch"]
        step = a.warmup
        t0 = time.perf_counter()
        done = 0
        while done < a.table_steps:
            k = min(chunk, a.table_steps - done)
            batch.run(k, step0=step + done, **lg)
            done += k
        t = time.perf_counter() - t0
        e = batch.get_state()[3]
        if clocked:
            ticks = batch.phase_clock() / 100.0  # microseconds
            ev = 1 + a.warmup + done  # force evaluations: the first launch's initial one, then one per step
            log(f"{mode} mode, phase us/eval, mean replica: " + " ".join(
                f"{nm}={v / ev / a.table_replicas:.1f}" for nm, v in zip(PHASE_NAMES, ticks[16:32]) if nm))
        res[mode] = dict(value=a.table_replicas * done / t, ms_per_step=1e3 * t / done, context_s=t_ctx,
                         engine=tuned["engine"], steps_per_launch=chunk, skin=tuned["skin"],
                         finite=bool(np.all(np.isfinite(e))))
        log(f"fm table line, {mode} mode: {res[mode]['value']:.0f} replica-steps/s ({tuned['engine']}, "
            f"context {t_ctx:.1f} s)")
        batch.close()
        fix.close()
    tb, mm = res["table"], res["memory"]
    n = s.n
    pairs = sum(max(0, n - sep) for sep in range(3, 10))
    gly_skip = 0  # glycine CB pairs have no curve
    seq = s.seq
    for i in range(n):
        for j in range(i + 3, min(n, i + 10)):
            gly_skip += (seq[i] == "G") * 2 + (seq[j] == "G") * 2 - (seq[i] == "G" and seq[j] == "G")
    interp = 4 * pairs - gly_skip
    return {"metric": "replica-steps/s, fragment-memory table mode", "value": tb["value"], "unit": "replica-steps/s",
            "n_gpus": 1, "steps": a.table_steps, "warmup": a.warmup, "ms_per_step": tb["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 pair math and curves, f64 accumulation/integration", "data": "synthetic",
            "config": {"workload": f"{a.table_replicas} replicas of one {n}-residue chain, 20 memories per 9-residue "
                                   "window, [Fragment_Memory_Table] 0 50 0.01, well width 0.1, sigma exponent 0.15, "
                                   "Langevin 300 K, dt 2 fs",
                       "replicas": a.table_replicas, "residues_per_replica": n, "engine": tb["engine"],
                       "steps_per_launch": tb["steps_per_launch"], "verlet_skin_A": tb["skin"]},
            "memory_mode": {"value": mm["value"], "ms_per_step": mm["ms_per_step"], "engine": mm["engine"],
                            "steps_per_launch": mm["steps_per_launch"]},
            "speedup_vs_memory_mode": tb["value"] / mm["value"],
            "interpolations_per_replica_step": interp,
            "curve_bytes_read_per_replica_step": 16 * interp,
            "curve_store_bytes": int(4 * n * 7 * 5002 * 8),
            "context_build_s": {"table": tb["context_s"], "memory": mm["context_s"]},
            "final_energies_finite": tb["finite"] and mm["finite"]}


# the frustratometer line: (name, section text with {m} decoys / calculations, GPU m, CPU m)
FRUST_MODES = [
    ("tertiary configurational", "[Tertiary_Frustratometer]\n9.5\n{m}\n1\nconfigurational", 1000, 1000),
    ("tertiary mutational", "[Tertiary_Frustratometer]\n9.5\n{m}\n1\nmutational", 1000, 200),
    ("tertiary singleresidue", "[Tertiary_Frustratometer]\n9.5\n{m}\n1\nsingleresidue", 1000, 200),
    ("nmer pairwise", "[Nmer_Frustratometer]\n6\n9.5\n10\n{m}\n1\n1.7 -1 0\n0 0 1 0.6\npairwise", 1000, 200),
    ("nmer singlenmer", "[Nmer_Frustratometer]\n6\n9.5\n10\n{m}\n1\n1.7 -1 0\n0 0 1 0.6\nsinglenmer", 1000, 200),
    ("fragment shuffle", "[Fragment_Frustratometer]\nshuffle\nfragsLAMW.mem\n{m}\n1", 101, 41),
]


def _frust_call(fix_or_oracle, kind, x, oracle=False):
    """One frustratometer call; returns its row count and decoy entries."""
    if kind.startswith("tertiary"):
        r = fix_or_oracle.tert_frust(x)
    elif kind.startswith("nmer"):
        r = fix_or_oracle.nmer_frust(x)[0]
    else:
        r = fix_or_oracle.frag_frust(x, 0)[0] if oracle else fix_or_oracle.frag_frust(x)[0]
    return len(r)


def _frust_systems(W, seed):
    """Two compact chains of 150 residues (tertiary and n-mer modes) and one
    compact 300-residue chain whose 5,840 memories are also the shuffled
    decoy fragments (fragment mode: decoys placed across chains are an error
    in the reference)."""
    return W.make_system([150, 150], seed=seed, compact=True), W.make_system([300], seed=seed + 1, compact=True)


def _frust_workdir(W, s, section, d):
    p = W.write_system(s, d)
    with open(p["coeff"], "a") as f:
        f.write("\n" + section + "\n\n")
    return p


def _entries(kind, rows, m, n):
    """Decoy energies one call evaluates (the natives included)."""
    if kind == "tertiary configurational":
        return rows  # the decoy statistics are drawn once, at the first call
    if kind == "fragment shuffle":
        return m * n  # m calculations of every residue's memory energy
    return rows * (m + 1)


def frustratometer_cpu_baseline(W, seed):
    """The oracle (FP64 C restatement, 1 core) on the same systems with fewer
    decoys (CPU m per mode), one call per mode after a warm-up call: decoy
    energies per second."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    sysA, sysB = _frust_systems(W, seed)
    out = {}
    for kind, sec, _, m in FRUST_MODES:
        s = sysB if kind.startswith("fragment") else sysA
        p = _frust_workdir(W, s, sec.format(m=m), tempfile.mkdtemp(prefix="awsem_fcpu_"))
        o = oracle.Oracle(p["coeff"], p["seq"], p["workdir"])
        x = s.coords()
        _frust_call(o, kind, x, oracle=True)
        t0 = time.perf_counter()
        rows = _frust_call(o, kind, x, oracle=True)
        t = time.perf_counter() - t0
        out[kind] = {"entries_per_s": _entries(kind, rows, m, s.n) / t, "decoys": m, "s_per_call": t}
    return out


def run_frustratometers(a, A, W):
    """The analysis workload: each frustratometer mode on one GPU, FP64, the
    steady-state wall time of one call (host decoy draws, uploads, kernels,
    copies back) averaged over a.frust_calls calls after a warm-up call, and
    the decoy energies per second, beside the oracle on one host core."""
    sysA, sysB = _frust_systems(W, a.seed)
    res = {}
    for kind, sec, m, _ in FRUST_MODES:
        s = sysB if kind.startswith("fragment") else sysA
        p = _frust_workdir(W, s, sec.format(m=m), tempfile.mkdtemp(prefix="awsem_frust_"))
        fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
        x = s.coords()
        _frust_call(fix, kind, x)
        t0 = time.perf_counter()
        for _ in range(a.frust_calls):
            rows = _frust_call(fix, kind, x)
        t = (time.perf_counter() - t0) / a.frust_calls
        fix.close()
        res[kind] = {"ms_per_call": 1e3 * t, "rows": rows, "decoys": m, "residues": s.n,
                     "entries_per_s": _entries(kind, rows, m, s.n) / t}
        log(f"frustratometer {kind}: {1e3 * t:.2f} ms per call, {rows} rows, {m} decoys")
    cpu = None if a.no_cpu else frustratometer_cpu_baseline(W, a.seed)
    if cpu:
        for kind in res:
            res[kind]["cpu_entries_per_s"] = cpu[kind]["entries_per_s"]
            res[kind]["speedup_vs_cpu"] = res[kind]["entries_per_s"] / cpu[kind]["entries_per_s"]
    return {"workload": "frustratometers on 2 x 150 compact chains (tertiary, n-mer: cutoff 9.5 A, 1000 decoys; "
                        "n-mers of 6 with 10 contacts) and one 300-residue chain (fragment shuffle: its 5,840 "
                        "memories as decoy fragments, 100 decoy calculations)",
            "unit": "decoy energies/s (and ms per call)", "dtype": "f64", "modes": res,
            "cpu_baseline": {"cores": 1, "kind": "port",
                             "sample": "oracle, same systems, 200 decoys per row (configurational 1000, its "
                                       "statistics drawn once; fragment shuffle 40 decoy calculations), one call "
                                       "after a warm-up call"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=640)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--chunk", type=int, default=0, help="batch-steps per resident launch (0: autotuned)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default Verlet skin (2 A) and --engine")
    ap.add_argument("--engine", choices=["resident", "wide"], default="resident",
                    help="batch engine when not autotuned")
    ap.add_argument("--plain-sample", type=int, default=32)
    ap.add_argument("--plain-steps", type=int, default=50)
    ap.add_argument("--cpu-budget", type=float, default=30.0, help="CPU seconds of the host baseline, all cores")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--serving-duration", type=float, default=5.0,
                    help="seconds of arrivals per offered rate of the serving mix (0: skip)")
    ap.add_argument("--serving-rates", default="200,400,800,1200,1600",
                    help="offered Poisson arrival rates of the serving mix (jobs/s), comma-separated")
    ap.add_argument("--single-steps", type=int, default=500,
                    help="steps of the N=1000 FP64 drop-in line in the default run, checked every 100 (0: skip)")
    ap.add_argument("--multimer-steps", type=int, default=20,
                    help="timed steps of the 64 x (8 x 300) multimer line in the default run (0: skip)")
    ap.add_argument("--table-steps", type=int, default=1000,
                    help="timed steps of the fragment-memory table-mode line in the default run (0: skip)")
    ap.add_argument("--table-replicas", type=int, default=1024)
    ap.add_argument("--table-len", type=int, default=63)
    ap.add_argument("--frust-calls", type=int, default=5,
                    help="timed calls per frustratometer mode in the default run (0: skip the line)")
    ap.add_argument("--workload", choices=["batch", "single", "multimer", "table", "frust"], default="batch",
                    help="batch: 256 chains (default, the headline); single: one N=1000 chain through the "
                         "FP64 drop-in; multimer: 64 x (8 x 300) annealing; table: fragment-memory table mode "
                         "on replicas of one 63-residue chain; frust: the frustratometer modes")
    ap.add_argument("--n-single", type=int, default=1000)
    ap.add_argument("--check-every", type=int, default=100)
    ap.add_argument("--anneal-steps", type=int, default=20000)
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run a device-to-device copy of this size (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every replica is independent, so a multi-GPU run "
              "would only shard replicas across GPUs; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import awsemmd_amd as A
    from awsemmd_amd import workload as W
    A.lib()

    if a.workload == "single":
        if a.steps == 640:
            a.steps = 10000
        print(json.dumps(run_single_chain(a, A, W)), flush=True)
        return
    if a.workload == "multimer":
        if a.replicas == 256:
            a.replicas = 64
        print(json.dumps(run_multimer(a, A, W)), flush=True)
        return
    if a.workload == "frust":
        print(json.dumps(run_frustratometers(a, A, W)), flush=True)
        return
    if a.workload == "table":
        print(json.dumps(run_fm_table(a, A, W)), flush=True)
        return

    t0 = time.perf_counter()
    systems = make_workload(a.replicas, a.seed)
    t_gen = time.perf_counter() - t0
    log(f"generated {len(systems)} chains in {t_gen:.1f}s")
    coeff_dir = tempfile.mkdtemp(prefix="awsem_bench_")
    W.write_system(systems[0], coeff_dir, fm=False)  # parameter tables
    coeff = os.path.join(coeff_dir, "fix_backbone_coeff.data")
    with open(coeff, "w") as f:
        f.write(W.coeff_text(fm=True, mem_file="unused.mem", gamma_file="uniform.gamma"))

    def ctx(s, prec):
        return A.FixBackbone.from_arrays(coeff, coeff_dir, precision=prec, **W.arrays(s))

    t0 = time.perf_counter()
    fixes = [ctx(s, A.PREC_FP32) for s in systems]
    batch = A.Batch(fixes, A.PREC_FP32)
    del fixes
    t_build = time.perf_counter() - t0
    log(f"batch built in {t_build:.1f}s")
    x0 = A.pack_states([s.coords() for s in systems])
    v0 = A.pack_states([velocities(s, 300.0, 1000 + k) for k, s in enumerate(systems)])
    batch.set_engine(a.engine)
    batch.set_state(x0, v0)
    lg = dict(dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0, seed=a.seed)

    # warm-up, then the autotuner (skin, steps per launch) from the warmed state
    batch.run(a.warmup, step0=0, **lg)
    log(f"warm-up done ({a.warmup} steps)")
    tuned = None
    if not a.no_autotune:
        t0 = time.perf_counter()
        # 160-step probes: each candidate is timed over several list-rebuild
        # cycles (a short probe from freshly built lists favours small skins)
        tuned = batch.autotune(probe_steps=160, step0=a.warmup, **lg)
        tuned["tuning_s"] = time.perf_counter() - t0
        log(f"autotuned: {tuned['engine']} engine, skin {tuned['skin']} A, {tuned['steps_per_launch']} steps per launch "
            f"({tuned['ms_per_step']:.3f} ms/step) in {tuned['tuning_s']:.1f}s")
    if a.chunk <= 0:
        a.chunk = tuned["steps_per_launch"] if tuned else 10
    # at least four timed launches, so that the per-launch figures (latency
    # percentiles, kernel time) come from several samples
    a.chunk = max(1, min(a.chunk, a.steps // 4 if a.steps >= 4 else a.steps))
    # one untimed launch at the chosen settings: the timed launches start
    # from lists built with the chosen skin, as in a long run
    batch.run(a.chunk, step0=a.warmup, **lg)
    step = a.warmup + a.chunk
    # timed region: K batch-steps in launches of `chunk` steps; each launch's
    # device time comes from HIP events recorded around it on the batch's
    # stream (awsem_batch_timing)
    chunks = []
    kernel_ms = []
    left = a.steps
    t_start = time.perf_counter()
    while left > 0:
        k = min(a.chunk, left)
        c0 = time.perf_counter()
        batch.run(k, step0=step, **lg)
        chunks.append((time.perf_counter() - c0, k))
        kernel_ms.append((batch.timing()[0], k))
        step += k
        left -= k
    t_total = time.perf_counter() - t_start
    replica_steps = a.replicas * a.steps
    value = replica_steps / t_total
    # per batch-step latency of each launch, from its HIP-event device time
    per_step_ms = np.array([ms / k for ms, k in kernel_ms])
    res_total = sum(s.n for s in systems)
    log(f"timed {a.steps} steps: {value:.1f} replica-steps/s")
    nb, npairs, nbad = batch.list_stats()
    log(f"Verlet lists: {nb} builds over {a.replicas} replicas, {npairs} pairs now, {nbad} invalid")

    if os.environ.get("AWSEM_PHASE_PROF"):
        ticks = batch.phase_clock() / 100.0  # microseconds
        names = PHASE_NAMES
        nsteps_all = (a.warmup + 1) + a.chunk + a.steps + (a.steps + a.chunk - 1) // a.chunk  # force evaluations
        log("phase us/eval, longest replica: " + " ".join(f"{n}={t / nsteps_all:.1f}" for n, t in zip(names, ticks[:16]) if n))
        log("phase us/eval, mean replica:    " + " ".join(f"{n}={t / nsteps_all / a.replicas:.1f}" for n, t in zip(names, ticks[16:32]) if n))
    # force error of the batched FP32 path against the FP64 path (sample)
    batch.compute()
    xs, vs, fs, es = batch.get_state()
    xs_l, fs_l = batch.split(xs), batch.split(fs)
    rng = np.random.default_rng(5)
    sample = sorted(rng.choice(a.replicas, size=min(a.plain_sample, a.replicas), replace=False).tolist())
    errs = []
    plain_time, plain_steps = 0.0, 0
    for r in sample:
        fx = ctx(systems[r], A.PREC_FP64)
        f64, e64 = fx.compute_res(xs_l[r])
        errs.append(rel_force_err(fs_l[r], f64))
        # plain FP64 drop-in throughput: device-resident dynamics, one chain
        xv = xs_l[r].copy()
        vv = batch.split(vs)[r].copy()
        fx.run(xv, vv, a.plain_steps, step0=0, **lg)  # warm-up: list build, graph capture
        c0 = time.perf_counter()
        fx.run(xv, vv, a.plain_steps, step0=a.plain_steps, **lg)
        plain_time += time.perf_counter() - c0
        plain_steps += a.plain_steps
        fx.close()
    plain_value = plain_steps / plain_time
    log(f"plain FP64 sample: {plain_value:.1f} replica-steps/s, force err {max(errs):.2e}")
    force_err = max(errs)

    bytes_step = algorithmic_bytes(systems)
    # achieved = algorithmic bytes of the steps the timed launches ran over
    # their summed device time (HIP events on the launch stream)
    launch_steps = float(np.mean([k for _, k in kernel_ms]))
    k_avg_ms = float(np.mean([ms for ms, _ in kernel_ms]))
    achieved = bytes_step * sum(k for _, k in kernel_ms) / (1e-3 * sum(ms for ms, _ in kernel_ms)) / 1e9
    traffic, valu_busy = profiled_traffic(launch_steps)
    if a.calib_copy_mb:
        calib_copy(a.calib_copy_mb << 20)
    serving = None
    if a.serving_duration > 0:
        rates = [float(r) for r in a.serving_rates.split(",") if r]
        serving = serving_sweep(A, ctx, rates, a.serving_duration, a.seed + 1)
    single = None
    if a.single_steps > 0:
        a1 = argparse.Namespace(**vars(a))
        a1.steps, a1.check_every = a.single_steps, 100
        single = run_single_chain(a1, A, W)
    multimer = None
    if a.multimer_steps > 0:
        a3 = argparse.Namespace(**vars(a))
        a3.replicas, a3.steps, a3.warmup, a3.chunk, a3.engine = 64, a.multimer_steps, 5, 0, "wide"
        multimer = run_multimer(a3, A, W)
    table = None
    if a.table_steps > 0:
        a4 = argparse.Namespace(**vars(a))
        a4.warmup = 20
        table = run_fm_table(a4, A, W)
    frust = None
    if a.frust_calls > 0:
        frust = run_frustratometers(a, A, W)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(systems, a.cpu_budget)

    out = {
        "metric": "replica-steps/s batched vs plain FP64 GPU at force err<=1e-5; p50/p99 lat",
        "value": value,
        "unit": "replica-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t_total / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair math, f64 accumulation/integration",
        "data": "synthetic",
        "config": {"workload": "256 chains, L~U[50,400], 20 memories per 9-residue window, Langevin 300 K, dt 2 fs",
                   "replicas": a.replicas, "residues": res_total, "memories": int(sum(len(s.memories) for s in systems)),
                   "steps_per_launch": a.chunk, "seed": a.seed,
                   "verlet_skin_A": tuned["skin"] if tuned else 2.0, "autotuned": tuned is not None,
                   "engine": tuned["engine"] if tuned else a.engine},
        "plain_fp64_gpu": {"value": plain_value, "unit": "replica-steps/s",
                           "sample": f"{len(sample)} chains x {a.plain_steps} steps through awsem_run (phase engine, "
                                     "captured graph), one chain at a time"},
        "speedup_vs_plain_fp64": value / plain_value,
        "force_err_max_vs_fp64": force_err,
        "latency_ms_per_batch_step": {"p50": float(np.percentile(per_step_ms, 50)),
                                      "p99": float(np.percentile(per_step_ms, 99)),
                                      "samples": len(per_step_ms),
                                      "source": "HIP events around each timed launch / its batch-steps"},
        "ns_per_day_per_replica": value / a.replicas * 2.0 * 86400 / 1e6,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/k_resident_256chains.md (rocprofv3 FETCH_SIZE x2.00 + WRITE_SIZE x1.00, per launch)",
                     "valu_busy_pct": valu_busy,
                     "kernel": "k_resident<float,LDS>", "kernel_avg_ms": k_avg_ms,
                     "steps_per_launch": launch_steps, "launches": len(kernel_ms),
                     "algorithmic_bytes_per_batch_step": bytes_step},
        "cpu_baseline": cpu,
        "serving": serving,
        "single_chain_fp64": single,
        "multimer": multimer,
        "fm_table": table,
        "frustratometers": frust,
        "setup_s": {"generate": t_gen, "build": t_build},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
