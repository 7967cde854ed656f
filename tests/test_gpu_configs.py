"""GPU parity at the benchmark configurations' own shapes, against the FP64
oracle, per term:
  c1  one N = 1000 chain through the FP64 drop-in (awsem_run: the phase
      engine's step sequence replayed from a captured graph), 1000 Langevin
      steps, every term checked every 100 steps;
  c2  chains of 300-400 residues (the upper end of L ~ U[50, 400]) through
      the resident engine with the replica in LDS, FP32 pair arithmetic; and
      the resident engine's global-slab variant, taken when a batch holds a
      replica too large for LDS;
  c3  one 8 x 300-residue complex (2400 residues) on the wide engine and on
      the resident engine's global-slab variant;
  c4  server jobs of 600 and 1200 residues (above the LDS cap: the wide
      engine's stream) against the oracle's integrator (FP64), and the same
      job alone and inside a mix of jobs (FP32).
Force error of a term: max over atoms |F_gpu - F_oracle| / max over atoms
|F_oracle,term| (floor 1e-9 of the total force for vanishing terms); energy
error |E_gpu - E_oracle| / max(|E_oracle|, 1e-8).  FP64 1e-9, FP32 1e-5."""
import numpy as np
import pytest

from conftest import make_case

pytestmark = pytest.mark.gpu

TERMS = ["Chain", "Chi", "Rama", "Excluded", "DSSP", "P_AP", "Water", "Burial", "Helix", "Frag_Mem", "Bond"]
TOL = {0: 1e-9, 1: 1e-5}


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


def _ferr(f, fo, ftot):
    d = np.linalg.norm((f - fo).reshape(-1, 3), axis=1).max()
    return d / max(np.linalg.norm(fo.reshape(-1, 3), axis=1).max(), 1e-9 * ftot)


def _check(label, prec, f, e, fo, eo, ftot):
    ferr = _ferr(f, fo, ftot)
    eerr = abs(e - eo) / max(abs(eo), 1e-8)
    print(f"{label:28s} E={e:.10g} Eref={eo:.10g} eerr={eerr:.1e} ferr={ferr:.1e}", flush=True)
    assert ferr <= TOL[prec] and eerr <= TOL[prec], (label, ferr, eerr)


def test_c1_single_chain_n1000_fp64_checked_every_100_steps(oracle_mod):
    A = _gpu()
    s, p = make_case([1000], 2024)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    x = s.coords()
    v = np.random.default_rng(3).normal(0.0, 0.003, size=x.shape)
    for k in range(10):
        x, v, _, _ = fix.run(x, v, 100, dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0, seed=5, step0=100 * k)
        assert np.all(np.isfinite(x))
        fo_all, eo_all = o.compute(x)
        ftot = np.linalg.norm(fo_all.reshape(-1, 3), axis=1).max()
        for t in TERMS + ["ALL"]:
            mask = 0xFFFFFFFF if t == "ALL" else (1 << A.ET[t])
            kk = 0 if t == "ALL" else A.ET[t]
            fix.set_term_mask(mask)
            f, e = fix.compute_res(x)
            fo, eo = (fo_all, eo_all) if t == "ALL" else o.compute(x, mask)
            _check(f"step {100 * (k + 1)} {t}", 0, f, e[kk], fo, eo[kk], ftot)
        fix.set_term_mask(0xFFFFFFFF)


def _batch_per_term(A, oracle_mod, cases, prec, engine, terms=TERMS + ["ALL"]):
    oracles = [oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"]) for _, p in cases]
    xs = A.pack_states([s.coords() for s, _ in cases])
    ftots = [np.linalg.norm(o.compute(s.coords())[0].reshape(-1, 3), axis=1).max() for o, (s, _) in zip(oracles, cases)]
    for t in terms:
        mask = 0xFFFFFFFF if t == "ALL" else (1 << A.ET[t])
        kk = 0 if t == "ALL" else A.ET[t]
        fixes = [A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec) for _, p in cases]
        fixes[0].set_term_mask(mask)
        b = A.Batch(fixes, prec)
        b.set_engine(engine)
        b.set_state(xs)
        b.compute()
        _, _, f, e = b.get_state()
        for r, ((s, _), o, fr) in enumerate(zip(cases, oracles, b.split(f))):
            fo, eo = o.compute(s.coords(), mask)
            _check(f"{engine} n={s.n} {t}", prec, fr, e[r, kk], fo, eo[kk], ftots[r])


def test_c2_resident_lds_chains_300_to_400(oracle_mod):
    """Eight chains of 300-400 residues, each with its own sequence, trace
    and 20-memory library; the batch's largest replica fits LDS, so this is
    the resident kernel with the slab in LDS (k_resident<float, true>)."""
    A = _gpu()
    lens = [300, 330, 345, 360, 372, 385, 392, 400]
    assert max(lens) <= 520  # the LDS cap is about 520 residues (39 doubles per residue in 160 KB)
    cases = [make_case([L], 4000 + L) for L in lens]
    _batch_per_term(A, oracle_mod, cases, A.PREC_FP32, "resident")


@pytest.mark.parametrize("prec", [0, 1])
def test_resident_global_slab_variant(oracle_mod, prec):
    """A batch with a replica above the LDS cap: the resident engine keeps
    every replica's slab in global memory (k_resident<T, false>)."""
    A = _gpu()
    cases = [make_case([640], 77), make_case([90], 78), make_case([40, 35], 79, compact=True)]
    _batch_per_term(A, oracle_mod, cases, prec, "resident", terms=["DSSP", "P_AP", "Water", "Frag_Mem", "ALL"])


@pytest.mark.parametrize("engine", ["wide", "resident"])
def test_c3_multimer_8x300(oracle_mod, engine):
    """One 8 x 300-residue compact complex (centroids 10-15 A apart),
    FP32 pair arithmetic, per term on the wide engine and on the resident
    engine (global-slab variant: 2400 residues exceed LDS)."""
    A = _gpu()
    cases = [make_case([300] * 8, 31 * 2024, compact=True)]
    terms = TERMS + ["ALL"] if engine == "wide" else ["DSSP", "P_AP", "Water", "Burial", "Frag_Mem", "ALL"]
    _batch_per_term(A, oracle_mod, cases, A.PREC_FP32, engine, terms=terms)


def _drain(srv, jids, n_max=100000):
    for _ in range(n_max):
        a, q = srv.pump()
        if all(srv.poll(j) == 2 for j in jids):
            return
        assert a + q > 0
    raise AssertionError("jobs did not finish")


def test_c4_server_large_jobs_match_oracle_fp64(oracle_mod):
    """FP64 jobs of 600 and 1200 residues (both above the LDS cap, so the
    server runs them on the wide engine's stream) next to a small LDS job,
    20 annealing steps in rounds of 7: each trajectory equals the oracle's
    integrator on the same counter-based noise."""
    A = _gpu()
    cases = [make_case([600], 601), make_case([700, 500], 1201, compact=True), make_case([60], 61)]
    ff = A.FixBackbone(cases[0][1]["coeff"], cases[0][1]["seq"], cases[0][1]["workdir"], A.PREC_FP64)
    srv = A.Server(ff, A.PREC_FP64, max_jobs=8, steps_per_round=7, dt=2.0, damp=100.0)
    rng = np.random.default_rng(2)
    jobs = []
    for k, (s, p) in enumerate(cases):
        fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
        v0 = rng.normal(0.0, 0.004, size=s.coords().shape)
        jobs.append((srv.submit(fix, s.coords(), v0, 20, t_start=600.0, t_stop=300.0, seed=11 + k), s, p, v0, k))
    _drain(srv, [j for j, *_ in jobs])
    for jid, s, p, v0, k in jobs:
        x, e, t = srv.result(jid, s.n)
        o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
        f0, _ = o.compute(s.coords())
        xo, vo, fo, eo = o.langevin(s.coords(), v0, f0, 20, dt=2.0, t_start=600.0, t_stop=300.0, damp=100.0,
                                    seed=11 + k, replica=0, step0=0)
        dx = np.abs(x - xo).max()
        de = np.max(np.abs(e - eo) / np.maximum(1.0, np.abs(eo)))
        print(f"job n={s.n}: max|dx|={dx:.2e} A, max rel dE={de:.2e}, device {t['device_ms']:.1f} ms", flush=True)
        assert dx < 1e-8 and de < 1e-7


def test_c4_server_large_job_alone_vs_mix():
    """FP32: a 600-residue job (wide stream) alone, and again submitted while
    a 1200-residue job and three small jobs run.  Tolerance after 60
    annealing steps: positions 1e-3 A, per-term energies 1e-4 relative (the
    runs differ only in the order of floating-point force accumulation)."""
    A = _gpu()
    big = make_case([600], 602)
    others = [make_case([700, 500], 1202, compact=True), make_case([80], 81), make_case([45], 82),
              make_case([30, 30], 83, compact=True)]
    fx = lambda c: A.FixBackbone(c[1]["coeff"], c[1]["seq"], c[1]["workdir"], A.PREC_FP32)
    fb = fx(big)
    job = dict(t_start=600.0, t_stop=300.0, seed=21)
    srv1 = A.Server(fb, A.PREC_FP32, max_jobs=8, steps_per_round=20)
    j1 = srv1.submit(fb, big[0].coords(), None, 60, **job)
    _drain(srv1, [j1])
    x1, e1, _ = srv1.result(j1, big[0].n)
    srv2 = A.Server(fb, A.PREC_FP32, max_jobs=8, steps_per_round=20)
    js = [srv2.submit(fx(c), c[0].coords(), None, 60, seed=40 + k) for k, c in enumerate(others[:2])]
    srv2.pump()
    j2 = srv2.submit(fb, big[0].coords(), None, 60, **job)
    js += [srv2.submit(fx(c), c[0].coords(), None, 60, seed=50 + k) for k, c in enumerate(others[2:])]
    _drain(srv2, js + [j2])
    x2, e2, _ = srv2.result(j2, big[0].n)
    dx = np.abs(x1 - x2).max()
    de = (np.abs(e1 - e2) / np.maximum(1.0, np.abs(e1))).max()
    print(f"alone vs mix: max|dx|={dx:.2e} A, max rel dE={de:.2e}, peak active {srv2.stats()['peak_active']}")
    assert dx <= 1e-3 and de <= 1e-4
    assert srv2.stats()["peak_active"] >= 4
