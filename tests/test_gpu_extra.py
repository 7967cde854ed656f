"""GPU parity of the coeff-toggled FixBackbone terms against the FP64 oracle:
AMH-Go (fb.cpp:3787-3992), vector fragment memory (3994-4077), contact
restraints (8137-8177), the membrane (6107-6235), the solvent-separated
barrier (6237-6305), Debye-Hueckel (6307-6363) and phosphorylation (its
water and burial gammas, 643-682, 2872-2900), each term alone and all of
them inside the full force field, on the single-system path, both batch
engines, Langevin dynamics and the server, and the list skin of the
list-limited terms.

These terms are evaluated in FP64 in both precisions; the FP32 tolerance of
the other terms (1e-5) applies to the mixed totals, and FP64 agrees to 1e-9.
The systems are compact multimers with every optional term switched on by
workload.write_extras (native structure = the structure plus 0.5 A noise,
contact list from the structure, charges on K/R/D/E, a membrane slab
through the system, phosphorylated serines)."""
import numpy as np
import pytest

from conftest import make_case

pytestmark = pytest.mark.gpu

TOL = {0: 1e-9, 1: 1e-5}
NAMES = ["AMH-Go", "Vec_FM", "Contact_Restraints", "Membrane", "SSB", "Electro", "Water", "Burial", "ALL"]
CASES = [([40, 30], 21, dict(compact=True)), ([70], 22, dict(compact=True)), ([30, 30, 25], 23, dict(compact=True))]


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


# one membrane for every system (the replicas of a batch share one
# parameter set): a slab through the compact systems' centre
MEMB = (1.0, 0.5, (150.0, 151.0, 148.0), 30.0, 6.0, 0.3, [[0.3, 0.5, 0.7, 1.1], [1.3, 0.7, 0.9, 0.4], [0.2, 0.6, 0.8, 1.2]])


def extra_case(lengths, seed, kw):
    from awsemmd_amd import workload as W
    s, p = make_case(lengths, seed, **kw)
    W.write_extras(s, p["workdir"], seed=seed, membrane=MEMB)
    return s, p


def term_err(f, fo, f_total_max):
    d = np.linalg.norm((f - fo).reshape(-1, 3), axis=1).max()
    m = np.linalg.norm(fo.reshape(-1, 3), axis=1).max()
    return d / max(m, 1e-4 * f_total_max)


def _mask(A, name):
    return 0xFFFFFFFF if name == "ALL" else 1 << A.ET[name]


def _check(A, tag, name, e, f, eo, fo, ftot, prec):
    k = 0 if name == "ALL" else A.ET[name]
    ferr = term_err(f, fo, ftot)
    eerr = abs(e[k] - eo[k]) / max(1.0, abs(eo[k]))
    print(f"{tag} {name:18s} E={e[k]:.10g} Eref={eo[k]:.10g} eerr={eerr:.2e} ferr={ferr:.2e}", flush=True)
    assert eo[k] != 0.0, (tag, name)
    assert eerr <= TOL[prec] and ferr <= TOL[prec], (tag, name)


@pytest.mark.parametrize("prec", [0, 1])
def test_extra_plain_per_term(oracle_mod, prec):
    A = _gpu()
    for lengths, seed, kw in CASES:
        s, p = extra_case(lengths, seed, kw)
        o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
        fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec)
        x = s.coords()
        ftot = np.linalg.norm(o.compute(x)[0].reshape(-1, 3), axis=1).max()
        for name in NAMES:
            m = _mask(A, name)
            fix.set_term_mask(m)
            f, e = fix.compute_res(x)
            fo, eo = o.compute(x, m)
            _check(A, str(lengths), name, e, f, eo, fo, ftot, prec)


@pytest.mark.parametrize("engine", ["resident", "wide"])
@pytest.mark.parametrize("prec", [0, 1])
def test_extra_batch_per_term(oracle_mod, engine, prec):
    A = _gpu()
    cases = [extra_case(lengths, seed, kw) for lengths, seed, kw in CASES]
    oracles = [oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"]) for _, p in cases]
    xs = [s.coords() for s, _ in cases]
    ftots = [np.linalg.norm(o.compute(x)[0].reshape(-1, 3), axis=1).max() for o, x in zip(oracles, xs)]
    for name in NAMES:
        m = _mask(A, name)
        fixes = [A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec) for _, p in cases]
        fixes[0].set_term_mask(m)
        b = A.Batch(fixes, prec)
        b.set_engine(engine)
        b.set_state(A.pack_states(xs))
        b.compute()
        _, _, f, e = b.get_state()
        fs = b.split(f)
        for r in range(len(cases)):
            fo, eo = oracles[r].compute(xs[r], m)
            _check(A, f"{engine} replica {r}", name, e[r], fs[r], eo, fo, ftots[r], prec)


def test_extra_batch_rejects_different_optional_terms():
    """The replicas of a batch share one parameter block: contexts whose
    optional-term parameters differ (here the membrane) are refused."""
    A = _gpu()
    from awsemmd_amd import workload as W
    s1, p1 = extra_case([30], 28, dict(compact=True))
    s2, p2 = make_case([30], 29, compact=True)
    W.write_extras(s2, p2["workdir"], seed=29)
    fixes = [A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64) for p in (p1, p2)]
    with pytest.raises(A.AwsemError) as e:
        A.Batch(fixes, A.PREC_FP64)
    assert e.value.code == -1 and "share one coeff parameter set" in str(e.value)


def test_extra_list_skin(oracle_mod):
    """awsem_set_list_skin: the list-limited terms (Debye-Hueckel, the
    barrier's CA-CA pairs, AMH-Go) count pairs up to pair_list_cutoff +
    skin; with skin 6 the Debye-Hueckel energy changes (pairs between 20 and
    26 A join) and still equals the oracle's with the same skin."""
    A = _gpu()
    s, p = extra_case([60, 50], 24, dict(compact=True))
    x = s.coords()
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    e0 = fix.compute_res(x)[1]
    o.set_list_skin(6.0)
    fix.set_list_skin(6.0)
    ftot = np.linalg.norm(o.compute(x)[0].reshape(-1, 3), axis=1).max()
    for name in ("Electro", "SSB", "AMH-Go", "ALL"):
        m = _mask(A, name)
        fix.set_term_mask(m)
        f, e = fix.compute_res(x)
        fo, eo = o.compute(x, m)
        _check(A, "skin 6", name, e, f, eo, fo, ftot, 0)
    fix.set_term_mask(0xFFFFFFFF)
    e6 = fix.compute_res(x)[1]
    print("DH skin 0 / 6:", e0[A.ET["Electro"]], e6[A.ET["Electro"]])
    assert abs(e6[A.ET["Electro"]] - e0[A.ET["Electro"]]) > 1e-6


def test_extra_langevin_fp64(oracle_mod):
    """awsem_run (the plain drop-in's dynamics) with every optional term on,
    against the oracle's integrator: 30 steps, FP64."""
    A = _gpu()
    s, p = extra_case([50, 30], 25, dict(compact=True))
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    rng = np.random.default_rng(7)
    x0 = s.coords()
    v0 = rng.normal(0.0, 0.005, size=x0.shape)
    lg = dict(dt=2.0, t_start=400.0, t_stop=300.0, damp=100.0, seed=3)
    x, v, f, e = fix.run(x0, v0, 30, **lg)
    f0, _ = o.compute(x0)
    xo, vo, fo, eo = o.langevin(x0, v0, f0, 30, replica=0, step0=0, **lg)
    dx = np.abs(x - xo).max()
    de = abs(e[0] - eo[0]) / max(1.0, abs(eo[0]))
    print(f"max|dx|={dx:.2e} A, rel dE={de:.2e}", flush=True)
    assert dx < 1e-8 and de < 1e-8


@pytest.mark.parametrize("engine", ["resident", "wide"])
def test_extra_batch_dynamics_fp64(oracle_mod, engine):
    """Both batch engines, FP64 Langevin dynamics with every optional term
    on, over two runs (the second continues from stored forces), against
    the oracle's integrator."""
    A = _gpu()
    cases = [extra_case([40, 20], 26, dict(compact=True)), extra_case([50], 27, dict(compact=True))]
    rng = np.random.default_rng(8)
    x0 = [s.coords() for s, _ in cases]
    v0 = [rng.normal(0.0, 0.005, size=x.shape) for x in x0]
    b = A.Batch([A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64) for _, p in cases], A.PREC_FP64)
    b.set_engine(engine)
    b.set_state(A.pack_states(x0), A.pack_states(v0))
    b.compute()
    lg = dict(dt=2.0, t_start=300.0, t_stop=300.0, damp=100.0, seed=5)
    b.run(12, step0=0, **lg)
    b.run(12, step0=12, **lg)
    x, v, f, e = b.get_state()
    xs = b.split(x)
    for r, (s, p) in enumerate(cases):
        o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
        fo, _ = o.compute(x0[r])
        xo, vo, fo, eo = o.langevin(x0[r], v0[r], fo, 12, replica=r, step0=0, **lg)
        xo, vo, fo, eo = o.langevin(xo, vo, fo, 12, replica=r, step0=12, **lg)
        dx = np.abs(xs[r] - xo).max()
        de = abs(e[r, 0] - eo[0]) / max(1.0, abs(eo[0]))
        print(f"{engine} replica {r}: max|dx|={dx:.2e} A, rel dE={de:.2e}", flush=True)
        assert dx < 1e-8 and de < 1e-8, r


def test_extra_server_jobs_fp64(oracle_mod):
    """Server jobs with every optional term on (FP64, 20 annealing steps in
    rounds of 7), one on the LDS stream and one above the LDS cap on the
    wide stream, each against the oracle's integrator."""
    A = _gpu()
    from test_gpu_server import _drain
    cases = [extra_case([40, 30], 31, dict(compact=True)), extra_case([330, 300], 32, dict(compact=True))]
    fix0 = A.FixBackbone(cases[0][1]["coeff"], cases[0][1]["seq"], cases[0][1]["workdir"], A.PREC_FP64)
    srv = A.Server(fix0, A.PREC_FP64, max_jobs=4, steps_per_round=7, dt=2.0, damp=100.0)
    rng = np.random.default_rng(9)
    jobs = []
    for s, p in cases:
        fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
        x0 = s.coords()
        v0 = rng.normal(0.0, 0.005, size=x0.shape)
        jobs.append((srv.submit(fix, x0, v0, 20, t_start=600.0, t_stop=300.0, seed=5), s, p, x0, v0))
    _drain(srv, [j[0] for j in jobs])
    for jid, s, p, x0, v0 in jobs:
        x, e, _ = srv.result(jid, s.n)
        o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
        f0, _ = o.compute(x0)
        xo, vo, fo, eo = o.langevin(x0, v0, f0, 20, dt=2.0, t_start=600.0, t_stop=300.0, damp=100.0, seed=5,
                                    replica=0, step0=0)
        dx = np.abs(x - xo).max()
        de = abs(e[0] - eo[0]) / max(1.0, abs(eo[0]))
        print(f"n={s.n}: max|dx|={dx:.2e} A, rel dE={de:.2e}", flush=True)
        assert dx < 1e-8 and de < 1e-7
