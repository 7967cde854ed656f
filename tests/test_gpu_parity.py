"""GPU parity: the HIP force path against the FP64 oracle, per term.

Force error of a term: max over atoms |F_gpu - F_oracle| divided by
max(max over atoms |F_oracle,term|, 1e-4 * max over atoms |F_oracle,total|);
the floor keeps terms whose forces are physically zero (e.g. water on a
chain with no |i-j|>=10 contacts, ~1e-12 kcal/mol/A) from dividing by
rounding noise.  Energy error: |E_gpu - E_oracle| / max(1, |E_oracle|).
"""
import os

import numpy as np
import pytest

from conftest import make_case

pytestmark = pytest.mark.gpu

TERMS = ["Chain", "Chi", "Rama", "Excluded", "DSSP", "P_AP", "Water", "Burial", "Helix", "Frag_Mem", "Bond"]
# FP64 path: rounding-level agreement; FP32 pair arithmetic with FP64
# accumulation: 1e-5 relative per term
TOL = {0: 1e-9, 1: 1e-5}


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


def term_err(f, fo, f_total_max):
    d = np.linalg.norm((f - fo).reshape(-1, 3), axis=1).max()
    m = np.linalg.norm(fo.reshape(-1, 3), axis=1).max()
    return d / max(m, 1e-4 * f_total_max)


@pytest.mark.parametrize("lengths,seed,compact", [([40, 25], 5, False), ([120], 11, False), ([60, 60], 17, True)])
@pytest.mark.parametrize("prec", [0, 1])
def test_plain_per_term(oracle_mod, lengths, seed, compact, prec):
    A = _gpu()
    s, p = make_case(lengths, seed, compact=compact)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec)
    x = s.coords()
    ftot = np.linalg.norm(o.compute(x)[0].reshape(-1, 3), axis=1).max()
    for name in TERMS + ["ALL"]:
        mask = 0xFFFFFFFF if name == "ALL" else (1 << A.ET[name])
        fix.set_term_mask(mask)
        f, e = fix.compute_res(x)
        fo, eo = o.compute(x, mask)
        k = 0 if name == "ALL" else A.ET[name]
        ferr = term_err(f, fo, ftot)
        eerr = abs(e[k] - eo[k]) / max(1.0, abs(eo[k]))
        print(f"{name:9s} E={e[k]:.9g} Eref={eo[k]:.9g} eerr={eerr:.2e} ferr={ferr:.2e}")
        assert eerr <= TOL[prec], name
        assert ferr <= TOL[prec], name


@pytest.mark.parametrize("engine", ["resident", "wide"])
def test_batch_fp32_per_term(oracle_mod, engine):
    A = _gpu()
    cases = [make_case([L], 100 + k) for k, L in enumerate([60, 37, 90])] + [make_case([30, 44], 7, compact=True)]
    xs = A.pack_states([s.coords() for s, _ in cases])
    oracles = [oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"]) for _, p in cases]
    ftots = [np.linalg.norm(o.compute(s.coords())[0].reshape(-1, 3), axis=1).max() for o, (s, _) in zip(oracles, cases)]
    for name in TERMS + ["ALL"]:
        mask = 0xFFFFFFFF if name == "ALL" else (1 << A.ET[name])
        fixes = [A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP32) for _, p in cases]
        fixes[0].set_term_mask(mask)  # the batch takes the parameters of its first system
        b = A.Batch(fixes, A.PREC_FP32)
        b.set_engine(engine)
        b.set_state(xs)
        b.compute()
        x, v, f, e = b.get_state()
        fs = b.split(f)
        k = 0 if name == "ALL" else A.ET[name]
        for r, (s, p) in enumerate(cases):
            fo, eo = oracles[r].compute(s.coords(), mask)
            ferr = term_err(fs[r], fo, ftots[r])
            eerr = abs(e[r, k] - eo[k]) / max(1.0, abs(eo[k]))
            print(f"replica {r} {name:9s} E={e[r, k]:.9g} Eref={eo[k]:.9g} eerr={eerr:.2e} ferr={ferr:.2e}")
            assert ferr <= TOL[1], (r, name)
            assert eerr <= TOL[1], (r, name)


def ragged_memory_case(seed):
    """Fragment-memory groups of every size from 1 to 29 memories, weights of
    both signs, and 6-residue windows beside the 9-residue ones: the table
    loads (first batch of 16 (FP32) or 8 (FP64) memories, a second batch of
    4, further batches of 4, clamped past the group's end) see full, partial
    and empty batches, and lanes of one wave see different group sizes."""
    import dataclasses
    import tempfile
    from awsemmd_amd import workload as W
    s = W.make_system([64, 30], seed=seed, n_struct=29)
    rng = np.random.default_rng(seed)
    keep = {}
    mems = []
    for m in s.memories:
        n = keep.setdefault(m.pos, int(rng.integers(1, 30)))
        if m.struct < n:
            mems.append(dataclasses.replace(m, weight=float(rng.uniform(-0.5, 2.0))))
    for m in s.memories[::53]:
        if m.pos + 6 <= s.n and m.struct < 5:
            mems.append(dataclasses.replace(m, length=6, weight=0.7))
    s.memories = mems
    d = tempfile.mkdtemp(prefix="awsem_case_")
    return s, W.write_system(s, d)


@pytest.mark.parametrize("engine", ["resident", "wide"])
@pytest.mark.parametrize("prec", [0, 1])
def test_batch_fragment_memory_ragged_groups(oracle_mod, engine, prec):
    A = _gpu()
    cases = [ragged_memory_case(91), ragged_memory_case(92)]
    mask = 1 << A.ET["Frag_Mem"]
    fixes = [A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec) for _, p in cases]
    fixes[0].set_term_mask(mask)
    b = A.Batch(fixes, prec)
    b.set_engine(engine)
    b.set_state(A.pack_states([s.coords() for s, _ in cases]))
    b.compute()
    x, v, f, e = b.get_state()
    fs = b.split(f)
    k = A.ET["Frag_Mem"]
    for r, (s, p) in enumerate(cases):
        fo, eo = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"]).compute(s.coords(), mask)
        ferr = term_err(fs[r], fo, np.linalg.norm(fo.reshape(-1, 3), axis=1).max())
        eerr = abs(e[r, k] - eo[k]) / max(1.0, abs(eo[k]))
        print(f"replica {r}: {len(s.memories)} memories, E={e[r, k]:.9g} Eref={eo[k]:.9g} eerr={eerr:.2e} ferr={ferr:.2e}")
        assert ferr <= TOL[prec] and eerr <= TOL[prec]


@pytest.mark.parametrize("kind,arg", [(0, 0), (0, 1), (1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (6, 0), (6, 1)])
@pytest.mark.parametrize("prec", [0, 1])
def test_lists_bit_exact(oracle_mod, kind, arg, prec):
    A = _gpu()
    for lengths, seed, compact in (([70, 30], 21, False), ([50, 50, 40], 4, True)):
        s, p = make_case(lengths, seed, compact=compact)
        o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
        fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec)
        x = s.coords()
        g = fix.list(kind, x, arg)
        r = o.list(kind, x, arg)
        assert np.array_equal(g, r), (kind, arg, len(g), len(r))


@pytest.mark.parametrize("engine", ["resident", "wide"])
def test_batch_lists_follow_dynamics(oracle_mod, engine):
    """Resident engine after Langevin dynamics: the Verlet pair lists are
    rebuilt when residues drift, and forces on the moved coordinates still
    match the oracle; replacing the state invalidates the lists."""
    A = _gpu()
    cases = [make_case([L], 300 + k) for k, L in enumerate([70, 45])] + [make_case([40, 30], 9, compact=True)]
    x0 = A.pack_states([s.coords() for s, _ in cases])
    oracles = [oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"]) for _, p in cases]
    b = A.Batch([A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP32) for _, p in cases], A.PREC_FP32)
    b.set_engine(engine)
    b.set_state(x0, np.zeros_like(x0))
    b.compute()
    builds0, pairs0, bad0 = b.list_stats()
    assert builds0 == len(cases) and bad0 == 0 and pairs0 > 0
    for rnd in range(2):
        b.run(400, dt=2.0, t_start=600.0, t_stop=600.0, damp=100.0, seed=11, step0=400 * rnd)
        b.compute()
        x, v, f, e = b.get_state()
        xs, fs = b.split(x), b.split(f)
        for r in range(len(cases)):
            fo, eo = oracles[r].compute(xs[r])
            ftot = np.linalg.norm(fo.reshape(-1, 3), axis=1).max()
            ferr = term_err(fs[r], fo, ftot)
            eerr = abs(e[r, 0] - eo[0]) / max(1.0, abs(eo[0]))
            print(f"round {rnd} replica {r} ferr={ferr:.2e} eerr={eerr:.2e} E={e[r, 0]:.6g}")
            assert ferr <= TOL[1] and eerr <= TOL[1], (rnd, r)
    builds, pairs, bad = b.list_stats()
    print("list builds", builds, "pairs", pairs)
    assert bad == 0 and builds >= len(cases) + 2
    # new coordinates: the lists are rebuilt before use
    b.set_state(x0, np.zeros_like(x0))
    b.compute()
    _, _, f, e = b.get_state()
    for r, (s, _) in enumerate(cases):
        fo, eo = oracles[r].compute(s.coords())
        ftot = np.linalg.norm(fo.reshape(-1, 3), axis=1).max()
        assert term_err(b.split(f)[r], fo, ftot) <= TOL[1]
    assert b.list_stats()[0] == builds + len(cases)


@pytest.mark.parametrize("engine", ["resident", "wide"])
def test_batch_fp64_dynamics_matches_oracle(oracle_mod, engine):
    """FP64 resident engine: Langevin dynamics (fix nve + fix langevin, the
    shared counter-based random stream) against the oracle's integrator,
    over two consecutive runs (the second continues from the stored forces)."""
    A = _gpu()
    cases = [make_case([40], 21), make_case([25, 20], 22)]
    rng = np.random.default_rng(4)
    x0 = [s.coords() for s, _ in cases]
    v0 = [rng.normal(0.0, 0.005, size=x.shape) for x in x0]
    b = A.Batch([A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64) for _, p in cases], A.PREC_FP64)
    b.set_engine(engine)
    b.set_state(A.pack_states(x0), A.pack_states(v0))
    b.compute()
    lg = dict(dt=2.0, t_start=300.0, t_stop=300.0, damp=100.0, seed=3)
    b.run(15, step0=0, **lg)
    b.run(15, step0=15, **lg)
    x, v, f, e = b.get_state()
    xs, vs = b.split(x), b.split(v)
    for r, (s, p) in enumerate(cases):
        o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
        fo, _ = o.compute(x0[r])
        xo, vo, fo, eo = o.langevin(x0[r], v0[r], fo, 15, replica=r, step0=0, **lg)
        xo, vo, fo, eo = o.langevin(xo, vo, fo, 15, replica=r, step0=15, **lg)
        dx = np.abs(xs[r] - xo).max()
        dv = np.abs(vs[r] - vo).max() / np.abs(vo).max()
        print(f"replica {r}: max|dx|={dx:.2e} A, rel dv={dv:.2e}")
        assert dx < 1e-8 and dv < 1e-7, r


def test_single_chain_n100_1000_steps_against_oracle(oracle_mod):
    """The CPU-runnable configuration end to end: one N=100 chain with every
    term on (fragment memory: 20 memories per 9-residue window), FP64
    Langevin for 1000 steps at 300 K, dt 2 fs, damp 1000 fs, through the
    plain drop-in (awsem_run), against the oracle integrating the same
    counter-based noise from the same start."""
    A = _gpu()
    s, p = make_case([100], 100)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    x0 = s.coords()
    v0 = np.zeros_like(x0)
    lg = dict(dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0, seed=11)
    x, v, f, e = fix.run(x0, v0, 1000, step0=0, **lg)
    fo, _ = o.compute(x0)
    xo, vo, fo, eo = o.langevin(x0, v0, fo, 1000, replica=0, step0=0, **lg)
    dx = np.abs(x - xo).max()
    dv = np.abs(v - vo).max() / np.abs(vo).max()
    de = np.max(np.abs(e - eo) / np.maximum(1.0, np.abs(eo)))
    print(f"after 1000 steps: max|dx|={dx:.2e} A, rel dv={dv:.2e}, max rel dE={de:.2e}")
    assert dx < 1e-6 and dv < 1e-6 and de < 1e-8


@pytest.mark.parametrize("engine", ["phase", "wide"])
def test_plain_fp64_run_checked_along_trajectory(oracle_mod, engine):
    """The FP64 drop-in (awsem_run on the phase engine, its default, and on
    the wide engine) over 500 Langevin steps of a 300-residue chain; every
    100 steps the GPU energies and forces at the trajectory's coordinates
    against the oracle."""
    A = _gpu()
    s, p = make_case([300], 81)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    x = s.coords()
    v = np.random.default_rng(8).normal(0.0, 0.004, size=x.shape)
    for k in range(5):
        os.environ["AWSEM_RUN_ENGINE"] = engine
        try:
            x, v, f, e = fix.run(x, v, 100, seed=2, step0=100 * k)
        finally:
            del os.environ["AWSEM_RUN_ENGINE"]
        assert np.all(np.isfinite(x))
        fg, eg = fix.compute_res(x)
        fo, eo = o.compute(x)
        ferr = term_err(fg, fo, np.linalg.norm(fo.reshape(-1, 3), axis=1).max())
        eerr = np.max(np.abs(eg - eo) / np.maximum(1.0, np.abs(eo)))
        print(f"step {100 * (k + 1)}: ferr {ferr:.2e} eerr {eerr:.2e}")
        assert ferr <= TOL[0] and eerr <= TOL[0]


def test_wide_engine_cell_list_build(oracle_mod):
    """A 600-residue two-chain complex (above the cell-list threshold) on the
    wide engine: lists built through the cell list, before and after
    Langevin dynamics that forces rebuilds, give the oracle's forces."""
    A = _gpu()
    os.environ["AWSEM_CELL_MIN_N"] = "512"  # force the cell-list build for this 600-residue system
    s, p = make_case([300, 300], 91, compact=True)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    try:
        b = A.Batch([A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP32)], A.PREC_FP32)
    finally:
        del os.environ["AWSEM_CELL_MIN_N"]
    b.set_engine("wide")
    x0 = s.coords()
    b.set_state(x0, np.zeros_like(x0))
    b.compute()
    for rnd in range(3):
        if rnd:
            b.run(150, dt=2.0, t_start=600.0, t_stop=600.0, damp=100.0, seed=5, step0=150 * rnd)
            b.compute()
        x, v, f, e = b.get_state()
        fo, eo = o.compute(x.reshape(3, -1, 3))
        ftot = np.linalg.norm(fo.reshape(-1, 3), axis=1).max()
        ferr = term_err(f.reshape(fo.shape), fo, ftot)
        eerr = np.max(np.abs(e[0] - eo) / np.maximum(1.0, np.abs(eo)))
        builds, pairs, bad = b.list_stats()
        print(f"round {rnd}: ferr {ferr:.2e} eerr {eerr:.2e} builds {builds} pairs {pairs}")
        assert ferr <= TOL[1] and eerr <= TOL[1] and bad == 0
    assert builds >= 2
