"""GPU runs of LAMMPS decks through the deck front end (awsem_deck_*),
against the FP64 oracle: the trajectory (fix nve + fix langevin with the
temperature ramped over each run, continued across the output intervals),
the fix print energy file, FixBackbone's energy.log (fb.cpp:101-106,
7661-7672: rows at the thermo steps, the fix's own excluded term 0), the
dump frames and the thermo temperature.

The oracle integrates the same steps in the same intervals: one force
evaluation at each run's start (the run's setup), then Langevin steps from
the stored forces, with the ramp's end temperatures of each interval."""
import os

import numpy as np
import pytest

import deckgen

pytestmark = pytest.mark.gpu

KB = 0.0019872067
MVV2E = 48.88821291 ** 2


def _A():
    import awsemmd_amd as A
    A.lib()
    return A


def oracle_run(o, x, v, s0, n, events, T0, T1, damp, seed, dt):
    """{step: (x, v, energies)} at s0 and every event of a run of n steps."""
    f, e = o.compute(x)
    out = {s0: (x, v, e)}
    cur = s0
    for ev in events:
        ts = T0 + (cur - s0) / n * (T1 - T0)
        te = T0 + (ev - s0) / n * (T1 - T0)
        x, v, f, e = o.langevin(x, v, f, ev - cur, dt=dt, t_start=ts, t_stop=te, damp=damp, seed=seed, replica=0,
                                step0=cur)
        out[ev] = (x, v, e)
        cur = ev
    return out


def events_of(s0, n, intervals):
    ev = set([s0 + n])
    for k in intervals:
        ev |= {s for s in range(s0 + 1, s0 + n + 1) if s % k == 0}
    return sorted(ev)


def masses(s):
    m = np.zeros((3, s.n))
    m[0] = 27.0
    m[1] = np.where(np.array(list(s.seq)) == "G", 2.0, 60.0)
    m[2] = 28.0
    return m


def check_energies(row, e, tol):
    """A fix print row (E_chain E_bond E_chi E_rama E_excl E_dssp E_pap
    E_water E_burial E_helix E_fmem E_P) against oracle energies."""
    want = [e[1], e[18], e[3], e[4], e[5], e[6], e[7], e[8], e[9], e[10], e[12]]
    want.append(sum(want))
    for name, g, w in zip(deckgen.TERMS, row, want):
        assert abs(g - w) <= tol * max(1.0, abs(w)), (name, g, w)


def test_deck_fp64_against_oracle(oracle_mod):
    A = _A()
    s, path, d = deckgen.make_deck([40, 30], 51, runs=(60,), thermo=20, dump=30, print_every=20)
    dk = A.Deck(path, A.PREC_FP64)
    x0, v0 = dk.residue_state()
    o = deckgen.deck_oracle(oracle_mod, d)
    ref = oracle_run(o, x0, v0, 0, 60, events_of(0, 60, [20, 30]), 500.0, 300.0, 100.0, 3456, 2.0)
    dk.run()
    x, v = dk.residue_state()
    _, _, _, e, step = dk.state()
    dx = np.abs(x - ref[60][0]).max()
    print(f"step {step}: max |dx| {dx:.2e} A", flush=True)
    assert step == 60 and dx < 1e-8 and np.abs(v - ref[60][1]).max() < 1e-10
    # fix print: title line, rows at 0, 20, 40, 60
    rows = open(os.path.join(d, "energy")).read().splitlines()
    assert rows[0] == "# Fix print output for fix energy"
    assert [int(r.split()[0]) for r in rows[1:]] == [0, 20, 40, 60]
    for r in rows[1:]:
        w = r.split()
        check_energies([float(t) for t in w[1:]], ref[int(w[0])][2], 1e-7)
    # energy.log: FixBackbone's rows at the thermo steps, %8.6f
    log = open(os.path.join(d, "energy.log")).read().splitlines()
    assert log[0].startswith("Step") and [int(r.split()[0]) for r in log[1:]] == [0, 20, 40, 60]
    for r in log[1:]:
        w = r.split()
        e_ref = ref[int(w[0])][2]
        cols = [float(t) for t in w[1:]]
        want = [0.0 if k == 5 else e_ref[k] for k in range(1, 18)]
        assert np.allclose(cols[:17], want, atol=2e-6, rtol=1e-7) and abs(cols[17] - sum(want)) < 1e-5
    # dump frames at 0, 30, 60: scaled coordinates in a box spanning the atoms
    frames = open(os.path.join(d, "dump.lammpstrj")).read().split("ITEM: TIMESTEP\n")[1:]
    assert [int(f.split("\n")[0]) for f in frames] == [0, 30, 60]
    ca, cb, ox = deckgen.atom_ids(s.n)
    for f in frames:
        ln = f.split("\n")
        st = int(ln[0])
        assert ln[3] == "ITEM: BOX BOUNDS ss ss ss" and ln[7] == "ITEM: ATOMS id type xs ys zs"
        box = np.array([[float(t) for t in ln[4 + k].split()] for k in range(3)])
        atoms = np.array([[float(t) for t in r.split()] for r in ln[8:8 + 3 * s.n]])
        assert np.array_equal(atoms[:, 0], np.arange(1, 3 * s.n + 1))
        xa = box[:, 0] + atoms[:, 2:5] * (box[:, 1] - box[:, 0])
        xr = ref[st][0]
        got = np.stack([xa[ca - 1], xa[cb - 1], xa[ox - 1]])
        assert np.abs(got - xr).max() < 2e-4, st
    # thermo: temperature of the final velocities, 3N - 3 degrees of freedom
    m = masses(s)
    T = MVV2E * (m[..., None] * ref[60][1] ** 2).sum() / ((9 * s.n - 3) * KB)
    th = [r.split() for r in open(os.path.join(d, "log.awsem")).read().splitlines()]
    assert th[0] == ["step", "temp", "epair", "emol", "etotal"]
    last = th[-1]
    assert int(last[0]) == 60 and abs(float(last[1]) - T) < 1e-6 * T
    print(f"final T {T:.3f} K, E_P {rows[-1].split()[-1]}", flush=True)


def test_deck_two_runs_and_step_cap(oracle_mod):
    """Two run commands (the second starts from a new setup evaluation and
    ramps again) under max_steps = 25: 50 steps in all."""
    A = _A()
    s, path, d = deckgen.make_deck([35], 52, runs=(1000, 1000), thermo=10, dump=25, print_every=10)
    dk = A.Deck(path, A.PREC_FP64, max_steps=25)
    x0, v0 = dk.residue_state()
    o = deckgen.deck_oracle(oracle_mod, d)
    r1 = oracle_run(o, x0, v0, 0, 25, events_of(0, 25, [10, 25]), 500.0, 300.0, 100.0, 3456, 2.0)
    r2 = oracle_run(o, r1[25][0], r1[25][1], 25, 25, events_of(25, 25, [10, 25]), 500.0, 300.0, 100.0, 3456, 2.0)
    dk.run()
    x, v = dk.residue_state()
    step = dk.state()[4]
    dx = np.abs(x - r2[50][0]).max()
    print(f"step {step}: max |dx| {dx:.2e} A", flush=True)
    assert step == 50 and dx < 1e-8
    # each run's setup evaluation writes a row on a thermo step (25 is none)
    log = open(os.path.join(d, "energy.log")).read().splitlines()[1:]
    assert [int(r.split()[0]) for r in log] == [0, 10, 20, 30, 40, 50]
    rows = open(os.path.join(d, "energy")).read().splitlines()[1:]
    assert [int(r.split()[0]) for r in rows] == [0, 10, 20, 30, 40, 50]
    ref = {**r1, **r2}
    for r in rows:
        w = r.split()
        check_energies([float(t) for t in w[1:]], ref[int(w[0])][2], 1e-7)


def test_deck_fp32_energies(oracle_mod):
    """FP32 pair arithmetic: the energies a deck reports at its final
    coordinates equal the oracle's there to 1e-5 relative per term."""
    A = _A()
    s, path, d = deckgen.make_deck([60], 53, runs=(40,), thermo=20, dump=40, print_every=20)
    dk = A.Deck(path, A.PREC_FP32)
    dk.run()
    x, v = dk.residue_state()
    o = deckgen.deck_oracle(oracle_mod, d)
    _, e_ref = o.compute(x)
    row = [float(t) for t in open(os.path.join(d, "energy")).read().splitlines()[-1].split()[1:]]
    print("E_P", row[-1], flush=True)
    check_energies(row, e_ref, 1e-5)


def test_deck_1r69_shipped_system(oracle_mod):
    """The shipped 1R69 system (fragment memory in table mode on the decks'
    0-50 A grid, a single native memory) with the example's settings
    (timestep 5 fs, Langevin 500 -> 300 K over the run, damping 1000 fs),
    the run capped at 300 steps: trajectory and energy rows against the
    oracle."""
    A = _A()
    path, d = deckgen.deck_1r69(thermo=100, print_every=100, dump=300)
    dk = A.Deck(path, A.PREC_FP64, max_steps=300)
    x0, v0 = dk.residue_state()
    o = deckgen.deck_oracle(oracle_mod, d, seq="extended.seq")
    ref = oracle_run(o, x0, v0, 0, 300, events_of(0, 300, [100, 300]), 500.0, 300.0, 1000.0, 3456, 5.0)
    dk.run()
    x, v = dk.residue_state()
    dx = np.abs(x - ref[300][0]).max()
    rows = open(os.path.join(d, "energy")).read().splitlines()[1:]
    print(f"1R69, 300 steps of 5 fs: max |dx| {dx:.2e} A; E_P {rows[0].split()[-1]} -> {rows[-1].split()[-1]}",
          flush=True)
    assert dx < 1e-7
    assert [int(r.split()[0]) for r in rows] == [0, 100, 200, 300]
    for r in rows:
        w = r.split()
        check_energies([float(t) for t in w[1:]], ref[int(w[0])][2], 1e-6)


def test_deck_frustratometers(oracle_mod):
    """A deck whose coeff file enables the fragment frustratometer (read
    mode, six decoy structures, every 10 steps), the tertiary frustratometer
    (mutational, 80 decoys, every 5 steps) and the n-mer frustratometer
    (single 4-mers, 30 decoys, every 10 steps): the run stops at their output
    steps and writes fragment_frustration.dat lines at 0, 10, 20,
    tertiary_frustration.dat blocks at 0, 5, ..., 20 and
    nmer_frustration.dat blocks at 0, 10, 20 whose rows equal the oracle's
    at the oracle's trajectory (one rand() stream, the tertiary call first
    at a shared step, continued across the blocks) to the printed
    precision."""
    A = _A()
    s, path, d = deckgen.make_deck([40, 30], 54, compact=True, runs=(20,), thermo=10, dump=20, print_every=10)
    with open(os.path.join(d, "decoys.mem"), "w") as f:
        f.write("[Target]\nquery\n\n[Memories]\n" + "".join("mem%03d.gro 1 1 70 1\n" % k for k in range(6)))
    with open(os.path.join(d, "fix_backbone_coeff.data"), "a") as f:
        f.write("\n[Fragment_Frustratometer]\nread\ndecoys.mem\n10\n1.0\n0 0.0\n0\n\n")
        f.write("\n[Tertiary_Frustratometer]\n9.5\n80\n5\nmutational\n\n")
        f.write("\n[Nmer_Frustratometer]\n4\n9.5\n1\n30\n10\n0.78 -1 0\n0 0 1 0.6\nsinglenmer\n\n")
    dk = A.Deck(path, A.PREC_FP64)
    x0, v0 = dk.residue_state()
    o = deckgen.deck_oracle(oracle_mod, d)
    ref = oracle_run(o, x0, v0, 0, 20, events_of(0, 20, [10, 20, 5]), 500.0, 300.0, 100.0, 3456, 2.0)
    dk.run()
    x, _ = dk.residue_state()
    assert np.abs(x - ref[20][0]).max() < 1e-8
    tert, nmer, frag = {}, {}, {}
    for st in (0, 5, 10, 15, 20):
        if st % 10 == 0:
            frag[st] = o.frag_frust(ref[st][0], st)[0]
        tert[st] = o.tert_frust(ref[st][0])
        if st % 10 == 0:
            nmer[st] = o.nmer_frust(ref[st][0])[0]
    fl = open(os.path.join(d, "fragment_frustration.dat")).read().splitlines()
    assert len(fl) == 3
    for ln, st in zip(fl, (0, 10, 20)):
        vals = np.array([float(t) for t in ln.split()])
        assert np.abs(vals - frag[st][:, 0]).max() <= 2e-6, st
    nd = open(os.path.join(d, "nmer_frustration.dat")).read().splitlines()
    nb = [k for k, ln in enumerate(nd) if ln.startswith("# timestep:")]
    assert [nd[k] for k in nb] == ["# timestep: %d" % k for k in (0, 10, 20)]
    for b in nb:
        st = int(nd[b].split()[-1])
        rows = [ln.split() for ln in nd[b + 1:b + 1 + len(nmer[st])]]
        assert len(rows) == len(nmer[st]) > 10
        for r, q in zip(rows, nmer[st]):
            assert int(r[0]) == int(q[0])
            assert max(abs(float(a) - c) for a, c in zip(r[2:6], q[3:7])) <= 2e-6, (st, r, q)
    dat = open(os.path.join(d, "tertiary_frustration.dat")).read().splitlines()
    blocks = [k for k, ln in enumerate(dat) if ln.startswith("# timestep:")]
    assert [dat[k] for k in blocks] == ["# timestep: %d" % k for k in (0, 5, 10, 15, 20)]
    blocks.append(len(dat))
    for b, b1 in zip(blocks[:-1], blocks[1:]):
        st = int(dat[b].split()[-1])
        want = tert[st]
        rows = [ln.split() for ln in dat[b + 1:b1]]
        assert len(rows) == len(want) > 10, st
        for r, q in zip(rows, want):
            assert int(r[0]) == int(q[0]) and int(r[1]) == int(q[1])
            vals = [float(v) for v in r[10:13]] + [float(v) for v in r[15:19]]
            ref_v = [q[4], q[5], q[6], q[7], q[8], q[9], q[10]]
            assert max(abs(a - c) for a, c in zip(vals, ref_v)) <= 0.0015, (st, r, q)
        print(f"step {st}: {len(rows)} contacts, mean index {np.mean(want[:, 10]):.3f}", flush=True)
