"""GPU parity of fragment-memory table mode ([Fragment_Memory_Table],
fb.cpp:310-320; table build 4238-4323, evaluation 4325-4429 over the pair
path's loop 7513-7527) against the FP64 oracle: the term alone and inside
the full force field, on the single-system path, both batch engines, the
Langevin drop-in path and the server; tables read back from fm_table.*; and
the out-of-range error (the reference aborts: here AWSEM_E_RANGE, and the
next call in range succeeds).

Tolerances as for the other terms: FP64 1e-9 relative (the interpolation is
evaluated as v1 + (v2 - v1) t instead of the reference's two-point form,
equal up to rounding); FP32 1e-5 (curves stored in FP32, distances and the
interpolation weight in FP64 before the FP32 product)."""
import numpy as np
import pytest

from conftest import make_table_case

pytestmark = pytest.mark.gpu

TOL = {0: 1e-9, 1: 1e-5}
E_RANGE = -11  # AWSEM_E_RANGE
CASES = [([40, 25], 51, {}), ([90], 52, {}), ([50, 40], 53, dict(compact=True))]


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


def term_err(f, fo, f_total_max):
    d = np.linalg.norm((f - fo).reshape(-1, 3), axis=1).max()
    m = np.linalg.norm(fo.reshape(-1, 3), axis=1).max()
    return d / max(m, 1e-4 * f_total_max)


def _masks(A):
    return [("Frag_Mem", 1 << A.ET["Frag_Mem"]), ("ALL", 0xFFFFFFFF)]


@pytest.mark.parametrize("prec", [0, 1])
def test_table_plain_per_term(oracle_mod, prec):
    A = _gpu()
    for lengths, seed, kw in CASES:
        s, p = make_table_case(lengths, seed, **kw)
        o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
        fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec)
        x = s.coords()
        ftot = np.linalg.norm(o.compute(x)[0].reshape(-1, 3), axis=1).max()
        for name, mask in _masks(A):
            fix.set_term_mask(mask)
            f, e = fix.compute_res(x)
            fo, eo = o.compute(x, mask)
            k = 0 if name == "ALL" else A.ET[name]
            ferr = term_err(f, fo, ftot)
            eerr = abs(e[k] - eo[k]) / max(1.0, abs(eo[k]))
            print(f"{lengths} {name:8s} E={e[k]:.9g} Eref={eo[k]:.9g} eerr={eerr:.2e} ferr={ferr:.2e}", flush=True)
            assert eo[A.ET["Frag_Mem"]] < -1.0
            assert eerr <= TOL[prec] and ferr <= TOL[prec], (lengths, name)


@pytest.mark.parametrize("engine", ["resident", "wide"])
@pytest.mark.parametrize("prec", [0, 1])
def test_table_batch_per_term(oracle_mod, engine, prec):
    A = _gpu()
    cases = [make_table_case(lengths, seed, **kw) for lengths, seed, kw in CASES]
    oracles = [oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"]) for _, p in cases]
    xs = [s.coords() for s, _ in cases]
    ftots = [np.linalg.norm(o.compute(x)[0].reshape(-1, 3), axis=1).max() for o, x in zip(oracles, xs)]
    for name, mask in _masks(A):
        fixes = [A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec) for _, p in cases]
        fixes[0].set_term_mask(mask)
        b = A.Batch(fixes, prec)
        b.set_engine(engine)
        b.set_state(A.pack_states(xs))
        b.compute()
        _, _, f, e = b.get_state()
        fs = b.split(f)
        k = 0 if name == "ALL" else A.ET[name]
        for r in range(len(cases)):
            fo, eo = oracles[r].compute(xs[r], mask)
            ferr = term_err(fs[r], fo, ftots[r])
            eerr = abs(e[r, k] - eo[k]) / max(1.0, abs(eo[k]))
            print(f"{engine} replica {r} {name:8s} E={e[r, k]:.9g} Eref={eo[k]:.9g} eerr={eerr:.2e} ferr={ferr:.2e}",
                  flush=True)
            assert eerr <= TOL[prec] and ferr <= TOL[prec], (r, name)


def test_table_langevin_fp64(oracle_mod):
    """awsem_run (fix nve + fix langevin) with table mode against the
    oracle's integrator: 30 steps of the full force field."""
    A = _gpu()
    s, p = make_table_case([50], 56)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    rng = np.random.default_rng(6)
    x0 = s.coords()
    v0 = rng.normal(0.0, 0.005, size=x0.shape)
    x, v, f, e = fix.run(x0, v0, 30, dt=2.0, t_start=400.0, t_stop=300.0, damp=100.0, seed=3)
    f0, _ = o.compute(x0)
    xo, vo, fo, eo = o.langevin(x0, v0, f0, 30, dt=2.0, t_start=400.0, t_stop=300.0, damp=100.0, seed=3,
                                replica=0, step0=0)
    dx = np.abs(x - xo).max()
    de = abs(e[A.ET["Frag_Mem"]] - eo[A.ET["Frag_Mem"]]) / max(1.0, abs(eo[A.ET["Frag_Mem"]]))
    print(f"max|dx|={dx:.2e} A, rel dE_fm={de:.2e}", flush=True)
    assert dx < 1e-8 and de < 1e-8


def test_table_read_back(oracle_mod):
    """fm_use_table = 1: the first context builds the tables and writes
    fm_table.energy / fm_table.force; the second reads them (and no memory).
    Forces of the read tables equal the oracle's (which reads the same files)
    and, to the files' 12 decimals, those of the built ones."""
    A = _gpu()
    s, p = make_table_case([40, 25], 57, use_table=1)
    x = s.coords()
    built = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    fb, eb = built.compute_res(x)
    read = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    fr, er = read.compute_res(x)
    fo, eo = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"]).compute(x)
    ftot = np.linalg.norm(fo.reshape(-1, 3), axis=1).max()
    k = A.ET["Frag_Mem"]
    print(f"E built {eb[k]:.12g} read {er[k]:.12g} oracle {eo[k]:.12g}", flush=True)
    assert term_err(fr, fo, ftot) <= 1e-9 and abs(er[k] - eo[k]) <= 1e-9 * max(1.0, abs(eo[k]))
    assert term_err(fr, fb, ftot) <= 1e-8 and abs(er[k] - eb[k]) <= 1e-8 * max(1.0, abs(eb[k]))


def test_table_range_error():
    """A distance beyond the table's rmax (18 A here, while the sequence
    neighbours reach 25.5 A): AWSEM_E_RANGE from the single-system path,
    both batch engines and a server job; the same contexts on coordinates
    scaled into range (0.6: 1.7 to 15.3 A) succeed before and after."""
    A = _gpu()
    s, p = make_table_case([60], 55, tb=(1.0, 18.0, 0.02))
    x = s.coords()
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    f_ok, e_ok = fix.compute_res(0.6 * x)
    with pytest.raises(A.AwsemError) as ex:
        fix.compute_res(x)
    assert ex.value.code == E_RANGE and "range" in str(ex.value)
    # forces are summed with atomics (order-dependent rounding): equal to
    # rounding, and nothing of the failed call is left in them
    f2, e2 = fix.compute_res(0.6 * x)
    d = term_err(f2, f_ok, np.linalg.norm(f_ok.reshape(-1, 3), axis=1).max())
    print(f"after the error: rel force change {d:.2e}, energy {e2[0]:.12g} vs {e_ok[0]:.12g}", flush=True)
    assert d <= 1e-12 and abs(e2[0] - e_ok[0]) <= 1e-12 * max(1.0, abs(e_ok[0]))
    for engine in ("resident", "wide"):
        b = A.Batch([A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP32)], A.PREC_FP32)
        b.set_engine(engine)
        b.set_state(A.pack_states([0.6 * x]))
        b.compute()
        b.set_state(A.pack_states([x]))
        with pytest.raises(A.AwsemError) as ex:
            b.compute()
        assert ex.value.code == E_RANGE, engine
        b.set_state(A.pack_states([0.6 * x]))
        b.compute()
        print(engine, "range error raised and cleared", flush=True)
    fix32 = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP32)
    srv = A.Server(fix32, A.PREC_FP32, max_jobs=4, steps_per_round=5)
    bad = srv.submit(fix32, x, None, 5)
    good = srv.submit(fix32, 0.6 * x, None, 5)
    for _ in range(100):
        srv.pump()
        if srv.poll(bad) not in (0, 1) and srv.poll(good) not in (0, 1):
            break
    print("server states", srv.poll(bad), srv.poll(good), flush=True)
    assert srv.poll(bad) == E_RANGE and srv.poll(good) == 2


@pytest.mark.parametrize("engine", ["resident", "wide"])
@pytest.mark.parametrize("table", [True, False])
def test_batch_repeated_context(oracle_mod, engine, table):
    """One context given three times (replicas of one protein): the later
    replicas read the first one's fragment-memory library or table curves;
    each replica, on its own perturbed coordinates, against the oracle."""
    from conftest import make_case
    A = _gpu()
    s, p = make_table_case([45, 30], 58) if table else make_case([45, 30], 58)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP32)
    rng = np.random.default_rng(8)
    xs = [s.coords() + rng.normal(0.0, 0.05, size=s.coords().shape) for _ in range(3)]
    b = A.Batch([fix, fix, fix], A.PREC_FP32)
    b.set_engine(engine)
    b.set_state(A.pack_states(xs))
    b.compute()
    _, _, f, e = b.get_state()
    fs = b.split(f)
    k = A.ET["Frag_Mem"]
    for r in range(3):
        fo, eo = o.compute(xs[r])
        ferr = term_err(fs[r], fo, np.linalg.norm(fo.reshape(-1, 3), axis=1).max())
        eerr = abs(e[r, 0] - eo[0]) / max(1.0, abs(eo[0]))
        print(f"{engine} table={table} replica {r}: E_fm={e[r, k]:.9g} ref {eo[k]:.9g} eerr={eerr:.2e} ferr={ferr:.2e}",
              flush=True)
        assert abs(e[r, k] - eo[k]) / max(1.0, abs(eo[k])) <= TOL[1]
        assert ferr <= TOL[1] and eerr <= TOL[1], r
