"""GPU tests of the continuous-batching server (awsem_server_*): a job's
trajectory against the oracle integrator, the same job alone and inside a
mix of jobs that arrive while others run, and admission by device-memory
budget."""
import numpy as np
import pytest

from conftest import make_case

pytestmark = pytest.mark.gpu


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


def _drain(srv, jids, n_max=10000):
    for _ in range(n_max):
        a, q = srv.pump()
        if all(srv.poll(j) == 2 for j in jids):
            return
        assert a + q > 0
    raise AssertionError("jobs did not finish")


def test_server_job_matches_oracle_fp64(oracle_mod):
    """FP64 job of 20 annealing steps (600 -> 300 K) run in rounds of 7: the
    ramp and random stream continue across rounds exactly as one run."""
    A = _gpu()
    s, p = make_case([40], 31)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    srv = A.Server(fix, A.PREC_FP64, max_jobs=4, steps_per_round=7, dt=2.0, damp=100.0)
    rng = np.random.default_rng(2)
    x0 = s.coords()
    v0 = rng.normal(0.0, 0.005, size=x0.shape)
    jid = srv.submit(fix, x0, v0, 20, t_start=600.0, t_stop=300.0, seed=5)
    _drain(srv, [jid])
    x, e, t = srv.result(jid, s.n)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    f0, _ = o.compute(x0)
    xo, vo, fo, eo = o.langevin(x0, v0, f0, 20, dt=2.0, t_start=600.0, t_stop=300.0, damp=100.0, seed=5,
                                replica=0, step0=0)
    dx = np.abs(x - xo).max()
    de = abs(e[0] - eo[0]) / max(1.0, abs(eo[0]))
    print(f"max|dx|={dx:.2e} A, rel dE={de:.2e}, times {t}")
    assert dx < 1e-8 and de < 1e-7
    assert srv.stats()["rounds"] == 3


def test_server_alone_vs_mix():
    """FP32: job A alone, then A submitted while B and C run and next to D.
    Tolerance: positions 1e-3 A and per-term energies 1e-4 relative after 100
    steps (the jobs differ only in the order of floating-point accumulation)."""
    A = _gpu()
    cases = {"A": make_case([70], 41), "B": make_case([45], 42), "C": make_case([40, 30], 43, compact=True),
             "D": make_case([90], 44)}
    fixes = {k: A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP32) for k, (s, p) in cases.items()}
    job = dict(nsteps=100, t_start=600.0, t_stop=300.0)
    sA = cases["A"][0]

    srv1 = A.Server(fixes["A"], A.PREC_FP32, max_jobs=8, steps_per_round=25)
    ja = srv1.submit(fixes["A"], sA.coords(), None, seed=9, **job)
    _drain(srv1, [ja])
    xa1, ea1, _ = srv1.result(ja, sA.n)

    srv2 = A.Server(fixes["A"], A.PREC_FP32, max_jobs=8, steps_per_round=25)
    jb = srv2.submit(fixes["B"], cases["B"][0].coords(), None, seed=1, **job)
    jc = srv2.submit(fixes["C"], cases["C"][0].coords(), None, seed=2, **job)
    srv2.pump()
    assert srv2.poll(jb) == 1 and srv2.poll(jc) == 1
    ja2 = srv2.submit(fixes["A"], sA.coords(), None, seed=9, **job)
    jd = srv2.submit(fixes["D"], cases["D"][0].coords(), None, seed=3, **job)
    _drain(srv2, [jb, jc, ja2, jd])
    xa2, ea2, _ = srv2.result(ja2, sA.n)
    for j, k in ((jb, "B"), (jc, "C"), (jd, "D")):
        _, e, _ = srv2.result(j, cases[k][0].n)
        assert np.all(np.isfinite(e))
    assert srv2.stats()["peak_active"] == 4

    # the same job through the batch API in one launch (same stream: replica 0)
    b = A.Batch([fixes["A"]], A.PREC_FP32)
    b.set_state(sA.coords(), np.zeros_like(sA.coords()))
    b.run(100, dt=2.0, t_start=600.0, t_stop=300.0, damp=1000.0, seed=9, step0=0)
    xb, _, _, eb = b.get_state()

    dx = np.abs(xa1 - xa2).max()
    dxb = np.abs(xa1.reshape(-1) - xb).max()
    de = np.abs(ea1 - ea2) / np.maximum(1.0, np.abs(ea1))
    deb = np.abs(ea1 - eb[0]) / np.maximum(1.0, np.abs(ea1))
    print(f"alone vs mix: max|dx|={dx:.2e} A, max rel dE={de.max():.2e}; alone vs batch: {dxb:.2e} A, {deb.max():.2e}")
    assert dx <= 1e-3 and de.max() <= 1e-4
    assert dxb <= 1e-3 and deb.max() <= 1e-4


def test_server_admission_by_memory_budget():
    """A budget that holds one job at a time: the jobs run one after another
    (background scheduler), and a job larger than the budget is refused."""
    A = _gpu()
    cases = [make_case([60], 50 + k) for k in range(3)]  # equal images: one fits the budget at a time
    fixes = [A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP32) for _, p in cases]
    probe = A.Server(fixes[0], A.PREC_FP32, max_jobs=4, steps_per_round=10)
    j = probe.submit(fixes[0], cases[0][0].coords(), None, 20)
    probe.pump()  # admitted, 10 of 20 steps done: the image is still resident
    big = probe.stats()["bytes_in_use"]
    assert big > 0
    _drain(probe, [j])
    assert probe.stats()["bytes_in_use"] == 0

    srv = A.Server(fixes[0], A.PREC_FP32, max_jobs=4, steps_per_round=10, hbm_budget_bytes=1.5 * big)
    srv.start()
    jids = [srv.submit(f, s.coords(), None, 30, seed=k) for k, (f, (s, _)) in enumerate(zip(fixes, cases))]
    for jid in jids:
        srv.wait(jid, 60.0)
    srv.stop()
    st = srv.stats()
    print(st)
    assert st["jobs_done"] == 3 and st["peak_active"] == 1 and st["bytes_in_use"] == 0

    tiny = A.Server(fixes[0], A.PREC_FP32, max_jobs=4, steps_per_round=10, hbm_budget_bytes=1024.0)
    with pytest.raises(A.AwsemError) as e:
        tiny.submit(fixes[0], cases[0][0].coords(), None, 10)
    assert e.value.code == -8


def test_server_failed_round_is_reported(monkeypatch):
    """A scheduling round that fails (injected from the host before it
    launches anything, AWSEM_SERVER_FAIL_ROUND) fails the jobs it held; the
    worker stops, and submit, pump, start and wait (with and without a time
    limit) return the failure instead of queueing or blocking."""
    A = _gpu()
    s, p = make_case([30], 35)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP32)
    monkeypatch.setenv("AWSEM_SERVER_FAIL_ROUND", "1")
    srv = A.Server(fix, A.PREC_FP32, max_jobs=4, steps_per_round=5)
    pumped = A.Server(fix, A.PREC_FP32, max_jobs=4, steps_per_round=5)
    monkeypatch.delenv("AWSEM_SERVER_FAIL_ROUND")
    jid = srv.submit(fix, s.coords(), None, 10)
    srv.start()
    with pytest.raises(A.AwsemError) as e:
        srv.wait(jid, 30.0)
    assert e.value.code == -7 and "injected" in str(e.value)
    assert srv.poll(jid) == -7
    with pytest.raises(A.AwsemError) as e:
        srv.wait(jid, 0.0)  # no time limit: returns at once
    assert e.value.code == -7
    for call in (lambda: srv.submit(fix, s.coords(), None, 10), srv.pump, srv.start):
        with pytest.raises(A.AwsemError) as e:
            call()
        assert e.value.code == -7
    srv.stop()
    srv.close()
    # the same on the calling thread: pump reports the failure, then refuses
    j2 = pumped.submit(fix, s.coords(), None, 10)
    with pytest.raises(A.AwsemError) as e:
        pumped.pump()
    assert e.value.code == -7 and pumped.poll(j2) == -7
    with pytest.raises(A.AwsemError):
        pumped.submit(fix, s.coords(), None, 10)
    with pytest.raises(A.AwsemError):
        pumped.wait(j2, 0.0)
    pumped.close()
