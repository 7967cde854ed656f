"""GPU tests of the batch engine's launch machinery: graph replays against
eager launches, Verlet-skin invariance, and the autotuner's state handling."""
import numpy as np
import pytest

from conftest import make_case

pytestmark = pytest.mark.gpu


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


def _batch(A, cases, prec):
    return A.Batch([A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec) for _, p in cases], prec)


def test_graph_replays_match_eager_launches():
    """The captured resident launch, replayed several times (and recaptured
    when damp changes), gives what eager launches give, with other eager
    work and synchronisations between the replays."""
    A = _gpu()
    cases = [make_case([50], 61), make_case([35, 25], 62), make_case([80], 63)]
    x0 = A.pack_states([s.coords() for s, _ in cases])
    g, e = _batch(A, cases, A.PREC_FP32), _batch(A, cases, A.PREC_FP32)
    e.use_graph(False)
    for b in (g, e):
        b.set_state(x0, np.zeros_like(x0))
    for i in range(6):
        damp = 1000.0 if i < 3 else 500.0
        g.run(10, dt=2.0, t_start=300.0, t_stop=300.0, damp=damp, seed=4, step0=10 * i)
        e.run(10, dt=2.0, t_start=300.0, t_stop=300.0, damp=damp, seed=4, step0=10 * i)
        xg, vg, fg, eg = g.get_state()
        xe, ve, fe, ee = e.get_state()
        d = max(np.abs(xg - xe).max(), np.abs(vg - ve).max())
        print(f"replay {i}: max diff {d:.2e}")
        assert d <= 1e-10 and np.allclose(eg, ee, rtol=1e-10, atol=1e-10)


def test_skin_does_not_change_forces():
    A = _gpu()
    cases = [make_case([120], 64), make_case([60, 60], 65, compact=True)]
    x0 = A.pack_states([s.coords() for s, _ in cases])
    out = []
    for skin in (0.5, 2.0, 4.0):
        b = _batch(A, cases, A.PREC_FP32)
        b.set_skin(skin)
        b.set_state(x0, np.zeros_like(x0))
        b.compute()
        out.append(b.get_state())
    for k in (1, 2):
        df = np.abs(out[k][2] - out[0][2]).max() / np.abs(out[0][2]).max()
        de = np.abs(out[k][3] - out[0][3]).max() / max(1.0, np.abs(out[0][3]).max())
        print(f"skin variant {k}: rel dF {df:.2e}, rel dE {de:.2e}")
        assert df <= 1e-6 and de <= 1e-9


def test_autotune_restores_state_and_keeps_a_candidate_skin():
    A = _gpu()
    cases = [make_case([L], 70 + k) for k, L in enumerate([90, 40, 60])]
    x0 = A.pack_states([s.coords() for s, _ in cases])
    rng = np.random.default_rng(1)
    v0 = rng.normal(0.0, 0.003, size=x0.shape)
    b = _batch(A, cases, A.PREC_FP32)
    b.set_state(x0, v0)
    b.compute()
    before = b.get_state()
    res = b.autotune(probe_steps=10)
    print(res)
    after = b.get_state()
    assert res["skin"] in (1.0, 1.5, 2.0, 2.5, 3.0) and res["steps_per_launch"] in (5, 10, 20, 40, 80, 160)
    assert res["ms_per_step"] > 0
    for k in range(3):
        assert np.array_equal(before[k], after[k])
    # dynamics after tuning match a batch that was never tuned
    ref = _batch(A, cases, A.PREC_FP32)
    ref.set_state(x0, v0)
    ref.compute()
    for bb in (b, ref):
        bb.run(20, seed=3, step0=0)
    d = np.abs(b.get_state()[0] - ref.get_state()[0]).max()
    print(f"tuned vs untuned after 20 steps: {d:.2e} A")
    assert d <= 1e-6


@pytest.mark.parametrize("engine", ["resident", "wide"])
@pytest.mark.parametrize("prec", [0, 1])
def test_energies_after_run_are_the_final_states(engine, prec):
    """A launch sums energies in its last evaluation only: after run(n) the
    reported per-term energies equal a fresh evaluation of the final
    coordinates (FP64 to rounding, FP32 to 1e-6 relative; the stored forces
    also hold the Langevin drag and noise of the last step, so they are not
    compared)."""
    A = _gpu()
    cases = [make_case([L], 80 + k) for k, L in enumerate([70, 45, 110])]
    x0 = A.pack_states([s.coords() for s, _ in cases])
    b = _batch(A, cases, prec)
    b.set_engine(engine)
    b.set_state(x0, np.zeros_like(x0))
    b.run(23, seed=7, step0=0)
    x, v, f, e = b.get_state()
    ref = _batch(A, cases, prec)
    ref.set_engine(engine)
    ref.set_state(x, v)
    ref.compute()
    er = ref.get_state()[3]
    tol = 1e-12 if prec == 0 else 1e-6
    de = np.abs(e - er) / np.maximum(1.0, np.abs(er))
    print(f"{engine} prec {prec}: max rel dE {de.max():.2e}")
    assert np.all(np.isfinite(e)) and de.max() <= tol


@pytest.mark.parametrize("engine", ["resident", "wide"])
def test_split_launches_continue_one_run(engine):
    """Dynamics split over two runs at another step continue as one run: a
    launch starts from the forces the previous one left and re-reads the
    Verlet list's count and validity, so the trajectory is the same (FP64;
    only the order of atomic force adds may differ, at rounding level)."""
    A = _gpu()
    cases = [make_case([L], 90 + k) for k, L in enumerate([60, 85])]
    x0 = A.pack_states([s.coords() for s, _ in cases])
    one, two = _batch(A, cases, A.PREC_FP64), _batch(A, cases, A.PREC_FP64)
    for bb in (one, two):
        bb.set_engine(engine)
        bb.set_state(x0, np.zeros_like(x0))
    one.run(24, seed=5, step0=0)
    two.run(7, seed=5, step0=0)
    two.run(17, seed=5, step0=7)
    xa, va, _, ea = one.get_state()
    xb, vb, _, eb = two.get_state()
    dx, dv = np.abs(xa - xb).max(), np.abs(va - vb).max()
    de = (np.abs(ea - eb) / np.maximum(1.0, np.abs(ea))).max()
    print(f"{engine}: one run vs 7 + 17 steps: |dx| {dx:.2e} A, |dv| {dv:.2e}, rel dE {de:.2e}")
    assert np.all(np.isfinite(xa)) and dx <= 1e-9 and dv <= 1e-9 and de <= 1e-9
