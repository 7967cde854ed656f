"""GPU order parameters (compute qwolynes, compute qonuchic) against the
oracle: one trace through awsem_q_compute for every kind, and the replicas
of a batch in place (awsem_batch_q) after some dynamics.  FP64 pair sums in
a different order than the oracle's: 1e-12 relative."""
import numpy as np
import pytest

from conftest import make_case

pytestmark = pytest.mark.gpu


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


def _dist(x):
    return np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)


def _files(tmp_path, native, rc=8.0):
    n = len(native)
    rn = tmp_path / "rnative.dat"
    rn.write_text("\n".join(" ".join("%.10f" % v for v in row) for row in _dist(native)) + "\n")
    nat = tmp_path / "native.dat"
    nat.write_text("%d\n" % n + "\n".join("%.10f %.10f %.10f" % tuple(v) for v in native) + "\n")
    d = _dist(native)
    lines = [(i, j, d[i, j]) for i in range(n) for j in range(i + 4, n) if d[i, j] < rc]
    cm = tmp_path / "map.dat"
    cm.write_text("".join("%d %d %.10f\n" % (i + 1, j + 1, r) for i, j, r in lines))
    # what the oracle reads back from the files
    rnf = np.loadtxt(rn)
    natf = np.loadtxt(nat, skiprows=1)
    return rn, nat, cm, rnf, natf, [(i, j, float("%.10f" % r)) for i, j, r in lines]


def _kinds(A, tmp_path, native):
    n = len(native)
    rn, nat, cm, rnf, natf, lines = _files(tmp_path, native)
    return [("wolynes", A.OrderParameter.wolynes(str(rn), n, 2, 0.15), lambda o, x: o.q_wolynes(x, rnf, 2, 0.15)),
            ("cutoff", A.OrderParameter.onuchic("cutoff", str(nat), n, 8.0, 1.2),
             lambda o, x: o.q_onuchic(x, "cutoff", native_xyz=natf, r_contact=8.0, factor=1.2)),
            ("cutoff/gauss", A.OrderParameter.onuchic("cutoff/gauss", str(nat), n, 8.0),
             lambda o, x: o.q_onuchic(x, "cutoff/gauss", native_xyz=natf, r_contact=8.0)),
            ("shadow", A.OrderParameter.onuchic("shadow", str(cm), n, 0.0, 1.2),
             lambda o, x: o.q_onuchic(x, "shadow", contacts=lines, factor=1.2)),
            ("shadow/gauss", A.OrderParameter.onuchic("shadow/gauss", str(cm), n),
             lambda o, x: o.q_onuchic(x, "shadow/gauss", contacts=lines))]


def test_q_single_trace(oracle_mod, tmp_path):
    A = _gpu()
    s, p = make_case([120], 61, compact=True)
    rng = np.random.default_rng(1)
    native = s.xca + rng.normal(scale=1.0, size=s.xca.shape)
    for name, q, ref in _kinds(A, tmp_path, native):
        got = q.compute(s.xca)
        want = ref(oracle_mod, s.xca)
        print(f"{name:13s} Q={got:.15f} oracle={want:.15f}", flush=True)
        assert 0.0 < want <= 1.0 and abs(got - want) <= 1e-12 * max(1.0, abs(want)), name
        assert abs(q.compute(native) - ref(oracle_mod, native)) <= 1e-12


def test_q_batch_in_place(oracle_mod, tmp_path):
    """Q of every replica of a batch at its device positions after 40
    Langevin steps at 500 K (replicas of one system, different seeds by
    replica index), against the oracle on the positions read back."""
    A = _gpu()
    s, p = make_case([90], 62, compact=True)
    fixes = [A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP32)] * 6
    b = A.Batch(fixes, A.PREC_FP32)
    x0 = s.coords()
    b.set_state(A.pack_states([x0] * 6))
    b.run(40, dt=2.0, t_start=500.0, t_stop=500.0, damp=100.0, seed=4, step0=0)
    x, _, _, _ = b.get_state()
    xs = b.split(x)
    assert np.abs(xs[0] - xs[1]).max() > 1e-3  # replicas diverged
    for name, q, ref in _kinds(A, tmp_path, s.xca):
        got = q.batch(b)
        want = np.array([ref(oracle_mod, xr[0]) for xr in xs])
        print(name, got, want, flush=True)
        assert np.abs(got - want).max() <= 1e-12
        sub = q.batch(b, 2, 3)
        assert np.array_equal(sub, got[2:5])
