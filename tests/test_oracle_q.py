"""Known-answer tests of the oracle's order parameters (CPU): compute
qwolynes (src/compute_q_wolynes.cpp:171-236) and compute qonuchic
(src/compute_q_onuchic.cpp:150-317), and the library's readers of their
input files (no device work).

Closed forms: a trace equal to its native has Q_W = 1 (every Gaussian 1,
and the normalisation 2/((n-sep)(n-sep-1)) is exactly the number of pairs
with |i-j| > sep, so a wrong separation predicate or normalisation shows);
native distances shifted by delta give Q_W = norm sum_d (n-d)
exp(-delta^2 / (2 (1+d)^(2 e))); Q_onuchic's contact counts are strict
(a native pair at exactly its native distance does not count unless
factor > 1); its Gaussian kinds give 1 at the native; shadow maps count
every line in the normalisation but only i < j pairs in the sum."""
import numpy as np
import pytest

from conftest import make_case


def _trace(n, seed):
    from awsemmd_amd import workload as W
    rng = np.random.default_rng(seed)
    return W.saw_trace(n, rng, start=np.array([100.0, 100.0, 100.0]))


def _dist(x):
    return np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)


@pytest.mark.parametrize("sep,e", [(2, 0.15), (3, 0.2), (0, 0.15)])
def test_q_wolynes_known(oracle_mod, sep, e):
    n = 37
    x = _trace(n, 3)
    rn = _dist(x)
    assert abs(oracle_mod.q_wolynes(x, rn, sep, e) - 1.0) < 1e-14
    delta = 0.7
    q = oracle_mod.q_wolynes(x, rn + delta, sep, e)
    want = sum((n - d) * np.exp(-delta ** 2 / (2 * (1 + d) ** (2 * e))) for d in range(sep + 1, n))
    want *= 2.0 / ((n - sep) * (n - sep - 1))
    assert abs(q - want) < 1e-14


def test_q_onuchic_known(oracle_mod):
    n = 30
    x = _trace(n, 5)
    d = _dist(x)
    rc = 9.0
    native_pairs = [(i, j) for i in range(n) for j in range(i + 4, n) if d[i, j] < rc]
    assert native_pairs
    # a contact at exactly its native distance fails the strict test
    assert oracle_mod.q_onuchic(x, "cutoff", native_xyz=x, r_contact=rc, factor=1.0) == 0.0
    assert oracle_mod.q_onuchic(x, "cutoff", native_xyz=x, r_contact=rc, factor=1.0 + 1e-9) == 1.0
    assert abs(oracle_mod.q_onuchic(x, "cutoff/gauss", native_xyz=x, r_contact=rc) - 1.0) < 1e-14
    # a native compressed by 0.9: every current distance is 1/0.9 = 1.11 of
    # its native one, so the contacts count with factor 1.2 and none with 1.05
    assert oracle_mod.q_onuchic(x, "cutoff", native_xyz=0.9 * x, r_contact=rc, factor=1.2) == 1.0
    assert oracle_mod.q_onuchic(x, "cutoff", native_xyz=0.9 * x, r_contact=rc, factor=1.05) == 0.0
    # shadow: three lines, one pair twice (the last distance wins), one reversed
    i0, j0 = native_pairs[0]
    i1, j1 = native_pairs[-1]
    lines = [(i0, j0, 1.0), (i0, j0, d[i0, j0] * 1.01), (j1, i1, d[i1, j1] * 2)]
    q = oracle_mod.q_onuchic(x, "shadow", contacts=lines, factor=1.0)
    assert abs(q - 1.0 / 3.0) < 1e-15
    qg = oracle_mod.q_onuchic(x, "shadow/gauss", contacts=[(i0, j0, d[i0, j0]), (i1, j1, d[i1, j1] + 0.5)])
    s2 = lambda i, j: (1 + abs(i - j)) ** 0.3
    assert abs(qg - (1.0 + np.exp(-0.25 / (2 * s2(i1, j1)))) / 2) < 1e-14


def test_q_files_are_read(tmp_path):
    """The library's readers: the native distance matrix, native
    coordinates and contact maps; the reference's messages for a missing
    file, a count mismatch and a bad map index (no device work)."""
    import awsemmd_amd as A
    n = 12
    x = _trace(n, 7)
    rn = tmp_path / "rnative.dat"
    rn.write_text("\n".join(" ".join("%.6f" % v for v in row) for row in _dist(x)) + "\n")
    q = A.OrderParameter.wolynes(str(rn), n, 2, 0.15)
    q.close()
    with pytest.raises(A.AwsemError) as e:
        A.OrderParameter.wolynes(str(tmp_path / "none.dat"), n)
    assert e.value.code == -2 and "Native distance file can't be read" in str(e.value)
    nat = tmp_path / "native.dat"
    nat.write_text("%d\n" % (n + 1) + "\n".join("%.4f %.4f %.4f" % tuple(v) for v in np.vstack([x, x[:1]])) + "\n")
    with pytest.raises(A.AwsemError) as e:
        A.OrderParameter.onuchic("cutoff", str(nat), n, 8.0, 1.2)
    assert "atom number mismatch" in str(e.value)
    cm = tmp_path / "map.dat"
    cm.write_text("1 3 5.0\n")
    with pytest.raises(A.AwsemError) as e:
        A.OrderParameter.onuchic("shadow", str(cm), n, 0.0, 1.2)
    assert "wrong residue index in shadow contact map file" in str(e.value)
