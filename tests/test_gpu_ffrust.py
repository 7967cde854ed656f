"""GPU fragment frustratometer ([Fragment_Frustratometer], fb.cpp:4431-4832)
against the oracle: shuffle mode (decoy fragments placed from the rand()
stream after every calculation, two calls in a row) and read mode (decoy
structures, one covering part of the chain; the statistics computed once),
the five output files, and the reference's error for a decoy placed across
two chains.

The per-residue energies are sums of Gaussians added with FP64 atomics
(another order than the oracle's): 1e-10 relative on energies and gaps,
1e-8 on the index."""
import os

import numpy as np
import pytest

from conftest import make_case

pytestmark = pytest.mark.gpu


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


def ff_case(lengths, seed, section, decoys):
    """A synthetic system with fragment memories; decoys: .mem lines
    (structure, target pos, structure pos, length), 0-based."""
    s, p = make_case(lengths, seed, compact=True)
    with open(os.path.join(p["workdir"], "decoys.mem"), "w") as f:
        f.write("[Target]\nquery\n\n[Memories]\n")
        for st, pos, fpos, ln in decoys:
            f.write("mem%03d.gro %d %d %d 1\n" % (st, pos + 1, fpos + 1, ln))
    with open(p["coeff"], "a") as f:
        f.write("\n[Fragment_Frustratometer]\n" + section + "\n\n")
    return s, p


def _close(got, want, tol):
    scale = np.maximum(1.0, np.abs(want))
    return (np.abs(got - want) / scale).max() <= tol


def test_shuffle_against_oracle(oracle_mod):
    A = _gpu()
    n = 80
    decoys = [(k % 6, (7 * k) % (n - 9), (7 * k) % (n - 9), 9) for k in range(30)] + [(2, 0, 10, 14)]
    s, p = ff_case([n], 91, "shuffle\ndecoys.mem\n40\n1", decoys)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    assert fix.frag_frust_info() == (0, 40, len(decoys), 1)
    x = s.coords()
    for call in range(2):
        got, g = fix.frag_frust(x)
        want, _ = o.frag_frust(x, call)
        assert g is None
        print("shuffle call", call, "index", got[:4, 0], "std", got[:4, 2], flush=True)
        assert np.all(want[:, 2] > 0)
        assert _close(got[:, 1:], want[:, 1:], 1e-10) and _close(got[:, 0], want[:, 0], 1e-8)
        x = x + np.random.default_rng(call).normal(scale=0.1, size=x.shape)


def test_read_against_oracle(oracle_mod):
    A = _gpu()
    n = 70
    decoys = [(k, 0, 0, n) for k in range(12)] + [(13, 20, 20, 30)]
    s, p = ff_case([40, 30], 92, "read\ndecoys.mem\n1\n1.3\n1 0.4\n0", decoys)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    assert fix.frag_frust_info() == (1, 14, 13, 1)
    x = s.coords()
    got, g = fix.frag_frust(x)
    want, wg = o.frag_frust(x, 0)
    assert g is not None and wg is not None
    assert _close(g[0], wg[0], 1e-10) and _close(g[1], wg[1], 1e-10)
    assert _close(got[:, 1:], want[:, 1:], 1e-10) and _close(got[:, 0], want[:, 0], 1e-8)
    print("read: native", g[0][:3], "decoy", g[1][0, :3], "index", got[:3, 0], flush=True)
    x2 = x + np.random.default_rng(5).normal(scale=0.1, size=x.shape)
    got2, g2 = fix.frag_frust(x2)
    want2, wg2 = o.frag_frust(x2, 1)
    assert g2 is None and wg2 is None
    assert np.array_equal(got2[:, 2], got[:, 2])
    assert _close(got2[:, 1:], want2[:, 1:], 1e-10) and _close(got2[:, 0], want2[:, 0], 1e-8)


def test_files(oracle_mod):
    A = _gpu()
    n = 50
    decoys = [(k, 0, 0, n) for k in range(5)]
    s, p = ff_case([n], 93, "read\ndecoys.mem\n2\n1.0\n0 0.0\n0", decoys)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    x = s.coords()
    want = []
    for step in range(4):
        fix.frag_frust_write(x, step)  # steps 1 and 3 are not output steps
        if step % 2 == 0:
            want.append(o.frag_frust(x, step))
    d = p["workdir"]
    read = lambda name: open(os.path.join(d, name)).read()
    lines = read("fragment_frustration.dat").splitlines()
    gap = read("fragment_frustration_gap.dat").splitlines()
    var = read("fragment_frustration_variance.dat").splitlines()
    assert len(lines) == len(gap) == len(var) == 2
    for k, (rows, _) in enumerate(want):
        for txt, col in ((lines[k], 0), (gap[k], 1), (var[k], 2)):
            vals = np.array([float(t) for t in txt.split()])
            assert len(vals) == n and np.abs(vals - rows[:, col]).max() <= 1e-6 * max(1.0, np.abs(rows[:, col]).max())
    nat = np.array([float(t) for t in read("fragment_frustration_native.dat").split()])
    dec = np.array([float(t) for t in read("fragment_frustration_decoy.dat").split()])
    g = want[0][1]
    assert np.abs(nat - g[0]).max() <= 1e-6 and np.abs(dec - g[1].ravel()).max() <= 1e-6


def test_decoy_across_chains_is_an_error():
    """Shuffled 15-residue decoys over two 20-residue chains: some placement
    crosses the chain boundary, which the reference rejects
    (fb.cpp:4501)."""
    A = _gpu()
    s, p = ff_case([20, 20], 94, "shuffle\ndecoys.mem\n30\n1", [(0, 0, 0, 15), (1, 20, 20, 15)])
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    with pytest.raises(A.AwsemError) as e:
        fix.frag_frust(s.coords())
    assert "Interaction between residues of different chains" in str(e.value)
