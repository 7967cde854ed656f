"""GPU n-mer frustratometer ([Nmer_Frustratometer], fb.cpp:5600-6061)
against the oracle: pairwise with the trap search and single n-mers, with
and without Debye-Hueckel and phosphorylated residues, two calls in a row
(the rand() stream continues), and the nmer_frustration.dat / .tcl and
nmer_traps.dat writer.

Energies are FP64 sums in the reference's order on the device: 1e-10
relative on energies, 1e-8 on the frustration index."""
import os

import numpy as np
import pytest

from conftest import make_case

pytestmark = pytest.mark.gpu


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


def nmer_case(lengths, seed, size, contacts, nd, trap, mode, extras=False, draw=1):
    from awsemmd_amd import workload as W
    s, p = make_case(lengths, seed, compact=True)
    if extras:
        W.write_extras(s, p["workdir"], terms=("dh", "phosph"), seed=seed)
    with open(p["coeff"], "a") as f:
        f.write("\n[Nmer_Frustratometer]\n%d\n9.5\n%d\n%d\n1\n0.78 -1 0\n%d %d 0.5 0.5\n%s\n\n"
                % (size, contacts, nd, trap, draw, mode))
    return s, p


def _rel(a, b):
    return np.abs(a - b) / np.maximum(1.0, np.abs(b))


def _compare(got, want):
    (gr, gt), (wr, wt) = got, want
    assert gr.shape == wr.shape and gt.shape == wt.shape, (gr.shape, wr.shape, gt.shape, wt.shape)
    assert np.array_equal(gr[:, :3], wr[:, :3])
    assert _rel(gr[:, 3:6], wr[:, 3:6]).max() <= 1e-10
    assert _rel(gr[:, 6], wr[:, 6]).max() <= 1e-8
    if len(wt):
        assert np.array_equal(gt[:, :5], wt[:, :5])
        assert _rel(gt[:, 5:], wt[:, 5:]).max() <= 1e-10


@pytest.mark.parametrize("mode,extras", [("pairwise", False), ("pairwise", True), ("singlenmer", False),
                                         ("singlenmer", True)])
def test_nmer_frust_against_oracle(oracle_mod, mode, extras):
    A = _gpu()
    size, contacts = (4, 3) if mode == "pairwise" else (5, 1)
    s, p = nmer_case([40, 30], 81 + extras, size, contacts, 60, int(mode == "pairwise"), mode, extras)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    x = s.coords()
    for call in range(2):
        got = fix.nmer_frust(x)
        want = o.nmer_frust(x)
        print(mode, extras, call, len(got[0]), "rows,", len(got[1]), "traps; native", got[0][:2, 3], flush=True)
        assert len(want[0]) > 10 and (mode != "pairwise" or len(want[1]) > 10)
        _compare(got, want)
        x = x + np.random.default_rng(call).normal(scale=0.05, size=x.shape)


def test_nmer_frust_files(oracle_mod):
    """The writer: headers, '# timestep: k' blocks, rows in the reference's
    %f format against the oracle's values, trap lines (direction arrows,
    sequences read backward for '<--'), the tcl drawing commands."""
    A = _gpu()
    s, p = nmer_case([35, 25], 85, 3, 2, 40, 1, "pairwise")
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    x = s.coords()
    want = []
    for step in (0, 1):
        fix.nmer_frust_write(x, step)
        want.append(o.nmer_frust(x))
    d = p["workdir"]
    dat = open(os.path.join(d, "nmer_frustration.dat")).read().splitlines()
    tcl = open(os.path.join(d, "nmer_frustration.tcl")).read().splitlines()
    trp = open(os.path.join(d, "nmer_traps.dat")).read().splitlines()
    assert dat[0] == "# i j ncontacts a_i a_j native_energy <decoy_energies> std(decoy_energies) f_ij"
    assert trp[0] == "# i a_i ss_i j a_j ss_j threshold_energy k a_k ss_k direction trap_energy"
    blocks = [k for k, ln in enumerate(dat) if ln.startswith("# timestep:")]
    assert [dat[k] for k in blocks] == ["# timestep: 0", "# timestep: 1"]
    seq = s.seq
    for b, (rows, traps) in zip(blocks, want):
        for ln, q in zip(dat[b + 1:b + 1 + len(rows)], rows):
            w = ln.split()
            i, j = int(w[0]), int(w[1])
            assert (i, j, int(w[2])) == (int(q[0]), int(q[1]), int(q[2]))
            assert w[3] == seq[i - 1:i + 2] and w[4] == seq[j - 1:j + 2]
            assert max(abs(float(a) - c) for a, c in zip(w[5:9], q[3:7])) <= 2e-6
    tl = trp[1:]
    allt = np.vstack([t for _, t in want])
    assert len(tl) == len(allt) > 0
    for ln, q in zip(tl, allt):
        w = ln.split()
        f, g, k, bw = int(q[1]), int(q[2]), int(q[3]), int(q[4])
        assert (int(w[0]), int(w[3]), int(w[7])) == (f, g, k)
        assert w[9] == ("<--" if bw else "-->")
        kseq = "".join(seq[k - 1 + 3 - o] for o in range(3)) if bw else seq[k - 1:k + 2]
        assert w[8] == kseq
        assert abs(float(w[6]) - q[5]) <= 2e-6 and abs(float(w[11]) - q[6]) <= 2e-6
    ndraw = sum(1 for rows, _ in want for q in rows if q[6] > 0.78 or q[6] < -1)
    ntrap_lines = sum(1 for q in allt if q[1] != q[3])
    assert sum(1 for ln in tcl if ln == "draw color purple") == ntrap_lines
    assert sum(1 for ln in tcl if ln.startswith("draw line")) == ndraw + ntrap_lines
    assert tcl.count("mol modstyle 0 top newcartoon") == 2


def test_nmer_contact_blocks(oracle_mod, monkeypatch):
    """The contact counts computed a few start rows at a time (the bound on
    the count buffer, here 300 pairs per block) give the same rows, traps
    and rand() stream as the oracle."""
    A = _gpu()
    monkeypatch.setenv("AWSEM_NMER_BLOCK", "300")
    s, p = nmer_case([40, 30], 86, 4, 3, 30, 1, "pairwise")
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    x = s.coords()
    for call in range(2):
        got, want = fix.nmer_frust(x), o.nmer_frust(x)
        assert len(want[0]) > 10
        _compare(got, want)
