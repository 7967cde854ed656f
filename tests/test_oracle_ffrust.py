"""Known-answer tests of the oracle's fragment frustratometer
([Fragment_Frustratometer], fb.cpp:359-384, 794-855, 4431-4832,
7537-7576) (CPU).

One chain whose fragment memories are windows of its own structure, so
every native Gaussian is 1 and the native energy of residue i is the
closed form -eps k_fm sum_mem w gamma(|i-j|) n_live over the pairs that
hold i (compute_decoy_memory_potential adds each Gaussian to both
residues, 4555-4560).  Read mode: decoy structures (one equal to the
native, so its energies equal the natives; one perturbed; one covering
only part of the chain, whose missing distances the reference reads as 0)
with the well width and the sequence-separation scaling
(|i-j| - minSep + 1)^-gamma (4510-4516); the per-residue statistics
divide by the number of decoys (4804-4830), computed once.  Shuffle mode:
the decoy fragments placed at rand() % (n - len + 1) after every
calculation, the calculation before the first placement being the native
(4565-4611, 7539-7554); natives divided by the number of memories, decoys
by the number of decoy fragments, statistics over calculations 1.. (4613-
4689).  The decoy positions are drawn from the C library's rand()."""
import os

import numpy as np

import kat
from kat import section
from test_oracle_frust import LIBC

EPS, KFM = 1.3, 2.5
SEQ = "AGSLKGAVEWYTGDMKLAP"
GAMMA = "3 5 7\nALL ALL 1 1.0\nALL ALL 2 0.4\n"
MEMS = [(0, 0, 0, 9, 1.0), (0, 5, 5, 9, 0.5), (0, 10, 10, 9, 2.0)]


def _gam(d):
    return 1.0 if 3 <= d < 5 else (0.4 if 5 <= d <= 7 else 0.0)


def _grid(rng, n, scale=1.0):
    return np.round(rng.normal(size=(n, 3)) * scale * 100.0) / 100.0


def _structs(rng, n):
    """The native (index 0) and two perturbed structures, CA/CB in nm with
    3 decimals, and the doubles read back from them (A)."""
    xca = np.round((np.cumsum(_grid(rng, n, 2.2), axis=0) + 50.0) * 100.0) / 100.0
    xcb = np.round((xca + _grid(rng, n, 1.0)) * 100.0) / 100.0
    out = []
    for k in range(3):
        a = xca + (_grid(rng, n, 0.6) if k else 0.0)
        b = xcb + (_grid(rng, n, 0.6) if k else 0.0)
        out.append(np.round(np.stack([a, b], axis=1) * 100.0) / 1000.0)
    return out


def _write(d, structs, decoys, ff):
    from awsemmd_amd import workload as W
    os.makedirs(d, exist_ok=True)
    for s, st in enumerate(structs):
        W.write_gro(os.path.join(d, "mem%03d.gro" % s), SEQ, st)
    for name, lst in (("frags.mem", MEMS), ("decoys.mem", decoys)):
        with open(os.path.join(d, name), "w") as f:
            f.write("[Target]\nquery\n\n[Memories]\n")
            for st, pos, fpos, ln, w in lst:
                f.write("mem%03d.gro %d %d %d %g\n" % (st, pos + 1, fpos + 1, ln, w))
    with open(os.path.join(d, "fm.gamma"), "w") as f:
        f.write(GAMMA)
    coeff = (section("[Epsilon]", EPS) + kat.abc_section() + "[Fragment_Memory]\n%r\nfrags.mem\nfm.gamma\n\n" % KFM
             + "[Fragment_Frustratometer]\n" + ff + "\n\n")
    return kat.write_case(d, [SEQ], coeff)


def _atoms(st):
    """CA (0) and CB (1) coordinates in A of a structure."""
    return st[:, 0] * 10.0, st[:, 1] * 10.0


def _energy(n, lib, rfun, ww=1.0, seqsep=None):
    """Per-residue memory energies: lib = [(pos, len, w, struct, fpos)],
    rfun(i, a, j, b) the distance the memory is compared with."""
    E = np.zeros(n)
    pairs = ((0, 0), (0, 1), (1, 0), (1, 1))
    for pos, ln, w, st, fpos in lib:
        mca, mcb = _atoms(st)
        for i in range(pos, pos + ln - 3):
            for j in range(i + 3, min(pos + ln - 1, i + 7) + 1):
                g = _gam(j - i)
                if seqsep is not None:
                    g *= (j - i - 3 + 1) ** (-seqsep)
                s2 = (j - i) ** 0.3 * ww * ww
                mi, mj = fpos + i - pos, fpos + j - pos
                for a, b in pairs:
                    if (a and (SEQ[i] == "G" or SEQ[mi] == "G")) or (b and (SEQ[j] == "G" or SEQ[mj] == "G")):
                        continue
                    rf = np.linalg.norm((mcb if a else mca)[mi] - (mcb if b else mca)[mj])
                    V = -EPS * KFM * w * g * np.exp(-(rfun(i, a, j, b) - rf) ** 2 / (2 * s2))
                    E[i] += V
                    E[j] += V
    return E


def _closed_native(n):
    E = np.zeros(n)
    for pos, _, ln, w in [(m[1], m[2], m[3], m[4]) for m in MEMS]:
        for i in range(pos, pos + ln - 3):
            for j in range(i + 3, min(pos + ln - 1, i + 7) + 1):
                gi, gj = SEQ[i] == "G", SEQ[j] == "G"
                live = 1 + (not gj) + (not gi) + (not gi and not gj)
                E[i] += -EPS * KFM * w * _gam(j - i) * live
                E[j] += -EPS * KFM * w * _gam(j - i) * live
    return E


def test_read_mode(oracle_mod, tmp_path):
    n = len(SEQ)
    structs = _structs(np.random.default_rng(21), n)
    decoys = [(0, 0, 0, n, 1.0), (1, 0, 0, n, 1.0), (2, 4, 4, 10, 1.0)]
    ww, sg = 1.5, 0.5
    p = _write(str(tmp_path), structs, decoys, "read\ndecoys.mem\n1\n%r\n1 %r\n0" % (ww, sg))
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    ca, cb = _atoms(structs[0])
    x = np.stack([ca, cb, ca + 2.0])
    rows, gen = o.frag_frust(x, 0)
    mlib = [(pos, ln, w, structs[0], fpos) for _, pos, fpos, ln, w in MEMS]
    cur = lambda i, a, j, b: np.linalg.norm((cb if a else ca)[i] - (cb if b else ca)[j])
    native = _energy(n, mlib, cur, ww, sg)
    # the memories are the structure itself: every Gaussian 1 (with the
    # scaling the closed form is not needed: check the plain case below)
    dec = []
    for st, pos, fpos, ln, _ in decoys:
        dca, dcb = _atoms(structs[st])

        def rdec(i, a, j, b, dca=dca, dcb=dcb, pos=pos, fpos=fpos, ln=ln):
            if not (pos <= i < pos + ln and pos <= j < pos + ln):
                return 0.0
            return np.linalg.norm((dcb if a else dca)[fpos + i - pos] - (dcb if b else dca)[fpos + j - pos])
        dec.append(_energy(n, mlib, rdec, ww, sg))
    dec = np.array(dec).T  # n x n_decoys
    assert gen is not None
    assert np.allclose(gen[0], native, rtol=1e-12, atol=1e-12)
    assert np.allclose(gen[1], dec, rtol=1e-12, atol=1e-14)
    assert np.allclose(dec[:, 0], native, rtol=1e-12)  # the decoy equal to the native
    mean, std = dec.mean(axis=1), dec.std(axis=1)
    live = std > 0
    assert live.sum() >= n - 2
    assert np.allclose(rows[live, 0], (native - mean)[live] / std[live], rtol=1e-10)
    assert np.allclose(rows[:, 1], native - mean, rtol=1e-12, atol=1e-12)
    assert np.allclose(rows[:, 2], std, rtol=1e-10, atol=1e-14)
    # the statistics are computed once: a second call only recomputes the natives
    x2 = x + np.random.default_rng(3).normal(scale=0.2, size=x.shape)
    rows2, gen2 = o.frag_frust(x2, 1)
    assert gen2 is None and np.allclose(rows2[:, 2], rows[:, 2], rtol=1e-15)


def test_read_mode_plain_closed_form(oracle_mod, tmp_path):
    """Width 1, no scaling: the native energies are the closed form."""
    n = len(SEQ)
    structs = _structs(np.random.default_rng(22), n)
    p = _write(str(tmp_path), structs, [(0, 0, 0, n, 1.0), (1, 0, 0, n, 1.0)], "read\ndecoys.mem\n1\n1.0\n0 0.0\n1")
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    ca, cb = _atoms(structs[0])
    rows, gen = o.frag_frust(np.stack([ca, cb, ca + 2.0]), 0)
    assert np.allclose(gen[0], _closed_native(n), rtol=1e-12, atol=1e-13)
    assert np.allclose(gen[1][:, 0], _closed_native(n), rtol=1e-12, atol=1e-13)


def test_shuffle_mode(oracle_mod, tmp_path):
    n = len(SEQ)
    structs = _structs(np.random.default_rng(23), n)
    decoys = [(1, 2, 2, 6, 1.0), (2, 8, 8, 8, 0.7), (1, 0, 4, 5, 1.3)]
    nc = 12
    p = _write(str(tmp_path), structs, decoys, "shuffle\ndecoys.mem\n%d\n1" % nc)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    ca, cb = _atoms(structs[1])  # current coordinates: a perturbed structure
    x = np.stack([ca, cb, ca + 2.0])
    cur = lambda i, a, j, b: np.linalg.norm((cb if a else ca)[i] - (cb if b else ca)[j])
    mlib = [(pos, ln, w, structs[0], fpos) for _, pos, fpos, ln, w in MEMS]
    LIBC.srand(1)
    for call in range(2):
        rows, gen = o.frag_frust(x, call)
        assert gen is None
        E = [_energy(n, mlib, cur) / len(MEMS)]
        pos = [d[1] for d in decoys]
        for d in range(nc):
            if d:
                lib = [(pos[k], ln, w, structs[st], fpos) for k, (st, _, fpos, ln, w) in enumerate(decoys)]
                E.append(_energy(n, lib, cur) / len(decoys))
            pos = [LIBC.rand() % (n - ln + 1) for (_, _, _, ln, _) in decoys]
        E = np.array(E)
        mean, var = E[1:].mean(axis=0), E[1:].var(axis=0)
        assert np.allclose(rows[:, 1], E[0] - mean, rtol=1e-11, atol=1e-12)
        assert np.allclose(rows[:, 2], np.sqrt(var), rtol=1e-10, atol=1e-13)
        live = var > 0
        assert live.sum() >= n - 3
        assert np.allclose(rows[live, 0], (E[0] - mean)[live] / np.sqrt(var[live]), rtol=1e-9)
    assert o.rng_next() == LIBC.rand()
