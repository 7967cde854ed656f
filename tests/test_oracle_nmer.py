"""Known-answer tests of the oracle's n-mer frustratometer
([Nmer_Frustratometer], fb.cpp:5600-6061) (CPU).

Designed systems of one residue per chain on a planar 3 x 3 grid (5 A
spacing: every distance, density and contact known).  The expected values
follow the reference's loops with the energies of the cited formulas
(water 5444-5476, burial 5478-5500) and decoys drawn from the C library's
rand() in the reference's call order: pairwise n-mers (contacts with the
sequence-separation test only, 5895-5919; burials then the water of every
cross pair, 5947-5988; decoy starts swapped into order and redrawn until
both fit and do not overlap, 5990-6060), the trap search (5786-5893: the
second n-mer's types taken from residues o or s - o, not from k, so one
energy per direction and orientation; the overlap test's precedence; the
secondary-structure distance from the ssweight helix and strand columns)
and single n-mers (5659-5769: each member's water with lower residues
only)."""
import numpy as np

import kat
from kat import section
from test_oracle_frust import LIBC, Ref, _stats, _tables, KW, KAPPA, KSIG, RHO0, WELLS, KB, BKAPPA, BRHO, EPS

CC = 2
GRID = np.array([[5.0 * (k % 3), 5.0 * (k // 3), 0.0] for k in range(9)])


def _coeff(size, cut, ncont, nd, thr, trap, mode):
    return (section("[Epsilon]", EPS) + kat.abc_section()
            + section("[Water]", KW, (KAPPA, KSIG), RHO0, CC, 2, (WELLS[0][0], WELLS[0][1], 1), (WELLS[1][0], WELLS[1][1], 1))
            + section("[Burial]", KB, BKAPPA, *BRHO)
            + section("[SSWeight]", (0, 0, 0, 1, 1, 0), (0, 0, 0, 0, 0, 0))
            + "[Nmer_Frustratometer]\n%d\n%r\n%d\n%d\n1\n%r %r 0\n%d 0 %r %r\n%s\n\n"
            % (size, cut, ncont, nd, thr[0], thr[1], trap[0], trap[1], trap[2], mode))


def _case(tmp_path, name, seq, coeff, ss):
    rng = np.random.default_rng(11)
    gam, bur = _tables(rng)
    ssw = np.array([[1.0 if c == "H" else 0.0, 1.0 if c == "E" else 0.0] for c in ss])
    p = kat.write_case(str(tmp_path / name), list(seq), coeff, ssw=ssw, gamma=gam, burial=bur)
    site = GRID[:len(seq)]
    return p, Ref(seq, site, gam, bur), np.stack([site, site, site + 20.0])


def _pair_energy(R, s, a, ta, b, tb):
    e = 0.0
    for o in range(s):
        e += R.burial(ta[o], R.rho[a + o])
    for o in range(s):
        e += R.burial(tb[o], R.rho[b + o])
    for o in range(s):
        for p in range(s):
            e += R.water(R.dist(a + o, b + p), ta[o], tb[p], R.rho[a + o], R.rho[b + p])
    return e


def test_pairwise_and_traps(oracle_mod, tmp_path):
    seq, ss = "ADKWLGVEF", "HH-EE-H-E"
    s, cut, ncont, nd, sigma, frac = 2, 9.0, 1, 40, 0.5, 0.5
    p, R, x = _case(tmp_path, "pair", seq, _coeff(s, cut, ncont, nd, (0.78, -1.0), (1, sigma, frac), "pairwise"), ss)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    rows, traps = o.nmer_frust(x)
    n = len(seq)
    t = R.t
    LIBC.srand(1)
    want_rows, want_traps = [], []
    for i in range(n - s):
        for j in range(i + 1, n - s):
            cnt = sum(1 for a in range(i, i + s) for b in range(j, j + s) if abs(a - b) >= CC and R.dist(a, b) < cut)
            if not (cnt >= ncont and j - i >= s):
                continue
            native = _pair_energy(R, s, i, t[i:i + s], j, t[j:j + s])
            dec = []
            for _ in range(nd):
                ir, jr = sorted((LIBC.rand() % n, LIBC.rand() % n))
                while ir + s > n or jr + s > n or jr - ir < s:
                    ir, jr = sorted((LIBC.rand() % n, LIBC.rand() % n))
                dec.append(_pair_energy(R, s, i, t[ir:ir + s], j, t[jr:jr + s]))
            mean, std = _stats(dec)
            thr = native - sigma * std
            for f, g in ((i, j), (j, i)):
                for bw in (0, 1):
                    tg = [t[(s - q) if bw else q] for q in range(s)]
                    et = _pair_energy(R, s, f, t[f:f + s], g, tg)
                    for k in range(n - s):
                        if abs(k - g) <= s or (abs(k - f) <= s and f != k):
                            continue
                        if sum(ss[g + q] != ss[k + q] for q in range(s)) > s * (1 - frac):
                            continue
                        if et < thr:
                            want_traps.append((len(want_rows), f + 1, g + 1, k + 1, bw, thr, et))
            want_rows.append((i + 1, j + 1, cnt, native, mean, std, (mean - native) / std))
    print(len(want_rows), "rows,", len(want_traps), "traps")
    assert len(want_rows) >= 5 and len(want_traps) >= 1
    assert rows.shape == (len(want_rows), 7) and traps.shape == (len(want_traps), 7)
    assert np.allclose(rows, np.array(want_rows), rtol=1e-12, atol=1e-12)
    assert np.allclose(traps, np.array(want_traps), rtol=1e-12, atol=1e-12)
    # the stream continues where the C library's is
    assert o.rng_next() == LIBC.rand()


def test_pairwise_homopolymer_has_no_decoy_spread(oracle_mod, tmp_path):
    """One residue type throughout: every decoy is the native sequence, so
    each contact's decoy mean equals its native energy and the spread is 0."""
    p, R, x = _case(tmp_path, "homo", "AAAAAAAAA", _coeff(2, 9.0, 1, 25, (0.78, -1.0), (0, 1.0, 0.5), "pairwise"), "-" * 9)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    rows, traps = o.nmer_frust(x)
    assert len(rows) >= 5 and len(traps) == 0
    assert np.allclose(rows[:, 4], rows[:, 3], rtol=1e-13) and np.all(rows[:, 5] < 1e-12)


def test_singlenmer(oracle_mod, tmp_path):
    seq = "RDKWLGVEY"
    s, cut, nd = 3, 9.0, 30
    p, R, x = _case(tmp_path, "single", seq, _coeff(s, cut, 1, nd, (0.78, -1.0), (0, 1.0, 0.5), "singlenmer"), "-" * 9)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    rows, traps = o.nmer_frust(x)
    n = len(seq)

    def single(j, tj):
        w = sum(R.water(R.dist(j, k), tj, R.t[k], R.rho[j], R.rho[k]) for k in range(j) if R.dist(j, k) < cut)
        return w + R.burial(tj, R.rho[j])
    LIBC.srand(1)
    assert rows.shape == (n - s + 1, 7) and len(traps) == 0
    for i, row in enumerate(rows):
        native = sum(single(j, R.t[j]) for j in range(i, i + s))
        dec = []
        for _ in range(nd):
            jr = LIBC.rand() % n
            while jr + s > n:
                jr = LIBC.rand() % n
            dec.append(sum(single(j, R.t[jr + j - i]) for j in range(i, i + s)))
        mean, std = _stats(dec)
        assert tuple(row[:3]) == (i + 1, 0, 0)
        assert abs(row[3] - native) < 1e-12 and abs(row[4] - mean) < 1e-12 and abs(row[5] - std) < 1e-12
