"""Known-answer tests of the oracle's tertiary frustratometer
([Tertiary_Frustratometer], fb.cpp:5057-5412, energies 5444-5548) (CPU).

The decoys draw residues with rand() % n from the C library's unseeded
stream (fb.cpp:5550-5556).  The oracle restates glibc's generator; it is
pinned here against the C library itself (ctypes: srand/rand), and every
expected decoy set below is drawn from the C library in the reference's
call order.  The energies are the reference's formulas at small designed
systems (two or three residues on separate chains, so every density and
distance is known): water with the direct well's mean gamma and the
mediated well's protein/water mix (5444-5476), burial (5478-5500), the
mutational environment (5211-5239) and the single-residue sum
(5346-5393); the frustration index is (mean - native) / std with the
population standard deviation (5413-5442, 5591-5598)."""
import ctypes

import numpy as np
import pytest

import kat
from kat import section, unit

EPS = 1.3
KW, KAPPA, KSIG, RHO0, CC = 1.0, 5.0, 7.0, 2.6, 10
WELLS = ((4.5, 6.5), (6.5, 9.5))
KB, BKAPPA, BRHO = 1.0, 4.0, ((0.0, 3.0), (3.0, 6.0), (6.0, 9.0))

LIBC = ctypes.CDLL("libc.so.6")


def _coeff(cut, nd, mode):
    return (section("[Epsilon]", EPS) + kat.abc_section()
            + section("[Water]", KW, (KAPPA, KSIG), RHO0, CC, 2, (WELLS[0][0], WELLS[0][1], 1), (WELLS[1][0], WELLS[1][1], 1))
            + section("[Burial]", KB, BKAPPA, *BRHO)
            + "[Tertiary_Frustratometer]\n%r\n%d\n1\n%s\n\n" % (cut, nd, mode))


def _theta(r, w):
    lo, hi = WELLS[w]
    return 0.25 * (1 + np.tanh(KAPPA * (r - lo))) * (1 + np.tanh(KAPPA * (hi - r)))


class Ref:
    """The reference's frustratometer energies for a designed system."""

    def __init__(self, seq, site, gam, bur):
        self.seq, self.site, self.gam, self.bur = seq, site, gam, bur
        self.n = len(seq)
        self.t = [kat.tidx(c) for c in seq]
        d = self.dist
        self.rho = [sum(_theta(d(i, j), 0) for j in range(self.n)
                        if j != i and WELLS[0][0] - 8 * 2.302585 / KAPPA < d(i, j) < WELLS[0][1] + 8 * 2.302585 / KAPPA)
                    for i in range(self.n)]

    def dist(self, i, j):
        return float(np.linalg.norm(self.site[i] - self.site[j]))

    def water(self, r, ti, tj, ri, rj):
        g = EPS * KW * self.gam
        sw = 0.25 * (1 - np.tanh(KSIG * (ri - RHO0))) * (1 - np.tanh(KSIG * (rj - RHO0)))
        sgd = (g[0, ti, tj, 0] + g[0, ti, tj, 1]) / 2
        sgm = (1 - sw) * g[1, ti, tj, 0] + sw * g[1, ti, tj, 1]
        return -(sgd * _theta(r, 0) + sgm * _theta(r, 1))

    def burial(self, ti, rho):
        return sum(-0.5 * EPS * KB * self.bur[ti, b] * (np.tanh(BKAPPA * (rho - lo)) + np.tanh(BKAPPA * (hi - rho)))
                   for b, (lo, hi) in enumerate(BRHO))


def _tables(rng):
    gam = rng.uniform(-1.0, 1.0, size=(2, 20, 20, 2))
    gam = (gam + gam.transpose(0, 2, 1, 3)) / 2
    return gam, rng.uniform(-1.0, 1.0, size=(20, 3))


def _stats(e):
    e = np.asarray(e)
    return e.mean(), np.sqrt(((e - e.mean()) ** 2).mean())


def test_rand_stream_is_the_c_library(oracle_mod):
    """The oracle's generator equals glibc's rand() after srand(seed),
    unseeded (seed 1) included."""
    from conftest import make_case
    s, p = make_case([10], 1)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    for seed in (1, 7, 2 ** 31 - 2):
        o.rng_seed(seed)
        LIBC.srand(seed)
        assert all(o.rng_next() == LIBC.rand() for _ in range(20000)), seed


def _case(tmp_path, name, seq, site, mode, nd, cut=9.0, rng=None):
    gam, bur = _tables(rng)
    chains = list(seq)  # one residue per chain: every pair is an inter-chain contact
    xca = np.array(site, float)
    xcb = xca.copy()
    p = kat.write_case(str(tmp_path / name), chains, _coeff(cut, nd, mode), gamma=gam, burial=bur)
    # the site is CB (CA for glycine); both placed at the site
    return p, Ref(seq, xca, gam, bur), np.stack([xca, xcb, xca + 20.0])


def test_configurational_known(oracle_mod, tmp_path):
    rng = np.random.default_rng(3)
    seq = "AVL"
    site = [[0, 0, 0], [5.5, 0, 0], [0, 30.0, 0]]
    p, R, x = _case(tmp_path, "conf", seq, site, "configurational", 40, rng=rng)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    rows = o.tert_frust(x)
    assert len(rows) == 1 and tuple(rows[0, :4]) == (1, 2, 1, 2)
    native = R.water(5.5, R.t[0], R.t[1], R.rho[0], R.rho[1]) + R.burial(R.t[0], R.rho[0]) + R.burial(R.t[1], R.rho[1])
    LIBC.srand(1)
    n = 3
    dec = []
    for _ in range(40):
        a, b = LIBC.rand() % n, LIBC.rand() % n
        while R.dist(a, b) > 9.0 or a == b:
            a, b = LIBC.rand() % n, LIBC.rand() % n
        r = R.dist(a, b)
        a2, b2 = LIBC.rand() % n, LIBC.rand() % n
        ri, rj = R.rho[a2], R.rho[b2]
        a3, b3 = LIBC.rand() % n, LIBC.rand() % n
        ta, tb = R.t[a3], R.t[b3]
        dec.append(R.water(r, ta, tb, ri, rj) + R.burial(ta, ri) + R.burial(tb, rj))
    mean, std = _stats(dec)
    print(rows[0], native, mean, std)
    assert abs(rows[0, 7] - native) < 1e-12 and abs(rows[0, 8] - mean) < 1e-12 and abs(rows[0, 9] - std) < 1e-12
    assert abs(rows[0, 10] - (mean - native) / std) < 1e-10
    # the statistics are computed once: a second call draws nothing
    rows2 = o.tert_frust(x)
    assert np.array_equal(rows2[:, 8:10], rows[:, 8:10])
    assert o.rng_next() == LIBC.rand()


def test_mutational_known(oracle_mod, tmp_path):
    """Contact (0, 1) with residue 2 in contact with 0 only: the native and
    every decoy add the environment water(0, 2) with 0's (decoy) type."""
    rng = np.random.default_rng(5)
    seq = "DKW"
    site = [[0, 0, 0], [6.0, 0, 0], [-7.0, 1.0, 0]]
    p, R, x = _case(tmp_path, "mut", seq, site, "mutational", 30, cut=8.0, rng=rng)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    rows = o.tert_frust(x)
    assert [tuple(r[:2]) for r in rows] == [(1, 2), (1, 3)]
    r01, r02 = R.dist(0, 1), R.dist(0, 2)
    assert r02 < 8.0 and R.dist(1, 2) > 8.0
    LIBC.srand(1)
    want = []
    for (i, j), k in (((0, 1), 2), ((0, 2), 1)):
        rij = R.dist(i, j)

        def energy(ti, tj):
            e = R.water(rij, ti, tj, R.rho[i], R.rho[j]) + R.burial(ti, R.rho[i]) + R.burial(tj, R.rho[j])
            for a, ta in ((i, ti), (j, tj)):
                if R.dist(a, k) < 8.0:
                    e += R.water(R.dist(a, k), ta, R.t[k], R.rho[a], R.rho[k])
            return e
        native = energy(R.t[i], R.t[j])
        dec = [energy(R.t[LIBC.rand() % 3], R.t[LIBC.rand() % 3]) for _ in range(30)]
        mean, std = _stats(dec)
        want.append((native, mean, std))
    for row, (native, mean, std) in zip(rows, want):
        print(row, native, mean, std)
        assert abs(row[7] - native) < 1e-12 and abs(row[8] - mean) < 1e-12 and abs(row[9] - std) < 1e-12


def test_singleresidue_known(oracle_mod, tmp_path):
    rng = np.random.default_rng(7)
    seq = "GEFY"
    site = [[0, 0, 0], [6.0, 0, 0], [0, 7.0, 0], [0, 0, 40.0]]
    p, R, x = _case(tmp_path, "single", seq, site, "singleresidue", 25, cut=9.0, rng=rng)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    rows = o.tert_frust(x)
    assert len(rows) == 4 and [int(r[0]) for r in rows] == [1, 2, 3, 4]
    LIBC.srand(1)
    for i, row in enumerate(rows):
        def energy(ti):
            e = R.burial(ti, R.rho[i])
            for j in range(4):
                if j != i and R.dist(i, j) < 9.0:
                    e += R.water(R.dist(i, j), ti, R.t[j], R.rho[i], R.rho[j])
            return e
        native = energy(R.t[i])
        dec = [energy(R.t[LIBC.rand() % 4]) for _ in range(25)]
        mean, std = _stats(dec)
        print(row, native, mean, std)
        assert abs(row[7] - native) < 1e-12 and abs(row[8] - mean) < 1e-12 and abs(row[9] - std) < 1e-12
        assert abs(row[5] - R.rho[i]) < 1e-14
