"""Pins the FP64 oracle (CPU, no GPU needed).

What fixes each part of the oracle (the reference has no automated tests
and LAMMPS is not available, so fb.cpp itself cannot be built here):
  * forces: central finite differences of the oracle's own energies, per
    term (a dropped term, wrong sign or index in a force fails this);
  * energies: closed forms on hand-built geometries for the switching
    functions, and the reference's own switching / distance / gamma /
    memory code compiled from src/smart_matrix_lib.h and
    src/fragment_memory.cpp (oracle/_ref) where that library is present;
  * integer lists: an independent numpy brute force of the predicates;
  * invariants: sum of forces and of torques vanish per term.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import make_case

TERMS = ["Chain", "Chi", "Rama", "Excluded", "DSSP", "P_AP", "Water", "Burial", "Helix", "Frag_Mem", "Bond"]
REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref_fm.so")


@pytest.fixture(scope="module")
def cases(oracle_mod):
    out = []
    for lengths, seed, compact in (([34, 21], 5, False), ([46], 9, True), ([24, 24, 20], 13, True)):
        s, p = make_case(lengths, seed, compact=compact)
        out.append((s, p, oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])))
    return out


@pytest.mark.parametrize("name", TERMS)
def test_forces_are_energy_gradients(cases, oracle_mod, name):
    rng = np.random.default_rng(1)
    m = 1 << oracle_mod.ET[name]
    k = oracle_mod.ET[name]
    checked = 0
    for s, p, o in cases:
        x = s.coords()
        f, e = o.compute(x, m)
        fmax = np.abs(f).max()
        if fmax == 0:
            continue
        h = 1e-5
        for _ in range(24):
            a, i, c = rng.integers(3), rng.integers(x.shape[1]), rng.integers(3)
            xp, xm = x.copy(), x.copy()
            xp[a, i, c] += h
            xm[a, i, c] -= h
            fd = -(o.compute(xp, m)[1][k] - o.compute(xm, m)[1][k]) / (2 * h)
            assert abs(fd - f[a, i, c]) <= 1e-6 * max(1.0, fmax) + 1e-7 * abs(e[k]), (name, a, i, c, fd, f[a, i, c])
            checked += 1
    assert checked > 0 or name in ("DSSP", "P_AP")


@pytest.mark.parametrize("name", TERMS)
def test_translation_and_rotation_invariance(cases, oracle_mod, name):
    m = 1 << oracle_mod.ET[name]
    for s, p, o in cases:
        x = s.coords()
        f, e = o.compute(x, m)
        scale = max(1.0, np.abs(f).max())
        assert np.abs(f.sum(axis=(0, 1))).max() <= 1e-9 * scale * x.shape[1]
        torque = np.cross(x.reshape(-1, 3), f.reshape(-1, 3)).sum(axis=0)
        assert np.abs(torque).max() <= 1e-8 * scale * x.shape[1] * 100
        # energy unchanged by a rigid rotation + translation
        th = 0.7
        R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
        xr = x @ R.T + np.array([3.0, -2.0, 1.0])
        e2 = o.compute(xr, m)[1]
        k = oracle_mod.ET[name]
        assert abs(e2[k] - e[k]) <= 1e-9 * max(1.0, abs(e[k]))


def test_total_is_sum_of_terms(cases):
    for s, p, o in cases:
        f, e = o.compute(s.coords())
        assert abs(e[0] - e[1:].sum()) <= 1e-9 * max(1.0, abs(e[0]))


def _site(s):
    x = s.coords()
    g = np.array([c == "G" for c in s.seq])
    return np.where(g[:, None], x[0], x[1])


def test_lists_match_numpy_brute_force(cases):
    """density (water/helix wells), P_AP and excluded-volume lists against an
    independent numpy statement of the predicates (fb.cpp:7760-7775, 8098,
    pev.cpp:104-113)."""
    for s, p, o in cases:
        x = s.coords()
        n = s.n
        ch = s.chain_index()
        site = _site(s)
        i, j = np.triu_indices(n, k=1)
        same = ch[i] == ch[j]
        d2 = ((site[i] - site[j]) ** 2).sum(-1)
        ok = ~same | (j - i > 1)
        for w, (rmin, rmax) in enumerate(((4.5, 6.5), (6.5, 9.5))):
            lo, hi = rmin - 8 * 2.302585 / 5.0, rmax + 8 * 2.302585 / 5.0
            exp = (i * n + j)[ok & (d2 > lo * lo) & (d2 < hi * hi)]
            assert np.array_equal(np.sort(exp), o.list(0, x, w)), w
        lo, hi = 4.5 - 8 * 2.302585 / 7.0, 6.5 + 8 * 2.302585 / 7.0
        assert np.array_equal(np.sort((i * n + j)[ok & (d2 > lo * lo) & (d2 < hi * hi)]), o.list(1, x))
        dca = ((x[0][i] - x[0][j]) ** 2).sum(-1)
        pap = (8.0 + np.arctanh(1 - 2e-12) / 7.0) ** 2
        assert np.array_equal(np.sort((i * n + j)[dca < pap]), o.list(3, x))
        # excluded volume: atoms 3*res+kind, types CA/CB/O, Gly CB = H-beta (lambda 0)
        allx = np.concatenate([x[0], x[1], x[2]])
        res = np.concatenate([np.arange(n)] * 3)
        kind = np.repeat([0, 1, 2], n)
        gly = np.array([c == "G" for c in s.seq])[res]
        typ = np.where(kind == 0, 1, np.where(kind == 2, 3, np.where(gly, 5, 4)))
        a, b = np.triu_indices(3 * n, k=1)
        ra, rb, ka, kb, ta, tb = res[a], res[b], kind[a], kind[b], typ[a], typ[b]
        lam = np.zeros((6, 6))
        for t1, t2 in ((1, 1), (1, 4), (4, 4), (3, 3)):
            lam[t1, t2] = lam[t2, t1] = 20.0
        cl = np.full((6, 6), 3.5)
        for t1, t2 in ((1, 1), (1, 4), (4, 4)):
            cl[t1, t2] = cl[t2, t1] = 4.5
        samec = ch[ra] == ch[rb]
        rc = np.where(samec & (np.abs(ra - rb) < 5), 3.5, cl[ta, tb])
        r = np.sqrt(((allx[a] - allx[b]) ** 2).sum(-1))
        lo_r, hi_r = np.minimum(ra, rb), np.maximum(ra, rb)
        k_lo = np.where(ra <= rb, ka, kb)
        k_hi = np.where(ra <= rb, kb, ka)
        same_res_bond = (ra == rb) & (((ka == 0) & ((kb == 1) | (kb == 2))) | ((kb == 0) & ((ka == 1) | (ka == 2))))
        nb = samec & (hi_r == lo_r + 1) & (k_hi == 0) & ((k_lo == 0) | (k_lo == 2))
        keep = (lam[ta, tb] != 0) & (r < rc) & ~same_res_bond & ~nb
        code_a = 3 * ra + ka
        code_b = 3 * rb + kb
        codes = np.where(code_a < code_b, code_a * 3 * n + code_b, code_b * 3 * n + code_a)[keep]
        assert np.array_equal(np.sort(codes), o.list(5, x))


def test_switching_closed_forms(cases):
    """theta (sml.h:594-610) and nu (sml.h:176-203) at points with known values."""
    s, p, o = cases[0]
    th, prd = o.theta_pair(4.5, well=0)  # r = r_min: 0.25 * 1 * (1 + tanh(kappa*(rmax-rmin)))
    assert abs(th - 0.25 * (1 + np.tanh(5.0 * 2.0))) < 1e-15
    assert abs(prd - 5.0 * th * (np.tanh(10.0) - 0.0) / 4.5) < 1e-15
    th, prd = o.theta_pair(5.5, well=0)  # symmetric point: t_min = t_max, derivative 0
    assert abs(prd) < 1e-15 and abs(th - 0.25 * (1 + np.tanh(5.0)) ** 2) < 1e-15
    assert o.theta_pair(0.5, well=0) == (0.0, 0.0)  # below rmin_theta
    nu, pr = o.pap_nu(8.0 ** 2)  # at P_AP_cut: 0.5, slope 0.5*pref/r
    assert abs(nu - 0.5) < 1e-15 and abs(pr - 0.5 * 7.0 / 8.0) < 1e-15
    assert o.pap_nu(12.0 ** 2) == (0.0, 0.0) and o.pap_nu(4.0 ** 2) == (1.0, 0.0)


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (reference sources absent)")
def test_against_reference_sources(cases, tmp_path):
    """Memories, gamma classes and switching functions against the
    reference's own code (src/fragment_memory.cpp, src/smart_matrix_lib.h)."""
    R = C.CDLL(REF)
    R.ref_mem_create.restype = C.c_void_p
    R.ref_mem_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_char_p]
    R.ref_mem_rf.restype = C.c_double
    R.ref_mem_rf.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    R.ref_mem_se.restype = C.c_char
    R.ref_mem_se.argtypes = [C.c_void_p, C.c_int]
    R.ref_mem_error.argtypes = [C.c_void_p]
    R.ref_mem_destroy.argtypes = [C.c_void_p]
    R.ref_gamma_create.restype = C.c_void_p
    R.ref_gamma_create.argtypes = [C.c_char_p]
    R.ref_gamma_get2.restype = C.c_double
    R.ref_gamma_get2.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    R.ref_gamma_minsep.argtypes = [C.c_void_p]
    R.ref_gamma_maxsep.argtypes = [C.c_void_p]
    R.ref_theta_pair.argtypes = [C.c_double] * 4 + [C.POINTER(C.c_double)] * 4
    R.ref_pap_nu.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double * 3), C.POINTER(C.c_double * 3),
                             C.POINTER(C.c_double), C.POINTER(C.c_double)]
    s, p, o = cases[0]
    # every memory: residue names and all four distance kinds (fm.cpp:160-202)
    mems = s.memories
    for k in range(0, len(mems), 7):
        m = mems[k]
        h = R.ref_mem_create(m.pos, m.fpos, m.length, m.weight, os.path.join(p["workdir"], "mem%03d.gro" % m.struct).encode())
        assert R.ref_mem_error(h) == 0
        pos, ln, w = o.fm_info(k)
        assert (pos, ln, w) == (m.pos, m.length, m.weight)
        for a in range(pos, pos + ln):
            assert R.ref_mem_se(h, a) == o.fm_se(k, a).encode()
            for b in range(pos, pos + ln):
                if a == b:
                    continue
                for ia in (1, 2):
                    for ja in (1, 2):
                        assert R.ref_mem_rf(h, a, ia, b, ja) == o.fm_rf(k, a, ia, b, ja)
        R.ref_mem_destroy(h)
    # gamma-class files: uniform (ALL), 4-class and 20-type tables
    gdir = tmp_path
    # every entry assigned: the reference leaves unassigned entries
    # uninitialised (new double[], fm.cpp:348); the oracle zeroes them
    (gdir / "four.gamma").write_text("# i_min i_max\n3 5 7 inf\nALL ALL 1 0.3\nSHL HPB 1 0.5\nALL ALL 2 1.5\n"
                                     "ALL ALL 3 -0.1\nBAS AHL 3 -0.25\nHPB HPB 1 2.0\n")
    (gdir / "twenty.gamma").write_text("3 9\nALL ALL 1 1.0\nA R 1 0.1\nG G 1 0.7\nW Y 1 -0.3\n")
    import oracle as O
    for gname in ("four.gamma", "twenty.gamma", "uniform.gamma"):
        path = str(gdir / gname) if gname != "uniform.gamma" else os.path.join(p["workdir"], gname)
        txt = open(p["coeff"]).read().replace("uniform.gamma", path)
        cf = gdir / ("coeff_" + gname)
        cf.write_text(txt)
        oo = O.Oracle(str(cf), p["seq"], p["workdir"])
        h = R.ref_gamma_create(path.encode())
        assert R.ref_gamma_minsep(h) == O.lib().or_fm_minsep(oo.h)
        assert R.ref_gamma_maxsep(h) == O.lib().or_fm_maxsep(oo.h)
        for it in range(20):
            for jt in range(20):
                for d in (2, 3, 4, 5, 6, 7, 8, 9, 10, 15):
                    assert R.ref_gamma_get2(h, it, jt, 10, 10 + d) == O.lib().or_fm_gamma(oo.h, it, jt, 10, 10 + d)
    # switching functions through the reference templates
    t, pr, rmin_th, rmax_th = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    for r in np.linspace(0.9, 13.1, 97):
        R.ref_theta_pair(5.0, 4.5, 6.5, r, C.byref(t), C.byref(pr), C.byref(rmin_th), C.byref(rmax_th))
        a, b = o.theta_pair(r, well=0)
        assert abs(t.value - a) <= 1e-15 and abs(pr.value - b) <= 1e-15
        R.ref_theta_pair(7.0, 4.5, 6.5, r, C.byref(t), C.byref(pr), C.byref(rmin_th), C.byref(rmax_th))
        a, b = o.theta_pair(r, helix=True)
        assert abs(t.value - a) <= 1e-15 and abs(pr.value - b) <= 1e-15
    xi = (C.c_double * 3)(0.0, 0.0, 0.0)
    for r in np.linspace(2.0, 12.0, 41):
        xj = (C.c_double * 3)(r, 0.0, 0.0)
        R.ref_pap_nu(8.0, 7.0, C.byref(xi), C.byref(xj), C.byref(t), C.byref(pr))
        a, b = o.pap_nu(r * r)
        assert abs(t.value - a) <= 1e-15 and abs(pr.value - b) <= 1e-15
