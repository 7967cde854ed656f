"""Known-answer tests of the oracle's fragment-memory table mode (CPU).

[Fragment_Memory_Table] (fb.cpp:310-320) replaces the per-memory
Gaussians by one curve per (residue i, residue j, atom pair k), summed over
the memories at the grid points r = rmin + ir dr (compute_fragment_memory_
table, fb.cpp:4238-4323) and interpolated linearly at the current distance
(table_fragment_memory, fb.cpp:4325-4429), over the pairs of the pair
path's loop (fb.cpp:7513-7527: j from i + minSep to ch_pos + ch_len - 2,
which with the 1-based chain starts of fb.cpp:546 is the last residue of
i's chain, at most i + maxSep).

The geometries put every CA and CB on one axis at multiples of 0.01 A, so
every atom-pair distance is a grid point (dr = 0.01) or, where a test
shifts an atom by half a step, the exact midpoint of two.  The expected
values are the reference's definitions at those points: with a memory
equal to the structure every curve value at the current distance is
-eps k_fm w gamma (Gaussian 1), otherwise the Gaussian of the displaced
distance with sigma^2 = |i-j|^(2 fm_sigma_exp) well_width^2, and at a
midpoint the mean of the two grid values (energy and force curve alike).
"""
import os

import numpy as np
import pytest

import kat
from kat import section

EPS = 1.3
K_FM = 2.5


def _o(oracle_mod, p):
    return oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])


def _table_case(d, chains, st, mems, gamma_text, tb=(0.0, 50.0, 0.01), ww=0.1, sigma_exp=0.15, use_table=0):
    from awsemmd_amd import workload as W
    seq = "".join(chains)
    os.makedirs(d, exist_ok=True)
    W.write_gro(os.path.join(d, "mem000.gro"), seq, st)
    with open(os.path.join(d, "frags.mem"), "w") as f:
        f.write("[Target]\nquery\n\n[Memories]\n")
        for pos, fpos, ln, w in mems:
            f.write("mem000.gro %d %d %d %g\n" % (pos + 1, fpos + 1, ln, w))
    with open(os.path.join(d, "fm.gamma"), "w") as f:
        f.write(gamma_text)
    coeff = (section("[Epsilon]", EPS) + kat.abc_section()
             + "[Fragment_Memory_Table]\n%r\nfrags.mem\nfm.gamma\n%r %r %r\n%r\n%d\n%r\n\n"
             % (K_FM, tb[0], tb[1], tb[2], ww, use_table, sigma_exp))
    return kat.write_case(d, chains, coeff)


def _axis_structure(seq, spacing=3.8, cb=1.53, seed=0):
    """CA of residue i at x = spacing i (+ a per-residue jitter on the 0.01
    grid), CB at CA + cb, both on the x axis; returned as the .gro nm array
    (3 decimals) and the doubles the reference reads back (x 10)."""
    n = len(seq)
    rng = np.random.default_rng(seed)
    x = np.round((np.arange(n) * spacing + rng.integers(-20, 21, n) * 0.01) * 100.0) / 100.0
    xca = np.stack([x + 50.0, np.full(n, 20.0), np.full(n, 30.0)], axis=1)
    xcb = xca + np.array([cb, 0.0, 0.0])
    st = np.round(np.stack([xca, xcb], axis=1) * 100.0) / 1000.0
    return st, st[:, 0] * 10.0, st[:, 1] * 10.0


def _coords(xca, xcb):
    return np.stack([xca, xcb, xca + np.array([0.0, 2.0, 0.0])]).astype(np.float64)


def _pairs(chains, minsep, maxsep):
    """(i, j) of the pair path's loop, fb.cpp:7513-7527 (0-based residues;
    ch_pos is 1-based, so ch_pos + ch_len - 2 is the chain's last residue)."""
    out = []
    start = 0
    for ch in chains:
        last = start + len(ch) - 1
        for i in range(start, last + 1):
            jrn = (start + 1) + len(ch) - 2
            if maxsep != -1:
                jrn = min(jrn, i + maxsep)
            for j in range(i + minsep, jrn + 1):
                out.append((i, j))
        start += len(ch)
    return out


def test_table_dims(oracle_mod, tmp_path):
    """tb_size = (int)((rmax - rmin) / dr) + 2 (fb.cpp:317); tb_nbrs = maxSep -
    minSep + 1 (fb.cpp:998-1002)."""
    chains = ["ALKVSTWDEF"]
    st, xca, xcb = _axis_structure(chains[0])
    p = _table_case(str(tmp_path), chains, st, [(0, 0, 10, 1.0)], "2 4 7\nALL ALL 1 1.0\nALL ALL 2 0.4\n",
                    tb=(1.0, 21.0, 0.05))
    assert _o(oracle_mod, p).fm_table_dims() == (int((21.0 - 1.0) / 0.05) + 2, 7 - 2 + 1)


def test_table_memory_equal_to_structure(oracle_mod, tmp_path):
    """Every memory equal to the structure, every distance a grid point: each
    curve reads -eps k_fm sum_mem w gamma at the current distance.  Counted
    per pair of the pair loop, atom pairs without a target glycine CB; a
    memory whose residue j is a glycine leaves the CA(i)-CB(j) curve missing
    when no other memory fills it, and a missing curve ends the pair (the
    reference returns, fb.cpp:4394), so CB(i)-CA(j) and CB-CB of that pair
    count 0 too."""
    chains = ["ALKVSTWDEFMLKA", "VKLSWE"]
    seq = "".join(chains)
    st, xca, xcb = _axis_structure(seq, seed=3)
    # memory residue 7 (target W) is a glycine in the memory structure
    mseq = seq[:7] + "G" + seq[8:]
    from awsemmd_amd import workload as W
    gamma = "2 4 7\nALL ALL 1 1.0\nALL ALL 2 0.4\n"
    mems = [(0, 0, 10, 1.0), (3, 3, 11, 0.5), (14, 14, 6, 2.0)]
    p = _table_case(str(tmp_path), chains, st, mems, gamma)
    W.write_gro(os.path.join(p["workdir"], "mem000.gro"), mseq, st)
    o = _o(oracle_mod, p)
    e = o.energy(_coords(xca, xcb), 1 << oracle_mod.ET["Frag_Mem"])[oracle_mod.ET["Frag_Mem"]]
    gam = lambda d: 1.0 if 2 <= d < 4 else (0.4 if 4 <= d <= 7 else 0.0)
    minsep, maxsep = 2, 7
    # curve contents per (i, j, k): sum of w gamma over memories covering (i, j)
    curve = {}
    for pos, _, ln, w in mems:
        for i in range(pos, pos + ln - minsep):
            for j in range(i + minsep, min(pos + ln - 1, i + maxsep) + 1):
                for k in range(4):
                    ci, cj = k >= 2, k in (1, 3)
                    if ci and (seq[i] == "G" or mseq[i] == "G"):
                        continue
                    if cj and (seq[j] == "G" or mseq[j] == "G"):
                        continue
                    curve[(i, j, k)] = curve.get((i, j, k), 0.0) + w * gam(j - i)
    expect = 0.0
    for i, j in _pairs(chains, minsep, maxsep):
        if (i, j, 0) not in curve:
            continue
        for k in range(4):
            ci, cj = k >= 2, k in (1, 3)
            if (ci and seq[i] == "G") or (cj and seq[j] == "G"):
                continue
            if (i, j, k) not in curve:
                break
            expect += curve[(i, j, k)]
    assert any((i, 7, 1) not in curve and (i, 7, 0) in curve for i in range(7)), "the glycine case is exercised"
    assert abs(e - (-EPS * K_FM * expect)) <= 1e-9 * abs(EPS * K_FM * expect), (e, -EPS * K_FM * expect)


@pytest.mark.parametrize("ww,sigma_exp", [(0.1, 0.15), (0.37, 0.22)])
def test_table_displaced_memory(oracle_mod, tmp_path, ww, sigma_exp):
    """A memory equal to the structure except two residues moved along the
    axis by whole grid steps: the curve at the current distance is
    -eps k_fm gamma exp(-(r - rf)^2 / (2 |i-j|^(2 s) ww^2)) (fb.cpp:4284-4319),
    which pins the well width and the sigma exponent."""
    chains = ["ALKVSTWDEFMLKA"]
    seq = chains[0]
    st, xca, xcb = _axis_structure(seq, seed=4)
    mst = st.copy()
    mst[5, :, 0] += 0.031  # nm: 0.31 A
    mst[9, :, 0] -= 0.017
    p = _table_case(str(tmp_path), chains, mst, [(0, 0, 14, 1.0)], "2 6\nALL ALL 1 1.0\n", ww=ww,
                    sigma_exp=sigma_exp)
    o = _o(oracle_mod, p)
    e = o.energy(_coords(xca, xcb), 1 << oracle_mod.ET["Frag_Mem"])[oracle_mod.ET["Frag_Mem"]]
    mca, mcb = mst[:, 0] * 10.0, mst[:, 1] * 10.0
    X = {0: xca, 1: xcb}
    M = {0: mca, 1: mcb}
    expect = 0.0
    for i, j in _pairs(chains, 2, 6):
        for k in range(4):
            a, b = (0 if k < 2 else 1), (0 if k in (0, 2) else 1)
            r = abs(X[a][i, 0] - X[b][j, 0])
            rf = abs(M[a][i, 0] - M[b][j, 0])
            sig2 = (j - i) ** (2 * sigma_exp) * ww * ww
            expect += np.exp(-(r - rf) ** 2 / (2.0 * sig2))
    assert abs(e - (-EPS * K_FM * expect)) <= 1e-9 * abs(EPS * K_FM * expect), (e, -EPS * K_FM * expect)


def test_table_midpoint_interpolation(oracle_mod, tmp_path):
    """One CA moved by half a grid step: its distances to every other atom
    fall at midpoints, where the energy is the mean of the two grid values
    and the force on the moved CA is ff dx with ff the mean of the two
    force-curve values V(r_g) (r_g - rf) / (sigma^2 r_g) (fb.cpp:4305-4318,
    4400-4417).  The other atoms keep their grid distances."""
    chains = ["ALKVSTWDEF"]
    seq = chains[0]
    st, xca, xcb = _axis_structure(seq, seed=5)
    ww, s = 0.25, 0.15
    p = _table_case(str(tmp_path), chains, st, [(0, 0, 10, 1.0)], "2 5\nALL ALL 1 1.0\n", ww=ww, sigma_exp=s)
    o = _o(oracle_mod, p)
    m = 4
    xca2 = xca.copy()
    xca2[m, 0] += 0.005
    f, e = o.compute(_coords(xca2, xcb), 1 << oracle_mod.ET["Frag_Mem"])
    A = EPS * K_FM
    X = {0: xca2, 1: xcb}
    M = {0: xca, 1: xcb}

    def gv(rg, rf, sig2):
        V = -A * np.exp(-(rg - rf) ** 2 / (2.0 * sig2))
        return V, V * (rg - rf) / (sig2 * rg)

    expect_e, fm = 0.0, np.zeros(3)
    for i, j in _pairs(chains, 2, 5):
        for k in range(4):
            a, b = (0 if k < 2 else 1), (0 if k in (0, 2) else 1)
            d = X[a][i] - X[b][j]
            r = np.linalg.norm(d)
            rf = abs(M[a][i, 0] - M[b][j, 0])
            sig2 = (j - i) ** (2 * s) * ww * ww
            moved = (a == 0 and i == m) or (b == 0 and j == m)
            if moved:
                r1, r2 = np.floor(r * 100.0 + 1e-6) / 100.0, np.floor(r * 100.0 + 1e-6) / 100.0 + 0.01
                v1, f1 = gv(r1, rf, sig2)
                v2, f2 = gv(r2, rf, sig2)
                expect_e += 0.5 * (v1 + v2)
                ff = 0.5 * (f1 + f2)
                fm += ff * d if (a == 0 and i == m) else -ff * d
            else:
                expect_e += gv(r, rf, sig2)[0]
    E = e[oracle_mod.ET["Frag_Mem"]]
    assert abs(E - expect_e) <= 1e-9 * abs(expect_e), (E, expect_e)
    assert np.abs(f[0][m] - fm).max() <= 1e-7 * max(1.0, np.abs(fm).max()), (f[0][m], fm)


def test_table_out_of_range_is_an_error(oracle_mod, tmp_path):
    """A distance outside [rmin, rmax] stops the evaluation with the
    reference's message (fb.cpp:4419-4427)."""
    chains = ["ALKVSTWDEF"]
    st, xca, xcb = _axis_structure(chains[0], seed=6)
    p = _table_case(str(tmp_path), chains, st, [(0, 0, 10, 1.0)], "2 5\nALL ALL 1 1.0\n", tb=(3.0, 30.0, 0.01))
    o = _o(oracle_mod, p)
    x = _coords(xca, xcb)
    o.compute(x)  # every FM distance within [3, 30]
    x[0][9, 0] += 100.0
    with pytest.raises(RuntimeError, match="r is out of computed range"):
        o.compute(x)


def test_table_converges_to_memory_mode(oracle_mod):
    """With well width 1 the table curves are the memory-mode Gaussians
    (fb.cpp:4279-4280 against 4126) sampled on the grid, so the table-mode
    energy and forces converge to the memory-mode ones at second order in dr
    (linear interpolation of smooth curves): the difference falls by about 4
    for each halving of dr, and is small at dr = 0.005 A.  A lost
    well-width factor, a shifted grid or a wrong interpolation weight breaks
    the order or the limit."""
    from conftest import make_case, make_table_case
    k = oracle_mod.ET["Frag_Mem"]
    fm = 1 << k
    s, p = make_case([30], 61)
    x = s.coords()
    f_mem, e_mem = _o(oracle_mod, p).compute(x, fm)
    diffs = []
    for dr in (0.04, 0.02, 0.01, 0.005):
        _, pt = make_table_case([30], 61, tb=(1.0, 30.0, dr), well_width=1.0)
        f_tb, e_tb = _o(oracle_mod, pt).compute(x, fm)
        de = abs(e_tb[k] - e_mem[k]) / abs(e_mem[k])
        dfv = np.linalg.norm((f_tb - f_mem).reshape(-1, 3), axis=1).max() / np.linalg.norm(
            f_mem.reshape(-1, 3), axis=1).max()
        diffs.append((de, dfv))
        print(f"dr={dr}: rel energy diff {de:.3e}, rel force diff {dfv:.3e}")
    assert e_mem[k] < -10.0
    for (e0, f0), (e1, f1) in zip(diffs, diffs[1:]):
        assert 2.8 < e0 / e1 < 5.5 and 2.8 < f0 / f1 < 5.5
    assert diffs[-1][0] < 1e-5 and diffs[-1][1] < 1e-4
