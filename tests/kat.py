"""Known-answer geometry and input files for pinning the FP64 oracle.

Every helper here builds inputs whose energy follows from the definitions of
the terms at special points (distances placed exactly at a well's centre or
a switch's midpoint, everything else far beyond every cutoff), so that the
expected value is a short closed form of parameters the test itself wrote.
Nothing here imports the oracle.

Residue type order (src/fix_backbone.cpp:54, se_map): ARNDCQEGHILKMFPSTWYV.
"""
from __future__ import annotations

import os
import shutil

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PARAMS = os.path.join(ROOT, "awsemmd_amd", "params")
AA = "ARNDCQEGHILKMFPSTWYV"


def tidx(c: str) -> int:
    return AA.index(c)


# ---------------------------------------------------------------------------
# coeff file: sections written from explicit values

def section(name: str, *rows) -> str:
    lines = [name]
    for r in rows:
        if isinstance(r, (list, tuple, np.ndarray)):
            lines.append(" ".join(repr(float(v)) if not isinstance(v, (int, np.integer)) else str(int(v)) for v in r))
        else:
            lines.append(repr(float(r)) if not isinstance(r, (int, np.integer)) else str(int(r)))
    return "\n".join(lines) + "\n\n"


ABC_ROWS = ((0.483, 0.703, -0.186), (0.444, 0.235, 0.321), (0.841, 0.893, -0.734))


def abc_section(abc=ABC_ROWS) -> str:
    return section("[ABC]", *abc)


class Abc:
    """Virtual-atom coefficients: N_i = an CA(i-1) + bn CA(i) + cn O(i-1),
    C'_i = ap CA(i) + bp CA(i+1) + cp O(i), H_i = ah CA(i-1) + bh CA(i) +
    ch O(i-1) (src/fix_backbone.cpp:7113-7138)."""

    def __init__(self, rows=ABC_ROWS):
        (self.an, self.bn, self.cn), (self.ap, self.bp, self.cp), (self.ah, self.bh, self.ch) = rows

    def virtual(self, xca, xo, chain_lens):
        n = len(xca)
        xn, xh, xcp = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        first = np.zeros(n, bool)
        last = np.zeros(n, bool)
        s = 0
        for L in chain_lens:
            first[s] = True
            last[s + L - 1] = True
            s += L
        for i in range(n):
            if not first[i]:
                xn[i] = self.an * xca[i - 1] + self.bn * xca[i] + self.cn * xo[i - 1]
                xh[i] = self.ah * xca[i - 1] + self.bh * xca[i] + self.ch * xo[i - 1]
            if not last[i]:
                xcp[i] = self.ap * xca[i] + self.bp * xca[i + 1] + self.cp * xo[i]
        return xn, xh, xcp


def write_tables(d, anti_HB=None, anti_NHB=None, para_HB=None, anti_one=None, para_one=None, gamma=None,
                 burial=None, extra=None):
    """The reference's side files in d: DSSP tables as [2][20][20] (k, i, j;
    written k-major like fb.cpp:568-585 reads them) and [20]; gamma.dat as
    [wells][20][20][2] upper triangles (fb.cpp:623-641); burial_gamma.dat
    [20][3].  None: the shipped file."""
    os.makedirs(d, exist_ok=True)

    def put(name, arr, writer):
        if arr is None:
            shutil.copyfile(os.path.join(PARAMS, name), os.path.join(d, name))
        else:
            with open(os.path.join(d, name), "w") as f:
                writer(f, np.asarray(arr, dtype=float))

    def w3(f, a):  # [20][20][2] -> k-major blocks of 20 rows
        for k in range(2):
            for i in range(20):
                f.write(" ".join("%.17g" % a[i, j, k] for j in range(20)) + "\n")

    def w1(f, a):
        f.write("\n".join("%.17g" % v for v in a) + "\n")

    def wg(f, a):
        for w in range(a.shape[0]):
            for i in range(20):
                for j in range(i, 20):
                    f.write("%.17g %.17g\n" % (a[w, i, j, 0], a[w, i, j, 1]))

    def wb(f, a):
        for i in range(20):
            f.write("%.17g %.17g %.17g\n" % tuple(a[i]))

    put("anti_HB", anti_HB, w3)
    put("anti_NHB", anti_NHB, w3)
    put("para_HB", para_HB, w3)
    put("anti_one", anti_one, w1)
    put("para_one", para_one, w1)
    put("gamma.dat", gamma, wg)
    put("burial_gamma.dat", burial, wb)
    for name, text in (extra or {}).items():
        with open(os.path.join(d, name), "w") as f:
            f.write(text)


def write_case(d, chains, coeff_text, ssw=None, **tables):
    """Coeff, sequence (one chain per line), ssweight ([n][k] values of the
    flagged [SSWeight] columns, in column order) and tables in d."""
    os.makedirs(d, exist_ok=True)
    write_tables(d, **tables)
    coeff = os.path.join(d, "fix_backbone_coeff.data")
    with open(coeff, "w") as f:
        f.write(coeff_text)
    seq = os.path.join(d, "protein.seq")
    with open(seq, "w") as f:
        f.write("\n".join(chains) + "\n")
    if ssw is not None:
        with open(os.path.join(d, "ssweight"), "w") as f:
            for row in np.atleast_2d(ssw):
                f.write(" ".join("%.17g" % v for v in row) + "\n")
    return dict(coeff=coeff, seq=seq, workdir=d)


# ---------------------------------------------------------------------------
# geometry helpers

def unit(v):
    return v / np.linalg.norm(v)


def two_sphere_point(P, Q, rp, rq, hint):
    """A point X with |X - P| = rp and |X - Q| = rq: on the circle where the
    spheres meet, in the direction of `hint` from the circle's axis."""
    P, Q = np.asarray(P, float), np.asarray(Q, float)
    d = np.linalg.norm(Q - P)
    e = (Q - P) / d
    a = (rp * rp - rq * rq + d * d) / (2 * d)
    h2 = rp * rp - a * a
    if h2 < 0:
        raise ValueError("spheres do not meet")
    u = np.asarray(hint, float) - P
    u = u - np.dot(u, e) * e
    if np.linalg.norm(u) < 1e-9:
        u = np.cross(e, [1.0, 0.0, 0.0]) if abs(e[0]) < 0.9 else np.cross(e, [0.0, 1.0, 0.0])
    u = unit(u)
    return P + a * e + np.sqrt(h2) * u


def sparse_points(n, rng, min_sep=45.0, box=600.0):
    """n points with pairwise distances >= min_sep."""
    pts = []
    while len(pts) < n:
        c = rng.uniform(0.0, box, 3)
        if all(np.linalg.norm(c - p) >= min_sep for p in pts):
            pts.append(c)
    return np.array(pts)


def rotate_about(points, origin, axis, angle):
    """Rotation of points by angle (right hand) about the axis through origin."""
    k = unit(np.asarray(axis, float))
    p = np.atleast_2d(points) - origin
    c, s = np.cos(angle), np.sin(angle)
    out = p * c + np.cross(k, p) * s + np.outer(p @ k, k) * (1 - c)
    return out + origin


def dihedral(p0, p1, p2, p3):
    """IUPAC dihedral of four points (atan2 form)."""
    b1, b2, b3 = p1 - p0, p2 - p1, p3 - p2
    n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
    return np.arctan2(np.linalg.norm(b2) * np.dot(b1, n2), np.dot(n1, n2))
