"""Test systems with real secondary structure: an ideal alpha helix,
antiparallel and parallel two-chain beta sheets, and beta hairpins (a short
turn and a long loop), built from ideal backbone internal coordinates.  On
these, the DSSP hydrogen-bond, P_AP and helix terms carry a sizeable share
of the force (random-walk chains form almost no hydrogen bonds).

Backbone: N, CA, C, O by the natural-extension reference frame method
(NeRF) with ideal bond lengths and angles (N-CA 1.458, CA-C 1.525, C-N
1.329, C=O 1.231 A; N-CA-C 111.2, CA-C-N 116.2, C-N-CA 121.7, CA-C-O
120.5 degrees; omega 180), CB from N, C, CA at 1.53 A.  AWSEM keeps CA, CB
(H-beta for glycine) and O; N, H and C' are rebuilt from them.
"""
from __future__ import annotations

import numpy as np

from awsemmd_amd import workload as W

DEG = np.pi / 180.0


def _place(a, b, c, bond, angle, torsion):
    """Atom d with |cd| = bond, angle(b, c, d) = angle, dihedral(a, b, c, d) = torsion."""
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    n /= np.linalg.norm(n)
    m = np.cross(n, bc)
    d2 = np.array([-bond * np.cos(angle), bond * np.sin(angle) * np.cos(torsion), bond * np.sin(angle) * np.sin(torsion)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def backbone(phi, psi):
    """CA, CB, O of a chain with the given torsions (radians, per residue)."""
    L = len(phi)
    N = np.zeros((L, 3))
    CA = np.zeros((L, 3))
    C = np.zeros((L, 3))
    N[0] = [0.0, 0.0, 0.0]
    CA[0] = [1.458, 0.0, 0.0]
    C[0] = CA[0] + 1.525 * np.array([np.cos(np.pi - 111.2 * DEG), np.sin(np.pi - 111.2 * DEG), 0.0])
    for i in range(1, L):
        N[i] = _place(N[i - 1], CA[i - 1], C[i - 1], 1.329, 116.2 * DEG, psi[i - 1])
        CA[i] = _place(CA[i - 1], C[i - 1], N[i], 1.458, 121.7 * DEG, np.pi)
        C[i] = _place(C[i - 1], N[i], CA[i], 1.525, 111.2 * DEG, phi[i])
    O = np.zeros((L, 3))
    CB = np.zeros((L, 3))
    for i in range(L):
        if i < L - 1:
            O[i] = _place(N[i + 1], CA[i], C[i], 1.231, 120.5 * DEG, np.pi)
        else:
            O[i] = _place(N[i], CA[i], C[i], 1.231, 120.5 * DEG, psi[i] + np.pi)
        CB[i] = _place(C[i], N[i], CA[i], 1.53, 110.5 * DEG, -122.5 * DEG)
    return CA, CB, O


def helix(L):
    return backbone(np.full(L, -57.0 * DEG), np.full(L, -47.0 * DEG))


def strand(L, phi=-120.0, psi=130.0):
    return backbone(np.full(L, phi * DEG), np.full(L, psi * DEG))


def _rot(axis, ang):
    k = axis / np.linalg.norm(axis)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def _hbond_score(ca_a, o_a, ca_b, o_b):
    """Sum over acceptor O / donor (N, H) pairs of both strands of
    exp(-(dNO - 2.98)^2 / (2 0.68^2) - (dHO - 2.06)^2 / (2 0.76^2)), with N and
    H rebuilt like AWSEM's virtual atoms (shipped [ABC] coefficients); minus a
    clash penalty for CA-CA below 4.2 A."""
    def nh(ca, o):
        n = 0.483 * ca[:-1] + 0.703 * ca[1:] - 0.186 * o[:-1]
        h = 0.841 * ca[:-1] + 0.893 * ca[1:] - 0.734 * o[:-1]
        return n, h
    s = 0.0
    for (ca1, o1), (ca2, o2) in (((ca_a, o_a), (ca_b, o_b)), ((ca_b, o_b), (ca_a, o_a))):
        n, h = nh(ca2, o2)
        dno = np.linalg.norm(o1[:, None] - n[None], axis=-1)
        dho = np.linalg.norm(o1[:, None] - h[None], axis=-1)
        s += np.exp(-0.5 * ((dno - 2.98) / 0.68) ** 2 - 0.5 * ((dho - 2.06) / 0.76) ** 2).sum()
    dca = np.linalg.norm(ca_a[:, None] - ca_b[None], axis=-1)
    return s - 10.0 * np.clip(4.2 - dca, 0.0, None).sum()


def pair_strands(L, parallel, rng=None):
    """Two ideal strands of L residues, the second placed (rotation about the
    first's axis, shift along it, lateral distance) where the inter-strand
    hydrogen-bond score is largest; antiparallel strands run in opposite
    directions."""
    ca, cb, o = strand(L)
    c0 = ca.mean(0)
    ax = ca[-1] - ca[0]
    ax /= np.linalg.norm(ax)
    # sheet-plane normal: perpendicular to the axis, in the plane of the O atoms' spread
    v = (o - ca) - np.outer((o - ca) @ ax, ax)
    side = np.linalg.svd(v - v.mean(0))[2][0]
    lat = np.cross(ax, side)
    lat /= np.linalg.norm(lat)
    best = None
    base = [ca - c0, cb - c0, o - c0]
    if not parallel:
        Rf = _rot(lat, np.pi)  # reverse the direction
        base = [p @ Rf.T for p in base]
    for th in np.arange(0.0, 360.0, 10.0) * DEG:
        R = _rot(ax, th)
        rot = [p @ R.T for p in base]
        for sft in np.arange(-1.9, 1.91, 0.19):
            for D in np.arange(4.2, 5.61, 0.2):
                for sgn in (1.0, -1.0):
                    off = sgn * D * side + sft * ax
                    cand = [p + c0 + off for p in rot]
                    sc = _hbond_score(ca, o, cand[0], cand[2])
                    if best is None or sc > best[0]:
                        best = (sc, cand)
    return (ca, cb, o), tuple(best[1])


def _system(chains, xca, xcb, xo, seed, beta_all=False, n_struct=3):
    """A workload System with these coordinates, an ssweight pattern and a
    fragment library of n_struct memories per 9-residue window (trace + 1 A
    noise), like the synthetic ones; fewer memories than the benchmark's 20
    keep fragment memory from dominating the forces of these small systems."""
    s = W.make_system([len(c) for c in chains], seed=seed, n_struct=n_struct)
    s.chains = list(chains)
    s.xca, s.xcb, s.xo = np.asarray(xca, float), np.asarray(xcb, float), np.asarray(xo, float)
    rng = np.random.default_rng(seed + 1)
    n = s.n
    if beta_all:
        s.ssw = np.tile([0.0, 1.0], (n, 1))
    structs = []
    for _ in range(len(s.mem_structs)):
        ca = s.xca + rng.normal(scale=1.0, size=s.xca.shape)
        cb = s.xcb + rng.normal(scale=1.0, size=s.xcb.shape)
        structs.append(np.round(np.stack([ca, cb], axis=1) * 100.0) / 1000.0)
    s.mem_structs = structs
    return s


def _seq(rng, L, avoid="P"):
    letters = [c for c in W.AA1 if c not in avoid]
    return "".join(rng.choice(letters, size=L))


def helix_system(L=26, seed=1):
    rng = np.random.default_rng(seed)
    ca, cb, o = helix(L)
    return _system([_seq(rng, L)], ca + 100.0, cb + 100.0, o + 100.0, seed, beta_all=True)


def sheet_system(L=10, parallel=False, seed=2, spread=0.0):
    """Two chains forming a sheet; spread moves the second strand that much
    further out (P_AP's switch is steep near its 8 A cut, flat below 5 A)."""
    rng = np.random.default_rng(seed)
    (a_ca, a_cb, a_o), (b_ca, b_cb, b_o) = pair_strands(L, parallel)
    if spread:
        out = b_ca.mean(0) - a_ca.mean(0)
        ax = a_ca[-1] - a_ca[0]
        out -= np.dot(out, ax) / np.dot(ax, ax) * ax
        out = spread * out / np.linalg.norm(out)
        b_ca, b_cb, b_o = b_ca + out, b_cb + out, b_o + out
    chains = [_seq(rng, L), _seq(rng, L)]
    xca, xcb, xo = np.vstack([a_ca, b_ca]), np.vstack([a_cb, b_cb]), np.vstack([a_o, b_o])
    return _system(chains, xca + 100.0, xcb + 100.0, xo + 100.0, seed)


def hairpin_system(L=10, loop=2, seed=3, beta_all=False):
    """One chain: strand, loop, the antiparallel partner strand.  The loop
    residues are laid on an arc from the first strand's end to the second
    strand's start (their local terms are strained; the strands' pairing is
    what the test is for)."""
    rng = np.random.default_rng(seed)
    (a_ca, a_cb, a_o), (b_ca, b_cb, b_o) = pair_strands(L, False)
    # the antiparallel partner runs from the far end back: its residue order
    # already starts next to the first strand's end
    p0, p1 = a_ca[-1], b_ca[0]
    bulge = a_ca[-1] - a_ca[-2]
    bulge /= np.linalg.norm(bulge)
    t = np.linspace(0.0, 1.0, loop + 2)[1:-1]
    r = max(3.8 * (loop + 1) / np.pi, 0.5 * np.linalg.norm(p1 - p0))
    lc = np.array([(1 - u) * p0 + u * p1 + bulge * r * np.sin(np.pi * u) for u in t])
    lo, lb = lc + np.array([0.0, 0.0, 2.4]), lc + np.array([0.0, 1.53, 0.0])
    xca = np.vstack([a_ca, lc, b_ca])
    xcb = np.vstack([a_cb, lb, b_cb])
    xo = np.vstack([a_o, lo, b_o])
    chains = [_seq(rng, 2 * L + loop)]
    return _system(chains, xca + 100.0, xcb + 100.0, xo + 100.0, seed, beta_all=beta_all)


def all_systems():
    """name -> System: the structured test set."""
    return {
        "helix": helix_system(),
        "sheet_antiparallel": sheet_system(parallel=False),
        "sheet_parallel": sheet_system(parallel=True, seed=4),
        "sheet_antiparallel_loose": sheet_system(parallel=False, seed=7, spread=2.6),
        "sheet_parallel_loose": sheet_system(parallel=True, seed=8, spread=2.6),
        "hairpin_turn": hairpin_system(L=9, loop=2, seed=5, beta_all=True),
        "hairpin_long_loop": hairpin_system(L=10, loop=16, seed=6),
    }
