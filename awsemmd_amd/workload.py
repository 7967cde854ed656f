"""Seeded synthetic AWSEM-MD workloads and their reference-format file sets.

A system is one or more protein chains with:
  * a sequence drawn uniformly from the 20 amino acids;
  * a C-alpha trace that is a self-avoiding random walk with 3.8 A bonds;
  * CB (H-beta for glycine) and O placed by ideal virtual-bond geometry
    (the harmonic-bond lengths of the reference deck, CA-O 2.41 A, O-CA(i+1)
    2.50 A, CA-CB 1.54 A: tools/create_project_tools/CoordinatesToWorkLammpsDataFile.py:289-294);
  * an ssweight column pair per residue drawn from {(1,0),(0,1),(0,0)} so the
    rama / DSSP / P_AP ssweight branches are all exercised;
  * a fragment-memory library: n_struct copies of the trace perturbed by
    Gaussian noise (sigma 1 A) written as .gro files, and one memory per copy
    at every 9-residue window start (the .mem format of
    src/fix_backbone.cpp:1966-2042).

`write_system` writes the exact file set the reference's FixBackbone
constructor reads (coeff, sequence, ssweight, gamma.dat, burial_gamma.dat,
DSSP tables, fragment-memory gamma, .mem, .gro).
"""
from __future__ import annotations

import dataclasses
import os
import shutil
from typing import List, Optional, Sequence

import numpy as np

AA1 = "ARNDCQEGHILKMFPSTWYV"
AA3 = {"A": "ALA", "R": "ARG", "N": "ASN", "D": "ASP", "C": "CYS", "Q": "GLN", "E": "GLU",
       "G": "GLY", "H": "HIS", "I": "ILE", "L": "LEU", "K": "LYS", "M": "MET", "F": "PHE",
       "P": "PRO", "S": "SER", "T": "THR", "W": "TRP", "Y": "TYR", "V": "VAL"}

# virtual-atom coefficients of the shipped coeff file [ABC] section
ABC = dict(an=0.483, bn=0.703, cn=-0.186, ap=0.444, bp=0.235, cp=0.321)
PARAM_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "params")


@dataclasses.dataclass
class Memory:
    """One fragment memory: target window [pos, pos+length) taken from
    structure `struct` (0-based), residues fpos.. of that structure."""
    struct: int
    pos: int
    fpos: int
    length: int
    weight: float = 1.0


@dataclasses.dataclass
class System:
    chains: List[str]                    # one sequence string per chain
    xca: np.ndarray                      # (n, 3) float64, Angstrom
    xcb: np.ndarray
    xo: np.ndarray
    ssw: np.ndarray                      # (n, 2) float64: ssweight columns 4 and 5 of [SSWeight]
    mem_structs: List[np.ndarray]        # per structure (n, 2, 3) CA/CB in nm, rounded to 1e-3
    memories: List[Memory]
    seed: int = 0

    @property
    def n(self) -> int:
        return len(self.seq)

    @property
    def seq(self) -> str:
        return "".join(self.chains)

    def chain_index(self) -> np.ndarray:
        return np.concatenate([np.full(len(c), k, dtype=np.int32) for k, c in enumerate(self.chains)])

    def coords(self) -> np.ndarray:
        """(3, n, 3) array [kind][residue][xyz] with kind 0 CA, 1 CB, 2 O."""
        return np.stack([self.xca, self.xcb, self.xo]).astype(np.float64)


def _unit(v):
    return v / np.linalg.norm(v)


def _rand_unit(rng):
    while True:
        v = rng.normal(size=3)
        nv = np.linalg.norm(v)
        if nv > 1e-8:
            return v / nv


def _rand_units(rng, m):
    u = rng.normal(size=(m, 3))
    return u / np.linalg.norm(u, axis=1, keepdims=True)


def _min_dist(c, pts):
    """min over pts of |c_k - p| for each candidate c_k."""
    d2 = (c * c).sum(1)[:, None] + (pts * pts).sum(1)[None, :] - 2.0 * c @ pts.T
    return np.sqrt(np.maximum(d2.min(axis=1), 0.0))


def saw_trace(n: int, rng, start=None, bond=3.8, excl=4.6, excl13=5.0, others: Optional[np.ndarray] = None,
              center=None, radius=None) -> np.ndarray:
    """Self-avoiding C-alpha walk: |CA_i - CA_j| >= excl for |i-j|>=3 (and
    against `others`), |CA_i - CA_{i-2}| >= excl13, optional confinement."""
    pts = np.zeros((n, 3))
    pts[0] = np.zeros(3) if start is None else np.asarray(start, dtype=float)
    i = 1
    fails = 0
    while i < n:
        ok = False
        if radius is not None:
            # confined (compact) walks: candidates tested in growing batches,
            # the first admissible one taken
            for m in (4, 16, 64, 256):
                c = pts[i - 1] + bond * _rand_units(rng, m)
                good = np.linalg.norm(c - center, axis=1) <= radius
                if i >= 2:
                    good &= np.linalg.norm(c - pts[i - 2], axis=1) >= excl13
                if i >= 3 and good.any():
                    good &= _min_dist(c, pts[: i - 2]) >= excl
                if others is not None and len(others) and good.any():
                    good &= _min_dist(c, others) >= excl
                k = np.flatnonzero(good)
                if len(k):
                    cand, ok = c[k[0]], True
                    break
        for _ in range(0 if radius is not None else 200):
            cand = pts[i - 1] + bond * _rand_unit(rng)
            if radius is not None and np.linalg.norm(cand - center) > radius:
                continue
            if i >= 2 and np.linalg.norm(cand - pts[i - 2]) < excl13:
                continue
            if i >= 3 and np.min(np.linalg.norm(pts[: i - 2] - cand, axis=1)) < excl:
                continue
            if others is not None and len(others) and np.min(np.linalg.norm(others - cand, axis=1)) < excl:
                continue
            ok = True
            break
        if ok:
            pts[i] = cand
            i += 1
            fails = 0
        else:
            fails += 1
            if fails > 40:
                # trapped (crowded confinement): restart the chain from a
                # shifted start with a 5% larger confining sphere
                pts[0] = pts[0] + rng.normal(size=3)
                if radius is not None:
                    radius *= 1.05
                i, fails = 1, 0
                continue
            i = max(1, i - 1 - min(fails, 8))  # backtrack
    return pts


def place_side_atoms(xca: np.ndarray, seq: str, chi0=-0.83):
    """Ideal O and CB placement for one chain's CA trace."""
    n = len(xca)
    xo = np.zeros_like(xca)
    for i in range(n):
        if i < n - 1:
            e = xca[i + 1] - xca[i]
            d = np.linalg.norm(e)
            e = e / d
            a = (2.41 ** 2 - 2.50 ** 2 + d * d) / (2 * d)
            h = np.sqrt(max(2.41 ** 2 - a * a, 0.25))
            ref = (xca[i - 1] - xca[i]) if i > 0 else np.array([0.3, 1.0, -0.2])
            q = ref - np.dot(ref, e) * e
            if np.linalg.norm(q) < 1e-6:
                q = np.cross(e, [1.0, 0.0, 0.0])
            q = -_unit(q)
            xo[i] = xca[i] + a * e + h * q
        else:
            e = _unit(xca[i] - xca[i - 1]) if n > 1 else np.array([1.0, 0.0, 0.0])
            perp = _unit(np.cross(e, [0.2, 1.0, 0.3]))
            xo[i] = xca[i] + 2.41 * _unit(0.5 * e + perp)
    xcb = np.zeros_like(xca)
    for i in range(n):
        if i > 0:
            pn = ABC["an"] * xca[i - 1] + ABC["bn"] * xca[i] + ABC["cn"] * xo[i - 1]
        else:
            pn = xca[i] + 1.46 * _unit(xca[i] - xca[min(i + 1, n - 1)] + np.array([1e-3, 0, 0]))
        if i < n - 1:
            pc = ABC["ap"] * xca[i] + ABC["bp"] * xca[i + 1] + ABC["cp"] * xo[i]
        else:
            pc = xca[i] + 1.52 * _unit(xca[i] - xca[max(i - 1, 0)] + np.array([0, 1e-3, 0]))
        b1 = _unit(pn - xca[i])
        b2 = _unit(pc - xca[i])
        bis = b1 + b2
        bis = _unit(bis) if np.linalg.norm(bis) > 1e-6 else _unit(np.cross(b1, [0.0, 0.0, 1.0]))
        nrm = np.cross(b1, b2)
        nrm = _unit(nrm) if np.linalg.norm(nrm) > 1e-6 else _unit(np.cross(bis, [0.0, 1.0, 0.0]))
        best = None
        for s in (1.0, -1.0):
            th = np.deg2rad(54.75)
            cb = xca[i] + 1.54 * _unit(-bis * np.cos(th) + s * nrm * np.sin(th))
            a = pc - xca[i]
            b = xca[i] - pn
            c = xca[i] - cb
            chi = np.dot(np.cross(a, b), c) / (np.linalg.norm(a) * np.linalg.norm(b) * np.linalg.norm(c))
            if best is None or abs(chi - chi0) < best[0]:
                best = (abs(chi - chi0), cb)
        xcb[i] = best[1]
    return xo, xcb


def make_system(lengths: Sequence[int], seed: int, n_struct: int = 20, mem_len: int = 9,
                noise: float = 1.0, compact: bool = False, spacing=(10.0, 15.0)) -> System:
    """Seeded system of len(lengths) chains.  compact=True confines each chain
    to a sphere and places chain centroids `spacing` apart (multimers)."""
    rng = np.random.default_rng(seed)
    chains = ["".join(rng.choice(list(AA1), size=L)) for L in lengths]
    traces = []
    all_pts = np.zeros((0, 3))
    centers = []
    for k, L in enumerate(lengths):
        if compact:
            if k == 0:
                c = np.zeros(3)
            else:
                for _ in range(1000):
                    c = centers[rng.integers(len(centers))] + rng.uniform(*spacing) * _rand_unit(rng)
                    if all(np.linalg.norm(c - cc) >= spacing[0] for cc in centers):
                        break
            centers.append(c)
            radius = 3.8 * (L ** (1.0 / 3.0)) * 1.6 + 8.0
            start = c + rng.normal(size=3)
            tr = saw_trace(L, rng, start=start, others=all_pts, center=c, radius=radius)
        else:
            start = np.array([100.0, 100.0, 100.0]) + 40.0 * k * np.array([1.0, 0.0, 0.0])
            tr = saw_trace(L, rng, start=start, others=all_pts)
        traces.append(tr)
        all_pts = np.vstack([all_pts, tr])
    xca_l, xcb_l, xo_l = [], [], []
    for seq, tr in zip(chains, traces):
        xo, xcb = place_side_atoms(tr, seq)
        xca_l.append(tr)
        xcb_l.append(xcb)
        xo_l.append(xo)
    xca = np.vstack(xca_l)
    xcb = np.vstack(xcb_l)
    xo = np.vstack(xo_l)
    n = len(xca)
    if compact:  # shift into a positive box
        shift = 150.0 - xca.mean(axis=0)
        xca, xcb, xo = xca + shift, xcb + shift, xo + shift
    choices = np.array([[1.0, 0.0], [0.0, 1.0], [0.0, 0.0]])
    ssw = choices[rng.integers(0, 3, size=n)]
    mem_structs = []
    for _ in range(n_struct):
        ca = xca + rng.normal(scale=noise, size=xca.shape)
        cb = xcb + rng.normal(scale=noise, size=xcb.shape)
        # nm with 3 decimals, exactly what a .gro file stores: k/1000 with
        # integer k is the double that "%8.3f" prints and strtod reads back
        s = np.round(np.stack([ca, cb], axis=1) * 100.0) / 1000.0
        mem_structs.append(s)
    memories = []
    off = 0
    for seq in chains:
        L = len(seq)
        for p in range(0, L - mem_len + 1):
            for s in range(n_struct):
                memories.append(Memory(struct=s, pos=off + p, fpos=off + p, length=mem_len, weight=1.0))
        off += L
    return System(chains=chains, xca=xca, xcb=xcb, xo=xo, ssw=ssw, mem_structs=mem_structs,
                  memories=memories, seed=seed)


def arrays(sys: System) -> dict:
    """Keyword arguments of FixBackbone.from_arrays for `sys` (same content
    as the file set write_system produces)."""
    n = sys.n
    aps = np.zeros((n, 12))
    aps[:, 3] = sys.ssw[:, 0]
    aps[:, 4] = sys.ssw[:, 1]
    seq = sys.seq
    return dict(chains=sys.chains, aps=aps, structs_nm=sys.mem_structs, struct_seqs=[seq] * len(sys.mem_structs),
                mem_struct=np.array([m.struct for m in sys.memories], dtype=np.int32),
                mem_tpos=np.array([m.pos for m in sys.memories], dtype=np.int32),
                mem_fpos=np.array([m.fpos for m in sys.memories], dtype=np.int32),
                mem_len=np.array([m.length for m in sys.memories], dtype=np.int32),
                mem_weight=np.array([m.weight for m in sys.memories], dtype=np.float64))


def coeff_text(fm: bool = True, mem_file: str = "fragsLAMW.mem", gamma_file: str = "uniform.gamma",
               base: Optional[str] = None, table: Optional[Sequence[float]] = None) -> str:
    """The shipped coeff file with the [Fragment_Memory] section enabled, or
    with table=(rmin, rmax, dr, well_width, use_table, sigma_exp) the
    [Fragment_Memory_Table] section in its place (same strength, memories
    and gamma file; fix_backbone.cpp:310-320 reads these lines)."""
    base = base or os.path.join(PARAM_DIR, "fix_backbone_coeff.data")
    txt = open(base).read()
    if fm:
        sec = "[Fragment_Memory]\n1.0\n%s\n%s" % (mem_file, gamma_file)
        if table is not None:
            rmin, rmax, dr, ww, use, sexp = table
            sec = "[Fragment_Memory_Table]\n1.0\n%s\n%s\n%r %r %r\n%r\n%d\n%r" % (
                mem_file, gamma_file, rmin, rmax, dr, ww, int(use), sexp)
        txt = txt.replace("[Fragment_Memory]-\n1.0\n1r69.mem\n1r69.gamma", sec)
        assert sec in txt
    return txt


def write_gro(path: str, seq: str, struct_nm: np.ndarray) -> None:
    """.gro with CA and CB lines (no CB for glycine), coordinates in nm."""
    lines = []
    k = 0
    for i, aa in enumerate(seq):
        name = AA3[aa]
        atoms = [("CA", struct_nm[i, 0])]
        if aa != "G":
            atoms.append(("CB", struct_nm[i, 1]))
        for an, xyz in atoms:
            k += 1
            lines.append("%5d  %s %-3s %5d %8.3f %8.3f %8.3f" % (i + 1, name, an, k, xyz[0], xyz[1], xyz[2]))
    with open(path, "w") as f:
        f.write(" Synthetic memory structure\n")
        f.write("%10d\n" % len(lines))
        f.write("\n".join(lines) + "\n")
        f.write("   0.00000   0.00000   0.00000\n")


def write_system(sys: System, workdir: str, fm: bool = True, coeff_base: Optional[str] = None,
                 table: Optional[Sequence[float]] = None) -> dict:
    """Writes the reference file set for `sys` into workdir; returns paths.
    table: fragment memory in table mode (see coeff_text)."""
    os.makedirs(workdir, exist_ok=True)
    for name in ("anti_HB", "anti_NHB", "para_HB", "anti_one", "para_one", "gamma.dat",
                 "burial_gamma.dat", "uniform.gamma"):
        shutil.copyfile(os.path.join(PARAM_DIR, name), os.path.join(workdir, name))
    coeff = os.path.join(workdir, "fix_backbone_coeff.data")
    with open(coeff, "w") as f:
        f.write(coeff_text(fm=fm, base=coeff_base, table=table))
    seqf = os.path.join(workdir, "protein.seq")
    with open(seqf, "w") as f:
        f.write("\n".join(sys.chains) + "\n")
    with open(os.path.join(workdir, "ssweight"), "w") as f:
        for a, b in sys.ssw:
            f.write("%.1f %.1f\n" % (a, b))
    if fm:
        seq = sys.seq
        for s, st in enumerate(sys.mem_structs):
            write_gro(os.path.join(workdir, "mem%03d.gro" % s), seq, st)
        with open(os.path.join(workdir, "fragsLAMW.mem"), "w") as f:
            f.write("[Target]\nquery\n\n[Memories]\n")
            for m in sys.memories:
                f.write("mem%03d.gro %d %d %d %g\n" % (m.struct, m.pos + 1, m.fpos + 1, m.length, m.weight))
    return dict(coeff=coeff, seq=seqf, workdir=workdir)


# ---------------------------------------------------------------------------
# coeff-toggled terms (src/fix_backbone.cpp:297-358, 412-421, 466-478)

EXTRA_TERMS = ("amhgo", "cont_rest", "vec_fm", "ssb", "membrane", "phosph", "dh")
# residue radii of the solvent barrier in the shipped coeff file
SSB_RSHIFT = (0.00, 2.04, 0.57, 0.57, 0.36, 1.11, 1.17, -1.52, 0.87, 0.67, 0.79, 1.47, 1.03, 1.00, -0.10, 0.26,
              0.37, 1.21, 1.15, 0.39)


def _sec(name: str, *rows) -> str:
    out = [name]
    for r in rows:
        out.append(" ".join(str(v) for v in r) if isinstance(r, (list, tuple)) else str(r))
    return "\n".join(out) + "\n\n"


def extras_text(terms: Sequence[str], amhgo=(1.0, 2, 8.0, 0), cont_rest=(2.0, 1.0), vec_fm=(1.0, 0.6),
                ssb=(1.0, 4.5, 6.5, 1.0, 6.0, 7.0, 5.0, 13, 1), membrane=None, phosph_k=1.2, phosph_res=(),
                dh=(1.0, 1.0, 1.0, 1.0, 10.0, 10)) -> str:
    """Coeff sections of the given optional terms (appended to a coeff file:
    the reader takes a section wherever its tag appears).  membrane:
    (k, k_bin, centre xyz, length, rho0_max, rho0_distor, 3x4 gammas)."""
    txt = ""
    for t in terms:
        if t == "amhgo":
            txt += _sec("[AMH-Go]", *amhgo)
        elif t == "cont_rest":
            txt += _sec("[Contact_Restraints]", cont_rest[0], cont_rest[1], "contact.list")
        elif t == "vec_fm":
            txt += _sec("[Vector_Fragment_Memory]", *vec_fm)
        elif t == "ssb":
            k1, r1a, r1b, k2, r2a, r2b, kap, sep, cor = ssb
            txt += _sec("[Solvent_Barrier]", k1, (r1a, r1b), k2, (r2a, r2b), kap, sep, cor, tuple(SSB_RSHIFT))
        elif t == "membrane":
            k, kb, c, ln, r0, dist, g = membrane
            txt += _sec("[Membrane]", k, kb, tuple(c), 0, ln, r0, dist, *[tuple(row) for row in g])
        elif t == "phosph":
            txt += _sec("[Phosphorylation]", phosph_k, len(phosph_res), *[int(r) for r in phosph_res])
        elif t == "dh":
            kpp, kmm, kpm, ks, lam, sep = dh
            txt += _sec("[DebyeHuckel]", (kpp, kmm, kpm), ks, lam, sep)
        else:
            raise ValueError("unknown optional term %r" % t)
    return txt


def write_extras(sys: System, workdir: str, terms: Sequence[str] = EXTRA_TERMS, seed: int = 0,
                 native_noise: float = 0.5, **kw) -> dict:
    """Enables optional FixBackbone terms on a system written by write_system
    (appends their coeff sections and writes their side files):
      amhgo      amh-go.gro = the structure plus Gaussian noise (native_noise,
                 A), amh-go.gamma with two sequence-separation classes
                 (3-4: 0.125, >= 5: 0.5);
      cont_rest  contact.list: residue pairs whose density sites are 4-9 A
                 apart (|i - j| >= 3), r0 = that distance plus noise, weights
                 0.5-1.5;
      vec_fm     the fragment memories (write_system with fm=True);
      ssb        the shipped barrier parameters;
      membrane   a slab through the system's centroid, zim classes 1-3;
      phosph     up to three serines (else the first residues) phosphorylated;
      dh         charge_on_residues.dat: K, R +1, D, E -1.
    Returns the membrane centre and the phosphorylated residues."""
    rng = np.random.default_rng(10007 * seed + 17)
    n, seq = sys.n, sys.seq
    site = np.where(np.array([c == "G" for c in seq])[:, None], sys.xca, sys.xcb)
    out = {}
    if "amhgo" in terms:
        ca = sys.xca + rng.normal(scale=native_noise, size=sys.xca.shape)
        cb = sys.xcb + rng.normal(scale=native_noise, size=sys.xcb.shape)
        st = np.round(np.stack([ca, cb], axis=1) * 100.0) / 1000.0
        write_gro(os.path.join(workdir, "amh-go.gro"), seq, st)
        with open(os.path.join(workdir, "amh-go.gamma"), "w") as f:
            f.write("# i_min i_med_min i_med_max\n3 5 inf\n#\n# ires_code   jres_code   class      gamma\n"
                    "ALL ALL 1 0.125\nALL ALL 2 0.5\n")
    if "cont_rest" in terms:
        d = np.linalg.norm(site[:, None, :] - site[None, :, :], axis=2)
        ii, jj = np.nonzero((d > 4.0) & (d < 9.0) & (np.abs(np.arange(n)[:, None] - np.arange(n)[None, :]) >= 3))
        keep = ii < jj
        ii, jj = ii[keep], jj[keep]
        if len(ii) > 2 * n:
            pick = rng.choice(len(ii), size=2 * n, replace=False)
            ii, jj = ii[pick], jj[pick]
        with open(os.path.join(workdir, "contact.list"), "w") as f:
            for a, b in zip(ii, jj):
                f.write("%d %d %.4f %.4f\n" % (b + 1, a + 1, rng.uniform(0.5, 1.5), d[a, b] + rng.normal(scale=0.7)))
    if "membrane" in terms:
        c = sys.xca.mean(axis=0) + rng.normal(scale=2.0, size=3)
        g = rng.uniform(0.2, 1.5, size=(3, 4))
        kw.setdefault("membrane", (1.0, 0.5, tuple(np.round(c, 3)), 30.0, 6.0, 0.3, g.round(3).tolist()))
        with open(os.path.join(workdir, "zim"), "w") as f:
            f.write("\n".join(str(int(z)) for z in rng.integers(1, 4, size=n)) + "\n")
        out["membrane"] = kw["membrane"]
    if "phosph" in terms:
        ser = [i + 1 for i, c in enumerate(seq) if c == "S"][:3]
        res = ser if ser else [1, 2]
        kw.setdefault("phosph_res", tuple(res))
        out["phosph_res"] = kw["phosph_res"]
    if "dh" in terms:
        q = {"K": 1.0, "R": 1.0, "D": -1.0, "E": -1.0}
        charged = [(i + 1, q[c]) for i, c in enumerate(seq) if c in q]
        with open(os.path.join(workdir, "charge_on_residues.dat"), "w") as f:
            f.write("%d\n" % len(charged))
            for r, v in charged:
                f.write("%d %.1f\n" % (r, v))
    with open(os.path.join(workdir, "fix_backbone_coeff.data"), "a") as f:
        f.write("\n" + extras_text(terms, **kw))
    return out
