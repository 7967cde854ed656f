"""ctypes wrapper of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product path never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_fm.so")

ET_NAMES = ["Total", "Chain", "Shake", "Chi", "Rama", "Excluded", "DSSP", "P_AP", "Water", "Burial",
            "Helix", "AMH-Go", "Frag_Mem", "Vec_FM", "Contact_Restraints", "Membrane", "SSB", "Electro",
            "Bond"]
N_ET = 19
ET = {name: k for k, name in enumerate(ET_NAMES)}
ALL_TERMS = 0xFFFFFFFF
BRANCHES = ["dssp_class1", "dssp_class2", "dssp_class3", "dssp_class4", "dssp_beta", "dssp_rep", "dssp_ap", "dssp_p",
            "dssp_nu_i", "dssp_nu_j", "dssp_nu_exit", "pap_medium", "pap_long", "pap_long_interchain", "pap_parallel",
            "pap_parallel_interchain", "pap_beta", "helix"]

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.or_create.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
        L.or_create.restype = C.c_int
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_n_residues.argtypes = [C.c_void_p]
        L.or_compute.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, dp, dp, C.c_uint]
        L.or_compute.restype = C.c_int
        L.or_list.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp, dp,
                              np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"), C.c_int]
        L.or_list.restype = C.c_int
        L.or_pair_list_cutoff.argtypes = [C.c_void_p]
        L.or_pair_list_cutoff.restype = C.c_double
        L.or_fm_count.argtypes = [C.c_void_p]
        L.or_fm_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.or_fm_rf.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_fm_rf.restype = C.c_double
        L.or_fm_se.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.or_fm_se.restype = C.c_char
        L.or_fm_gamma.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_fm_gamma.restype = C.c_double
        L.or_fm_minsep.argtypes = [C.c_void_p]
        L.or_fm_maxsep.argtypes = [C.c_void_p]
        L.or_fm_map.argtypes = [C.c_void_p, C.c_int, np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"), C.c_int]
        L.or_theta_pair.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_pap_nu.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_set_vexcluded.argtypes = [C.c_void_p, C.c_int, dp, dp, dp]
        L.or_set_bonds.argtypes = [C.c_void_p, C.c_double, dp]
        L.or_set_list_skin.argtypes = [C.c_void_p, C.c_double]
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.or_rng_seed.argtypes = [C.c_void_p, C.c_uint]
        L.or_rng_next.argtypes = [C.c_void_p]
        L.or_rng_next.restype = C.c_int
        L.or_tert_frust.argtypes = [C.c_void_p, dp, dp, dp, dp, C.c_int]
        L.or_tert_frust.restype = C.c_int
        L.or_nmer_frust.argtypes = [C.c_void_p, dp, dp, dp, dp, C.c_int, dp, C.c_int, C.POINTER(C.c_int)]
        L.or_nmer_frust.restype = C.c_int
        L.or_frag_frust.argtypes = [C.c_void_p, dp, dp, dp, C.c_longlong, dp, dp, C.POINTER(C.c_int)]
        L.or_frag_frust.restype = C.c_int
        L.or_frag_frust_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_q_wolynes.argtypes = [C.c_int, dp, dp, C.c_int, C.c_double]
        L.or_q_wolynes.restype = C.c_double
        L.or_q_onuchic.argtypes = [C.c_int, dp, C.c_int, dp, C.c_double, C.c_int, ip, ip, dp, C.c_double]
        L.or_q_onuchic.restype = C.c_double
        L.or_langevin_run.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_ulonglong, C.c_int, C.c_longlong, dp, dp, dp, dp]
        L.or_langevin_run.restype = C.c_int
        L.or_branch_counts.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"), C.c_int,
                                       C.c_int]
        L.or_last_error.argtypes = [C.c_void_p]
        L.or_last_error.restype = C.c_char_p
        L.or_fm_table_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_fm_table.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.or_fm_table.restype = C.c_int
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


class Oracle:
    """FP64 FixBackbone restatement on one system (all residues local)."""

    def __init__(self, coeff: str, seq: str, workdir: str):
        L = lib()
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = L.or_create(coeff.encode(), seq.encode(), workdir.encode(), C.byref(h), err, 512)
        if rc != 0:
            raise OracleError(err.value.decode())
        self.h = h
        self.n = L.or_n_residues(h)

    def close(self):
        if self.h:
            lib().or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bonds(self, k: float, r0):
        """bond_style harmonic + bond_coeff: K, and r0 of CA-CA, CA-O, O-CA(+1), CA-CB, CA-H-beta."""
        lib().or_set_bonds(self.h, float(k), np.ascontiguousarray(r0, dtype=np.float64))

    def set_list_skin(self, skin: float):
        """Neighbour-list skin of the list-limited terms (Debye-Hueckel, the
        solvent barrier's CA-CA pairs, AMH-Go): pairs closer than
        pair_list_cutoff + skin count (a list just built by LAMMPS)."""
        lib().or_set_list_skin(self.h, float(skin))

    def set_vexcluded(self, p: int, lam, cut_short, cut_long):
        """pair_style vexcluded p + pair_coeff: [6][6] tables by atom type (1 CA, 2 N, 3 O, 4 CB, 5 H-beta)."""
        f = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
        lib().or_set_vexcluded(self.h, int(p), f(lam), f(cut_short), f(cut_long))

    @property
    def pair_list_cutoff(self) -> float:
        return lib().or_pair_list_cutoff(self.h)

    def compute(self, coords: np.ndarray, mask: int = ALL_TERMS):
        """coords (3, n, 3) [CA, CB, O]; returns forces (3, n, 3), energy (19,)."""
        x = np.ascontiguousarray(coords, dtype=np.float64)
        f = np.zeros_like(x)
        e = np.zeros(N_ET)
        fca, fcb, fo = np.zeros((self.n, 3)), np.zeros((self.n, 3)), np.zeros((self.n, 3))
        rc = lib().or_compute(self.h, np.ascontiguousarray(x[0]), np.ascontiguousarray(x[1]),
                              np.ascontiguousarray(x[2]), fca, fcb, fo, e, mask & 0xFFFFFFFF)
        if rc:
            raise RuntimeError(lib().or_last_error(self.h).decode())
        f[0], f[1], f[2] = fca, fcb, fo
        return f, e

    def rng_seed(self, seed: int):
        lib().or_rng_seed(self.h, seed)

    def rng_next(self) -> int:
        return lib().or_rng_next(self.h)

    def tert_frust(self, coords: np.ndarray) -> np.ndarray:
        """One tertiary-frustratometer call: rows (i, j, i_chain, j_chain,
        r_ij, rho_i, rho_j, native, decoy mean, decoy std, index)."""
        x = np.ascontiguousarray(coords, dtype=np.float64)
        cap = 4096
        while True:
            out = np.zeros(11 * cap)
            k = lib().or_tert_frust(self.h, np.ascontiguousarray(x[0]), np.ascontiguousarray(x[1]),
                                    np.ascontiguousarray(x[2]), out, cap)
            if k >= 0:
                return out[:11 * k].reshape(k, 11)
            cap = -k + 16

    def nmer_frust(self, coords: np.ndarray):
        """One n-mer-frustratometer call: (rows, traps); rows (i, j,
        contacts, native, decoy mean, decoy std, index), trap records (row,
        first, second, k, backward, threshold, energy)."""
        x = np.ascontiguousarray(coords, dtype=np.float64)
        cap, tcap = 4096, 4096
        while True:
            out, tr = np.zeros(7 * cap), np.zeros(7 * tcap)
            nt = C.c_int()
            k = lib().or_nmer_frust(self.h, np.ascontiguousarray(x[0]), np.ascontiguousarray(x[1]),
                                    np.ascontiguousarray(x[2]), out, cap, tr, tcap, C.byref(nt))
            if k >= 0 and nt.value >= 0:
                return out[:7 * k].reshape(k, 7), tr[:7 * nt.value].reshape(nt.value, 7)
            cap, tcap = max(cap, -k + 16), max(tcap, -nt.value + 16)

    def frag_frust(self, coords: np.ndarray, step: int = 0):
        """One fragment-frustratometer calculation: (rows, gen); rows n x 3
        (index, native - mean, std per residue); gen (read mode, at the call
        that computes the decoy statistics) (native energies n, decoy
        energies n x n_decoys), else None."""
        x = np.ascontiguousarray(coords, dtype=np.float64)
        n = x.shape[1]
        mode, nc, nd = C.c_int(), C.c_int(), C.c_int()
        lib().or_frag_frust_dims(self.h, C.byref(mode), C.byref(nc), C.byref(nd))
        rows = np.zeros(3 * n)
        gen = np.zeros(n + n * max(nd.value, 1))
        w = C.c_int()
        rc = lib().or_frag_frust(self.h, np.ascontiguousarray(x[0]), np.ascontiguousarray(x[1]),
                                 np.ascontiguousarray(x[2]), int(step), rows, gen, C.byref(w))
        if rc:
            raise RuntimeError("or_frag_frust: error %d" % rc)
        g = (gen[:n].copy(), gen[n:].reshape(n, nd.value)) if w.value else None
        return rows.reshape(n, 3), g

    def fm_table_dims(self):
        """(grid points per curve, neighbours per residue) of fragment-memory
        table mode; (0, 0) without it."""
        a, b = C.c_int(), C.c_int()
        lib().or_fm_table_dims(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def fm_table(self, itb: int):
        """curve itb = 4 tb_nbrs i + 4 (j - i - minSep) + k: (energy, force)
        on the grid, or None where no memory contributes."""
        size, _ = self.fm_table_dims()
        e, f = np.zeros(size), np.zeros(size)
        return (e, f) if lib().or_fm_table(self.h, itb, e, f) else None

    def energy(self, coords, mask=ALL_TERMS):
        return self.compute(coords, mask)[1]

    def list(self, kind: int, coords: np.ndarray, arg: int = 0) -> np.ndarray:
        x = np.ascontiguousarray(coords, dtype=np.float64)
        cap = 1 << 16
        while True:
            out = np.zeros(cap, dtype=np.int64)
            k = lib().or_list(self.h, kind, arg, np.ascontiguousarray(x[0]), np.ascontiguousarray(x[1]),
                              np.ascontiguousarray(x[2]), out, cap)
            if k >= 0:
                return out[:k].copy()
            cap = -k + 16

    def branches(self, reset=True) -> dict:
        """Branch counters since the last reset (see awsem_oracle.h)."""
        out = np.zeros(64, dtype=np.int64)
        n = lib().or_branch_counts(self.h, out, 64, int(reset))
        return dict(zip(BRANCHES, out[:n].tolist()))

    def theta_pair(self, r, helix=False, well=0):
        t, p = C.c_double(), C.c_double()
        lib().or_theta_pair(self.h, int(helix), well, r, C.byref(t), C.byref(p))
        return t.value, p.value

    def pap_nu(self, drsq):
        a, b = C.c_double(), C.c_double()
        lib().or_pap_nu(self.h, drsq, C.byref(a), C.byref(b))
        return a.value, b.value

    def fm_count(self):
        return lib().or_fm_count(self.h)

    def fm_info(self, k):
        p, l, w = C.c_int(), C.c_int(), C.c_double()
        lib().or_fm_info(self.h, k, C.byref(p), C.byref(l), C.byref(w))
        return p.value, l.value, w.value

    def fm_rf(self, k, ires, iatom, jres, jatom):
        return lib().or_fm_rf(self.h, k, ires, iatom, jres, jatom)

    def fm_se(self, k, resno):
        return lib().or_fm_se(self.h, k, resno).decode()

    def fm_map(self, i):
        out = np.zeros(4096, dtype=np.int32)
        k = lib().or_fm_map(self.h, i, out, 4096)
        return out[:k].copy()

    def langevin(self, coords, vel, force, nsteps, dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0,
                 seed=1, replica=0, step0=0):
        x = np.ascontiguousarray(coords, dtype=np.float64).copy()
        v = np.ascontiguousarray(vel, dtype=np.float64).copy()
        f = np.ascontiguousarray(force, dtype=np.float64).copy()
        e = np.zeros(N_ET)
        rc = lib().or_langevin_run(self.h, nsteps, dt, t_start, t_stop, damp, seed, replica, step0,
                                   x.reshape(-1), v.reshape(-1), f.reshape(-1), e)
        if rc:
            raise RuntimeError(lib().or_last_error(self.h).decode())
        return x, v, f, e


def q_wolynes(xca, rnative, sep=2, sigma_exp=0.15):
    """compute qwolynes of a CA trace (n x 3) against n x n native distances."""
    x = np.ascontiguousarray(xca, dtype=np.float64)
    rn = np.ascontiguousarray(rnative, dtype=np.float64)
    return lib().or_q_wolynes(len(x), x.reshape(-1), rn.reshape(-1), int(sep), float(sigma_exp))


def q_onuchic(xca, kind, native_xyz=None, r_contact=0.0, contacts=None, factor=1.0):
    """compute qonuchic; kind 'cutoff', 'cutoff/gauss' (native_xyz, r_contact)
    or 'shadow', 'shadow/gauss' (contacts: (i, j, rn) 0-based)."""
    t = {"cutoff": 0, "cutoff/gauss": 1, "shadow": 2, "shadow/gauss": 3}[kind]
    x = np.ascontiguousarray(xca, dtype=np.float64)
    nat = np.ascontiguousarray(native_xyz if native_xyz is not None else np.zeros((1, 3)), dtype=np.float64)
    c = contacts if contacts is not None else []
    mi = np.ascontiguousarray([a for a, _, _ in c] or [0], dtype=np.int32)
    mj = np.ascontiguousarray([b for _, b, _ in c] or [0], dtype=np.int32)
    mr = np.ascontiguousarray([r for _, _, r in c] or [0.0], dtype=np.float64)
    return lib().or_q_onuchic(len(x), x.reshape(-1), t, nat.reshape(-1), float(r_contact), len(c), mi, mj, mr,
                              float(factor))
