"""awsemmd_amd -- MI355X (gfx950) FixBackbone force path of AWSEM-MD.

Python mirror of the reference's FixBackbone interface (src/fix_backbone.h:
33-49) over the C ABI in include/awsem.h, implemented by the in-tree HIP
library libawsem_hip.so.  There is no CPU fallback: loading fails loudly if
the library is missing.

    fix = FixBackbone("fix_backbone_coeff.data", "protein.seq", workdir)
    fix.set_topology(group, mol, res)        # Construct_Computational_Arrays
    e = fix.pre_force(x, f)                  # f += forces; per-term energies
    fix.compute_vector(k)                    # == energy[k+1] (fb.cpp:8223-8232)
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# AWSEM_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("AWSEM_LIB") or os.path.join(HERE, "libawsem_hip.so")

ET_NAMES = ["Total", "Chain", "Shake", "Chi", "Rama", "Excluded", "DSSP", "P_AP", "Water", "Burial",
            "Helix", "AMH-Go", "Frag_Mem", "Vec_FM", "Contact_Restraints", "Membrane", "SSB", "Electro",
            "Bond"]
ET = {name: k for k, name in enumerate(ET_NAMES)}
N_ET = len(ET_NAMES)
PREC_FP64, PREC_FP32 = 0, 1

ERRORS = {-1: "AWSEM_E_ARG", -2: "AWSEM_E_FILE", -3: "AWSEM_E_FORMAT", -4: "AWSEM_E_UNSUPPORTED",
          -5: "AWSEM_E_TOPOLOGY", -6: "AWSEM_E_MISSING_ATOM", -7: "AWSEM_E_DEVICE", -8: "AWSEM_E_NOMEM",
          -9: "AWSEM_E_NONFINITE", -10: "AWSEM_E_STATE", -11: "AWSEM_E_RANGE"}

EXPORTS = ["awsem_create", "awsem_create_arrays", "awsem_set_topology", "awsem_compute", "awsem_compute_res",
           "awsem_set_vexcluded", "awsem_set_bonds", "awsem_set_list_skin", "awsem_set_term_mask", "awsem_list", "awsem_run",
           "awsem_n_residues", "awsem_pair_list_cutoff", "awsem_last_error", "awsem_destroy",
           "awsem_batch_create", "awsem_batch_size", "awsem_batch_set_state", "awsem_batch_get_state",
           "awsem_batch_compute", "awsem_batch_run", "awsem_batch_timing", "awsem_batch_phase_clock",
           "awsem_batch_list_stats", "awsem_batch_share_stats", "awsem_batch_last_error", "awsem_batch_use_graph", "awsem_batch_set_skin",
           "awsem_batch_autotune", "awsem_batch_set_engine",
           "awsem_server_create", "awsem_server_submit", "awsem_server_pump", "awsem_server_start",
           "awsem_server_stop", "awsem_server_poll", "awsem_server_wait", "awsem_server_result",
           "awsem_server_release", "awsem_server_stats", "awsem_server_last_error", "awsem_server_destroy",
           "awsem_batch_destroy", "awsem_compute_res_noalloc", "awsem_batch_lists", "awsem_set_box",
           "awsem_compute_image", "awsem_energy_log_open", "awsem_energy_log_write", "awsem_energy_log_close",
           "awsem_deck_open", "awsem_deck_run", "awsem_deck_natoms", "awsem_deck_state",
           "awsem_deck_residue_atoms", "awsem_deck_variable", "awsem_deck_last_error", "awsem_deck_destroy",
           "awsem_q_wolynes_create", "awsem_q_onuchic_create", "awsem_q_compute", "awsem_batch_q",
           "awsem_q_last_error", "awsem_q_destroy", "awsem_tert_frust", "awsem_tert_frust_write",
           "awsem_tert_frust_every", "awsem_nmer_frust", "awsem_nmer_frust_write", "awsem_nmer_frust_every",
           "awsem_frag_frust", "awsem_frag_frust_write", "awsem_frag_frust_info"]


class AwsemError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Langevin(C.Structure):
    _fields_ = [("dt", C.c_double), ("t_start", C.c_double), ("t_stop", C.c_double), ("damp", C.c_double),
                ("seed", C.c_uint64), ("step0", C.c_int64)]


class DeckOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("quiet", C.c_int), ("max_steps", C.c_int64)]


class ServerOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("max_jobs", C.c_int), ("steps_per_round", C.c_int),
                ("hbm_budget_bytes", C.c_double), ("dt", C.c_double), ("damp", C.c_double)]


_lib = None


def lib():
    """Loads libawsem_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `python -m awsemmd_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.awsem_create.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.POINTER(vp)]
        L.awsem_create_arrays.argtypes = [C.c_char_p, C.c_char_p, C.c_int, ip, C.c_char_p, dp, C.c_int, ip, dp,
                                          C.c_char_p, C.c_int, ip, ip, ip, ip, dp, C.c_int, C.POINTER(vp)]
        L.awsem_set_topology.argtypes = [vp, C.c_int, ip, ip, ip]
        L.awsem_compute.argtypes = [vp, dp, dp, dp]
        L.awsem_compute_res.argtypes = [vp, dp, dp, dp]
        L.awsem_compute_res_noalloc.argtypes = [vp, dp, dp, dp]
        L.awsem_set_box.argtypes = [vp, ip, dp]
        L.awsem_compute_image.argtypes = [vp, dp, C.POINTER(C.c_int32), dp, dp]
        L.awsem_energy_log_open.argtypes = [vp, C.c_char_p]
        L.awsem_energy_log_write.argtypes = [vp, C.c_int64, dp]
        L.awsem_energy_log_close.argtypes = [vp]
        L.awsem_set_vexcluded.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp]
        L.awsem_set_bonds.argtypes = [vp, C.c_int, C.c_double, dp]
        L.awsem_set_term_mask.argtypes = [vp, C.c_uint32]
        L.awsem_set_list_skin.argtypes = [vp, C.c_double]
        L.awsem_list.argtypes = [vp, C.c_int, C.c_int, dp, C.POINTER(C.c_int64), C.c_int, ip]
        L.awsem_run.argtypes = [vp, C.c_int, C.POINTER(Langevin), dp, dp, dp, dp]
        L.awsem_n_residues.argtypes = [vp, ip]
        L.awsem_pair_list_cutoff.argtypes = [vp]
        L.awsem_pair_list_cutoff.restype = C.c_double
        L.awsem_last_error.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.awsem_destroy.argtypes = [vp]
        L.awsem_batch_create.argtypes = [C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp)]
        L.awsem_batch_size.argtypes = [vp, ip, ip]
        L.awsem_batch_set_state.argtypes = [vp, dp, dp]
        L.awsem_batch_get_state.argtypes = [vp, dp, dp, dp, dp]
        L.awsem_batch_compute.argtypes = [vp]
        L.awsem_batch_run.argtypes = [vp, C.c_int, C.POINTER(Langevin)]
        L.awsem_batch_timing.argtypes = [vp, dp, dp, ip]
        L.awsem_batch_last_error.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.awsem_batch_phase_clock.argtypes = [vp, C.POINTER(C.c_uint64)]
        L.awsem_batch_list_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), ip]
        L.awsem_batch_share_stats.argtypes = [vp, ip, C.POINTER(C.c_longlong)]
        L.awsem_batch_use_graph.argtypes = [vp, C.c_int]
        L.awsem_batch_set_skin.argtypes = [vp, C.c_double]
        L.awsem_batch_autotune.argtypes = [vp, C.POINTER(Langevin), C.c_int, dp, ip, ip, dp]
        L.awsem_batch_set_engine.argtypes = [vp, C.c_int]
        L.awsem_batch_lists.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int64), C.c_int, ip]
        u64 = C.c_uint64
        L.awsem_server_create.argtypes = [vp, C.POINTER(ServerOpts), C.POINTER(vp)]
        L.awsem_server_submit.argtypes = [vp, vp, dp, dp, C.c_int, C.c_double, C.c_double, u64, C.POINTER(u64)]
        L.awsem_server_pump.argtypes = [vp, ip, ip]
        L.awsem_server_start.argtypes = [vp]
        L.awsem_server_stop.argtypes = [vp]
        L.awsem_server_poll.argtypes = [vp, u64, ip]
        L.awsem_server_wait.argtypes = [vp, u64, C.c_double]
        L.awsem_server_result.argtypes = [vp, u64, dp, dp, dp]
        L.awsem_server_release.argtypes = [vp, u64]
        L.awsem_server_stats.argtypes = [vp, dp]
        L.awsem_server_last_error.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.awsem_server_destroy.argtypes = [vp]
        L.awsem_batch_destroy.argtypes = [vp]
        i64p = C.POINTER(C.c_int64)
        L.awsem_deck_open.argtypes = [C.c_char_p, C.POINTER(DeckOpts), C.POINTER(vp)]
        L.awsem_deck_run.argtypes = [vp]
        L.awsem_deck_natoms.argtypes = [vp, ip, ip]
        L.awsem_deck_state.argtypes = [vp, i64p, dp, dp, dp, i64p]
        L.awsem_deck_residue_atoms.argtypes = [vp, i64p]
        L.awsem_deck_variable.argtypes = [vp, C.c_char_p, dp]
        L.awsem_deck_last_error.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.awsem_deck_destroy.argtypes = [vp]
        L.awsem_q_wolynes_create.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, C.POINTER(vp)]
        L.awsem_q_onuchic_create.argtypes = [C.c_int, C.c_char_p, C.c_int, C.c_double, C.c_double, C.POINTER(vp)]
        L.awsem_q_compute.argtypes = [vp, dp, dp]
        L.awsem_batch_q.argtypes = [vp, vp, C.c_int, C.c_int, dp]
        L.awsem_q_last_error.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.awsem_q_destroy.argtypes = [vp]
        L.awsem_tert_frust.argtypes = [vp, dp, dp, C.c_int, C.POINTER(C.c_int)]
        L.awsem_tert_frust_write.argtypes = [vp, dp, C.c_int64]
        L.awsem_tert_frust_every.argtypes = [vp, C.POINTER(C.c_int64)]
        L.awsem_nmer_frust.argtypes = [vp, dp, dp, C.c_int, C.POINTER(C.c_int), dp, C.c_int, C.POINTER(C.c_int)]
        L.awsem_nmer_frust_write.argtypes = [vp, dp, C.c_int64]
        L.awsem_nmer_frust_every.argtypes = [vp, C.POINTER(C.c_int64)]
        L.awsem_frag_frust.argtypes = [vp, dp, dp, dp, ip]
        L.awsem_frag_frust_write.argtypes = [vp, dp, C.c_int64]
        L.awsem_frag_frust_info.argtypes = [vp, ip, ip, ip, C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def _dp(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


class FixBackbone:
    """One system (the LAMMPS `fix ID CA backbone CB O coeff seq` instance)."""

    def __init__(self, coeff: str, seq: str, workdir: str = ".", precision: int = PREC_FP64):
        h = C.c_void_p()
        rc = lib().awsem_create(coeff.encode(), seq.encode(), workdir.encode(), precision, C.byref(h))
        self._finish(h, rc)

    @classmethod
    def from_arrays(cls, coeff: str, param_dir: str, chains: Sequence[str], aps: Optional[np.ndarray],
                    structs_nm: Sequence[np.ndarray], struct_seqs: Sequence[str], mem_struct, mem_tpos, mem_fpos,
                    mem_len, mem_weight, precision: int = PREC_FP32):
        """Arrays in place of the sequence/ssweight/.mem/.gro files:
        structs_nm[s] is (len_s, 2, 3) CA/CB in nm, memories reference
        (structure, target pos, structure pos, length, weight), 0-based."""
        self = cls.__new__(cls)
        chain_len = np.array([len(c) for c in chains], dtype=np.int32)
        seq = "".join(chains).encode()
        a = None if aps is None else np.ascontiguousarray(aps, dtype=np.float64)
        slen = np.array([len(q) for q in struct_seqs], dtype=np.int32)
        sx = np.ascontiguousarray(np.concatenate([np.asarray(q, dtype=np.float64).reshape(-1) for q in structs_nm])
                                  if len(structs_nm) else np.zeros(1))
        sseq = "".join(struct_seqs).encode()
        arrs = [np.ascontiguousarray(v, dtype=np.int32) for v in (mem_struct, mem_tpos, mem_fpos, mem_len)]
        mw = np.ascontiguousarray(mem_weight, dtype=np.float64)
        h = C.c_void_p()
        rc = lib().awsem_create_arrays(coeff.encode(), param_dir.encode(), len(chains), _ip(chain_len), seq, _dp(a),
                                       len(slen), _ip(slen), _dp(sx), sseq, len(mw), *[_ip(v) for v in arrs], _dp(mw),
                                       precision, C.byref(h))
        self._finish(h, rc)
        return self

    def _finish(self, h, rc):
        self.h = h
        if rc != 0:
            msg = self.last_error() if h.value else ""
            if h.value:
                lib().awsem_destroy(h)
            self.h = C.c_void_p()
            raise AwsemError(rc, msg)
        n = C.c_int()
        lib().awsem_n_residues(h, C.byref(n))
        self.n = n.value
        self.energy = np.zeros(N_ET)

    def last_error(self) -> str:
        buf = C.create_string_buffer(1024)
        lib().awsem_last_error(self.h, buf, 1024)
        return buf.value.decode()

    def _check(self, rc):
        if rc != 0:
            raise AwsemError(rc, self.last_error())

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().awsem_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def pair_list_cutoff(self) -> float:
        return lib().awsem_pair_list_cutoff(self.h)

    def set_topology(self, group, mol, res):
        g = np.ascontiguousarray(group, dtype=np.int32)
        m = np.ascontiguousarray(mol, dtype=np.int32)
        r = np.ascontiguousarray(res, dtype=np.int32)
        self._check(lib().awsem_set_topology(self.h, len(g), _ip(g), _ip(m), _ip(r)))

    def set_term_mask(self, mask: int):
        self._check(lib().awsem_set_term_mask(self.h, mask & 0xFFFFFFFF))

    def tert_frust(self, xres: np.ndarray) -> np.ndarray:
        """The tertiary frustratometer at xres (3, n, 3): rows (i, j, i_chain,
        j_chain, r_ij, rho_i, rho_j, native, decoy mean, decoy std, index)."""
        x = np.ascontiguousarray(xres, dtype=np.float64)
        cap = 4 * self.n * self.n + 16
        out = np.zeros(11 * cap)
        k = C.c_int()
        self._check(lib().awsem_tert_frust(self.h, _dp(x), _dp(out), cap, C.byref(k)))
        return out[:11 * k.value].reshape(k.value, 11)

    def tert_frust_write(self, xres: np.ndarray, step: int):
        x = np.ascontiguousarray(xres, dtype=np.float64)
        self._check(lib().awsem_tert_frust_write(self.h, _dp(x), step))

    def nmer_frust(self, xres: np.ndarray):
        """The n-mer frustratometer at xres (3, n, 3): (rows, traps); rows (i,
        j, contacts, native, decoy mean, decoy std, index), traps (row,
        first, second, k, backward, threshold, energy)."""
        x = np.ascontiguousarray(xres, dtype=np.float64)
        cap, tcap = self.n * self.n + 16, 4096
        while True:
            out, tr = np.zeros(7 * cap), np.zeros(7 * tcap)
            k, t = C.c_int(), C.c_int()
            rc = lib().awsem_nmer_frust(self.h, _dp(x), _dp(out), cap, C.byref(k), _dp(tr), tcap, C.byref(t))
            if rc == -1 and (k.value > cap or t.value > tcap):
                cap, tcap = max(cap, k.value), max(tcap, t.value)
                continue
            self._check(rc)
            return out[:7 * k.value].reshape(k.value, 7), tr[:7 * t.value].reshape(t.value, 7)

    def nmer_frust_write(self, xres: np.ndarray, step: int):
        x = np.ascontiguousarray(xres, dtype=np.float64)
        self._check(lib().awsem_nmer_frust_write(self.h, _dp(x), step))

    def nmer_frust_every(self) -> int:
        """Output interval of the [Nmer_Frustratometer] section (0: none)."""
        k = C.c_int64()
        self._check(lib().awsem_nmer_frust_every(self.h, C.byref(k)))
        return k.value

    def frag_frust_info(self):
        """(mode 0 shuffle / 1 read / -1 none, decoy calculations, decoys,
        output interval) of the [Fragment_Frustratometer] section."""
        m, nc, nd, ev = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        self._check(lib().awsem_frag_frust_info(self.h, C.byref(m), C.byref(nc), C.byref(nd), C.byref(ev)))
        return m.value, nc.value, nd.value, ev.value

    def frag_frust(self, xres: np.ndarray):
        """One fragment-frustratometer calculation at xres (3, n, 3): (rows,
        gen); rows n x 3 (index, native - mean, std per residue); gen (read
        mode, at the call that computes the decoy statistics) (natives n,
        decoy energies n x n_decoys), else None."""
        x = np.ascontiguousarray(xres, dtype=np.float64)
        _, _, nd, _ = self.frag_frust_info()
        rows = np.zeros(3 * self.n)
        gen = np.zeros(self.n * (nd + 1))
        w = C.c_int()
        self._check(lib().awsem_frag_frust(self.h, _dp(x), _dp(rows), _dp(gen), C.byref(w)))
        g = (gen[:self.n].copy(), gen[self.n:].reshape(self.n, nd)) if w.value else None
        return rows.reshape(self.n, 3), g

    def frag_frust_write(self, xres: np.ndarray, step: int):
        x = np.ascontiguousarray(xres, dtype=np.float64)
        self._check(lib().awsem_frag_frust_write(self.h, _dp(x), step))

    def tert_frust_every(self) -> int:
        """Output interval of the [Tertiary_Frustratometer] section (0: none)."""
        k = C.c_int64()
        self._check(lib().awsem_tert_frust_every(self.h, C.byref(k)))
        return k.value

    def set_list_skin(self, skin: float):
        """Neighbour-list skin of the list-limited terms (Debye-Hueckel, the
        solvent barrier, AMH-Go): pairs closer than pair_list_cutoff + skin."""
        self._check(lib().awsem_set_list_skin(self.h, float(skin)))

    def set_vexcluded(self, enable: bool = True, p: int = 2, lam=None, cut_short=None, cut_long=None):
        arr = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(36) for a in (lam, cut_short, cut_long)]
        self._check(lib().awsem_set_vexcluded(self.h, int(enable), p, *[_dp(a) for a in arr]))

    def set_bonds(self, enable: bool = True, k: float = 200.0, r0=None):
        r = None if r0 is None else np.ascontiguousarray(r0, dtype=np.float64)
        self._check(lib().awsem_set_bonds(self.h, int(enable), k, _dp(r)))

    def set_box(self, periodic, prd=None):
        """Periodic box: flags per dimension and box lengths (domain->prd)."""
        pf = np.ascontiguousarray(periodic, dtype=np.int32)
        pr = None if prd is None else np.ascontiguousarray(prd, dtype=np.float64)
        self._check(lib().awsem_set_box(self.h, _ip(pf), _dp(pr)))

    def pre_force(self, x: np.ndarray, f: np.ndarray, image: Optional[np.ndarray] = None) -> np.ndarray:
        """FixBackbone::pre_force: f[n_atoms,3] += forces (atom layout); image
        (LAMMPS image flags per atom) is needed in a periodic box."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert f.dtype == np.float64 and f.flags.c_contiguous
        e = np.zeros(N_ET)
        if image is None:
            self._check(lib().awsem_compute(self.h, _dp(x), _dp(f), _dp(e)))
        else:
            im = np.ascontiguousarray(image, dtype=np.int32)
            self._check(lib().awsem_compute_image(self.h, _dp(x), im.ctypes.data_as(C.POINTER(C.c_int32)), _dp(f),
                                                  _dp(e)))
        self.energy = e
        return e

    def energy_log_open(self, path: str):
        self._check(lib().awsem_energy_log_open(self.h, path.encode()))

    def energy_log_write(self, step: int, energy: Optional[np.ndarray] = None):
        """Appends one energy.log line: the given 19 energies, or (None) the
        last evaluation's."""
        e = None if energy is None else np.ascontiguousarray(energy, dtype=np.float64)
        self._check(lib().awsem_energy_log_write(self.h, int(step), _dp(e)))

    def energy_log_close(self):
        self._check(lib().awsem_energy_log_close(self.h))

    def compute_res(self, xres: np.ndarray):
        """Residue layout (3, n, 3) [CA, CB, O]; returns (forces, energies)."""
        x = np.ascontiguousarray(xres, dtype=np.float64)
        f = np.zeros_like(x)
        e = np.zeros(N_ET)
        self._check(lib().awsem_compute_res(self.h, _dp(x), _dp(f), _dp(e)))
        self.energy = e
        return f, e

    def compute_res_noalloc(self, xres: np.ndarray, f: Optional[np.ndarray] = None):
        """compute_res without any device allocation in the call (per-call
        buffers in a fixed 4 MiB device array; the context's device image
        must exist, e.g. after one compute_res).  Forces are added into f."""
        x = np.ascontiguousarray(xres, dtype=np.float64)
        if f is None:
            f = np.zeros_like(x)
        assert f.dtype == np.float64 and f.flags.c_contiguous and f.shape == x.shape
        e = np.zeros(N_ET)
        self._check(lib().awsem_compute_res_noalloc(self.h, _dp(x), _dp(f), _dp(e)))
        self.energy = e
        return f, e

    def compute_scalar(self) -> float:
        return float(self.energy[0])

    def compute_vector(self, k: int) -> float:
        return float(self.energy[k + 1])

    def list(self, kind: int, xres: np.ndarray, arg: int = 0, cap: int = 1 << 20) -> np.ndarray:
        x = np.ascontiguousarray(xres, dtype=np.float64)
        out = np.zeros(cap, dtype=np.int64)
        cnt = C.c_int()
        self._check(lib().awsem_list(self.h, kind, arg, _dp(x), out.ctypes.data_as(C.POINTER(C.c_int64)), cap,
                                     C.byref(cnt)))
        return out[: cnt.value].copy()

    def run(self, xres, vres, nsteps, dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0, seed=1, step0=0):
        x = np.ascontiguousarray(xres, dtype=np.float64).copy()
        v = np.ascontiguousarray(vres, dtype=np.float64).copy()
        f = np.zeros_like(x)
        e = np.zeros(N_ET)
        lg = Langevin(dt, t_start, t_stop, damp, seed, step0)
        self._check(lib().awsem_run(self.h, nsteps, C.byref(lg), _dp(x), _dp(v), _dp(f), _dp(e)))
        self.energy = e
        return x, v, f, e


class Batch:
    """Independent replicas evaluated by the resident (workgroup-per-replica) engine."""

    def __init__(self, fixes: Sequence[FixBackbone], precision: int = PREC_FP32):
        arr = (C.c_void_p * len(fixes))(*[fx.h.value for fx in fixes])
        h = C.c_void_p()
        rc = lib().awsem_batch_create(len(fixes), arr, precision, C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.last_error() if h.value else ""
            raise AwsemError(rc, msg)
        nrep, nres = C.c_int(), C.c_int()
        lib().awsem_batch_size(h, C.byref(nrep), C.byref(nres))
        self.nrep, self.nres = nrep.value, nres.value
        self.sizes = [fx.n for fx in fixes]

    def last_error(self) -> str:
        buf = C.create_string_buffer(1024)
        lib().awsem_batch_last_error(self.h, buf, 1024)
        return buf.value.decode()

    def _check(self, rc):
        if rc != 0:
            raise AwsemError(rc, self.last_error())

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().awsem_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, x: Optional[np.ndarray] = None, v: Optional[np.ndarray] = None):
        xs = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        vs = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        self._check(lib().awsem_batch_set_state(self.h, _dp(xs), _dp(vs)))

    def get_state(self):
        n9 = self.nres * 9
        x, v, f = np.zeros(n9), np.zeros(n9), np.zeros(n9)
        e = np.zeros(self.nrep * N_ET)
        self._check(lib().awsem_batch_get_state(self.h, _dp(x), _dp(v), _dp(f), _dp(e)))
        return x, v, f, e.reshape(self.nrep, N_ET)

    def compute(self):
        self._check(lib().awsem_batch_compute(self.h))

    def run(self, nsteps, dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0, seed=1, step0=0):
        lg = Langevin(dt, t_start, t_stop, damp, seed, step0)
        self._check(lib().awsem_batch_run(self.h, nsteps, C.byref(lg)))

    def timing(self):
        a, b, k = C.c_double(), C.c_double(), C.c_int()
        lib().awsem_batch_timing(self.h, C.byref(a), C.byref(b), C.byref(k))
        return a.value, b.value, k.value

    def phase_clock(self):
        """Per-phase wall-clock ticks (100 MHz) when run with AWSEM_PHASE_PROF=1."""
        out = (C.c_uint64 * 32)()
        self._check(lib().awsem_batch_phase_clock(self.h, out))
        return np.array(out[:], dtype=np.float64)

    def use_graph(self, on: bool = True):
        self._check(lib().awsem_batch_use_graph(self.h, 1 if on else 0))

    def set_skin(self, skin: float):
        self._check(lib().awsem_batch_set_skin(self.h, skin))

    def autotune(self, probe_steps=40, dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0, seed=1, step0=0):
        """Fastest (engine, skin, steps per launch) on this batch: dict with
        skin, steps_per_launch, engine, ms_per_step.  The batch keeps the
        engine and the skin."""
        lg = Langevin(dt, t_start, t_stop, damp, seed, step0)
        sk, ch, en, ms = C.c_double(), C.c_int(), C.c_int(), C.c_double()
        self._check(lib().awsem_batch_autotune(self.h, C.byref(lg), probe_steps, C.byref(sk), C.byref(ch),
                                               C.byref(en), C.byref(ms)))
        return {"skin": sk.value, "steps_per_launch": ch.value, "engine": ["resident", "wide"][en.value],
                "ms_per_step": ms.value}

    def set_engine(self, engine: str):
        """'resident' (default) or 'wide'."""
        self._check(lib().awsem_batch_set_engine(self.h, {"resident": 0, "wide": 1}[engine]))

    def list_stats(self):
        """(list builds so far, pairs in the current lists, replicas without a valid list)."""
        b, p, k = C.c_longlong(), C.c_longlong(), C.c_int()
        self._check(lib().awsem_batch_list_stats(self.h, C.byref(b), C.byref(p), C.byref(k)))
        return b.value, p.value, k.value

    def share_stats(self):
        """(replicas whose fragment-memory chunks idle workgroups may take,
        chunks evaluated by another workgroup than the replica's own so far)."""
        n, c = C.c_int(), C.c_longlong()
        self._check(lib().awsem_batch_share_stats(self.h, C.byref(n), C.byref(c)))
        return n.value, c.value

    LIST_KINDS = {"verlet": 0, "dssp": 1, "pap": 2, "excluded": 3, "density": 4}

    def lists(self, replica: int, kind: str, cap: int = 1 << 22) -> np.ndarray:
        """The production pair selection of the last evaluation (sorted codes;
        see awsem_batch_lists)."""
        out = np.zeros(cap, dtype=np.int64)
        cnt = C.c_int()
        self._check(lib().awsem_batch_lists(self.h, replica, self.LIST_KINDS[kind],
                                            out.ctypes.data_as(C.POINTER(C.c_int64)), cap, C.byref(cnt)))
        return out[: cnt.value].copy()

    def split(self, flat: np.ndarray):
        """Splits a concatenated residue-layout array into per-replica (3, n, 3) arrays."""
        out, off = [], 0
        for n in self.sizes:
            out.append(flat[off * 9:(off + n) * 9].reshape(3, n, 3))
            off += n
        return out


class Server:
    """Continuous batching of annealing jobs (awsem_server_*): jobs are
    admitted between rounds as job slots and device memory free up; every
    round advances all active jobs in one resident launch."""

    STATES = {0: "queued", 1: "running", 2: "done"}

    def __init__(self, forcefield: FixBackbone, precision: int = PREC_FP32, max_jobs: int = 512,
                 steps_per_round: int = 50, hbm_budget_bytes: float = 0.0, dt: float = 2.0, damp: float = 1000.0):
        o = ServerOpts(precision, max_jobs, steps_per_round, hbm_budget_bytes, dt, damp)
        h = C.c_void_p()
        rc = lib().awsem_server_create(forcefield.h, C.byref(o), C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.last_error()
            self.close()
            raise AwsemError(rc, msg)

    def last_error(self) -> str:
        if not self.h:
            return ""
        buf = C.create_string_buffer(1024)
        lib().awsem_server_last_error(self.h, buf, 1024)
        return buf.value.decode()

    def _check(self, rc):
        if rc != 0:
            raise AwsemError(rc, self.last_error())

    def close(self):
        if getattr(self, "h", None):
            lib().awsem_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, system: FixBackbone, xres: np.ndarray, vres: Optional[np.ndarray], nsteps: int,
               t_start: float = 600.0, t_stop: float = 300.0, seed: int = 1) -> int:
        x = np.ascontiguousarray(xres, dtype=np.float64)
        v = None if vres is None else np.ascontiguousarray(vres, dtype=np.float64)
        jid = C.c_uint64()
        self._check(lib().awsem_server_submit(self.h, system.h, _dp(x), _dp(v), nsteps, t_start, t_stop,
                                              seed, C.byref(jid)))
        return jid.value

    def pump(self):
        """One scheduling round on this thread: (active, queued) afterwards."""
        a, q = C.c_int(), C.c_int()
        self._check(lib().awsem_server_pump(self.h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def start(self):
        self._check(lib().awsem_server_start(self.h))

    def stop(self):
        self._check(lib().awsem_server_stop(self.h))

    def poll(self, jid: int) -> int:
        st = C.c_int()
        self._check(lib().awsem_server_poll(self.h, jid, C.byref(st)))
        return st.value

    def wait(self, jid: int, timeout: float = 0.0):
        self._check(lib().awsem_server_wait(self.h, jid, timeout))

    def result(self, jid: int, n: int):
        """(coordinates (3, n, 3), 19 energies, {queued, device, total} ms)."""
        x = np.zeros((3, n, 3))
        e = np.zeros(N_ET)
        t = np.zeros(3)
        self._check(lib().awsem_server_result(self.h, jid, _dp(x), _dp(e), _dp(t)))
        return x, e, {"queued_ms": t[0], "device_ms": t[1], "total_ms": t[2]}

    def release(self, jid: int):
        self._check(lib().awsem_server_release(self.h, jid))

    def stats(self) -> dict:
        s = np.zeros(8)
        self._check(lib().awsem_server_stats(self.h, _dp(s)))
        keys = ["rounds", "jobs_done", "bytes_in_use", "peak_active", "replica_steps", "device_ms", "budget_bytes",
                "active"]
        return dict(zip(keys, s.tolist()))


class Deck:
    """A LAMMPS input deck of the FixBackbone path (awsem_deck_*): the script
    is executed up to its first `run` when the object is made (no device
    work), `run()` executes the rest on the device.  File names in the
    script are relative to its directory, where energy.log and log.awsem
    are written."""

    def __init__(self, path: str, precision: int = PREC_FP64, max_steps: int = 0, quiet: bool = True):
        o = DeckOpts(precision, 1 if quiet else 0, max_steps)
        h = C.c_void_p()
        rc = lib().awsem_deck_open(path.encode(), C.byref(o), C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.last_error()
            self.close()
            raise AwsemError(rc, msg)

    def last_error(self) -> str:
        if not self.h:
            return ""
        buf = C.create_string_buffer(1024)
        lib().awsem_deck_last_error(self.h, buf, 1024)
        return buf.value.decode()

    def _check(self, rc):
        if rc != 0:
            raise AwsemError(rc, self.last_error())

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().awsem_deck_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self):
        self._check(lib().awsem_deck_run(self.h))

    def natoms(self):
        na, nr = C.c_int(), C.c_int()
        self._check(lib().awsem_deck_natoms(self.h, C.byref(na), C.byref(nr)))
        return na.value, nr.value

    def state(self):
        """(ids, x (n, 3), v (n, 3), energies, step), atoms in id order."""
        na, _ = self.natoms()
        ids = np.zeros(na, dtype=np.int64)
        x, v, e = np.zeros((na, 3)), np.zeros((na, 3)), np.zeros(N_ET)
        st = C.c_int64()
        self._check(lib().awsem_deck_state(self.h, ids.ctypes.data_as(C.POINTER(C.c_int64)), _dp(x), _dp(v), _dp(e),
                                           C.byref(st)))
        return ids, x, v, e, st.value

    def residue_atoms(self) -> np.ndarray:
        """Atom ids of the residue layout slots, (3, n_res): CA, CB/H-beta, O."""
        _, nr = self.natoms()
        ids = np.zeros(3 * nr, dtype=np.int64)
        self._check(lib().awsem_deck_residue_atoms(self.h, ids.ctypes.data_as(C.POINTER(C.c_int64))))
        return ids.reshape(3, nr)

    def residue_state(self):
        """(x, v) in the residue layout (3, n_res, 3)."""
        ids, x, v, _, _ = self.state()
        pos = {int(i): k for k, i in enumerate(ids)}
        slots = self.residue_atoms()
        idx = np.vectorize(lambda i: pos[int(i)])(slots)
        return x[idx], v[idx]

    def variable(self, formula: str) -> float:
        out = C.c_double()
        self._check(lib().awsem_deck_variable(self.h, formula.encode(), C.byref(out)))
        return out.value


Q_KINDS = {"wolynes": 0, "cutoff": 1, "cutoff/gauss": 2, "shadow": 3, "shadow/gauss": 4}


class OrderParameter:
    """compute qwolynes / compute qonuchic of n CAs on the device
    (src/compute_q_wolynes.cpp, src/compute_q_onuchic.cpp).

    OrderParameter.wolynes("rnative.dat", n, sep=2, sigma_exp=0.15)
    OrderParameter.onuchic("cutoff", "native.dat", n, r_contact=8.0, factor=1.2)
    OrderParameter.onuchic("shadow/gauss", "contacts.map", n)"""

    def __init__(self, h):
        self.h = h

    @classmethod
    def wolynes(cls, path: str, n: int, sep: int = 2, sigma_exp: float = 0.15):
        h = C.c_void_p()
        rc = lib().awsem_q_wolynes_create(path.encode(), n, sep, sigma_exp, C.byref(h))
        q = cls(h)
        if rc != 0:
            raise AwsemError(rc, q.last_error())
        return q

    @classmethod
    def onuchic(cls, kind: str, path: str, n: int, r_contact: float = 0.0, factor: float = 1.0):
        h = C.c_void_p()
        rc = lib().awsem_q_onuchic_create(Q_KINDS[kind], path.encode(), n, r_contact, factor, C.byref(h))
        q = cls(h)
        if rc != 0:
            raise AwsemError(rc, q.last_error())
        return q

    def last_error(self) -> str:
        buf = C.create_string_buffer(512)
        if self.h.value:
            lib().awsem_q_last_error(self.h, buf, 512)
        return buf.value.decode()

    def _check(self, rc):
        if rc != 0:
            raise AwsemError(rc, self.last_error())

    def compute(self, xca: np.ndarray) -> float:
        x = np.ascontiguousarray(xca, dtype=np.float64)
        out = np.zeros(1)
        self._check(lib().awsem_q_compute(self.h, _dp(x), _dp(out)))
        return float(out[0])

    def batch(self, b: "Batch", first: int = 0, count: Optional[int] = None) -> np.ndarray:
        count = b.nrep - first if count is None else count
        out = np.zeros(count)
        rc = lib().awsem_batch_q(b.h, self.h, first, count, _dp(out))
        if rc != 0:
            raise AwsemError(rc, b.last_error())
        return out

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().awsem_q_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pack_states(states: Sequence[np.ndarray]) -> np.ndarray:
    """Concatenates per-replica (3, n, 3) arrays in batch order."""
    return np.concatenate([np.ascontiguousarray(s, dtype=np.float64).reshape(-1) for s in states])
