"""GPU parity on systems with real secondary structure (tests/secondary.py:
an alpha helix, tight and loose two-chain sheets, hairpins), where the DSSP,
P_AP and helix terms carry real force: the HIP path against the FP64 oracle
per term, through the plain drop-in and both batch engines.

Force error of a term: max over atoms |F_gpu - F_oracle| / max over atoms
|F_oracle,term| (the term's own force; floor 1e-9 of the total only for
terms that vanish).  Per-atom error, also reported: max over atoms of
|F_gpu - F_oracle| / max(|F_oracle|, 1e-3 max |F_oracle,term|).  Energy error:
|E_gpu - E_oracle| / max(|E_oracle|, 1e-8).  Tolerances: FP64 1e-9, FP32
pair arithmetic 1e-5 (the per-atom error of FP32 is reported, and bounded
by 1e-4: atoms whose net force is a near-cancellation of larger
contributions carry the contributions' rounding)."""
import tempfile

import numpy as np
import pytest

import secondary

pytestmark = pytest.mark.gpu

TERMS = ["Chain", "Chi", "Rama", "Excluded", "DSSP", "P_AP", "Water", "Burial", "Helix", "Frag_Mem", "Bond"]
TOL = {0: 1e-9, 1: 1e-5}
ATOM_TOL = {0: 1e-9, 1: 1e-4}


@pytest.fixture(scope="module")
def cases(oracle_mod):
    from awsemmd_amd import workload as W
    out = []
    for name, s in secondary.all_systems().items():
        d = tempfile.mkdtemp(prefix="awsem_ss_")
        p = W.write_system(s, d)
        out.append((name, s, p, oracle_mod.Oracle(p["coeff"], p["seq"], d)))
    return out


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


def _errs(f, fo, ftot_max):
    d = np.linalg.norm((f - fo).reshape(-1, 3), axis=1)
    a = np.linalg.norm(fo.reshape(-1, 3), axis=1)
    m = a.max()
    ferr = d.max() / max(m, 1e-9 * ftot_max)
    aerr = (d / np.maximum(a, max(1e-3 * m, 1e-9 * ftot_max))).max()
    return ferr, aerr


def _check(A, name, label, prec, f, e, fo, eo, ftot, k):
    ferr, aerr = _errs(f, fo, ftot)
    eerr = abs(e - eo) / max(abs(eo), 1e-8)
    print(f"{name:26s} {label:9s} E={e:.10g} Eref={eo:.10g} eerr={eerr:.1e} ferr={ferr:.1e} atom={aerr:.1e}")
    assert eerr <= TOL[prec] and ferr <= TOL[prec] and aerr <= ATOM_TOL[prec], (name, label, eerr, ferr, aerr)


@pytest.mark.parametrize("prec", [0, 1])
def test_structured_plain_per_term(cases, prec):
    A = _gpu()
    for name, s, p, o in cases:
        print(f"system {name} ({s.n} residues), precision {prec}", flush=True)
        fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec)
        x = s.coords()
        ftot = np.linalg.norm(o.compute(x)[0].reshape(-1, 3), axis=1).max()
        for t in TERMS + ["ALL"]:
            mask = 0xFFFFFFFF if t == "ALL" else (1 << A.ET[t])
            k = 0 if t == "ALL" else A.ET[t]
            fix.set_term_mask(mask)
            f, e = fix.compute_res(x)
            fo, eo = o.compute(x, mask)
            _check(A, name, t, prec, f, e[k], fo, eo[k], ftot, k)


@pytest.mark.parametrize("engine", ["resident", "wide"])
@pytest.mark.parametrize("prec", [0, 1])
def test_structured_batch_per_term(cases, engine, prec):
    """All structured systems as one batch, per term, on both engines."""
    A = _gpu()
    xs = A.pack_states([s.coords() for _, s, _, _ in cases])
    ftots = [np.linalg.norm(o.compute(s.coords())[0].reshape(-1, 3), axis=1).max() for _, s, _, o in cases]
    for t in ["DSSP", "P_AP", "Helix", "ALL"]:
        mask = 0xFFFFFFFF if t == "ALL" else (1 << A.ET[t])
        k = 0 if t == "ALL" else A.ET[t]
        fixes = [A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec) for _, _, p, _ in cases]
        fixes[0].set_term_mask(mask)  # the batch takes the parameters of its first system
        b = A.Batch(fixes, prec)
        b.set_engine(engine)
        b.set_state(xs)
        b.compute()
        _, _, f, e = b.get_state()
        fs = b.split(f)
        for r, (name, s, p, o) in enumerate(cases):
            fo, eo = o.compute(s.coords(), mask)
            _check(A, name, f"{engine}:{t}", prec, fs[r], e[r, k], fo, eo[k], ftots[r], k)
