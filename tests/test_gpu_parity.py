"""GPU parity: the HIP force path against the FP64 oracle, per term."""
import numpy as np
import pytest

from conftest import make_case, rel_force_err

pytestmark = pytest.mark.gpu

TERMS = ["Chain", "Chi", "Rama", "Excluded", "DSSP", "P_AP", "Water", "Burial", "Helix", "Frag_Mem", "Bond"]
# tolerances: FP64 path agrees to rounding; FP32 pair arithmetic with FP64
# accumulation must stay within 1e-5 relative (force: max-norm over atoms)
TOL = {0: 1e-9, 1: 1e-5}


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


@pytest.mark.parametrize("lengths,seed", [([40, 25], 5), ([120], 11)])
def test_plain_fp64_per_term(oracle_mod, lengths, seed):
    A = _gpu()
    s, p = make_case(lengths, seed)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    x = s.coords()
    for name in TERMS + ["ALL"]:
        mask = 0xFFFFFFFF if name == "ALL" else (1 << A.ET[name])
        fix.set_term_mask(mask)
        f, e = fix.compute_res(x)
        fo, eo = o.compute(x, mask)
        k = 0 if name == "ALL" else A.ET[name]
        print(name, e[k], eo[k])
        assert abs(e[k] - eo[k]) <= TOL[0] * max(1.0, abs(eo[k])), name
        if np.abs(fo).max() > 0:
            assert rel_force_err(f, fo) <= TOL[0], name


def test_batch_fp32_resident(oracle_mod):
    A = _gpu()
    cases = [make_case([L], 100 + k) for k, L in enumerate([60, 37, 90])]
    fixes = [A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP32) for _, p in cases]
    b = A.Batch(fixes, A.PREC_FP32)
    b.set_state(A.pack_states([s.coords() for s, _ in cases]))
    b.compute()
    x, v, f, e = b.get_state()
    fs = b.split(f)
    for r, (s, p) in enumerate(cases):
        o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
        fo, eo = o.compute(s.coords())
        err = rel_force_err(fs[r], fo)
        print("replica", r, "E", e[r, 0], eo[0], "ferr", err)
        assert err <= TOL[1]
        for name in TERMS:
            k = A.ET[name]
            assert abs(e[r, k] - eo[k]) <= TOL[1] * max(1.0, abs(eo[k])), (r, name, e[r, k], eo[k])


@pytest.mark.parametrize("kind,arg", [(0, 0), (0, 1), (1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (6, 0), (6, 1)])
def test_lists_bit_exact(oracle_mod, kind, arg):
    A = _gpu()
    s, p = make_case([70, 30], 21)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    x = s.coords()
    g = fix.list(kind, x, arg)
    r = o.list(kind, x, arg)
    assert np.array_equal(g, r), (kind, arg, len(g), len(r))
