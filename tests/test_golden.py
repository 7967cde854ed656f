"""The oracle against vectors the reference's own code produced
(tests/golden/reference_vectors.npz, written by tests/golden/make_golden.py
from src/smart_matrix_lib.h and src/fragment_memory.cpp compiled into
oracle/_ref): switching functions, the P_AP switch, gamma-class tables and
fragment-memory Rf tables.  Runs without the reference present."""
import os
import shutil

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden", "reference_vectors.npz")
TABLES = ["anti_HB", "anti_NHB", "para_HB", "anti_one", "para_one", "gamma.dat", "burial_gamma.dat", "uniform.gamma"]
SEPS = (2, 3, 4, 5, 6, 7, 8, 9, 10, 15)


@pytest.fixture(scope="module")
def gold(tmp_path_factory, oracle_mod):
    g = np.load(GOLD, allow_pickle=False)
    d = tmp_path_factory.mktemp("golden")
    for t in TABLES:
        shutil.copy(os.path.join(ROOT, "awsemmd_amd", "params", t), d / t)
    for key in g.files:
        if key.startswith("file_"):
            name = key[5:]
            for ext in ("_seq", "_data", "_mem", "_gro"):
                if name.endswith(ext):
                    name = name[: -len(ext)] + "." + ext[1:]
            (d / name).write_text(str(g[key]))
    o = oracle_mod.Oracle(str(d / "fix_backbone_coeff.data"), str(d / "protein.seq"), str(d))
    return g, d, o


def test_switching_functions_match_reference(gold):
    g, d, o = gold
    for k, r in enumerate(g["theta_r"]):
        a, b = o.theta_pair(r, well=0)
        assert abs(a - g["theta_water0"][k, 0]) <= 1e-15 and abs(b - g["theta_water0"][k, 1]) <= 1e-15
        a, b = o.theta_pair(r, helix=True)
        assert abs(a - g["theta_helix"][k, 0]) <= 1e-15 and abs(b - g["theta_helix"][k, 1]) <= 1e-15
    for k, r in enumerate(g["pap_r"]):
        nu, pr = o.pap_nu(r * r)
        assert abs(nu - g["pap_nu"][k, 0]) <= 1e-15 and abs(pr - g["pap_nu"][k, 1]) <= 1e-15


def test_fragment_memories_match_reference(gold):
    g, d, o = gold
    desc, rf, se = g["mem_desc"], g["mem_rf"], g["mem_se"]
    assert o.fm_count() == len(desc)
    for k, (pos, fpos, ln, st) in enumerate(desc):
        assert o.fm_info(k)[:2] == (pos, ln)
        assert "".join(o.fm_se(k, pos + a) for a in range(ln)) == str(se[k])
        for a in range(ln):
            for b in range(ln):
                if a == b:
                    continue
                for ka in (1, 2):
                    for kb in (1, 2):
                        assert o.fm_rf(k, pos + a, ka, pos + b, kb) == rf[k, a, ka - 1, b, kb - 1]


@pytest.mark.parametrize("name", ["four", "twenty", "uniform"])
def test_gamma_tables_match_reference(gold, oracle_mod, name):
    g, d, o = gold
    gp = d / f"g_{name}.gamma"
    gp.write_text(str(g["gamma_text_" + name]))
    coeff = (d / "fix_backbone_coeff.data").read_text().replace("uniform.gamma", str(gp))
    cf = d / f"coeff_{name}"
    cf.write_text(coeff)
    oo = oracle_mod.Oracle(str(cf), str(d / "protein.seq"), str(d))
    L = oracle_mod.lib()
    assert (L.or_fm_minsep(oo.h), L.or_fm_maxsep(oo.h)) == tuple(g["gamma_seps_" + name])
    tab = g["gamma_" + name]
    for it in range(20):
        for jt in range(20):
            for q, sp in enumerate(SEPS):
                assert L.or_fm_gamma(oo.h, it, jt, 10, 10 + sp) == tab[it, jt, q]
