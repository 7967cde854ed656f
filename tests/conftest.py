import os
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


def make_case(lengths, seed, fm=True, **kw):
    """Generates a system and writes its reference file set into a temp dir."""
    from awsemmd_amd import workload as W
    s = W.make_system(lengths, seed=seed, **kw)
    d = tempfile.mkdtemp(prefix="awsem_case_")
    paths = W.write_system(s, d, fm=fm)
    return s, paths


def make_table_case(lengths, seed, tb=(1.0, 30.0, 0.02), well_width=0.1, use_table=0, sigma_exp=0.15, **kw):
    """make_case with the fragment memory in table mode: the coeff's
    [Fragment_Memory] section replaced by [Fragment_Memory_Table] with the
    same strength, memories and gamma file (fb.cpp:310-320)."""
    s, p = make_case(lengths, seed, **kw)
    txt = open(p["coeff"]).read()
    old = "[Fragment_Memory]\n1.0\nfragsLAMW.mem\nuniform.gamma"
    assert old in txt
    txt = txt.replace(old, "[Fragment_Memory_Table]\n1.0\nfragsLAMW.mem\nuniform.gamma\n%r %r %r\n%r\n%d\n%r"
                      % (tb[0], tb[1], tb[2], well_width, use_table, sigma_exp))
    with open(p["coeff"], "w") as f:
        f.write(txt)
    return s, p


def rel_force_err(f, fref):
    """max_atom |f - fref| / max_atom |fref| (per-term force error measure)."""
    d = np.linalg.norm((f - fref).reshape(-1, 3), axis=1).max()
    m = np.linalg.norm(fref.reshape(-1, 3), axis=1).max()
    return d / m if m > 0 else d


@pytest.fixture
def case_factory():
    return make_case
