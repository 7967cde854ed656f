"""Fragment-memory table mode on the host (CPU, no GPU): the context built
from a [Fragment_Memory_Table] coeff holds the same curves as the oracle,
writes them to fm_table.energy / fm_table.force in the reference's format
when fm_use_table is set (output_fragment_memory_table, fb.cpp:6062-6106),
and reads them back instead of the memories when both files exist
(fb.cpp:746-758, read_fragment_memory_table 4179-4236)."""
import os

import numpy as np
import pytest

from conftest import make_table_case


def _A():
    import awsemmd_amd as A
    A.lib()
    return A


def test_written_tables_equal_the_oracles(oracle_mod):
    """4 n tb_nbrs rows of tb_size values each, "%.12f " per value: every row
    equals the oracle's curve (zeros where the oracle has none) to the
    printed precision."""
    A = _A()
    s, p = make_table_case([30, 24], 31, tb=(1.0, 25.0, 0.05), use_table=1)
    A.FixBackbone(p["coeff"], p["seq"], p["workdir"]).close()
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    size, nb = o.fm_table_dims()
    assert size == int((25.0 - 1.0) / 0.05) + 2 and nb == 9 - 3 + 1
    rows = {}
    for name in ("energy", "force"):
        lines = open(os.path.join(p["workdir"], "fm_table." + name)).read().split("\n")
        assert lines[-1] == "" and len(lines) - 1 == 4 * s.n * nb
        assert all(ln.endswith(" ") for ln in lines[:-1])
        rows[name] = np.array([[float(v) for v in ln.split()] for ln in lines[:-1]])
        assert rows[name].shape == (4 * s.n * nb, size)
    present = 0
    for itb in range(4 * s.n * nb):
        t = o.fm_table(itb)
        if t is None:
            assert not rows["energy"][itb].any() and not rows["force"][itb].any(), itb
            continue
        present += 1
        for got, want in ((rows["energy"][itb], t[0]), (rows["force"][itb], t[1])):
            want = np.where(np.isinf(want), 0.0, want)
            assert np.abs(got - want).max() <= 6e-13, (itb, np.abs(got - want).max())
    assert present > 100


def test_stored_tables_replace_the_memories(tmp_path):
    """With fm_use_table set and both files present the context reads them
    and no memory file: a broken .mem then does not matter; without the
    flag the memories are read and the broken file is an error."""
    A = _A()
    s, p = make_table_case([26], 32, tb=(1.0, 25.0, 0.05), use_table=1)
    A.FixBackbone(p["coeff"], p["seq"], p["workdir"]).close()  # writes the tables
    with open(os.path.join(p["workdir"], "fragsLAMW.mem"), "w") as f:
        f.write("[Target]\nquery\n\n[Memories]\nmissing.gro 1 1 9 1\n")
    A.FixBackbone(p["coeff"], p["seq"], p["workdir"]).close()
    txt = open(p["coeff"]).read()
    assert "0.1\n1\n0.15" in txt
    txt = txt.replace("0.1\n1\n0.15", "0.1\n0\n0.15")
    off = tmp_path / "coeff_off"
    off.write_text(txt)
    with pytest.raises(A.AwsemError):
        A.FixBackbone(str(off), p["seq"], p["workdir"])


def test_truncated_table_file_is_a_format_error():
    """A table file that ends inside a row before the last curve is refused
    (the end check of fb.cpp:4210, 4235)."""
    A = _A()
    s, p = make_table_case([26], 33, tb=(1.0, 25.0, 0.05), use_table=1)
    A.FixBackbone(p["coeff"], p["seq"], p["workdir"]).close()
    path = os.path.join(p["workdir"], "fm_table.energy")
    vals = open(path).read().split()
    with open(path, "w") as f:
        f.write(" ".join(vals[: len(vals) // 2 + 7]) + "\n")
    with pytest.raises(A.AwsemError) as e:
        A.FixBackbone(p["coeff"], p["seq"], p["workdir"])
    assert e.value.code == -3 and "table" in str(e.value)
