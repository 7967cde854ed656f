"""The LAMMPS deck front end on the host (CPU, no device work): a deck in the
shipped examples' grammar is executed up to its first run (read_data of an
atom_style awsemmd file, groups, pair and bond coefficients, velocity
create, fix backbone), and the state and formulas it holds are checked
against the data written and the definitions LAMMPS documents."""
import os

import numpy as np
import pytest

import deckgen

KB = 0.0019872067
MVV2E = 48.88821291 ** 2


def _A():
    import awsemmd_amd as A
    A.lib()
    return A


def _masses(s):
    m = np.zeros((3, s.n))
    m[0] = 27.0
    m[1] = np.where(np.array(list(s.seq)) == "G", 2.0, 60.0)
    m[2] = 28.0
    return m


def test_deck_setup_state():
    """Coordinates of the data file in the residue layout; velocity create
    (uniform draws, momentum removed, rescaled to T with 3N - 3 degrees of
    freedom); energy.log opened with the reference's header."""
    A = _A()
    s, path, d = deckgen.make_deck([30, 20], 41)
    dk = A.Deck(path)
    na, nr = dk.natoms()
    assert (na, nr) == (3 * s.n, s.n)
    x, v = dk.residue_state()
    assert np.array_equal(x, s.coords())
    m = _masses(s)
    p = (m[..., None] * v).sum(axis=(0, 1))
    ke = 0.5 * MVV2E * (m[..., None] * v * v).sum()
    T = 2 * ke / ((3 * na - 3) * KB)
    print("momentum", p, "T", T)
    assert np.abs(p).max() < 1e-12 and abs(T - 300.0) < 1e-9
    assert abs(dk.variable("ke") - ke) < 1e-9 * ke and abs(dk.variable("temp") - 300.0) < 1e-9
    ids = dk.residue_atoms()
    assert np.array_equal(ids, np.array(deckgen.atom_ids(s.n)))
    with open(os.path.join(d, "energy.log")) as f:
        assert f.readline().startswith("Step   \tChain   \tShake")
    # the same seed draws the same velocities; another seed others
    _, v2 = A.Deck(path).residue_state()
    assert np.array_equal(v, v2)
    with open(path) as f:
        txt = f.read().replace("velocity all create 300 2349852", "velocity all create 300 777")
    p2 = os.path.join(d, "deck2.in")
    with open(p2, "w") as f:
        f.write(txt)
    _, v3 = A.Deck(p2).residue_state()
    assert np.abs(v3 - v).max() > 1e-3


def test_deck_groups_by_list_and_range():
    """group id by explicit lists and by a:b:c ranges select the same atoms."""
    A = _A()
    s, path, d = deckgen.make_deck([25], 42, ranges=False)
    a = A.Deck(path).residue_atoms()
    s2, path2, _ = deckgen.make_deck([25], 42, ranges=True)
    b = A.Deck(path2).residue_atoms()
    assert np.array_equal(a, b)


def test_deck_formulas():
    """Equal-style formulas: precedence, parentheses, unary minus, powers,
    v_ references and the thermo keywords, with LAMMPS's definitions."""
    A = _A()
    s, path, d = deckgen.make_deck([20], 43, extra="variable a equal 2*(3+4)^2/7\nvariable b equal -v_a+1.5e1")
    dk = A.Deck(path)
    assert dk.variable("v_a") == 14.0
    assert dk.variable("v_b") == 1.0
    assert dk.variable("2^3^2") == 512.0
    assert dk.variable("step") == 0.0 and dk.variable("atoms") == 3 * s.n
    assert dk.variable("f_2[12]") == 0.0 and dk.variable("v_E_P") == 0.0  # no evaluation yet
    assert dk.variable("etotal") == dk.variable("ke")
    with pytest.raises(A.AwsemError) as e:
        dk.variable("press")
    assert e.value.code == -4 and "press" in str(e.value)
    with pytest.raises(A.AwsemError):
        dk.variable("1/(v_a-14)")
    with pytest.raises(A.AwsemError):
        dk.variable("f_2[18]")


@pytest.mark.parametrize("edit,code,word", [
    ("fix 1 all nve", "fix 1 all nvt temp 300.0 300.0 100.0", "fix style nvt"),
    ("units real", "units metal", "units metal"),
    ("special_bonds fene", "special_bonds lj/coul 0.0 0.0 0.5", "special_bonds"),
    ("pair_style vexcluded 2 3.5 3.5", "pair_style lj/cut 10.0", "pair_style lj/cut"),
    ("read_data data.sys", "read_data missing.data", "missing.data"),
    ("atom_style awsemmd", "atom_style full", "atom_style full"),
])
def test_deck_refuses(edit, code, word):
    """A command outside the subset, or a missing file, is an error naming
    it and its line; nothing is ignored silently."""
    A = _A()
    s, path, d = deckgen.make_deck([20], 44)
    txt = open(path).read()
    assert edit in txt
    with open(path, "w") as f:
        f.write(txt.replace(edit, code))
    with pytest.raises(A.AwsemError) as e:
        A.Deck(path)
    print(e.value)
    assert word in str(e.value) and "deck.in:" in str(e.value)


def test_deck_refuses_foreign_data():
    """Data the engine cannot represent: other masses, an extra bond, two
    bond constants."""
    A = _A()
    for old, new, word in (("1 27.0", "1 12.0", "masses"), ("1 60.0 3.816", "1 61.0 3.816", "different K")):
        s, path, d = deckgen.make_deck([20], 45)
        data = os.path.join(d, "data.sys")
        txt = open(data).read()
        assert old in txt
        with open(data, "w") as f:
            f.write(txt.replace(old, new, 1))
        with pytest.raises(A.AwsemError) as e:
            A.Deck(path)
        assert e.value.code == -4 and word in str(e.value), str(e.value)
    s, path, d = deckgen.make_deck([20], 46)
    data = os.path.join(d, "data.sys")
    txt = open(data).read().replace(f"{4 * 20 - 2} bonds", f"{4 * 20 - 1} bonds")
    with open(data, "w") as f:
        f.write(txt.rstrip("\n") + f"\n{4 * 20 - 1} 1 1 7\n")
    with pytest.raises(A.AwsemError) as e:
        A.Deck(path)
    assert e.value.code == -4 and "bond" in str(e.value)


def test_deck_1r69_setup(oracle_mod):
    """The shipped 1R69 system (data file, coeff with fragment memory in table
    mode, single native memory) sets up through the front end: 189 atoms of
    63 residues in the data file's order (CA, O, CB / H-beta), the data
    file's coordinates, and finite oracle energies there with the table term
    present."""
    A = _A()
    path, d = deckgen.deck_1r69()
    dk = A.Deck(path)
    assert dk.natoms() == (189, 63)
    ids = dk.residue_atoms()
    assert np.array_equal(ids, np.array(deckgen.atom_ids(63)))
    x, v = dk.residue_state()
    rows = {}
    txt = open(os.path.join(d, "data.1r69")).read()
    atoms = txt.split("Atoms")[1].split("Bond Coeffs")[0].split("\n")
    for ln in atoms:
        w = ln.split()
        if len(w) >= 8:
            rows[int(w[0])] = [float(t) for t in w[5:8]]
    want = np.array([[rows[int(i)] for i in ids[k]] for k in range(3)])
    assert np.array_equal(x, want)
    o = deckgen.deck_oracle(oracle_mod, d, seq="extended.seq")
    f, e = o.compute(x)
    k = oracle_mod.ET["Frag_Mem"]
    print("1R69 energies at the data file's coordinates:", dict(zip(oracle_mod.ET_NAMES, np.round(e, 4))))
    assert np.all(np.isfinite(e)) and e[k] < -1.0
