"""LAMMPS decks for the tests of the deck front end (awsem_deck_*).

A synthetic system is written as the shipped examples hold theirs: an
atom_style awsemmd data file (id molecule residue type q x y z,
src/atom_vec_awsemmd.cpp:56) with three atoms per residue in the order CA
(type 1), O (type 3), CB (type 4, or 5 for a glycine's H-beta), residue
tags numbered across chains, molecule = chain, masses 27 / 14 / 28 / 60 / 2,
the chain bonds with the shipped data files' coefficients (K 60; r0 3.816
CA-CA, 2.40 CA-O, 2.76 O-CA, 1.53 CA-CB, 1.09 CA-H-beta), and an input
script in the shipped decks' grammar (units real, special_bonds fene,
pair_style vexcluded 2 3.5 3.5 with the decks' pair_coeff lines, velocity
create, fix nve + fix langevin, fix backbone, thermo, dump atom, fix print of
equal-style variables)."""
import os
import tempfile

import numpy as np

from awsemmd_amd import workload as W

BOND_K = 60.0
BOND_R0 = [3.816, 2.40, 2.76, 1.53, 1.09]  # bond types 1..5: CA-CA, CA-O, O-CA(+1), CA-CB, CA-H-beta
MASSES = [27.0, 14.0, 28.0, 60.0, 2.0]
PAIR = [(1, 1, 20.0, 3.5, 4.5), (1, 4, 20.0, 3.5, 4.5), (4, 4, 20.0, 3.5, 4.5), (3, 3, 20.0, 3.5, 3.5)]
TERMS = ["E_chain", "E_bond", "E_chi", "E_rama", "E_excl", "E_dssp", "E_pap", "E_water", "E_burial", "E_helix",
         "E_fmem", "E_P"]


def atom_ids(n):
    """(CA, CB, O) atom ids of residue i: 3i+1, 3i+3, 3i+2."""
    i = np.arange(n)
    return 3 * i + 1, 3 * i + 3, 3 * i + 2


def write_data(s, path):
    n = s.n
    seq = s.seq
    chain = s.chain_index()
    x = s.coords()  # (3, n, 3): CA, CB, O
    lines = ["LAMMPS protein data file (synthetic)", "",
             f"{3 * n} atoms", f"{4 * n - 2 * len(s.chains)} bonds", "0 angles", "0 dihedrals", "0 impropers", "",
             "5 atom types", "5 bond types", "0 angle types", "0 dihedral types", "0 improper types", "",
             "-1000.0 1000.0 xlo xhi", "-1000.0 1000.0 ylo yhi", "-1000.0 1000.0 zlo zhi", "", "Masses", ""]
    lines += [f"{t + 1} {m}" for t, m in enumerate(MASSES)]
    lines += ["", "Atoms", ""]
    for i in range(n):
        cb_type = 5 if seq[i] == "G" else 4
        for aid, typ, k in ((3 * i + 1, 1, 0), (3 * i + 2, 3, 2), (3 * i + 3, cb_type, 1)):
            lines.append("%d %d %d %d 0.0 %.17g %.17g %.17g" % (aid, chain[i] + 1, i + 1, typ, *x[k, i]))
    lines += ["", "Bond Coeffs", ""]
    lines += [f"{t + 1} {BOND_K} {r}" for t, r in enumerate(BOND_R0)]
    lines += ["", "Bonds", ""]
    b = 0
    for i in range(n):
        ca, o, cb = 3 * i + 1, 3 * i + 2, 3 * i + 3
        bonds = [(2, ca, o), (5 if seq[i] == "G" else 4, ca, cb)]
        if i + 1 < n and chain[i + 1] == chain[i]:
            bonds += [(1, ca, 3 * (i + 1) + 1), (3, o, 3 * (i + 1) + 1)]
        for t, a1, a2 in bonds:
            b += 1
            lines.append(f"{b} {t} {a1} {a2}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def deck_text(n, *, runs=(60,), thermo=20, dump=30, print_every=20, timestep=2.0, langevin=(500.0, 300.0, 100.0, 3456),
              velocity=(300.0, 2349852), ranges=True, extra="", data="data.sys", seq="protein.seq"):
    ca, cb, o = atom_ids(n)
    if ranges:
        g = [f"group alpha_carbons id 1:{3 * n}:3", f"group beta_atoms id 3:{3 * n}:3", f"group oxygens id 2:{3 * n}:3"]
    else:
        g = ["group alpha_carbons id " + " ".join(map(str, ca)), "group beta_atoms id " + " ".join(map(str, cb)),
             "group oxygens id " + " ".join(map(str, o))]
    t = ["# synthetic deck in the grammar of the shipped examples", "units real", f"timestep {timestep}",
         "dimension 3", "boundary s s s", "neighbor 5 bin", "neigh_modify delay 5", "atom_modify sort 0 0.0",
         "special_bonds fene", "atom_style awsemmd", "bond_style harmonic", "pair_style vexcluded 2 3.5 3.5",
         f"read_data {data}", "pair_coeff * * 0.0"]
    t += ["pair_coeff %d %d %g %g %g" % p for p in PAIR]
    if velocity:
        t.append("velocity all create %g %d" % velocity)
    t += g
    t.append("fix 1 all nve")
    if langevin:
        t.append("fix 11 all langevin %g %g %g %d  # ramped thermostat" % langevin)
    t += [f"fix 2 alpha_carbons backbone beta_atoms oxygens fix_backbone_coeff.data {seq}",
          f"thermo {thermo}", f"dump 1 all atom {dump} dump.lammpstrj", "dump_modify 1 sort id",
          "variable E_bond  equal emol", "variable E_chain equal f_2[1]", "variable E_excl  equal epair",
          "variable E_chi   equal f_2[3]", "variable E_rama  equal f_2[4]", "variable E_dssp  equal f_2[6]",
          "variable E_pap   equal f_2[7]", "variable E_water equal f_2[8]", "variable E_burial equal f_2[9]",
          "variable E_helix equal f_2[10]", "variable E_fmem  equal f_2[12]",
          "variable E_P     equal v_E_chain+v_E_chi+v_E_rama+v_E_water+v_E_helix+v_E_fmem+v_E_excl+v_E_dssp"
          "+v_E_pap+v_E_bond+v_E_burial",
          "variable E_K     equal ke", "variable Step equal step",
          'fix energy all print %d "${Step} ${E_chain} ${E_bond} ${E_chi} ${E_rama} ${E_excl} ${E_dssp} ${E_pap} '
          '${E_water} ${E_burial} ${E_helix} ${E_fmem} ${E_P}" file energy screen no' % print_every]
    if extra:
        t.append(extra)
    t.append("reset_timestep 0")
    t += [f"run {r}" for r in runs]
    return "\n".join(t) + "\n"


def make_deck(lengths, seed, **kw):
    """(system, deck path, directory) of a synthetic deck."""
    s = W.make_system(lengths, seed=seed, **{k: kw.pop(k) for k in list(kw) if k in ("compact",)})
    d = tempfile.mkdtemp(prefix="awsem_deck_")
    W.write_system(s, d)
    write_data(s, os.path.join(d, "data.sys"))
    path = os.path.join(d, "deck.in")
    with open(path, "w") as f:
        f.write(deck_text(s.n, **kw))
    return s, path, d


FIXTURE_1R69 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "deck_1r69")
N_1R69 = 63


def deck_1r69(**kw):
    """(deck path, directory) of the shipped 1R69 system (tests/golden/deck_1r69,
    written by make_deck_1r69.py) with a deck of the example's settings:
    timestep 5 fs, Langevin 500 -> 300 K with 1000 fs damping, seed 3456,
    velocities at 300 K (seed 2349852), thermo and energy output every 1000
    steps, a run of 2,000,000 steps (callers cap it)."""
    import shutil
    d = tempfile.mkdtemp(prefix="awsem_1r69_")
    for f in os.listdir(FIXTURE_1R69):
        shutil.copyfile(os.path.join(FIXTURE_1R69, f), os.path.join(d, f))
    opts = dict(runs=(2000000,), thermo=1000, dump=2000, print_every=1000, timestep=5.0,
                langevin=(500.0, 300.0, 1000.0, 3456), data="data.1r69", seq="extended.seq")
    opts.update(kw)
    path = os.path.join(d, "deck.in")
    with open(path, "w") as f:
        f.write(deck_text(N_1R69, **opts))
    return path, d


def deck_oracle(oracle_mod, d, seq="protein.seq"):
    """Oracle of a deck's directory with the deck's bond and pair coefficients."""
    o = oracle_mod.Oracle(os.path.join(d, "fix_backbone_coeff.data"), os.path.join(d, seq), d)
    o.set_bonds(BOND_K, BOND_R0)
    lam, cs, cl = np.zeros((6, 6)), np.zeros((6, 6)), np.zeros((6, 6))
    for i, j, l, a, b in PAIR:
        lam[i, j] = lam[j, i] = l
        cs[i, j] = cs[j, i] = a
        cl[i, j] = cl[j, i] = b
    o.set_vexcluded(2, lam, cs, cl)
    return o
