"""Host-side tests of the C ABI (no GPU calls): exported symbols, the
reference-format readers, error codes and the topology map."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import make_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _A():
    import awsemmd_amd as A
    return A


def test_library_exports_every_declared_symbol():
    A = _A()
    hdr = open(os.path.join(ROOT, "include", "awsem.h")).read()
    declared = sorted(set(re.findall(r"\bint\s+(awsem_\w+)\s*\(", hdr) + re.findall(r"\bdouble\s+(awsem_\w+)\s*\(", hdr)))
    out = subprocess.run(["nm", "-D", "--defined-only", A.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (awsem_\w+)", out))
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    assert set(A.EXPORTS) <= exported


def test_create_reads_reference_file_set(oracle_mod):
    A = _A()
    s, p = make_case([30, 20], 3)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    assert fix.n == 50 == o.n
    # fb.cpp:1048-1066 with the shipped coeff: the water well-2 outer edge
    assert abs(fix.pair_list_cutoff - o.pair_list_cutoff) < 1e-12
    assert abs(fix.pair_list_cutoff - (9.5 + 8 * 2.302585 / 5.0)) < 1e-12


def test_create_from_arrays_matches_files():
    A = _A()
    from awsemmd_amd import workload as W
    s, p = make_case([40], 8)
    f1 = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    f2 = A.FixBackbone.from_arrays(p["coeff"], p["workdir"], precision=A.PREC_FP64, **W.arrays(s))
    assert f1.n == f2.n == 40


def test_error_codes(tmp_path):
    A = _A()
    s, p = make_case([20], 2)
    with pytest.raises(A.AwsemError) as e:
        A.FixBackbone(str(tmp_path / "nope.data"), p["seq"], p["workdir"])
    assert e.value.code == -2
    txt = open(p["coeff"]).read().replace("[Amylometer]-", "[Amylometer]", 1)
    bad = tmp_path / "coeff_amylometer"
    bad.write_text(txt)
    with pytest.raises(A.AwsemError) as e:
        A.FixBackbone(str(bad), p["seq"], p["workdir"])
    assert e.value.code == -4 and "Amylometer" in str(e.value)
    # gamma.dat truncated
    d2 = tmp_path / "w2"
    d2.mkdir()
    for f in os.listdir(p["workdir"]):
        data = open(os.path.join(p["workdir"], f), "rb").read()
        if f == "gamma.dat":
            data = data[: len(data) // 3]
        (d2 / f).write_bytes(data)
    with pytest.raises(A.AwsemError) as e:
        A.FixBackbone(p["coeff"], p["seq"], str(d2))
    assert e.value.code == -3 and "gamma.dat" in str(e.value)
    # memory fragment past the end of the sequence
    memf = d2 / "fragsLAMW.mem"
    (d2 / "gamma.dat").write_bytes(open(os.path.join(p["workdir"], "gamma.dat"), "rb").read())
    memf.write_text("[Target]\nquery\n\n[Memories]\nmem000.gro 15 15 9 1\n")
    with pytest.raises(A.AwsemError) as e:
        A.FixBackbone(p["coeff"], p["seq"], str(d2))
    assert e.value.code == -3


def _atoms(s, rng=None):
    """LAMMPS-style atom table: per residue CA, O, CB (the data-file order of
    CoordinatesToWorkLammpsDataFile.py), optionally shuffled."""
    n = s.n
    ch = s.chain_index()
    group, mol, res, xyz = [], [], [], []
    for i in range(n):
        for g, arr in ((0, s.xca), (2, s.xo), (1, s.xcb)):
            group.append(g)
            mol.append(ch[i] + 1)
            res.append(i + 1)
            xyz.append(arr[i])
    group, mol, res, xyz = map(np.array, (group, mol, res, xyz))
    if rng is not None:
        perm = rng.permutation(len(group))
        group, mol, res, xyz = group[perm], mol[perm], res[perm], xyz[perm]
    return group, mol, res, np.ascontiguousarray(xyz, dtype=np.float64)


def test_topology_errors():
    """Construct_Computational_Arrays error branches (fb.cpp:1640-1696)."""
    A = _A()
    s, p = make_case([12, 10], 4)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"])
    g, m, r, x = _atoms(s)
    fix.set_topology(g, m, r)
    bad = m.copy()
    bad[0] = 3
    with pytest.raises(A.AwsemError) as e:
        fix.set_topology(g, bad, r)
    assert e.value.code == -5 and "Chain tag is out of range" in str(e.value)
    bad = m.copy()
    bad[1] = 2  # oxygen of residue 1 on the other chain
    with pytest.raises(A.AwsemError) as e:
        fix.set_topology(g, bad, r)
    assert "different chain tag" in str(e.value)
    bad = r.copy()
    bad[0] = 0
    with pytest.raises(A.AwsemError) as e:
        fix.set_topology(g, m, bad)
    assert "must be positive" in str(e.value)
    gg = g.copy()
    gg[2] = -1  # drop one CB
    with pytest.raises(A.AwsemError) as e:
        fix.set_topology(gg, m, r)
    assert "same # of atoms" in str(e.value)
    rr = r.copy()
    rr[3 * 12] = 5  # a chain-2 residue claims a chain-1 tag
    with pytest.raises(A.AwsemError):
        fix.set_topology(g, m, rr)


def test_workload_generator_is_seeded():
    from awsemmd_amd import workload as W
    a = W.make_system([30, 12], seed=77)
    b = W.make_system([30, 12], seed=77)
    assert a.chains == b.chains and np.array_equal(a.coords(), b.coords())
    assert all(np.array_equal(u, v) for u, v in zip(a.mem_structs, b.mem_structs))
    # self-avoidance and bond lengths of the trace
    d = np.linalg.norm(np.diff(a.xca[:30], axis=0), axis=1)
    assert np.allclose(d, 3.8)
    # .gro round trip is exact: coordinates are k/1000 nm
    assert np.array_equal(np.round(a.mem_structs[0] * 1000) / 1000, a.mem_structs[0])


def test_non_finite_input_is_refused():
    """NaN/Inf coordinates are refused before any device work (AWSEM_E_NONFINITE)."""
    A = _A()
    s, p = make_case([20], 12)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"])
    x = s.coords().copy()
    x[1, 3, 2] = np.nan
    with pytest.raises(A.AwsemError) as e:
        fix.compute_res(x)
    assert e.value.code == -9
    x[1, 3, 2] = np.inf
    g, m, r, xa = _atoms(s)
    fix.set_topology(g, m, r)
    xa[5, 1] = np.inf
    with pytest.raises(A.AwsemError) as e:
        fix.pre_force(xa, np.zeros_like(xa))
    assert e.value.code == -9


def test_server_and_batch_symbols_are_bound():
    A = _A()
    L = A.lib()
    for name in ("awsem_server_create", "awsem_server_submit", "awsem_batch_autotune", "awsem_batch_set_engine",
                 "awsem_batch_use_graph", "awsem_batch_set_skin"):
        assert getattr(L, name).argtypes is not None


def test_noalloc_rejects_systems_beyond_its_array():
    """awsem_compute_res_noalloc refuses a system whose buffers exceed its
    fixed 4 MiB device array, before any device call, with a message naming
    the limit (9196 residues)."""
    A = _A()
    from awsemmd_amd import workload as W
    import tempfile
    s = W.make_system([9300], seed=1, n_struct=1)
    d = tempfile.mkdtemp(prefix="awsem_big_")
    p = W.write_system(s, d, fm=False)
    fix = A.FixBackbone(p["coeff"], p["seq"], d, A.PREC_FP64)
    x = s.coords()
    with pytest.raises(A.AwsemError) as e:
        fix.compute_res_noalloc(x, np.zeros_like(x))
    assert e.value.code == -8 and "9196 residues" in str(e.value)


def test_energy_log_byte_format(tmp_path):
    """energy.log as the reference writes it (fb.cpp:101-106, 7661-7672):
    the tab-separated header, then "%d " and "\\t%8.6f" per FixBackbone term
    and the total; the excluded-volume pair style and the bonds (slots 5 and
    18 here) belong to LAMMPS' own styles and stay out of the line."""
    A = _A()
    s, p = make_case([20], 2)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    path = tmp_path / "energy.log"
    fix.energy_log_open(str(path))
    rng = np.random.default_rng(1)
    rows = []
    for step in (0, 1000, 123456):
        e = rng.normal(scale=50.0, size=A.N_ET)
        e[0] = e[1:].sum()
        fix.energy_log_write(step, e)
        rows.append((step, e))
    fix.energy_log_close()
    head = ("Step   \tChain   \tShake   \tChi     \tRama    \tExcluded\tDSSP    \tP_AP    \tWater   \tBurial  \tHelix   "
            "\tAMH-Go  \tFrag_Mem\tVec_FM  \tContact_Restraints  \tMembrane\tSSB     \tElectro.\tVTotal\n")
    want = head
    for step, e in rows:
        cols = [0.0 if k == 5 else e[k] for k in range(1, 18)]
        want += "%d " % step + "".join("\t%8.6f" % v for v in cols) + "\t%8.6f\n" % sum(cols)
    assert path.read_text() == want


def test_periodic_box_needs_image_flags():
    """In a periodic box the drop-in needs image flags to unwrap the
    coordinates (fb.cpp:7067-7112): awsem_compute refuses, before any device
    work, and a periodic dimension needs a positive length."""
    A = _A()
    s, p = make_case([12], 3)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    n = s.n
    fix.set_topology(np.repeat([0, 1, 2], n), np.ones(3 * n, dtype=int), np.tile(np.arange(1, n + 1), 3))
    with pytest.raises(A.AwsemError) as e:
        fix.set_box([1, 0, 0], [0.0, 0.0, 0.0])
    assert e.value.code == -1
    fix.set_box([1, 1, 0], [50.0, 60.0, 0.0])
    x = s.coords().reshape(-1, 3)
    with pytest.raises(A.AwsemError) as e:
        fix.pre_force(x, np.zeros_like(x))
    assert e.value.code == -4 and "image" in str(e.value)


def test_extra_term_errors(tmp_path):
    """The coeff-toggled terms' side files: the reference's messages for a
    missing file (AWSEM_E_FILE) and for malformed content (AWSEM_E_FORMAT),
    and the same pair_list_cutoff as the oracle (fb.cpp:1048-1066 with the
    contact-restraint, Debye-Hueckel and barrier shares)."""
    import shutil
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import oracle
    from awsemmd_amd import workload as W
    A = _A()
    s, p = make_case([30, 20], 3, compact=True)
    W.write_extras(s, p["workdir"], seed=3)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"])
    o = oracle.Oracle(p["coeff"], p["seq"], p["workdir"])
    assert A.lib().awsem_pair_list_cutoff(fix.h) == o.pair_list_cutoff
    fix.set_list_skin(2.5)
    with pytest.raises(A.AwsemError) as e:
        fix.set_list_skin(-1.0)
    assert e.value.code == -1

    def variant(name, edit):
        d = tmp_path / name
        shutil.copytree(p["workdir"], d)
        edit(d)
        with pytest.raises(A.AwsemError) as e:
            A.FixBackbone(str(d / "fix_backbone_coeff.data"), p["seq"], str(d))
        return e.value

    err = variant("nocharge", lambda d: os.remove(d / "charge_on_residues.dat"))
    assert err.code == -2 and "charge_on_residues.dat" in str(err)
    err = variant("nozim", lambda d: os.remove(d / "zim"))
    assert err.code == -2 and "zim" in str(err)
    err = variant("zim0", lambda d: (d / "zim").write_text("0\n" * 50))
    assert err.code == -3
    err = variant("zimshort", lambda d: (d / "zim").write_text("1\n2\n"))
    assert err.code == -3 and "Membrane potential parameter file format error" in str(err)
    err = variant("cr", lambda d: (d / "contact.list").write_text("1 99 1.0 5.0\n"))
    assert err.code == -3 and "Indices out of range in Contact Restraints potential file" in str(err)
    err = variant("nogo", lambda d: os.remove(d / "amh-go.gro"))
    assert err.code == -2 and "Cannot read file amh-go.gro" in str(err)
    err = variant("nogamma", lambda d: os.remove(d / "amh-go.gamma"))
    assert err.code == -2 and "Cannot read file amh-go.gamma" in str(err)

    def phos_first(d):
        txt = open(d / "fix_backbone_coeff.data").read()
        (d / "fix_backbone_coeff.data").write_text("[Phosphorylation]\n1.2\n1\n3\n\n" + txt)
    err = variant("phos", phos_first)
    assert err.code == -3 and "phosphorylation without water" in str(err)


def test_frustratometer_sections_and_errors(tmp_path):
    """The three frustratometer sections (fb.cpp:359-412): their intervals
    and modes, the reference's messages for a bad mode and for a fragment
    frustratometer without fragment memory, a missing decoy library, and
    AWSEM_E_STATE from a call on a context without the section (no device
    work)."""
    A = _A()
    s, p = make_case([30, 20], 5)
    base = open(p["coeff"]).read()

    def fix_with(text, fm=True):
        c = tmp_path / ("coeff_%d" % (abs(hash(text)) % 10**8))
        c.write_text((base if fm else base.replace("[Fragment_Memory]\n", "[Fragment_Memory]-\n")) + "\n" + text + "\n\n")
        return A.FixBackbone(str(c), p["seq"], p["workdir"])

    f = fix_with("[Tertiary_Frustratometer]\n9.5\n10\n7\nmutational\n\n"
                 "[Nmer_Frustratometer]\n6\n9.5\n10\n10\n3\n1.7 -1 0\n0 0 1 0.6\nsinglenmer\n\n"
                 "[Fragment_Frustratometer]\nshuffle\nfragsLAMW.mem\n11\n5")
    assert f.tert_frust_every() == 7 and f.nmer_frust_every() == 3
    mode, nc, nd, every = f.frag_frust_info()
    assert (mode, nc, every) == (0, 11, 5) and nd == len(s.memories)
    f.close()
    g = fix_with("")
    assert g.tert_frust_every() == 0 and g.nmer_frust_every() == 0 and g.frag_frust_info()[0] == -1
    x = s.coords()
    for call in (lambda: g.tert_frust(x), lambda: g.nmer_frust(x), lambda: g.frag_frust(x)):
        with pytest.raises(A.AwsemError) as e:
            call()
        assert e.value.code == -10
    g.close()
    cases = [
        ("[Tertiary_Frustratometer]\n9.5\n10\n1\nsideways", True,
         'Only "configurational", "mutational", "singleresidue" are acceptable modes for the Tertiary_Frustratometer.'),
        ("[Nmer_Frustratometer]\n6\n9.5\n10\n10\n1\n1.7 -1 0\n0 0 1 0.6\ntriple", True,
         'Only "pairwise", "singlenmer" are acceptable modes for the Nmer_Frustratometer.'),
        ("[Fragment_Frustratometer]\nsideways\nfragsLAMW.mem\n1", True,
         'Only "shuffle" and "read" are acceptable modes for the Fragment_Frustratometer.'),
        ("[Fragment_Frustratometer]\nshuffle\nfragsLAMW.mem\n10\n1", False,
         "Cannot run Fragment_Frustratometer without Fragment_Memory or Fragment_Memory_Table."),
        ("[Fragment_Frustratometer]\nread\nnone.mem\n1\n1.0\n0 0\n0", True, "Error opening mem file"),
    ]
    for text, fm, msg in cases:
        with pytest.raises(A.AwsemError) as e:
            fix_with(text, fm)
        assert msg in str(e.value), (text, str(e.value))
