"""The drop-in boundary on the GPU, against the oracle: awsem_compute on a
LAMMPS-ordered atom table (shuffled, with atoms outside the fix's groups,
forces added into a non-zero incoming array: pre_force semantics,
fb.cpp:8181-8189, Construct_Computational_Arrays fb.cpp:1590-1698); a
periodic box with image flags (fb.cpp:7067-7112); energy.log written from
the last evaluation; and the [Shake] term (fb.cpp:2172-2237), off in the
shipped coeff, per term on every path."""
import numpy as np
import pytest

from conftest import make_case

pytestmark = pytest.mark.gpu


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


def _table(s, rng, n_extra=17):
    """A LAMMPS-like atom table: the CA, CB and O atoms of every residue plus
    n_extra atoms in no fix group, in shuffled order; returns the order
    (atom a holds table entry perm[a]) and the group, molecule and residue
    tags per atom."""
    n = s.n
    ch = s.chain_index()
    group = np.concatenate([np.repeat([0, 1, 2], n), np.full(n_extra, -1)])
    mol = np.concatenate([np.tile(ch + 1, 3), rng.integers(1, 5, n_extra)])
    res = np.concatenate([np.tile(np.arange(1, n + 1), 3), rng.integers(1, n + 1, n_extra)])
    perm = rng.permutation(3 * n + n_extra)
    return perm, group[perm], mol[perm], res[perm]


@pytest.mark.parametrize("prec", [0, 1])
def test_compute_on_shuffled_atom_table(oracle_mod, prec):
    A = _gpu()
    tol = 1e-9 if prec == 0 else 1e-5
    s, p = make_case([70, 45], 12, compact=True)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    rng = np.random.default_rng(4)
    perm, group, mol, res = _table(s, rng)
    n = s.n
    xres = s.coords().reshape(-1, 3)
    extra = rng.uniform(0.0, 300.0, size=(len(perm) - 3 * n, 3))
    xall = np.concatenate([xres, extra])[perm]
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec)
    fix.set_topology(group, mol, res)
    f0 = rng.normal(size=xall.shape)
    f = f0.copy()
    e = fix.pre_force(xall, f)
    fo, eo = o.compute(s.coords())
    want = np.concatenate([fo.reshape(-1, 3), np.zeros_like(extra)])[perm]
    d = np.linalg.norm(f - f0 - want, axis=1).max() / np.linalg.norm(want, axis=1).max()
    de = np.max(np.abs(e - eo) / np.maximum(1.0, np.abs(eo)))
    print(f"precision {prec}: rel dF {d:.2e}, rel dE {de:.2e}; compute_vector(5) = {fix.compute_vector(5):.6f}")
    assert d <= tol and de <= tol
    assert np.array_equal((f - f0)[group == -1], np.zeros_like(extra))  # atoms outside the groups untouched
    assert fix.compute_scalar() == e[0] and fix.compute_vector(A.ET["DSSP"] - 1) == e[A.ET["DSSP"]]


def test_periodic_box_with_image_flags(oracle_mod, tmp_path):
    """A compact two-chain system wrapped into a box smaller than itself
    (periodic in x and y), with LAMMPS image flags: the forces and energies
    are those of the unwrapped system; energy.log written from the last
    evaluation holds them."""
    A = _gpu()
    s, p = make_case([60, 60], 17, compact=True)
    o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
    n = s.n
    xu = s.coords().reshape(-1, 3)
    L = np.array([23.0, 29.0, 0.0])
    box = np.zeros((3 * n, 3), dtype=np.int64)
    xw = xu.copy()
    for l in (0, 1):
        box[:, l] = np.floor(xu[:, l] / L[l]).astype(np.int64)
        xw[:, l] = xu[:, l] - box[:, l] * L[l]
    assert len(np.unique(box[:, 0])) > 1 and len(np.unique(box[:, 1])) > 1  # the chains cross faces
    image = ((box[:, 0] + 512) & 1023) | (((box[:, 1] + 512) & 1023) << 10) | ((box[:, 2] + 512) << 20)
    fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], A.PREC_FP64)
    fix.set_topology(np.repeat([0, 1, 2], n), np.tile(s.chain_index() + 1, 3), np.tile(np.arange(1, n + 1), 3))
    fix.set_box([1, 1, 0], L)
    f = np.zeros_like(xw)
    e = fix.pre_force(xw, f, image=image.astype(np.int32))
    fo, eo = o.compute(s.coords())
    d = np.abs(f - fo.reshape(-1, 3)).max() / np.abs(fo).max()
    de = np.max(np.abs(e - eo) / np.maximum(1.0, np.abs(eo)))
    print(f"periodic: rel dF {d:.2e}, rel dE {de:.2e}")
    assert d <= 1e-9 and de <= 1e-9
    log = tmp_path / "energy.log"
    fix.energy_log_open(str(log))
    fix.energy_log_write(0)
    fix.energy_log_close()
    line = log.read_text().splitlines()[1].split("\t")
    cols = [0.0 if k == 5 else e[k] for k in range(1, 18)]
    assert line[0] == "0 " and [float(v) for v in line[1:18]] == [float("%8.6f" % c) for c in cols]
    assert float(line[18]) == float("%8.6f" % sum(cols))


def _shake_case(oracle_mod):
    from awsemmd_amd import workload as W
    import tempfile
    s = W.make_system([55, 40], seed=27)
    d = tempfile.mkdtemp(prefix="awsem_shake_")
    p = W.write_system(s, d)
    # rest lengths away from the generator's (3.8, 2.41, 2.50 A): a sizeable term
    txt = open(p["coeff"]).read() + "\n[Shake]\n7.0 3.7 2.3 2.65\n"
    open(p["coeff"], "w").write(txt)
    return s, p, oracle_mod.Oracle(p["coeff"], p["seq"], d)


@pytest.mark.parametrize("prec", [0, 1])
def test_shake_per_term(oracle_mod, prec):
    """[Shake] enabled: the term alone and with everything, plain drop-in and
    both batch engines."""
    A = _gpu()
    tol = 1e-9 if prec == 0 else 1e-5
    s, p, o = _shake_case(oracle_mod)
    x = s.coords()
    for mask in (1 << A.ET["Shake"], 0xFFFFFFFF):
        k = A.ET["Shake"]
        fo, eo = o.compute(x, mask)
        assert abs(eo[k]) > 1.0
        fix = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec)
        fix.set_term_mask(mask)
        f, e = fix.compute_res(x)
        outs = [("plain", f, e)]
        for engine in ("resident", "wide"):
            fx = A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec)
            fx.set_term_mask(mask)
            b = A.Batch([fx], prec)
            b.set_engine(engine)
            b.set_state(x)
            b.compute()
            _, _, fb, eb = b.get_state()
            outs.append((engine, fb.reshape(x.shape), eb[0]))
        for label, f, e in outs:
            d = np.abs(f - fo).max() / np.abs(fo).max()
            de = max(abs(e[k] - eo[k]) / abs(eo[k]), abs(e[0] - eo[0]) / max(1.0, abs(eo[0])))
            print(f"{label} mask {mask:#x}: rel dF {d:.2e}, rel dE {de:.2e}")
            assert d <= tol and de <= tol, label
