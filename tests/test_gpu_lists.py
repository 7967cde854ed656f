"""The production pair selection against the oracle's lists, bit for bit.

After one force evaluation (and again after Langevin dynamics that rebuilds
the Verlet lists), the lists the force passes actually read are exported
(awsem_batch_lists) from both batch engines and both precisions, and from
the wide engine's cell-list build:
  * DSSP calls (acceptor, donor) and P_AP pairs: equal to the oracle's
    lists (fb.cpp:8064-8103, 2904);
  * the Verlet list, the excluded-volume residue pairs and the density
    pairs of the water-xi and force passes: supersets of every pair the
    oracle selects (density wells fb.cpp:7729-7786, water xi 7879-7937,
    excluded volume pev.cpp:104-113), since those passes apply the exact
    cutoffs themselves;
  * the Verlet list is the same set from every engine and list builder.
The geometries carry pairs planted at 2e-7 relative (in r^2) inside and
outside the DSSP O-N cut, the P_AP list cut and the outer edge of the water
well's theta range, where the FP32 distance estimate cannot decide and the
FP64 re-check must."""
import os

import numpy as np
import pytest

import kat
from conftest import make_case

pytestmark = pytest.mark.gpu

PAP_CUT2 = (8.0 + np.arctanh(1.0 - 2e-12) / 7.0) ** 2
W_MAX2 = (6.5 + 8 * np.log(10.0) / 5.0) ** 2
DSSP_CUT = 7.0


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


def planted_case(lengths, seed, compact, n_plant=10, rel=2e-7):
    """A generated system whose coordinates are then moved so that n_plant
    pairs per cut sit at cut^2 (1 +- rel): CA-CA at the P_AP list cut, density
    sites (CB) at the water well's outer theta edge, O-N at the DSSP cut."""
    s, p = make_case(lengths, seed, compact=compact)
    x = s.coords().copy()
    n, seq = s.n, s.seq
    ch = s.chain_index()
    lens = [len(c) for c in s.chains]
    starts = np.cumsum([0] + lens)
    first = np.zeros(n, bool)
    last = np.zeros(n, bool)
    first[starts[:-1]] = True
    last[starts[1:] - 1] = True
    rng = np.random.default_rng(seed)
    used = set()
    planted = {"pap": [], "dens": [], "dssp": []}

    def pick(cands):
        rng.shuffle(cands)
        out = []
        for i, j in cands:
            if i in used or j in used:
                continue
            used.update((i, j))
            out.append((i, j))
            if len(out) == n_plant:
                break
        return out

    iu, ju = np.triu_indices(n, k=1)
    # P_AP: move CA(j) along CA(i) -> CA(j)
    d = np.linalg.norm(x[0][iu] - x[0][ju], axis=1)
    for k, (i, j) in enumerate(pick([(a, b) for a, b, r in zip(iu, ju, d) if 8.5 < r < 11.5])):
        sgn = 1.0 if k % 2 else -1.0
        v = x[0][j] - x[0][i]
        shift = v * (np.sqrt(PAP_CUT2 * (1 + sgn * rel)) / np.linalg.norm(v) - 1.0)
        x[:, j] += shift  # the whole residue: its CA moves, its shape stays
        planted["pap"].append((i, j, sgn))
    # density sites (non-glycine CB) at the outer theta edge of water well 0
    site = x[1]
    d = np.linalg.norm(site[iu] - site[ju], axis=1)
    ok = lambda a, b: seq[a] != "G" and seq[b] != "G" and (ch[a] != ch[b] or b - a > 1)
    for k, (i, j) in enumerate(pick([(a, b) for a, b, r in zip(iu, ju, d) if 9.5 < r < 11.5 and ok(a, b)])):
        sgn = 1.0 if k % 2 else -1.0
        v = x[1][j] - x[1][i]
        x[1][j] = x[1][i] + v * np.sqrt(W_MAX2 * (1 + sgn * rel)) / np.linalg.norm(v)
        planted["dens"].append((i, j, sgn))
    # DSSP: O(i) against N(j) (O moves do not touch CA or CB)
    abc = kat.Abc()
    xn, xh, _ = abc.virtual(x[0], x[2], lens)
    cands = []
    for i in range(n):
        for j in range(n):
            if i == j or seq[j] == "P" or last[i] or first[j] or (ch[i] == ch[j] and abs(i - j) <= 2):
                continue
            r = np.linalg.norm(x[2][i] - xn[j])
            if 6.0 < r < 8.0:
                cands.append((i, j))
    rng.shuffle(cands)
    acc, don = set(), set()
    for i, j in cands:
        # O(i) sets N(i+1): no planted donor may be i + 1, no acceptor j - 1
        if i in acc or i + 1 in don or j - 1 in acc or i == j - 1:
            continue
        acc.add(i)
        don.add(j)
        planted["dssp"].append((i, j, 1.0 if len(planted["dssp"]) % 2 else -1.0))
        if len(planted["dssp"]) == n_plant:
            break
    for i, j, sgn in planted["dssp"]:
        v = x[2][i] - xn[j]
        x[2][i] = xn[j] + v * DSSP_CUT * np.sqrt(1 + sgn * rel) / np.linalg.norm(v)
    xn, xh, _ = abc.virtual(x[0], x[2], lens)
    for i, j, sgn in planted["dssp"]:  # the plants survived each other
        r2 = ((x[2][i] - xn[j]) ** 2).sum()
        assert abs(r2 / DSSP_CUT ** 2 - (1 + sgn * rel)) < 1e-9
    return s, p, x, planted


def _res_pair(code, n):
    i, j = divmod(int(code), n)
    return (min(i, j), max(i, j))


def _check_replica(b, r, x, o, s, planted):
    n = s.n
    g_dssp, o_dssp = b.lists(r, "dssp"), o.list(2, x)
    g_pap, o_pap = b.lists(r, "pap"), o.list(3, x)
    # the plants: inside pairs selected, outside ones not (by the oracle)
    so_d, so_p = set(o_dssp.tolist()), set(o_pap.tolist())
    for i, j, sgn in planted["dssp"]:
        assert ((i * n + j) in so_d) == (sgn < 0), ("dssp plant", i, j, sgn)
    for i, j, sgn in planted["pap"]:
        assert ((min(i, j) * n + max(i, j)) in so_p) == (sgn < 0), ("pap plant", i, j, sgn)
    assert np.array_equal(g_dssp, o_dssp), ("DSSP", len(g_dssp), len(o_dssp),
                                            sorted(set(g_dssp.tolist()) ^ set(o_dssp.tolist()))[:10])
    assert np.array_equal(g_pap, o_pap), ("P_AP", len(g_pap), len(o_pap),
                                          sorted(set(g_pap.tolist()) ^ set(o_pap.tolist()))[:10])
    verlet = {_res_pair(c, n) for c in b.lists(r, "verlet")}
    dens = {_res_pair(c, n) for c in b.lists(r, "density")}
    ev = {_res_pair(c, n) for c in b.lists(r, "excluded")}
    # every density pair the oracle selects (both wells, helix well, xi pairs)
    want_dens = set()
    for kind, arg in ((0, 0), (0, 1), (1, 0), (6, 0), (6, 1)):
        want_dens |= {_res_pair(c, n) for c in o.list(kind, x, arg)}
    for i, j, sgn in planted["dens"]:
        assert ((min(i, j), max(i, j)) in {_res_pair(c, n) for c in o.list(0, x, 0)}) == (sgn < 0)
    assert want_dens <= dens, sorted(want_dens - dens)[:10]
    want_ev = set()
    for c in o.list(5, x):
        a, bb = divmod(int(c), 3 * n)
        ra, rb = a // 3, bb // 3
        if ra != rb:
            want_ev.add((min(ra, rb), max(ra, rb)))
    assert want_ev <= ev, sorted(want_ev - ev)[:10]
    want_all = want_dens | want_ev | {_res_pair(c, n) for c in o_dssp} | {_res_pair(c, n) for c in o_pap}
    assert want_all <= verlet, sorted(want_all - verlet)[:10]
    return verlet, len(g_dssp), len(g_pap), len(dens), len(ev)


def _batch(A, cases, prec, engine, cells=False):
    if cells:
        os.environ["AWSEM_CELL_MIN_N"] = "16"  # every replica through the cell-list build
    try:
        b = A.Batch([A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec) for _, p, _, _ in cases], prec)
    finally:
        os.environ.pop("AWSEM_CELL_MIN_N", None)
    b.set_engine(engine)
    return b


@pytest.fixture(scope="module")
def cases():
    return [planted_case([60, 60], 17, True), planted_case([120], 11, False), planted_case([50, 50, 40], 4, True)]


@pytest.mark.parametrize("prec", [0, 1])
def test_production_lists_bit_exact(oracle_mod, cases, prec):
    """Both engines (and the wide engine's cell-list build): DSSP and P_AP
    lists equal the oracle's, the other lists contain every oracle pair, and
    the Verlet lists are one set; then again after 300 steps of dynamics at
    600 K (lists rebuilt on drift) on the final coordinates."""
    A = _gpu()
    oracles = [oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"]) for _, p, _, _ in cases]
    x0 = A.pack_states([x for _, _, x, _ in cases])
    results = {}
    for engine, cells in (("resident", False), ("wide", False), ("wide", True)):
        label = engine + (" (cell list)" if cells else "")
        print(f"precision {prec}, {label}", flush=True)
        b = _batch(A, cases, prec, engine, cells)
        b.set_state(x0, np.zeros_like(x0))
        b.compute()
        vs = []
        for r, ((s, p, x, planted), o) in enumerate(zip(cases, oracles)):
            verlet, nd, npap, nden, nev = _check_replica(b, r, x, o, s, planted)
            print(f"  replica {r}: verlet {len(verlet)}, dssp {nd}, pap {npap}, density {nden}, ev {nev}")
            vs.append(verlet)
        results[label] = vs
        # dynamics: the lists follow the coordinates
        b.run(300, dt=2.0, t_start=600.0, t_stop=600.0, damp=100.0, seed=7, step0=0)
        b.compute()
        xf, _, _, _ = b.get_state()
        for r, (xr, (s, p, x, planted), o) in enumerate(zip(b.split(xf), cases, oracles)):
            _check_replica(b, r, xr, o, s, {"dssp": [], "pap": [], "dens": []})
        print(f"  after 300 steps: list builds {b.list_stats()[0]}", flush=True)
    for label, vs in results.items():
        assert all(a == c for a, c in zip(vs, results["resident"])), label
