"""Fragment-memory sharing of the resident engine: the fragment-memory items
of large replicas are cut into chunks that workgroups done with their own
replicas evaluate against a copy of the replica's CA/CB positions, adding
the forces into the replica's accumulator.

Every chunk's items are evaluated with the same arithmetic wherever they
run, so a shared batch differs from the same batch with sharing off
(AWSEM_SHARE=0) only in the order of FP64 force additions: forces and
energies agree to 1e-10 relative, trajectories to 1e-8 A after 150 steps.
The fragment-memory term of a shared batch also equals the oracle's."""
import os

import numpy as np
import pytest

from conftest import make_case

pytestmark = pytest.mark.gpu

LENS = [40, 50, 60, 70, 80, 90, 100, 120, 140, 170, 200, 250, 300, 350, 400]


def _gpu():
    import awsemmd_amd as A
    A.lib()
    return A


def _batch(A, cases, prec, share, min_n=None, mask=None):
    if not share:
        os.environ["AWSEM_SHARE"] = "0"
    if min_n is not None:
        os.environ["AWSEM_SHARE_MIN_N"] = str(min_n)
    try:
        fixes = [A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec) for _, p in cases]
        if mask is not None:
            fixes[0].set_term_mask(mask)
        b = A.Batch(fixes, prec)
    finally:
        os.environ.pop("AWSEM_SHARE", None)
        os.environ.pop("AWSEM_SHARE_MIN_N", None)
    b.set_engine("resident")
    return b


@pytest.fixture(scope="module")
def cases():
    return [make_case([L], 9000 + L) for L in LENS]


def _rel(a, b):
    fa, fb = a.reshape(-1, 3), b.reshape(-1, 3)
    return np.linalg.norm(fa - fb, axis=1).max() / max(np.linalg.norm(fb, axis=1).max(), 1e-30)


@pytest.mark.parametrize("min_n", [None, 0])
@pytest.mark.parametrize("prec", [0, 1])
def test_shared_batch_equals_unshared(cases, prec, min_n):
    """One evaluation, then 150 steps of dynamics in one launch (the small
    replicas finish early and help the large ones for most of it); sharing
    for replicas above the mean size (default) and for every replica."""
    A = _gpu()
    x0 = A.pack_states([s.coords() for s, _ in cases])
    rng = np.random.default_rng(4)
    v0 = rng.normal(0.0, 0.003, size=x0.shape)
    out = {}
    for share in (False, True):
        b = _batch(A, cases, prec, share, min_n if share else None)
        b.set_state(x0, v0)
        b.compute()
        _, _, f1, e1 = b.get_state()
        b.run(150, dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0, seed=9, step0=0)
        x2, _, f2, e2 = b.get_state()
        out[share] = (f1, e1, x2, f2, e2, b.share_stats())
        b.close()
    f1, e1, x2, f2, e2, st0 = out[False]
    g1, d1, y2, g2, d2, st1 = out[True]
    print(f"prec {prec}, min_n {min_n}: shared replicas {st1[0]}, chunks evaluated elsewhere {st1[1]}", flush=True)
    assert st0 == (0, 0)
    assert st1[0] == (len(LENS) if min_n == 0 else sum(L >= np.mean(LENS) for L in LENS))
    assert st1[1] > 0, "no chunk was taken by another workgroup"
    for r in range(len(LENS)):
        sl = slice(9 * sum(LENS[:r]), 9 * sum(LENS[:r + 1]))
        assert _rel(g1[sl], f1[sl]) < 1e-10, ("forces", r, _rel(g1[sl], f1[sl]))
        assert _rel(g2[sl], f2[sl]) < 1e-8, ("forces after dynamics", r, _rel(g2[sl], f2[sl]))
    assert np.max(np.abs(d1 - e1) / np.maximum(1.0, np.abs(e1))) < 1e-10
    assert np.abs(y2 - x2).max() < 1e-8
    assert np.max(np.abs(d2 - e2) / np.maximum(1.0, np.abs(e2))) < 1e-8


@pytest.mark.parametrize("term", ["Chain", "Water"])
def test_fragment_memory_masked_off_is_not_shared(cases, term):
    """With fragment memory masked off no chunk is opened: the forces of a
    single other term equal those of the unshared batch."""
    A = _gpu()
    mask = 1 << A.ET[term]
    x0 = A.pack_states([s.coords() for s, _ in cases])
    out = []
    for share in (False, True):
        b = _batch(A, cases, A.PREC_FP64, share, min_n=0 if share else None, mask=mask)
        b.set_state(x0, np.zeros_like(x0))
        b.run(20, dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0, seed=3, step0=0)
        b.compute()
        out.append((b.get_state(), b.share_stats()))
        b.close()
    (x1, _, f1, e1), st1 = out[0]
    (x2, _, f2, e2), st2 = out[1]
    assert st2[1] == 0, st2
    assert _rel(f2, f1) < 1e-10 and np.abs(x2 - x1).max() < 1e-10
    assert np.max(np.abs(e2 - e1) / np.maximum(1.0, np.abs(e1))) < 1e-10


def test_shared_fragment_memory_matches_oracle(cases, oracle_mod):
    """FP64, fragment memory alone, every replica sharing: per-replica
    forces and energy against the oracle after 40 steps of dynamics."""
    A = _gpu()
    mask = 1 << A.ET["Frag_Mem"]
    b = _batch(A, cases, A.PREC_FP64, True, min_n=0, mask=mask)
    x0 = A.pack_states([s.coords() for s, _ in cases])
    b.set_state(x0, np.zeros_like(x0))
    b.run(40, dt=2.0, t_start=300.0, t_stop=300.0, damp=1000.0, seed=3, step0=0)
    b.compute()
    x, _, f, e = b.get_state()
    assert b.share_stats()[1] > 0
    for r, ((s, p), xr, fr) in enumerate(zip(cases, b.split(x), b.split(f))):
        o = oracle_mod.Oracle(p["coeff"], p["seq"], p["workdir"])
        fo, eo = o.compute(xr, mask)
        kk = A.ET["Frag_Mem"]
        assert _rel(fr, fo) < 1e-9, (s.n, _rel(fr, fo))
        assert abs(e[r, kk] - eo[kk]) <= 1e-9 * max(1.0, abs(eo[kk])), (s.n, e[r, kk], eo[kk])
