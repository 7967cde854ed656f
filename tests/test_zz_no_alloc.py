"""awsem_compute_res_noalloc (the residue-layout evaluation with its per-call
device buffers in a fixed device array, no allocation in the call) against
awsem_compute_res on the same coordinates: two no-alloc calls in a row, the
second on a smaller system, both precisions, forces added into a non-zero
incoming array.  FP64: rounding-level agreement (1e-9 relative); FP32 pair
arithmetic: 1e-5, the path's stated tolerance."""
import ctypes as C

import numpy as np
import pytest

from conftest import make_case

pytestmark = pytest.mark.gpu

TOL = {0: 1e-9, 1: 1e-5}


def _noalloc(A, fix, x, f, e):
    dp = C.POINTER(C.c_double)
    return A.lib().awsem_compute_res_noalloc(fix.h, x.ctypes.data_as(dp), f.ctypes.data_as(dp), e.ctypes.data_as(dp))


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_compute_res(prec):
    import awsemmd_amd as A
    A.lib()
    print(f"no-alloc vs compute_res, precision {prec}", flush=True)
    cases = [make_case([150, 60], 501), make_case([40], 502)]
    fixes = [A.FixBackbone(p["coeff"], p["seq"], p["workdir"], prec) for _, p in cases]
    # a context without a device image is refused (no allocation in the call)
    x0 = np.ascontiguousarray(cases[1][0].coords())
    fresh = A.FixBackbone(cases[1][1]["coeff"], cases[1][1]["seq"], cases[1][1]["workdir"], prec)
    assert _noalloc(A, fresh, x0, np.zeros_like(x0), np.zeros(A.N_ET)) == -10
    refs = []
    for (s, _), fix in zip(cases, fixes):
        refs.append(fix.compute_res(s.coords()))  # builds each context's device image
    rng = np.random.default_rng(3)
    for (s, _), fix, (fr, er) in zip(cases, fixes, refs):  # the larger system first, then the smaller
        x = np.ascontiguousarray(s.coords())
        f0 = rng.normal(size=x.shape)
        f = f0.copy()
        e = np.zeros(A.N_ET)
        rc = _noalloc(A, fix, x, f, e)
        assert rc == 0, fix.last_error()
        df = np.abs((f - f0) - fr).max() / np.abs(fr).max()
        de = np.abs(e - er).max() / max(1.0, np.abs(er).max())
        print(f"n={s.n}: rel dF {df:.2e}, rel dE {de:.2e}", flush=True)
        assert df <= TOL[prec] and de <= TOL[prec]
    # the context's other entry points still use their own buffers afterwards
    s, _ = cases[0]
    f2, e2 = fixes[0].compute_res(s.coords())
    assert np.abs(f2 - refs[0][0]).max() <= TOL[prec] * np.abs(refs[0][0]).max()
    lst = fixes[0].list(3, s.coords())
    assert len(lst) > 0
