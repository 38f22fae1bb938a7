"""bastr_step_noalloc (the step without device allocation) against bastr_step
on twin handles with the same field, through the C-ABI: two calls in a row,
the second on a smaller grid with fewer iterations (its scratch holds the
first call's contents beforehand), in 2-D and 3-D, FP64 and FP32, with band
forcing, method-5 forcing and a NaN stop; and a call whose scratch would
exceed the fixed array is rejected before it launches anything."""
import ctypes as C

import numpy as np
import pytest

from bastr_amd import abi
from bastr_amd.solver import BastrError, Solver, diag_table, make_params

pytestmark = pytest.mark.gpu

FIELDS = abi.Diag.SPECALL + abi.Diag.VELGRAD + ("force_factor", "force_E", "force_energy")


def step_noalloc(s, nsteps):
    diags = (abi.Diag * nsteps)()
    spec = np.zeros((nsteps, 3, s.allkmax), dtype=np.float64)
    rc = s._check(s._lib.bastr_step_noalloc(s._h, nsteps, diags, spec.ctypes.data_as(C.POINTER(C.c_double))))
    return rc, diags, spec


def same(a, b):
    ta, tb = diag_table(a, FIELDS), diag_table(b, FIELDS)
    assert np.array_equal(ta, tb, equal_nan=True), np.nanmax(np.abs(ta - tb))
    assert [(d.nstep, d.time, d.nan_flag) for d in a] == [(d.nstep, d.time, d.nan_flag) for d in b]


def twin_run(hip, p, nsteps, seed, zero=False):
    shape = (p.ia, p.ja) if p.ka == 0 else (p.ia, p.ja, p.ka)
    g = Solver(hip, "bastr", p)
    r = Solver(hip, "bastr", p)
    if zero:
        u = [np.zeros(shape, order="F") for _ in range(len(shape))]
        g.set_fields(*u)
        r.set_fields(*u)
    else:
        g.init_random_phase(k0=4.0, seed=seed, energy=1.0)
        r.init_random_phase(k0=4.0, seed=seed, energy=1.0)
    rg, dg, sg = step_noalloc(g, nsteps)
    rr, dr, sr = r.step(nsteps)
    assert rg == rr
    same(dg, dr)
    assert np.array_equal(sg, sr, equal_nan=True)
    assert all(np.array_equal(x, y, equal_nan=True) for x, y in zip(g.get_fields(), r.get_fields()))
    assert g.get_time() == r.get_time()
    # a second call on the same handle continues the run like bastr_step
    if rg == 0:
        rg2, dg2, sg2 = step_noalloc(g, 2)
        rr2, dr2, sr2 = r.step(2)
        assert rg2 == rr2
        same(dg2, dr2)
        assert np.array_equal(sg2, sr2, equal_nan=True)
    g.close()
    r.close()
    return rg


@pytest.mark.parametrize("case", [
    # (first: grid, params, nsteps), (second, smaller)
    (((256, 256, 0), dict(forcemethod=2, lproject=True), 6), ((128, 64, 0), dict(forcemethod=5, lproject=True), 3)),
    (((64, 64, 64), dict(forcemethod=2, lproject=True), 3), ((32, 32, 32), dict(forcemethod=0), 2)),
    (((256, 256, 0), dict(forcemethod=2, precision=abi.PREC_F32), 5), ((64, 64, 0), dict(forcemethod=1), 4)),
])
def test_noalloc_equals_step(hip, case):
    for (ia, ja, ka), kw, n in case:
        print(f"case noalloc {ia}x{ja}x{ka} {kw} n={n}", flush=True)
        p = make_params(ia, ja, ka=ka, kmax=128, target_energy=0.9 if kw.get("forcemethod") in (1, 5) else 0.075,
                        **kw)
        assert twin_run(hip, p, n, seed=3) == 0


def test_noalloc_nan_stop(hip):
    print("case noalloc nan stop 32^2", flush=True)
    p = make_params(32, 32, forcemethod=2)
    assert twin_run(hip, p, 4, seed=0, zero=True) == abi.BASTR_STOP_NAN


def test_noalloc_rejects_oversized_calls(hip):
    print("case noalloc oversized", flush=True)
    p = make_params(2048, 2048, forcemethod=2, lproject=True, kmax=200)
    g = Solver(hip, "bastr", p)
    g.init_random_phase(k0=4.0, seed=1, energy=1.0)
    with pytest.raises(BastrError) as e:
        step_noalloc(g, 4000)
    assert e.value.code == abi.BASTR_EINVAL and "scratch" in str(e.value)
    assert g.get_time() == (0, 0.0)
    rc, d, _ = step_noalloc(g, 1)  # still usable
    assert rc == 0 and d[0].nstep == 0
    g.close()
