"""Radix plans of the column passes (bastr_autotune / bastr_tune_set; the
reference's analogue is FFTW_MEASURE planning, fftwlink.F90:284-303) and
graphs over several handles (bastr_step_group).

Both plans of a length are checked against the oracle; switching plans remakes
the captured step graphs; the autotuner reports its timings and keeps the
faster plan.  A group graph stays correct when one of its handles is replaced.
"""
import numpy as np
import pytest

import bastr_amd
from bastr_amd import abi
from bastr_amd.fields import hitgen
from bastr_amd.solver import Solver, make_params, step_group

from test_gpu_parity import check_diags, mixed_field, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def default_plans():
    yield
    for n in (32, 64, 128, 256, 512, 1024, 2048):
        for prec in (0, 1):
            bastr_amd.tune_set(n, 16, prec)


PLANS = [(16, 1), (8, 1), (16, 2), (8, 2)]


@pytest.mark.parametrize("shape,prec", [((256, 256), 0), ((512, 2048), 0), ((512, 512), 1), ((512, 2048), 1)])
def test_every_plan_matches_the_oracle(hip, orc, shape, prec):
    print(f"case column plans {shape} prec={prec}", flush=True)
    orc.set_threads(16)
    n1 = shape[1]
    p = make_params(*shape, forcemethod=2, lproject=True, kmax=128, precision=prec)
    u = mixed_field(shape, seed=31)
    res = []
    plans = []
    for pts, wid in PLANS:
        try:
            bastr_amd.tune_set(n1, pts, prec, wid)
        except Exception:
            continue  # not a plan of this length and precision
        plans.append((pts, wid))
        assert bastr_amd.tune_get(n1, prec) == (pts, wid)
        g = Solver(hip, "bastr", p)
        g.set_fields(*u)
        rg, dg, sg = g.step(4)
        assert rg == 0
        res.append((g.get_fields(), dg, sg))
        g.close()
    o = orc.solver(make_params(*shape, forcemethod=2, lproject=True, kmax=128))
    o.set_fields(*u)
    ro, do, so = o.step(4)
    tol = 1e-11 if prec == 0 else 1e-5
    for f, dg, sg in res:
        assert rel_l2(f, o.get_fields()) < tol
        if prec == 0:
            check_diags(dg, do)
    assert len(plans) >= 2
    for (pa, wa), r in zip(plans, res):
        # the 8-point plans differ in rounding only; the width changes nothing
        if pa == 16:
            assert all(np.array_equal(x, y) for x, y in zip(r[0], res[0][0])) if wa == 2 else True
        assert rel_l2(r[0], res[0][0]) < (1e-13 if prec == 0 else 1e-5)


def test_plan_change_remakes_captured_graphs(hip):
    print("case radix plan switch under graph replay", flush=True)
    p = make_params(256, 256, forcemethod=2, lproject=True, kmax=128)
    u = hitgen((256, 256), k0=4, seed=8)
    a, b, d = (Solver(hip, "bastr", p) for _ in range(3))
    for s in (a, b, d):
        s.set_fields(*u)
    bastr_amd.tune_set(256, 16)
    for s in (a, b, d):
        s.step(4)  # captures each handle's graph with the 16-point plan
    bastr_amd.tune_set(256, 8)
    a.step(4)
    b.step(4)
    bastr_amd.tune_set(256, 16)
    d.step(4)
    fa, fb, fd = a.get_fields(), b.get_fields(), d.get_fields()
    assert all(np.array_equal(x, y) for x, y in zip(fa, fb))
    assert not all(np.array_equal(x, y) for x, y in zip(fa, fd))  # the switch took effect
    assert rel_l2(fa, fd) < 1e-13


@pytest.mark.parametrize("n,prec", [(2048, 0), (512, 1)])
def test_autotune_keeps_the_fastest_plan(hip, n, prec):
    print(f"case autotune {n} prec={prec}", flush=True)
    plan, ms = bastr_amd.autotune(n, prec)
    print("autotune", n, prec, plan, ms, flush=True)
    assert bastr_amd.tune_get(n, prec) == plan
    assert np.isfinite(ms[0]) and ms[0] > 0
    times = {p: t for p, t in zip(PLANS, ms) if np.isfinite(t)}
    fast = {p: t for p, t in times.items() if t < 0.97 * times[(16, 1)]}
    assert plan == (min(fast, key=fast.get) if fast else (16, 1))


def test_group_graph_with_a_replaced_handle(hip):
    print("case step_group one graph, handle replaced", flush=True)

    def mk(shape, seed):
        s = Solver(hip, "bastr", make_params(*shape, forcemethod=2, lproject=True, kmax=128))
        s.set_fields(*hitgen(shape, k0=4, seed=seed))
        return s

    a, b = mk((128, 128), 1), mk((64, 64), 2)
    a2, b2 = mk((128, 128), 1), mk((64, 64), 2)
    rc, _ = step_group([a, b], 5)
    assert rc == 0
    a2.step(5)
    b2.step(5)
    b.close()
    b2.close()
    c, c2 = mk((64, 64), 3), mk((64, 64), 3)  # a new handle, possibly at b's address
    for _ in range(2):  # capture, then replay
        rc, out = step_group([a, c], 5)
        assert rc == 0
        a2.step(5)
        c2.step(5)
    for x, y in ((a, a2), (c, c2)):
        assert all(np.array_equal(f, g) for f, g in zip(x.get_fields(), y.get_fields()))


@pytest.mark.parametrize("shape,prec", [((256, 256), 0), ((512, 512), 1), ((32, 32, 32), 0)])
def test_autotune_step_restores_the_state(hip, shape, prec):
    """bastr_autotune_step times the plans on the handle's own step and puts
    the state back: the handle then steps exactly as a fresh handle with the
    chosen plan set by hand (fields, diagnostics, nstep, time)."""
    print(f"case autotune_step {shape} prec={prec}", flush=True)
    nd = len(shape)
    p = make_params(*shape, forcemethod=2, lproject=True, kmax=128, precision=prec)
    u = mixed_field(shape, seed=41) if nd == 2 else hitgen(shape, k0=4, seed=41)
    lens = sorted(set(shape[1:]))
    for n in lens:
        bastr_amd.tune_set(n, 16, prec)
    a = Solver(hip, "bastr", p)
    a.set_fields(*u)
    a.step(3)
    plan, ms = bastr_amd.autotune_step(a, iters=6)
    print("autotune_step", shape, prec, plan, ms, flush=True)
    assert np.isfinite(ms[0]) and ms[0] > 0
    times = {q: t for q, t in zip(PLANS, ms) if np.isfinite(t)}
    fast = {q: t for q, t in times.items() if t < 0.97 * times[(16, 1)]}
    assert plan == (min(fast, key=fast.get) if fast else (16, 1))
    for n in lens:
        assert bastr_amd.tune_get(n, prec) == plan
    ra, da, sa = a.step(4)
    # the same calls on a fresh handle: 3 iterations with the default plan, 4 with the chosen one
    for n in lens:
        bastr_amd.tune_set(n, 16, prec)
    b = Solver(hip, "bastr", p)
    b.set_fields(*u)
    b.step(3)
    for n in lens:
        bastr_amd.tune_set(n, plan[0], prec, plan[1])
    rb, db, sb = b.step(4)
    assert ra == 0 and rb == 0
    assert all(np.array_equal(x, y) for x, y in zip(a.get_fields(), b.get_fields()))
    assert [(d.nstep, d.time, d.Esspe, d.Edspe) for d in da] == [(d.nstep, d.time, d.Esspe, d.Edspe) for d in db]
    assert np.array_equal(sa, sb)
    a.close()
    b.close()


def test_autotune_step_on_an_ensemble(hip):
    print("case autotune_step ensemble 3 x 128^2", flush=True)
    p = make_params(128, 128, forcemethod=2, lproject=True, kmax=64)
    e = bastr_amd.ensemble(p, 3, forced=[1, 0, 1])
    f = bastr_amd.ensemble(p, 3, forced=[1, 0, 1])
    for m in range(3):
        u = hitgen((128, 128), k0=3 + m, seed=50 + m)
        e.set_member_fields(m, *u)
        f.set_member_fields(m, *u)
    bastr_amd.tune_set(128, 16)
    assert e.step(1)[0] == 0 and f.step(1)[0] == 0  # both entry transforms with the default plan
    plan, ms = bastr_amd.autotune_step(e, iters=4)
    print("autotune_step ensemble", plan, ms, flush=True)
    assert np.isfinite(ms[0])
    assert bastr_amd.tune_get(128, 0) == plan
    re, de, se = e.step(3)
    rf, df, sf = f.step(3)
    assert re == 0 and rf == 0
    for m in range(3):
        assert all(np.array_equal(x, y) for x, y in zip(e.get_member_fields(m), f.get_member_fields(m)))
    assert np.array_equal(se, sf)
    e.close()
    f.close()


def test_autotune_step_refuses_a_nan_run(hip):
    """u = 0 with band forcing reaches NaN in its first iteration: the tuner
    fails and leaves the plans as they were."""
    print("case autotune_step NaN", flush=True)
    n = 256
    p = make_params(n, n, forcemethod=2, lproject=False, kmax=32)
    g = Solver(hip, "bastr", p)
    g.set_fields(np.zeros((n, n), order="F"), np.zeros((n, n), order="F"))
    bastr_amd.tune_set(n, 8)
    with pytest.raises(bastr_amd.BastrError):
        bastr_amd.autotune_step(g, iters=2)
    assert bastr_amd.tune_get(n, 0) == (8, 1)
    g.close()
    bastr_amd.tune_set(n, 16)


def test_bastr_tune_env_tunes_at_the_first_step(hip, monkeypatch):
    """BASTR_TUNE=1 at bastr_create: the first bastr_step tunes the handle's
    column length on its own step and then steps exactly as a handle created
    with the chosen plan already set (fields, diagnostics, nstep, time)."""
    print("case BASTR_TUNE=1 first-step tuning 256^2", flush=True)
    p = make_params(256, 256, forcemethod=2, lproject=True, kmax=128)
    u = mixed_field((256, 256), seed=61)
    bastr_amd.tune_set(256, 16, 0)
    monkeypatch.setenv("BASTR_TUNE", "1")
    a = Solver(hip, "bastr", p)
    monkeypatch.delenv("BASTR_TUNE")
    a.set_fields(*u)
    ra, da, sa = a.step(5)
    plan = bastr_amd.tune_get(256, 0)
    print("BASTR_TUNE plan", plan, flush=True)
    b = Solver(hip, "bastr", p)  # the plan table is process-wide: b steps with the chosen plan
    b.set_fields(*u)
    rb, db, sb = b.step(5)
    assert ra == 0 and rb == 0
    assert [(d.nstep, d.time) for d in da] == [(d.nstep, d.time) for d in db]
    if plan[0] == 16:
        # a's entry transform ran with the default plan, b's with the chosen one:
        # the same arithmetic for 16-point plans, so the runs agree bit for bit
        assert all(np.array_equal(x, y) for x, y in zip(a.get_fields(), b.get_fields()))
        assert np.array_equal(sa, sb)
    else:
        # an 8-point plan rounds the entry transform differently
        assert rel_l2(a.get_fields(), b.get_fields()) < 1e-13
    a.close()
    b.close()
