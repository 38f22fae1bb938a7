"""GPU ensembles: nmembers independent 2-D runs in one handle
(bastr_create_ensemble).  The reference has no ensemble (it advances one
field set); each member must be exactly the run a one-field handle makes with
the same parameters, so the checks are bitwise against one-field handles (the
one-field path is itself checked against the oracle in test_gpu_parity.py),
plus one member against the oracle directly.
"""
import numpy as np
import pytest

from bastr_amd import abi
from bastr_amd.fields import hitgen
from bastr_amd.solver import Ensemble, Solver, diag_table, make_params

pytestmark = pytest.mark.gpu

FIELDS = abi.Diag.SPECALL + abi.Diag.VELGRAD + ("force_factor", "force_E", "force_energy")


def member_params(p, forced):
    q = make_params(p.ia, p.ja, forcemethod=p.forcemethod if forced else 0, lproject=bool(p.lproject),
                    kmax=p.kmax, timemethod=p.timemethod, precision=p.precision, forcek=p.forcek,
                    target_energy=p.target_energy)
    return q


def same_diags(a, b):
    ta, tb = diag_table(a, FIELDS), diag_table(b, FIELDS)
    assert np.array_equal(ta, tb, equal_nan=True), np.nanmax(np.abs(ta - tb))
    assert [(d.nstep, d.time, d.nan_flag) for d in a] == [(d.nstep, d.time, d.nan_flag) for d in b]


@pytest.mark.parametrize("shape,fm,tm,prec", [
    ((64, 64), 2, 1, abi.PREC_F64),
    ((128, 64), 5, 2, abi.PREC_F64),
    ((64, 128), 1, 1, abi.PREC_F32),
    ((256, 256), 2, 1, abi.PREC_F32),
])
def test_members_equal_one_field_runs(hip, shape, fm, tm, prec):
    print(f"case ensemble {shape} fm={fm} tm={tm} prec={prec}", flush=True)
    forced = [1, 0, 1, 1, 0]
    M = len(forced)
    p = make_params(*shape, forcemethod=fm, lproject=True, kmax=128, timemethod=tm, precision=prec)
    ens = Ensemble(hip, p, M, forced)
    assert ens.nmembers == M
    fields = [hitgen(shape, k0=3 + m, seed=11 + m) for m in range(M)]
    for m in range(M):
        ens.set_member_fields(m, *fields[m])
    singles = []
    for m in range(M):
        s = Solver(hip, "bastr", member_params(p, forced[m]))
        s.set_fields(*fields[m])
        singles.append(s)
    for n in (3, 4):  # two calls: the second replays a new graph length
        re_, de, se = ens.step(n)
        assert re_ == 0
        for m in range(M):
            rs, ds, ss = singles[m].step(n)
            assert rs == 0
            same_diags(de[m], ds)
            assert np.array_equal(se[m], ss)
    for m in range(M):
        a = ens.get_member_fields(m)
        b = singles[m].get_fields()
        assert all(np.array_equal(x, y) for x, y in zip(a, b)), m
        assert ens.get_member_time(m) == singles[m].get_time()
        assert ens.get_member_time(m)[0] == 7


def test_member_matches_oracle(hip, orc):
    print("case ensemble member vs oracle 128x128", flush=True)
    shape = (128, 128)
    p = make_params(*shape, forcemethod=2, lproject=True, kmax=128)
    ens = Ensemble(hip, p, 3, [1, 0, 1])
    fields = [hitgen(shape, k0=4, seed=s) for s in (5, 6, 7)]
    for m in range(3):
        ens.set_member_fields(m, *fields[m])
    _, de, _ = ens.step(4)
    for m in (1, 2):
        o = orc.solver(member_params(p, m != 1))
        o.set_fields(*fields[m])
        _, do, _ = o.step(4)
        a = np.concatenate([x.ravel() for x in ens.get_member_fields(m)])
        b = np.concatenate([x.ravel() for x in o.get_fields()])
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-11
        ta, tb = diag_table(de[m], abi.Diag.SPECALL), diag_table(do, abi.Diag.SPECALL)
        assert np.max(np.abs(ta - tb) / np.abs(tb).max()) < 1e-10


def test_random_phase_members_and_nan_isolation(hip):
    """Members built by the device generator equal one-field generator runs;
    a member whose field is zero under band forcing goes NaN at iteration 1
    (E = 0, solution.F90:370) and stops there while the others advance."""
    print("case ensemble random phase + NaN member 64x64", flush=True)
    shape = (64, 64)
    p = make_params(*shape, forcemethod=2, lproject=True, kmax=128)
    ens = Ensemble(hip, p, 3)
    ens.init_random_phase_member(0, k0=4, seed=3)
    ens.set_member_fields(1, np.zeros(shape, order="F"), np.zeros(shape, order="F"))
    ens.init_random_phase_member(2, k0=6, seed=4)
    rc, de, _ = ens.step(4)
    assert rc == abi.BASTR_STOP_NAN
    assert [d.nan_flag for d in de[1][:2]] == [0, 1]
    assert ens.get_member_time(1) == (1, 1e-4)
    for m, (k0, seed) in ((0, (4, 3)), (2, (6, 4))):
        s = Solver(hip, "bastr", p)
        s.init_random_phase(k0=k0, seed=seed)
        rs, ds, _ = s.step(4)
        assert rs == 0
        same_diags(de[m], ds)
        assert all(np.array_equal(x, y) for x, y in zip(ens.get_member_fields(m), s.get_fields()))
        assert ens.get_member_time(m) == s.get_time()
        assert s.get_time()[0] == 4


def test_single_member_flags(hip):
    """A one-member ensemble keeps its forcing flag: forced=[0] under band
    forcing is the free-decay run."""
    print("case ensemble of one, unforced", flush=True)
    shape = (64, 64)
    p = make_params(*shape, forcemethod=2, lproject=True, kmax=128)
    u = hitgen(shape, k0=4, seed=8)
    e = Ensemble(hip, p, 1, [0])
    e.set_member_fields(0, *u)
    s = Solver(hip, "bastr", member_params(p, False))
    s.set_fields(*u)
    _, de, se = e.step(3)
    _, ds, ss = s.step(3)
    same_diags(de[0], ds)
    assert np.array_equal(se[0], ss)


def test_one_field_entry_points_reject_ensembles(hip):
    print("case ensemble argument checks", flush=True)
    p = make_params(32, 32)
    ens = Ensemble(hip, p, 2)
    u = np.zeros((32, 32), order="F")
    with pytest.raises(Exception):
        ens.set_fields(u, u)
    with pytest.raises(Exception):
        ens.step(1)  # members not set
    with pytest.raises(Exception):
        ens.set_member_fields(2, u, u)
    with pytest.raises(Exception):
        Ensemble(hip, make_params(16, 16, ka=16), 2)


def test_handles_stepped_from_threads(hip):
    """Handles on their own streams stepped from concurrent host threads (each
    call captures its iteration graph) give the same results as one after
    another."""
    print("case ensemble threads", flush=True)
    from concurrent.futures import ThreadPoolExecutor

    def build():
        hs = []
        for shape, M in (((64, 64), 3), ((128, 128), 2)):
            p = make_params(*shape, forcemethod=2, lproject=True, kmax=128, precision=abi.PREC_F32)
            e = Ensemble(hip, p, M, [m % 2 for m in range(M)])
            for m in range(M):
                e.init_random_phase_member(m, k0=3 + m, seed=20 + m)
            hs.append(e)
        s = Solver(hip, "bastr", make_params(128, 64, forcemethod=2, lproject=True))
        s.set_fields(*hitgen((128, 64), k0=4, seed=9))
        hs.append(s)
        return hs

    serial = build()
    ref = [[h.step(n)[1:] for n in (5, 5, 3)] for h in serial]
    conc = build()
    with ThreadPoolExecutor(max_workers=len(conc)) as pool:
        out = list(pool.map(lambda h: [h.step(n) for n in (5, 5, 3)], conc))
    for h, r, o in zip(conc, ref, out):
        for (rd, rs), (rc, od, os_) in zip(r, o):
            assert rc == 0
            if isinstance(h, Ensemble):
                for a, b in zip(od, rd):
                    same_diags(a, b)
            else:
                same_diags(od, rd)
            assert np.array_equal(os_, rs)


def test_step_group_equals_separate_steps(hip):
    """bastr_step_group (every handle launched before any is waited for)
    gives each handle exactly the results of its own bastr_step calls."""
    print("case step_group", flush=True)
    from bastr_amd.solver import step_group

    def build():
        hs = []
        for shape, M in (((64, 64), 3), ((128, 128), 2)):
            p = make_params(*shape, forcemethod=2, lproject=True, kmax=128, precision=abi.PREC_F32)
            e = Ensemble(hip, p, M, [m % 2 for m in range(M)])
            for m in range(M):
                e.init_random_phase_member(m, k0=3 + m, seed=40 + m)
            hs.append(e)
        s = Solver(hip, "bastr", make_params(128, 64, forcemethod=2, lproject=True))
        s.set_fields(*hitgen((128, 64), k0=4, seed=19))
        hs.append(s)
        return hs

    sep = build()
    ref = [[h.step(n)[1:] for n in (4, 4, 6)] for h in sep]
    grp = build()
    got = []
    for n in (4, 4, 6):
        rc, out = step_group(grp, n)
        assert rc == 0
        got.append(out)
    for i, h in enumerate(grp):
        for c in range(3):
            (od, os_), (rd, rs) = got[c][i], ref[i][c]
            if isinstance(h, Ensemble):
                for a, b in zip(od, rd):
                    same_diags(a, b)
            else:
                same_diags(od, rd)
            assert np.array_equal(os_, rs)
    for a, b in zip(grp, sep):
        if isinstance(a, Ensemble):
            for m in range(a.nmembers):
                assert all(np.array_equal(x, y) for x, y in zip(a.get_member_fields(m), b.get_member_fields(m)))
        else:
            assert all(np.array_equal(x, y) for x, y in zip(a.get_fields(), b.get_fields()))
