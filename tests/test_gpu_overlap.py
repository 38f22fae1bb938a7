"""With BASTR_OVERLAP=1 at bastr_create, a 2-D handle runs the spectra, the
dissipation moment's x pass and the diagnostics reduction on a second stream
beside the main chain (the default is one stream).  Both orders must give the
same run bit for bit: the same kernels with the same inputs, only scheduled
differently."""
import numpy as np
import pytest

from bastr_amd import abi
from bastr_amd.fields import hitgen
from bastr_amd.solver import Ensemble, Solver, diag_table, make_params

pytestmark = pytest.mark.gpu

FIELDS = abi.Diag.SPECALL + abi.Diag.VELGRAD + ("force_factor", "force_E", "force_energy")


def run(hip, monkeypatch, overlap, p, u, chunks):
    monkeypatch.setenv("BASTR_OVERLAP", overlap)
    s = Solver(hip, "bastr", p)
    s.set_fields(*u)
    out = [s.step(n) for n in chunks]
    return out, s.get_fields(), s.get_time()


@pytest.mark.parametrize("shape,tm,fm,prec", [
    ((256, 256), 1, 2, abi.PREC_F64),    # y-forward fused with the update (cf2)
    ((256, 2048), 1, 2, abi.PREC_F64),   # 2048-point columns: stage A's y-forward into WA
    ((128, 2048), 2, 5, abi.PREC_F64),   # CN, projected band forcing
    ((512, 512), 1, 1, abi.PREC_F32),    # rescale forcing, FP32
])
def test_overlap_equals_one_stream(hip, monkeypatch, shape, tm, fm, prec):
    print(f"case overlap {shape} tm={tm} fm={fm} prec={prec}", flush=True)
    p = make_params(*shape, forcemethod=fm, lproject=True, kmax=128, timemethod=tm, precision=prec)
    u = hitgen(shape, k0=4, seed=31)
    a, fa, ta = run(hip, monkeypatch, "1", p, u, (3, 5))
    b, fb, tb = run(hip, monkeypatch, "0", p, u, (3, 5))
    for (ra, da, sa), (rb, db, sb) in zip(a, b):
        assert ra == rb == 0
        assert np.array_equal(diag_table(da, FIELDS), diag_table(db, FIELDS), equal_nan=True)
        assert np.array_equal(sa, sb)
    assert all(np.array_equal(x, y) for x, y in zip(fa, fb))
    assert ta == tb


def test_overlap_nan_stop(hip, monkeypatch):
    """u = 0 under band forcing goes NaN at iteration 1 (E = 0,
    solution.F90:370): the field and clock stop as in the one-stream order."""
    print("case overlap nan 64x2048", flush=True)
    shape = (64, 2048)
    p = make_params(*shape, forcemethod=2, lproject=True, kmax=128)
    u = [np.zeros(shape, order="F"), np.zeros(shape, order="F")]
    a, fa, ta = run(hip, monkeypatch, "1", p, u, (4,))
    b, fb, tb = run(hip, monkeypatch, "0", p, u, (4,))
    assert a[0][0] == b[0][0] == abi.BASTR_STOP_NAN
    assert np.array_equal(diag_table(a[0][1], FIELDS), diag_table(b[0][1], FIELDS), equal_nan=True)
    assert [d.nan_flag for d in a[0][1]] == [d.nan_flag for d in b[0][1]]
    assert all(np.array_equal(x, y, equal_nan=True) for x, y in zip(fa, fb))
    assert ta == tb == (1, 1e-4)


def test_overlap_ensemble(hip, monkeypatch):
    print("case overlap ensemble 128x2048", flush=True)
    shape = (128, 2048)
    p = make_params(*shape, forcemethod=2, lproject=True, kmax=128)
    res = []
    for ov in ("1", "0"):
        monkeypatch.setenv("BASTR_OVERLAP", ov)
        e = Ensemble(hip, p, 3, [1, 0, 1])
        for m in range(3):
            e.init_random_phase_member(m, k0=3 + m, seed=50 + m)
        r, d, sp = e.step(4)
        assert r == 0
        res.append((d, sp, [e.get_member_fields(m) for m in range(3)]))
    (da, sa, fa), (db, sb, fb) = res
    for x, y in zip(da, db):
        assert np.array_equal(diag_table(x, FIELDS), diag_table(y, FIELDS), equal_nan=True)
    assert np.array_equal(sa, sb)
    for x, y in zip(fa, fb):
        assert all(np.array_equal(a, b) for a, b in zip(x, y))
