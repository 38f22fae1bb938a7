"""The 256^2 free-decay reference-check case (tests/golden/decay2d_256,
written by tests/golden/make_decay2d_256_case.py): 256^2 FP64 free decay, 201 loop iterations.
Its expected logs are pinned here against the oracle on the CPU (so the kit a
maintainer runs the reference binary on is itself checked), and the driver
reproduces them on the GPU."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from bastr_amd.solver import make_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASE = os.path.join(ROOT, "tests", "golden", "decay2d_256")
EXP = os.path.join(CASE, "expected", "log")
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
from compare_logs import compare, read_log  # noqa: E402
from test_driver import DRIVER, h5read  # noqa: E402


def test_expected_logs_match_oracle(orc, tmp_path):
    f = tmp_path / "flowini2d.h5"  # h5read writes its dumps beside the file it reads
    shutil.copy(os.path.join(CASE, "datin", "flowini2d.h5"), f)
    u = [h5read(f, v, (256, 256)) for v in ("u1", "u2")]
    # the file's field: u_rms = 1 (hitgen, solution.F90:1163)
    assert abs(np.mean(u[0] ** 2 + u[1] ** 2) - 1.0) < 1e-12
    orc.set_threads(min(8, os.cpu_count() or 1))
    o = orc.solver(make_params(256, 256, forcemethod=0, lproject=False, kmax=128))
    o.set_fields(*u)
    rc, d, spec = o.step(201)
    assert rc == 0
    hdr, sa = read_log(os.path.join(EXP, "stat2d_specall.dat"))
    assert list(sa[:, 0]) == list(range(201))
    tot = sa[:, 2] + sa[:, 3]
    for i in range(201):
        assert sa[i, 1] == pytest.approx(d[i].time, rel=1e-12, abs=1e-300)
        for j, f in ((2, "Esspe"), (3, "Edspe")):
            assert abs(sa[i, j] - getattr(d[i], f)) <= 1e-10 * tot[i]
        for j, f in ((4, "k2Edspe"), (5, "kMEdspe")):
            assert sa[i, j] == pytest.approx(getattr(d[i], f), rel=1e-10)
    _, vg = read_log(os.path.join(EXP, "stat2d_velgrad.dat"))
    names = ("umumtheta2", "umumijji", "u2theta", "dissp", "eta_min", "Taylength", "ReTay", "Kollength")
    got = np.array([[getattr(x, n) for n in names] for x in d])
    assert np.abs(vg[:, 2:] - got).max(axis=0).max() <= 1e-10 * np.abs(got).max()
    # spectra at nstep 0, 100, 200 (feqwspe = 100): k, Es, Ed per shell 1..85
    _, sp = read_log(os.path.join(EXP, "stat2d_spect.dat"))
    K = 85
    assert sp.shape[0] == 3 * K
    for b, ns in enumerate((0, 100, 200)):
        blk = sp[b * K:(b + 1) * K]
        assert (blk[:, 0] == ns).all()
        assert np.allclose(blk[:, 2], spec[ns, 0], rtol=1e-12)
        E_ref = spec[ns, 1] + spec[ns, 2]
        assert np.abs(blk[:, 3] + blk[:, 4] - E_ref).max() <= 1e-10 * E_ref.max()


@pytest.mark.gpu
def test_driver_reproduces_expected_logs(tmp_path, orc):
    """The driver's run of the case reproduces the expected logs, and its dump
    at nstep = 200 (written before that iteration's step, solution.F90:69-79)
    is the oracle's field after 200 steps to 1e-11 relative L2."""
    print("case golden decay2d_256 driver", flush=True)
    work = tmp_path / "decay2d_256"
    shutil.copytree(os.path.join(CASE, "datin"), work / "datin")
    r = subprocess.run([DRIVER], cwd=work, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    worst = compare(str(work / "log"), EXP, 1e-12)
    assert all(v <= 1e-12 for v in worst.values()), worst
    f = tmp_path / "flowini2d.h5"
    shutil.copy(os.path.join(CASE, "datin", "flowini2d.h5"), f)
    u0 = [h5read(f, v, (256, 256)) for v in ("u1", "u2")]
    got = [h5read(work / "outdat" / "flowfield0001.h5", v, (256, 256)) for v in ("u1", "u2")]
    orc.set_threads(min(16, os.cpu_count() or 1))
    o = orc.solver(make_params(256, 256, forcemethod=0, lproject=False, kmax=128))
    o.set_fields(*u0)
    assert o.step(200)[0] == 0
    ref = o.get_fields()
    num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(got, ref))
    err = np.sqrt(num / sum(float(np.sum(b * b)) for b in ref))
    print("256^2 decay 200-step field rel-L2", err)
    assert err < 1e-11
