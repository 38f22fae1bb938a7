"""The drop-in host driver bastr_amd/bin/bastr_run: datin/input.dat parsing
(readwrite.F90:19-112), HDF5 field files (hdf5io.F90:204-256, 752-802),
log files in Fortran E20.13E2 format (utility.F90:87-164), dump cadence and
restart (solution.F90:69-81, readwrite.F90:183-220, utility.F90:42-62).

Dumps are read back with the image's h5dump (an independent HDF5 reader)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from bastr_amd.fields import hitgen
from bastr_amd.solver import Solver, make_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "bastr_amd", "bin", "bastr_run")
H5DUMP = "/opt/conda/bin/h5dump"


def write_input(d, ia, ja, ka, maxstep, dt, lwsequ, feqwsequ, lwspectra, feqwspe, kmax, tm, fm, fk, et,
                lproj, init, re_="5000.d0"):
    os.makedirs(os.path.join(d, "datin"), exist_ok=True)
    L = lambda b: "t" if b else "f"  # noqa: E731
    txt = f"""########################################################
#              Bastr input                             #
########################################################

# ia,ja,ka
{ia},{ja},{ka}

# ref_t,reynolds
273.15d0, {re_}

# maxstep,deltat,lwsequ,feqwsequ,lwspectra,feqwspe,kmax,timemethod
{maxstep}, {dt}, {L(lwsequ)}, {feqwsequ}, {L(lwspectra)}, {feqwspe}, {kmax}, {tm}

# forcemethod, forcek, targetenergy, lproject
{fm},{fk},{et},{L(lproj)}

# initialmethod
{init}
"""
    with open(os.path.join(d, "datin", "input.dat"), "w") as f:
        f.write(txt)


def run(d, *args):
    r = subprocess.run([DRIVER, *args], cwd=d, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    return r


def test_parse_input_dat(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("driver not built; run __graft_entry__.build()")
    write_input(tmp_path, 256, 128, 0, 200, "1.d-4", True, 50, True, 20, 128, 2, 5, 5, "0.075d0", True, 1)
    r = run(tmp_path, "--parse")
    assert r.returncode == 0
    kv = dict(x.split("=") for x in r.stdout.strip().splitlines()[-1].split())
    assert (kv["ia"], kv["ja"], kv["ka"]) == ("256", "128", "0")
    assert float(kv["deltat"]) == 1e-4 and float(kv["target_energy"]) == 0.075
    assert kv["lwsequ"] == "1" and kv["lprojectd"] == "1" and kv["timemethod"] == "2"
    assert kv["forcemethod"] == "5" and kv["feqwspe"] == "20" and kv["initialmethod"] == "1"


def h5read(path, name, shape):
    """Read the solver block of a field dataset with h5dump (binary, LE)."""
    out = str(path) + "." + name + ".bin"
    subprocess.run([H5DUMP, "-d", "/" + name, "-b", "LE", "-o", out, str(path)], check=True,
                   capture_output=True)
    raw = np.fromfile(out, dtype="<f8")
    ia, ja = shape[0], shape[1]
    ka = shape[2] if len(shape) == 3 else 0
    a = raw.reshape(ka + 1, ja + 1, ia + 1)  # C extent of Fortran (ia+1, ja+1, ka+1)
    blk = a[: max(ka, 1), :ja, :ia]
    return np.asfortranarray(blk.transpose(2, 1, 0).reshape(shape, order="F") if ka else blk[0].T)


def h5scalar(path, name, dtype):
    out = str(path) + "." + name + ".bin"
    subprocess.run([H5DUMP, "-d", "/" + name, "-b", "LE", "-o", out, str(path)], check=True,
                   capture_output=True)
    return np.fromfile(out, dtype=dtype)[0]


FLOAT = r"(-| )0\.\d{13}E[+-]\d{2}"


def read_log(path):
    lines = open(path).read().splitlines()
    rows = []
    for l in lines[1:]:
        assert re.fullmatch(r" {1}[ \d]{19}\d( " + FLOAT + r")+", l), l
        parts = l.split()
        rows.append([int(parts[0])] + [float(x) for x in parts[1:]])
    return lines[0], rows


@pytest.mark.gpu
def test_driver_run_dumps_and_restart(tmp_path):
    print("case driver 64x64", flush=True)
    n, maxstep = 64, 30
    write_input(tmp_path, n, n, 0, maxstep, "1.d-4", True, 10, True, 10, 128, 1, 2, 5, "0.075d0", True, 1)
    assert run(tmp_path, "hitgen", "7", "4", "1.0").returncode == 0
    u0 = [h5read(tmp_path / "datin" / "flowini2d.h5", v, (n, n)) for v in ("u1", "u2")]
    assert abs(np.mean(u0[0] ** 2 + u0[1] ** 2) - 1.0) < 1e-12  # u_rms = 1
    # the same construction as the numpy generator and the device generator
    ref = hitgen((n, n), k0=4, seed=7, energy=1.0)
    assert max(np.abs(a - b).max() for a, b in zip(u0, ref)) < 1e-12 * max(np.abs(b).max() for b in ref)
    r = run(tmp_path)
    assert r.returncode == 0, r.stdout
    logs = tmp_path / "log"
    h, spec = read_log(logs / "stat2d_specall.dat")
    assert h == "".join(f" {x:>20s}" for x in "ns ti Es Ed k2Ed kMEd".split())
    assert [r_[0] for r_ in spec] == list(range(maxstep + 1))
    _, vg = read_log(logs / "stat2d_velgrad.dat")
    _, fo = read_log(logs / "factor_output.dat")
    _, sp = read_log(logs / "stat2d_spect.dat")
    assert [r_[0] for r_ in fo] == list(range(1, maxstep + 2))
    K = min(n // 3, 128)
    assert len(sp) == 4 * K and sorted({r_[0] for r_ in sp}) == [0, 10, 20, 30]
    # the same run through the library API
    p = make_params(n, n, forcemethod=2, lproject=True, kmax=128)
    import bastr_amd
    s = Solver(bastr_amd.load(), "bastr", p)
    s.set_fields(*u0)
    rc, d, spc = s.step(maxstep + 1)
    for i, row in enumerate(spec):
        ref = [d[i].time, d[i].Esspe, d[i].Edspe, d[i].k2Edspe, d[i].kMEdspe]
        assert np.allclose(row[1:], ref, rtol=1e-12, atol=1e-300), (i, row, ref)
    for i, row in enumerate(vg):
        ref = [d[i].umumtheta2, d[i].umumijji, d[i].u2theta, d[i].dissp, d[i].eta_min, d[i].Taylength,
               d[i].ReTay, d[i].Kollength]
        assert np.allclose(row[2:], ref, rtol=1e-12), (i, row, ref)
    for i, row in enumerate(fo):
        assert np.allclose(row[2:], [d[i].force_factor, d[i].force_E], rtol=1e-12)
    # dumps: files 1..3 at nstep 10, 20, 30 (solution.F90:69-79)
    for fnum, ns in ((1, 10), (2, 20), (3, 30)):
        fpath = tmp_path / "outdat" / f"flowfield{fnum:04d}.h5"
        assert h5scalar(fpath, "nstep", "<i4") == ns
        assert abs(h5scalar(fpath, "time", "<f8") - d[ns].time) == 0
    s2 = Solver(bastr_amd.load(), "bastr", p)
    s2.set_fields(*u0)
    s2.step(20)
    f20 = s2.get_fields()
    for v, ref in zip(("u1", "u2"), f20):
        got = h5read(tmp_path / "outdat" / "flowfield0002.h5", v, (n, n))
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    # restart from file 2: logs resume at nstep 20 and end like the full run
    full = {k: read_log(logs / k)[1] for k in ("stat2d_specall.dat", "stat2d_velgrad.dat", "factor_output.dat")}
    r = run(tmp_path, "0002")
    assert r.returncode == 0, r.stdout
    for k, rows in full.items():
        _, again = read_log(logs / k)
        if k == "factor_output.dat":
            # listinit resumes at the first record with ns >= nstep (utility.F90:42-62):
            # the line ns = 20, written by iteration 19's forcing, is overwritten and the
            # restarted loop's first forcing line is ns = 21, exactly as in the reference
            rows = [x for x in rows if x[0] != 20]
        assert [x[0] for x in again] == [x[0] for x in rows]
        # Es of this purely dilatational field is roundoff (~1e-32): an absolute floor at 1e-13
        assert np.allclose(np.array(again), np.array(rows), rtol=1e-11, atol=1e-13), k
    # filenumb = 2 + 1 after the restart (readwrite.F90:215-217): the nstep-30 dump is file 3 again
    assert not (tmp_path / "outdat" / "flowfield0004.h5").exists()
    f3 = tmp_path / "outdat" / "flowfield0003.h5"
    assert h5scalar(f3, "nstep", "<i4") == 30


@pytest.mark.gpu
def test_driver_3d_and_nan_stop(tmp_path):
    print("case driver 3d 16^3 nan", flush=True)
    n = 16
    # initialmethod 0 (u = 0) with band forcing: E = 0, the field becomes NaN
    write_input(tmp_path, n, n, n, 5, "1.d-4", False, 10, True, 2, 128, 1, 2, 5, "0.075d0", False, 0)
    r = run(tmp_path)
    assert r.returncode == 0 and 'STOP "E" is a NaN' in r.stdout
    lines = open(tmp_path / "log" / "stat3d_specall.dat").read().splitlines()
    assert len(lines) == 1 + 2 and "NaN" in lines[-1]
    assert (tmp_path / "outdat" / "flowfield0001.h5").exists()  # NaN dump (solution.F90:141-153)


@pytest.mark.gpu
def test_driver_nan_on_a_dump_iteration_writes_one_file(tmp_path):
    """The reference's single condition '(lwsequ .and. nstep==nxtwsequ) .or.
    isnan(E)' (solution.F90:69-79) writes one file when the NaN falls on a
    scheduled dump: u = 0 with band forcing is NaN from nstep 1, the first
    dump is at min(feqwsequ, maxstep) = 1."""
    print("case driver 2d 32^2 nan on dump", flush=True)
    n = 32
    write_input(tmp_path, n, n, 0, 5, "1.d-4", True, 1, False, 2, 128, 1, 2, 5, "0.075d0", False, 0)
    r = run(tmp_path)
    assert r.returncode == 0 and 'STOP "E" is a NaN' in r.stdout
    files = sorted(os.listdir(tmp_path / "outdat"))
    assert files == ["flowfield0001.h5"], files


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(64, 32), (16, 16, 16)])
def test_driver_postprocessing_commands(tmp_path, orc, shape):
    """spectra2d / transfer2d / transfer3d NNNN (Bastrpp.F90:174-599): read
    outdat/flowfieldNNNN.h5, write pp/*.dat with the dump's nstep and time."""
    print(f"case driver pp {shape}", flush=True)
    nd = len(shape)
    ka = shape[2] if nd == 3 else 0
    write_input(tmp_path, shape[0], shape[1], ka, 4, "1.d-4", True, 4, False, 10, 128, 1, 0, 5, "0.075d0",
                False, 1)
    assert run(tmp_path, "hitgen", "5", "3", "1.0").returncode == 0
    assert run(tmp_path).returncode == 0
    dump = tmp_path / "outdat" / "flowfield0001.h5"
    u = [h5read(dump, v, shape) for v in ("u1", "u2", "u3")[:nd]]
    ns, ti = int(h5scalar(dump, "nstep", "<i4")), float(h5scalar(dump, "time", "<f8"))
    assert ns == 4
    o = orc.solver(make_params(shape[0], shape[1], ka, kmax=128))
    K = min(shape[0] // 3, 128)
    cmds = ["transfer3d"] if nd == 3 else ["spectra2d", "transfer2d"]
    for cmd in cmds:
        r = run(tmp_path, cmd, "1")
        assert r.returncode == 0, r.stdout
    _, tr = read_log(tmp_path / "pp" / "Transfer0001.dat")
    assert len(tr) == K and all(x[0] == ns and x[1] == ti for x in tr)
    assert [x[2] for x in tr] == list(range(1, K + 1))
    ref = o.pp_transfer(*u)
    got = np.array([x[3:] for x in tr]).T
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()
    if nd == 2:
        h, sp = read_log(tmp_path / "pp" / "Spectra0001.dat")
        assert h == "".join(f" {x:>20s}" for x in "ns ti k Es Ed".split())
        sh, tot = o.pp_spectra(*u)
        assert np.abs(np.array([x[3:] for x in sp]).T - sh).max() <= 1e-11 * np.abs(sh).max()
        _, aux = read_log(tmp_path / "pp" / "Spectra_aux0001.dat")
        assert len(aux) == 1 and aux[0][0] == ns
        assert np.allclose(aux[0][2:], tot, rtol=1e-11)
    # a 2-D command on a 3-D case (and the reverse) stops with a message
    r = run(tmp_path, "transfer2d" if nd == 3 else "transfer3d", "1")
    assert r.returncode != 0 and "needs a" in r.stdout
