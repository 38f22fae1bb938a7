"""GPU parity of the 3-D step (mainloop3D / RK33D / CN3D / compute_ut3D /
spectra_compute3D / velgrad_calculate3D / forcing3D) against the CPU oracle.
Tolerances as in 2-D: fields 1e-11 rel-L2 (FP64), diagnostics 1e-10."""
import numpy as np
import pytest

from bastr_amd import abi
from bastr_amd.fields import hitgen
from bastr_amd.solver import Solver, make_params

from test_gpu_parity import check_diags, rel_l2

pytestmark = pytest.mark.gpu


def mixed3(n, seed, k0=4):
    d = hitgen((n, n, n), k0=k0, seed=seed)
    s = hitgen((n, n, n), k0=k0, seed=seed + 1000)
    # permuting the components of a dilatational field gives both parts
    return [np.asfortranarray(d[0] + 0.5 * s[1]), np.asfortranarray(d[1] + 0.5 * s[2]),
            np.asfortranarray(d[2] + 0.5 * s[0])]


@pytest.mark.parametrize("n", [16, 32])
def test_fft3d_compat_matches_numpy(hip, n):
    print(f"case fft3d {n}", flush=True)
    rng = np.random.default_rng(3)
    a = np.asfortranarray(rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n)))
    with Solver(hip, "bastr", make_params(n, n, ka=n)) as s:
        f = s.fft_forward(a)
        ref = np.fft.fftn(a) / a.size
        assert np.abs(f - ref).max() / np.abs(ref).max() < 1e-13
        b = s.fft_backward(f)
        assert np.abs(b - a).max() / np.abs(a).max() < 1e-13


@pytest.mark.parametrize("n,tm,force,proj,nsteps", [
    (32, 1, 0, False, 6),
    (32, 1, 2, True, 5),
    (64, 1, 2, True, 3),
    (32, 2, 0, False, 3),
    (32, 1, 1, False, 4),
    (32, 1, 5, True, 4),
    (128, 1, 2, True, 2),
    (256, 1, 2, True, 1),
])
def test_step3d_parity(hip, orc, n, tm, force, proj, nsteps):
    print(f"case step3d {n} tm={tm} fm={force} proj={proj}", flush=True)
    orc.set_threads(16)
    p = make_params(n, n, ka=n, timemethod=tm, forcemethod=force, lproject=proj, kmax=128,
                    target_energy=0.9 if force in (1, 5) else 0.075)
    u = mixed3(n, seed=21)
    g = Solver(hip, "bastr", p)
    o = orc.solver(p)
    g.set_fields(*u)
    o.set_fields(*u)
    assert rel_l2(g.get_fields(), o.get_fields()) < 1e-14
    rg, dg, sg = g.step(nsteps)
    ro, do, so = o.step(nsteps)
    assert rg == 0 and ro == 0
    err = rel_l2(g.get_fields(), o.get_fields())
    print("rel-L2", err)
    assert err < 1e-11
    check_diags(dg, do)
    assert np.abs(sg - so).max() / np.abs(so[:, 1:, :]).max() < 1e-10
    assert np.array_equal(g.shell_counts(), o.shell_counts())
    if force:
        ff = np.array([d.force_factor for d in dg])
        fo = np.array([d.force_factor for d in do])
        assert np.abs(ff - fo).max() <= 1e-9 * max(1.0, np.abs(fo).max())


def test_step3d_fp32(hip, orc):
    print("case step3d fp32 32", flush=True)
    n = 32
    p64 = make_params(n, n, ka=n, forcemethod=2, lproject=True)
    p32 = make_params(n, n, ka=n, forcemethod=2, lproject=True, precision=abi.PREC_F32)
    u = mixed3(n, seed=22)
    g = Solver(hip, "bastr", p32)
    o = orc.solver(p64)
    g.set_fields(*u)
    o.set_fields(*u)
    _, dg, sg = g.step(4)
    _, do, so = o.step(4)
    assert rel_l2(g.get_fields(), o.get_fields()) < 1e-5
    E_g = sg[:, 1, :] + sg[:, 2, :]
    E_o = so[:, 1, :] + so[:, 2, :]
    assert np.linalg.norm(E_g - E_o) / np.linalg.norm(E_o) < 1e-5


def test_nan_stop_3d(hip, orc):
    print("case nan3d 16", flush=True)
    n = 16
    u = [np.zeros((n, n, n), order="F") for _ in range(3)]
    p = make_params(n, n, ka=n, forcemethod=2)
    g = Solver(hip, "bastr", p)
    o = orc.solver(p)
    g.set_fields(*u)
    o.set_fields(*u)
    rg, dg, _ = g.step(4)
    ro, do, _ = o.step(4)
    assert rg == ro == abi.BASTR_STOP_NAN
    assert g.get_time() == o.get_time() == (1, 1e-4)


def test_device_random_phase_3d(hip, orc):
    print("case init_rp 3d 32", flush=True)
    n = 32
    p = make_params(n, n, ka=n)
    g = Solver(hip, "bastr", p)
    g.init_random_phase(k0=4.0, seed=5, energy=1.0)
    o = orc.solver(p)
    o.init_random_phase(k0=4.0, seed=5, energy=1.0)
    assert rel_l2(g.get_fields(), o.get_fields()) < 1e-12


def test_long_run_parity_3d(hip, orc):
    """30 loop iterations at 128^3 with band forcing and projection: field and
    shell spectrum within 1e-11 of the oracle after the whole run."""
    print("case step3d long 128^3 x 30", flush=True)
    orc.set_threads(16)
    n = 128
    p = make_params(n, n, ka=n, forcemethod=2, lproject=True, kmax=256)
    u = mixed3(n, seed=23)
    g = Solver(hip, "bastr", p)
    o = orc.solver(p)
    g.set_fields(*u)
    o.set_fields(*u)
    rg, dg, sg = g.step(30)
    ro, do, so = o.step(30)
    assert rg == 0 and ro == 0
    err = rel_l2(g.get_fields(), o.get_fields())
    print("128^3 30-step field rel-L2", err)
    assert err < 1e-11
    E_g, E_o = sg[-1, 1] + sg[-1, 2], so[-1, 1] + so[-1, 2]
    assert np.linalg.norm(E_g - E_o) / np.linalg.norm(E_o) < 1e-11
    check_diags(dg, do)
