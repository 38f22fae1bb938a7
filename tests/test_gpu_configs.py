"""The benchmark workloads at their full sizes on the GPU:

- 3-D 512^3 FP64, band forcing (k = 5, E_t = 0.075) with projection: one loop
  iteration against the oracle (mainloop3D / RK33D / forcing3D,
  solution.F90:100-170, 233-270, 542-640); FP32 at 512^3 against the FP64
  device run;
- 2-D 2048^2 FP32 against the FP64 device run (the FP32 variant of the
  2048^2 band-forced, projected workload);
- the 96-member ensemble of 2-D runs (48 x 256^2, 32 x 512^2, 16 x 1024^2,
  FP32, band forcing on every other member, k0 drawn from {3..8}, one seed
  per member) stepped through bastr_step_group, with one forced and one
  free-decay member of every size bitwise equal to its one-field run, and an
  FP64 ensemble with one member per size against the oracle.

Initial fields come from the device generator, which is itself checked
against the oracle's generator (test_gpu_parity.py, test_gpu_parity3d.py).
"""
import numpy as np
import pytest

from bastr_amd import abi
from bastr_amd.solver import Ensemble, Solver, make_params, step_group

from test_gpu_ensemble import member_params, same_diags
from test_gpu_parity import check_diags, rel_l2

pytestmark = pytest.mark.gpu


def shell_energy(sp):
    """E = Es + Ed per shell, iterations x shells."""
    return sp[:, 1, :] + sp[:, 2, :]


def check_spectra(sg, so, tol=1e-10):
    """kn (mean |k| per shell, host-summed in different orders) relative to
    itself; Es and Ed against the largest shell energy."""
    kn_err = float(np.abs(sg[:, 0, :] - so[:, 0, :]).max() / np.abs(so[:, 0, :]).max())
    e_err = float(np.abs(sg[:, 1:, :] - so[:, 1:, :]).max() / np.abs(so[:, 1:, :]).max())
    assert kn_err < 1e-12 and e_err < tol, (kn_err, e_err)


def ek_err(a, b):
    Ea, Eb = shell_energy(a), shell_energy(b)
    return float(np.linalg.norm(Ea - Eb) / np.linalg.norm(Eb))


@pytest.mark.timeout(900)
def test_512cubed_fp64_iteration_matches_oracle(hip, orc):
    print("case 3d 512^3 fp64 band+proj, 1 iteration vs oracle", flush=True)
    orc.set_threads(16)
    n = 512
    p = make_params(n, n, ka=n, forcemethod=2, forcek=5, target_energy=0.075, lproject=True, kmax=256)
    g = Solver(hip, "bastr", p)
    g.init_random_phase(k0=4.0, seed=1, energy=1.0)
    u = g.get_fields()
    o = orc.solver(p)
    o.set_fields(*u)
    del u
    rg, dg, sg = g.step(1)
    ro, do, so = o.step(1)
    assert rg == 0 and ro == 0
    err = rel_l2(g.get_fields(), o.get_fields())
    print("512^3 FP64 field rel-L2", err, flush=True)
    assert err < 1e-11
    check_diags(dg, do)
    check_spectra(sg, so)
    assert abs(dg[0].force_factor - do[0].force_factor) <= 1e-9 * max(1.0, abs(do[0].force_factor))
    o.close()


def test_512cubed_fp32_against_fp64(hip):
    print("case 3d 512^3 fp32 vs fp64 device, 5 iterations", flush=True)
    n = 512
    kw = dict(forcemethod=2, forcek=5, target_energy=0.075, lproject=True, kmax=256)
    g64 = Solver(hip, "bastr", make_params(n, n, ka=n, **kw))
    g64.init_random_phase(k0=4.0, seed=1, energy=1.0)
    u = g64.get_fields()
    g32 = Solver(hip, "bastr", make_params(n, n, ka=n, precision=abi.PREC_F32, **kw))
    g32.set_fields(*u)
    del u
    r64, d64, s64 = g64.step(5)
    r32, d32, s32 = g32.step(5)
    assert r64 == 0 and r32 == 0
    e = ek_err(s32, s64)
    f = rel_l2(g32.get_fields(), g64.get_fields())
    print(f"512^3 FP32 vs FP64: E(k) rel-L2 {e:.3e}, fields {f:.3e}", flush=True)
    assert e < 1e-5 and f < 1e-4
    for a, b in zip(d32, d64):
        assert abs((a.Esspe + a.Edspe) - (b.Esspe + b.Edspe)) < 1e-5 * (b.Esspe + b.Edspe)
    g32.close()
    g64.close()


def test_2048sq_fp32_against_fp64(hip):
    print("case 2d 2048^2 fp32 vs fp64 device, 20 iterations", flush=True)
    n = 2048
    kw = dict(forcemethod=2, forcek=5, target_energy=0.075, lproject=True, kmax=200)
    g64 = Solver(hip, "bastr", make_params(n, n, **kw))
    g64.init_random_phase(k0=4.0, seed=1, energy=1.0)
    u = g64.get_fields()
    g32 = Solver(hip, "bastr", make_params(n, n, precision=abi.PREC_F32, **kw))
    g32.set_fields(*u)
    r64, d64, s64 = g64.step(20)
    r32, d32, s32 = g32.step(20)
    assert r64 == 0 and r32 == 0
    e = ek_err(s32, s64)
    f = rel_l2(g32.get_fields(), g64.get_fields())
    print(f"2048^2 FP32 vs FP64: E(k) rel-L2 {e:.3e}, fields {f:.3e}", flush=True)
    assert e < 1e-5 and f < 1e-4
    ff = np.array([d.force_factor for d in d32])
    fo = np.array([d.force_factor for d in d64])
    assert np.abs(ff - fo).max() < 1e-5


SIZES = ((256, 48), (512, 32), (1024, 16))


def ensemble_groups(hip, prec, counts=None):
    groups = []
    rng = np.random.default_rng(2024)
    seed = 1
    for n, M in SIZES:
        M = counts.get(n, M) if counts else M
        p = make_params(n, n, forcemethod=2, forcek=5, target_energy=0.075, lproject=True, kmax=128,
                        precision=prec)
        forced = [1 - (m % 2) for m in range(M)]
        e = Ensemble(hip, p, M, forced)
        k0s = rng.integers(3, 9, size=M)
        for m in range(M):
            e.init_random_phase_member(m, k0=float(k0s[m]), seed=seed + m, energy=1.0)
        groups.append(dict(n=n, M=M, ens=e, forced=forced, k0=k0s, seed0=seed, params=p))
        seed += M
    return groups


def test_ensemble96_members_equal_one_field_runs(hip):
    print("case ensemble96 fp32 48x256^2 + 32x512^2 + 16x1024^2, step_group", flush=True)
    groups = ensemble_groups(hip, abi.PREC_F32)
    assert sum(g["M"] for g in groups) == 96
    nsteps = 4
    rc, out = step_group([g["ens"] for g in groups], nsteps)
    assert rc == 0
    for g, (de, se) in zip(groups, out):
        for m in (0, 1, g["M"] - 1):  # forced, free decay, the last member
            s = Solver(hip, "bastr", member_params(g["params"], g["forced"][m]))
            s.init_random_phase(k0=float(g["k0"][m]), seed=g["seed0"] + m, energy=1.0)
            rs, ds, ss = s.step(nsteps)
            assert rs == 0
            same_diags(de[m], ds)
            assert np.array_equal(se[m], ss), (g["n"], m)
            a = g["ens"].get_member_fields(m)
            b = s.get_fields()
            assert all(np.array_equal(x, y) for x, y in zip(a, b)), (g["n"], m)
            s.close()
        # forced members hold the band energy, free-decay ones report no forcing
        assert all(np.isfinite(de[0][i].force_factor) for i in range(nsteps))
    for g in groups:
        g["ens"].close()


def test_ensemble_fp64_members_match_oracle(hip, orc):
    print("case ensemble fp64 2 members per size vs oracle", flush=True)
    orc.set_threads(16)
    groups = ensemble_groups(hip, abi.PREC_F64, counts={256: 2, 512: 2, 1024: 2})
    nsteps = 3
    rc, out = step_group([g["ens"] for g in groups], nsteps)
    assert rc == 0
    for g, (de, se) in zip(groups, out):
        for m in range(g["M"]):
            u0 = Solver(hip, "bastr", member_params(g["params"], g["forced"][m]))
            u0.init_random_phase(k0=float(g["k0"][m]), seed=g["seed0"] + m, energy=1.0)
            o = orc.solver(member_params(g["params"], g["forced"][m]))
            o.set_fields(*u0.get_fields())
            u0.close()
            ro, do, so = o.step(nsteps)
            assert ro == 0
            err = rel_l2(g["ens"].get_member_fields(m), o.get_fields())
            print(f"  {g['n']}^2 member {m} field rel-L2 {err:.3e}", flush=True)
            assert err < 1e-11
            check_diags(de[m], do)
            check_spectra(se[m], so)
    for g in groups:
        g["ens"].close()
