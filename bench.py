#!/usr/bin/env python3
"""Benchmark of the Bastr pseudo-spectral time step on one MI355X.

Workload (default, --config 2d-2048): 2-D viscous Burgers on a 2048^2 periodic
grid, FP64, Re = 5000, dt = 1e-4, Kutta RK3, band forcing at forcek = 5 with
target energy 0.075, dilatational projection on, spectra to kmax = 200;
initial field: Bastrpp's hitgen random-phase construction with the spectrum
law E(k) = k^4 exp(-2 (k/4)^2) at every k (k0 = 4), seed 1, rescaled to
u_rms = 1, built on the device.  The default run also times the rest of the
metric ("legs"): the FP32 variant of the 2048^2 workload and the 3-D 512^3
workload (band forcing, projection, kmax = 256) in FP64 and FP32, each with
its own dominant-kernel roofline, and reports FP32 over FP64 throughput and
the FP32 E(k) error against the FP64 device run of the same iterations.
--config 3d-512 is the 3-D 512^3 case alone, 2d-2048-fp32 / 3d-512-fp32 the
FP32 variants, 2d-256-decay the 256^2 free-decay case, 2d-ensemble96 the
96-member ensemble of 2-D runs.
A "step" is one iteration of the reference main loop (solution.F90:54-97,
124-169): spectra + velocity-gradient diagnostics, RK3 step, forcing.  The
field stays resident in HBM for the whole timed region; bastr_step is called
in chunks of 200 iterations (the spectra cadence of the configuration) and
returns the per-iteration diagnostics to the host, as a log-writing driver
would.

Prints one JSON line.  `value` = grid points x steps / wall seconds of the
timed run, whose step calls run as a caller's do: replaying their captured HIP
graphs on grids of at most 2^21 points, with direct launches above (the
library's default, faster there).  `roofline` is the dominant kernel class: its compulsory HBM bytes per
launch over its mean duration, timed with HIP events on the library's stream
around each of its launches in a run of as many iterations from the same
field (events make those step calls launch directly; `direct_launch` is that
run's throughput).  `step_roofline` prices the whole step at the reference
algorithm's transform count (one compulsory real round trip, 2*s*N^d bytes,
per whole-field transform).  `cpu_baseline` is the FP64 CPU restatement
(oracle/) timed on this host's cores over a bounded sample of the same
workload.  `accuracy` gives the E(k) and field error of the FP64 device path
against that oracle sample and, for FP32 configurations, of the FP32 path
against the FP64 device path.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LAW_PURE = 1  # bastr_set_spectrum_law: E(k) = k^4 exp(-2 (k/k0)^2) at every k

METRIC = "grid-pt·steps/s at 2048² & 512³ (FP64 vs FP32), % HBM roofline, E(k) err"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # grid, precision, forcemethod, lproject, kmax, default timed steps, oracle sample steps
    "2d-256-decay": dict(grid=(256, 256), prec=0, force=0, proj=False, kmax=128, steps=2000, cpu=2000),
    "2d-2048": dict(grid=(2048, 2048), prec=0, force=2, proj=True, kmax=200, steps=1000, cpu=20),
    "2d-2048-fp32": dict(grid=(2048, 2048), prec=1, force=2, proj=True, kmax=200, steps=1000, cpu=20),
    "3d-512": dict(grid=(512, 512, 512), prec=0, force=2, proj=True, kmax=256, steps=100, cpu=1),
    "3d-512-fp32": dict(grid=(512, 512, 512), prec=1, force=2, proj=True, kmax=256, steps=100, cpu=1),
}


def prof_api(lib):
    lib.bastr_prof_enable.argtypes = [C.c_void_p, C.c_uint32]
    lib.bastr_prof_enable.restype = C.c_int
    lib.bastr_prof_read.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
    lib.bastr_prof_read.restype = C.c_int
    lib.bastr_prof_name.argtypes = [C.c_int]
    lib.bastr_prof_name.restype = C.c_char_p
    lib.bastr_prof_copy_gbps.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(C.c_double)]
    lib.bastr_prof_copy_gbps.restype = C.c_int
    lib.bastr_prof_graph.argtypes = [C.c_void_p]
    lib.bastr_prof_graph.restype = C.c_int


def cpu_model():
    """Host CPU model name and logical CPU count, as /proc/cpuinfo states them."""
    name, count = None, 0
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                count += 1
                name = name or line.split(":", 1)[1].strip()
    except OSError:
        pass
    return name, count or os.cpu_count()


NCLASS = 12  # BASTR_K_NCLASS (include/bastr_prof.h)


def read_prof(lib, h, ncls=NCLASS):
    out = {}
    for c in range(ncls):
        n = C.c_int64()
        ms = C.c_double()
        by = C.c_double()
        lib.bastr_prof_read(h, c, C.byref(n), C.byref(ms), C.byref(by))
        out[lib.bastr_prof_name(c).decode()] = (c, n.value, ms.value, by.value)
    return out


def traffic_from_profile(kclass, prec, grid):
    """HBM bytes of one invocation of the kernel class, from the committed
    counter summary of this workload (profiles/hbm_traffic_<tag>.txt, the
    `class:` lines: FETCH_SIZE + WRITE_SIZE calibrated against a device copy),
    or None if the workload was not profiled."""
    if grid == (2048, 2048):
        tag = "2d_2048"
    elif grid == (512, 512, 512):
        tag = "3d_512"
    else:
        return None
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{tag}_{'fp32' if prec == 1 else 'fp64'}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if parts and parts[0] == f"class:{kclass}":
            return float(parts[3]) * 1e6  # total_MB
    return None


def run_steps(s, n, chunk):
    """n iterations in calls of at most `chunk`; returns the last iteration's
    (kn, Es, Ed)."""
    done = 0
    last = None
    while done < n:
        k = min(chunk, n - done)
        rc, _, sp = s.step(k, want_spectra=True)
        if rc != 0:
            raise RuntimeError(f"bastr_step returned {rc}")
        last = sp[-1]
        done += k
    return last


def cpu_baseline(params, fields, sample_steps):
    """The oracle on this host's cores, timed over `sample_steps` iterations of
    the same workload (its entry dealias runs first, untimed).  Returns the
    rate, thread count, seconds, and the oracle's last-iteration spectrum and
    final fields for the accuracy check."""
    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = min(threads, 16)
    oracle.set_threads(threads)
    o = oracle.solver(params)
    o.set_fields(*fields)
    o.get_fields()  # runs the entry dealias
    t0 = time.perf_counter()
    _, _, spec = o.step(sample_steps)
    dt = time.perf_counter() - t0
    npts = params.ia * params.ja * max(params.ka, 1)
    out = o.get_fields()
    o.close()
    return npts * sample_steps / dt, threads, dt, spec[-1], out


def ek_err(test, ref):
    """E(k) = Es(k) + Ed(k) over shells 1..allkmax, from one iteration's
    (kn, Es, Ed): ||E_test - E_ref||_2 / ||E_ref||_2 over the shells, and the
    largest per-shell relative error over shells holding at least 1e-12 of the
    peak."""
    import numpy as np
    a, b = test[1] + test[2], ref[1] + ref[2]
    keep = np.abs(b) >= 1e-12 * np.abs(b).max()
    return (float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b * b))),
            float((np.abs(a - b)[keep] / np.abs(b)[keep]).max()))


def reference_transforms(nd, timemethod, forcemethod):
    """Whole-field transforms the reference performs per loop iteration
    (solution.F90): top-of-loop FFT, velgrad, the RHS evaluations, the stage
    FFTs, and forcing's transform pair."""
    top, vg, rhs, stage_ffts, force = (2, 7, 12, 4, 4) if nd == 2 else (3, 13, 21, 6, 6)
    if timemethod == 2:
        evals, stage_ffts = 6, stage_ffts // 2 * 5
    else:
        evals = 3
    return top + vg + evals * rhs + stage_ffts + (force if forcemethod in (2, 5) else 0)


def rel_l2(a, b):
    import numpy as np
    num = sum(float(np.sum((x - y) ** 2)) for x, y in zip(a, b))
    return float(np.sqrt(num / sum(float(np.sum(y * y)) for y in b)))


# --config 2d-ensemble96: the batched ensemble, one handle per member size, every member
# of a size advanced by the same launches (bastr_create_ensemble)
ENSEMBLE = [(256, 48), (512, 32), (1024, 16)]


def bench_ensemble(args):
    """96 independent 2-D runs (48 x 256^2, 32 x 512^2, 16 x 1024^2), FP32 with
    FP64 energy/spectrum reductions, band forcing (k=5, E_t=0.075) on every
    other member and free decay on the rest, projection on, each member its own
    seed with k0 drawn uniformly from {3..8}; dt = 1e-4."""
    import numpy as np

    import bastr_amd
    from bastr_amd.solver import make_params

    lib = bastr_amd.load()
    prof_api(lib)
    # with the pure law, the first member (1024^2, k0 = 7, free decay) reaches a
    # NaN energy at iteration 929: the timed iterations stay below that, so that
    # every member does every iteration's work
    steps = args.steps if args.steps is not None else 600

    def build_groups():
        rng = np.random.default_rng(2024)
        gs = []
        seed = 1
        for n, M in ENSEMBLE:
            p = make_params(n, n, reynolds=5000.0, deltat=1e-4, timemethod=1, forcemethod=2, forcek=5,
                            target_energy=0.075, lproject=True, kmax=n // 2, precision=1)
            forced = [m % 2 for m in range(M)]
            e = bastr_amd.ensemble(p, M, forced)
            e.set_spectrum_law(LAW_PURE)
            k0s = rng.integers(3, 9, size=M)
            for m in range(M):
                e.init_random_phase_member(m, k0=float(k0s[m]), seed=seed + m, energy=1.0)
            gs.append(dict(n=n, M=M, ens=e, forced=forced, k0=k0s, seed0=seed, params=p))
            seed += M
        return gs

    groups = build_groups()
    total_pts = sum(n * n * M for n, M in ENSEMBLE)
    # column plans of each size, timed on its own ensemble handle's step
    # (bastr_autotune_step; the state is restored, the timed run uses fresh handles)
    tune = {}
    if not args.no_tune:
        for g in groups:
            (pts, wid), ms = bastr_amd.autotune_step(g["ens"], 10)
            tune[f"{g['n']} fp32"] = {
                "points_per_thread": pts, "width": wid,
                "ms_per_iteration": {f"{p}x{w}": (None if m != m else round(m, 4))
                                     for (p, w), m in zip([(16, 1), (8, 1), (16, 2), (8, 2)], ms)},
                "timed_on": f"the {g['M']} x {g['n']}^2 handle alone: 10 iterations per plan and round, two rounds"}
        print(f"  autotune: {tune}", file=sys.stderr, flush=True)

    # each size's handle has its own HIP stream: bastr_step_group launches the
    # three sizes' iterations before waiting for any, so their kernels overlap
    # on the GPU; --serial steps them one after another with bastr_step
    from bastr_amd.solver import step_group as step_all

    def step_group(g, k):
        rc, _, sp = g["ens"].step(k, want_spectra=True)
        if rc != 0:
            raise RuntimeError(f"bastr_step returned {rc} ({g['n']}^2 ensemble)")
        return sp

    def advance(n_it, keep=False, groups=groups):
        out = []
        done = 0
        while done < n_it:
            k = min(args.chunk, n_it - done)
            if args.serial:
                sps = [step_group(g, k) for g in groups]
            else:
                rc, res = step_all([g["ens"] for g in groups], k)
                if rc != 0:
                    raise RuntimeError(f"bastr_step_group returned {rc}")
                sps = [sp for _, sp in res]
            if keep:
                out.extend((g["n"], done, sp) for g, sp in zip(groups, sps))
            done += k
        return out

    # the timed run, as a caller gets it, first: its own members, no events, so
    # bastr_step_group replays all three sizes' iterations as one captured
    # graph per call (--serial: each handle's graph)
    ggroups = build_groups()
    gsteps = min(steps, args.chunk)
    advance(gsteps, groups=ggroups)  # captures the graphs of the call length
    tg = time.perf_counter()
    advance(steps, groups=ggroups)
    wall = time.perf_counter() - tg
    for g in ggroups:
        g["ens"].close()

    # warm-up with every kernel class timed; its first iterations give the
    # FP32 spectra for the accuracy check
    for g in groups:
        lib.bastr_prof_enable(g["ens"]._h, (1 << NCLASS) - 1)
    nchk = max(args.warmup, 1)
    warm_sp = advance(nchk, keep=True)
    best = None
    for g in groups:
        w = read_prof(lib, g["ens"]._h)
        for name, (c, n_l, ms, by) in w.items():
            if n_l and (best is None or ms > best[3]):
                best = (g, name, c, ms)
    dg, dom_name, dom_cls, _ = best
    # the profiled run: HIP events around the dominant class's launches (its
    # handle then launches directly), as many iterations as the timed run
    for g in groups:
        lib.bastr_prof_enable(g["ens"]._h, (1 << dom_cls) if g is dg else 0)
    t0 = time.perf_counter()
    advance(steps)
    wall_direct = time.perf_counter() - t0
    _, n_l, ms_l, by_l = read_prof(lib, dg["ens"]._h)[dom_name]
    mean_s = ms_l / n_l * 1e-3
    achieved = by_l / mean_s / 1e9
    print(f"  timed {steps} ensemble steps in {wall:.2f} s (graph replay); profiled run {wall_direct:.2f} s",
          file=sys.stderr, flush=True)
    copy_gbps = C.c_double()
    if lib.bastr_prof_copy_gbps(groups[0]["ens"]._h, 1 << 30, 10, C.byref(copy_gbps)) != 0:
        copy_gbps.value = float("nan")

    # accuracy: member 0 (forced) and member 1 (free decay) of each size, FP32
    # ensemble spectra at the last warm-up iteration against one-field FP64 runs
    accuracy = {}
    for g in groups:
        for m in (0, 1):
            fp = make_params(g["n"], g["n"], reynolds=5000.0, deltat=1e-4, timemethod=1,
                             forcemethod=2 if g["forced"][m] else 0, forcek=5, target_energy=0.075,
                             lproject=True, kmax=g["n"] // 2, precision=0)
            ref = bastr_amd.solver(fp)
            ref.set_spectrum_law(LAW_PURE)
            ref.init_random_phase(k0=float(g["k0"][m]), seed=g["seed0"] + m, energy=1.0)
            _, _, rsp = ref.step(nchk)
            ref.close()
            sp32 = [sp for (n, d0, sp) in warm_sp if n == g["n"] and d0 + sp.shape[1] == nchk][0]
            e_abs, e_shell = ek_err(sp32[m, -1], rsp[-1])
            accuracy[f"{g['n']}^2 member {m} ({'forced' if g['forced'][m] else 'free decay'})"] = {
                "iterations": nchk, "ek_rel_l2_fp32_vs_fp64": e_abs, "ek_err_max_per_shell": e_shell}
    for g in groups:
        g["ens"].close()

    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        threads = min(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1), 16)
        oracle.set_threads(threads)
        t_iter, parts = 0.0, []
        for n, M in ENSEMBLE:
            its = {256: 1200, 512: 300, 1024: 75}[n]  # about 8 s of CPU work each
            fp = make_params(n, n, reynolds=5000.0, deltat=1e-4, timemethod=1, forcemethod=2, forcek=5,
                             target_energy=0.075, lproject=True, kmax=n // 2, precision=0)
            g0 = new_field(bastr_amd, fp)
            u = g0.get_fields()
            g0.close()
            o = oracle.solver(fp)
            o.set_fields(*u)
            o.get_fields()
            t1 = time.perf_counter()
            o.step(its)
            dt = time.perf_counter() - t1
            o.close()
            t_iter += M * dt / its
            parts.append(f"{its} iterations of one {n}^2 member ({dt:.1f} s)")
        model, ncpu = cpu_model()
        cpu = {"value": total_pts / t_iter, "unit": "grid-pt·steps/s", "cores": threads, "kind": "port",
               "cpu_model": model, "host_logical_cpus": ncpu,
               "sample": "; ".join(parts) + ", FP64 C restatement, OpenMP; ensemble rate = all points / "
                         "sum over sizes of members x seconds per member iteration"}
    out = {
        "metric": METRIC,
        "value": total_pts * steps / wall,
        "unit": "grid-pt·steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": nchk,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded random-phase fields, E(k) = k^4 exp(-2 (k/k0)^2) (pure law), k0 uniform in "
                "{3..8}, one seed per member, u_rms=1, built on the device",
        "autotune": tune or None,
        "config": {"workload": "ensemble of 96 independent 2-D Burgers runs: 48 x 256^2, 32 x 512^2, 16 x 1024^2 "
                               "(28,311,552 points), Re=5000, dt=1e-4, RK3, band forcing k=5 E_t=0.075 on every "
                               "other member, free decay on the rest, lproject=t, spectra kmax=n/2",
                   "members": [list(x) for x in ENSEMBLE], "precision": "fp32",
                   "launch": "one handle per size (every member of a size in each launch), the three on their "
                             "own streams, " + ("stepped one after another, each call replaying its handle's "
                                               "graph" if args.serial else
                                               "stepped together by bastr_step_group: one captured graph over the "
                                               "three sizes per call")},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom_name,
                     "ensemble_size": dg["n"], "bytes_per_launch": by_l, "mean_launch_us": round(mean_s * 1e6, 2),
                     "copy_measured": round(copy_gbps.value, 1),
                     "frac_of_copy": round(achieved / copy_gbps.value, 4)},
        "cpu_baseline": cpu,
        "accuracy": accuracy,
        "direct_launch": {"ms_per_step": wall_direct / steps * 1e3, "value": total_pts * steps / wall_direct,
                          "launch": "the profiled run: the dominant size's handle launches directly"},
    }
    print(json.dumps(out))


# --config 2d-closure: the learned subgrid closure (include/bastr_closure.h)
CLOSURE_FWD_FLOPS = 13056.0  # 2 (36*64 + 64*64 + 64*2) per point
MFMA_BF16_PEAK = 2500.0      # dense BF16 MFMA, TFLOP/s


def closure_kernels(lib, args):
    """--closure-kernels: the network's kernels alone at the shapes of the
    closure workload (training batches of 2048 tiles of 32 points, inference
    over a 256^2 grid), on pairs filtered from one short 512^2 run (their
    values do not change the kernels' work): bastr_prof_closure's event
    timings against the BF16 MFMA peak, then 150 closure-run iterations."""
    import bastr_amd
    from bastr_amd.closure import Closure, filter_to, fine_run, make_pairs
    from bastr_amd.solver import make_params

    n = 256
    fp = make_params(512, 512, reynolds=5000.0, deltat=1e-4, forcemethod=2, forcek=5, target_energy=0.075,
                     lproject=True, kmax=128)
    cp = make_params(n, n, reynolds=5000.0, deltat=4e-4, forcemethod=2, forcek=5, target_energy=0.075,
                     lproject=True, kmax=n // 3)
    ps = make_pairs(fp, n, [1], [200, 300], k0=4.0)
    net = Closure(seed=1)
    net.fit_scales(ps)
    net.train(ps, iters=200, batch=256, lr=2e-3, seed=3)
    tr_ms, fw_ms = C.c_double(), C.c_double()
    if lib.bastr_prof_closure(net._h, ps._h, 20, 2048, C.byref(tr_ms), C.byref(fw_ms)) != 0:
        raise RuntimeError("bastr_prof_closure failed")
    f = fine_run(fp, 2, 4.0)
    f.step(200, want_spectra=False)
    c = bastr_amd.solver(cp)
    filter_to(f, c)
    net.attach(c)
    rc = c.step(150, want_spectra=False)[0]
    c.close()
    f.close()
    train_pts = 2048 * 32
    train_tf = 3 * CLOSURE_FWD_FLOPS * train_pts / (tr_ms.value * 1e-3) / 1e12
    fwd_tf = CLOSURE_FWD_FLOPS * n * n / (fw_ms.value * 1e-3) / 1e12
    print(json.dumps({"closure_kernels": {
        "train_kernel": {"points": train_pts, "mean_launch_us": round(tr_ms.value * 1e3, 2),
                         "tflops": round(train_tf, 2), "frac_of_bf16_peak": round(train_tf / MFMA_BF16_PEAK, 5)},
        "forward_kernel": {"points": n * n, "mean_launch_us": round(fw_ms.value * 1e3, 2),
                           "tflops": round(fwd_tf, 2), "frac_of_bf16_peak": round(fwd_tf / MFMA_BF16_PEAK, 5)},
        "closure_run_status": rc}}))
def closure_prof_api(lib):
    lib.bastr_prof_closure.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
    lib.bastr_prof_closure.restype = C.c_int


def bench_closure(args):
    """Coarse 2-D 256^2 solve (dt = 4e-4) with the per-point MLP closure,
    trained on pairs from 2048^2 fine runs of this solver (FP64, Re = 5000,
    dt = 1e-4, band forcing k = 5 E_t = 0.075, projection; pure-law fields k0 = 4)
    at 20 seeds (snapshots after 800, 950, 1100, 1250, 1400 iterations,
    filtered at k_c = 85; Adam at learning rates 2e-3, 7e-4, 2e-4 for half,
    a quarter and a quarter of the steps), judged on 5 held-out seeds: the
    coarse run starts from the filtered fine field after 800 fine iterations
    and runs 150 iterations (600 fine ones); E(k) for k < k_c against the
    filtered fine run, with and without the network: rel-L2 over the shells,
    and the mean and largest per-shell relative error, with the force evaluated at
    every RK3 stage and once per time step.  value = the closure run's
    grid-pt.steps/s with the once-per-step force (the per-stage rate beside it);
    the fine run's rate beside them, and both in simulated time per second."""
    import numpy as np

    import bastr_amd
    from bastr_amd.closure import Closure, evaluate, filter_to, fine_run, make_pairs
    from bastr_amd.solver import make_params

    lib = bastr_amd.load()
    prof_api(lib)
    closure_prof_api(lib)
    if args.closure_kernels:
        return closure_kernels(lib, args)
    N, n = 2048, 256
    fp = make_params(N, N, reynolds=5000.0, deltat=1e-4, forcemethod=2, forcek=5, target_energy=0.075,
                     lproject=True, kmax=128)
    cp = make_params(n, n, reynolds=5000.0, deltat=4e-4, forcemethod=2, forcek=5, target_energy=0.075,
                     lproject=True, kmax=n // 3)
    train_seeds, test_seeds = list(range(1, 21)), list(range(21, 26))
    # snapshots well before the blow-up of these shocked runs (the first seeds'
    # fields reach NaN energies between 1450 and 1900 iterations)
    snaps = [800, 950, 1100, 1250, 1400]
    t0 = time.perf_counter()
    ps = make_pairs(fp, n, train_seeds, snaps, k0=4.0)
    t_pairs = time.perf_counter() - t0
    print(f"  [2d-closure] {len(ps)} training pairs in {t_pairs:.1f} s", file=sys.stderr, flush=True)
    net = Closure(seed=1)
    net.fit_scales(ps)
    iters = args.steps if args.steps is not None else 8000
    t1 = time.perf_counter()
    losses = np.concatenate([net.train(ps, iters=iters // 2, batch=256, lr=2e-3, seed=3),
                             net.train(ps, iters=iters // 4, batch=256, lr=7e-4, seed=4),
                             net.train(ps, iters=iters - iters // 2 - iters // 4, batch=256, lr=2e-4, seed=5)])
    t_train = time.perf_counter() - t1
    print(f"  [2d-closure] {iters} training steps in {t_train:.2f} s, batch loss {losses[0]:.4f} -> "
          f"{float(np.mean(losses[-50:])):.4f}", file=sys.stderr, flush=True)
    # E(k), k < k_c, on the held-out seeds
    res = {"without": [], "with": [], "with_per_step": []}
    for seed in test_seeds:
        for label, nn, per in (("without", None, False), ("with", net, False), ("with_per_step", net, True)):
            Ef, Ec = evaluate(fp, cp, seed=seed, t0_steps=800, fine_steps=1400, net=nn, k0=4.0, per_step=per)
            rel = np.abs(Ec - Ef) / Ef
            res[label].append({"seed": seed, "ek_rel_l2": float(np.linalg.norm(Ec - Ef) / np.linalg.norm(Ef)),
                               "ek_err_mean_per_shell": float(rel.mean()),
                               "ek_err_max_per_shell": float(rel.max())})
    summ = {k: {q: float(np.mean([r[q] for r in v]))
                for q in ("ek_rel_l2", "ek_err_mean_per_shell", "ek_err_max_per_shell")}
            for k, v in res.items()}
    print(f"  [2d-closure] held-out E(k): {summ}", file=sys.stderr, flush=True)
    # throughput: the closure run (graph replay) in both cadences and the fine run
    f = fine_run(fp, 21, 4.0)
    f.step(800, want_spectra=False)
    steps = 150
    # column plans timed on both runs' own steps (bastr_autotune_step; states restored)
    tune = {}
    if not args.no_tune:
        c = bastr_amd.solver(cp)
        filter_to(f, c)
        net.attach(c)
        Closure.per_step(c, True)
        for key, h in (("256 fp64 (closure run)", c), ("2048 fp64 (fine run)", f)):
            (pts, wid), ms = bastr_amd.autotune_step(h, 20)
            tune[key] = {"points_per_thread": pts, "width": wid,
                         "ms_per_iteration": {f"{p}x{w}": (None if m != m else round(m, 4))
                                              for (p, w), m in zip([(16, 1), (8, 1), (16, 2), (8, 2)], ms)}}
        c.close()
        print(f"  autotune: {tune}", file=sys.stderr, flush=True)
    t_cad, rc_cad = {}, {}
    for per in (True, False):
        c = bastr_amd.solver(cp)
        filter_to(f, c)
        net.attach(c)
        Closure.per_step(c, per)
        c.step(steps, want_spectra=False)  # captures the graph of the call length
        filter_to(f, c)
        tc = time.perf_counter()
        rc_cad[per] = c.step(steps, want_spectra=True)[0]
        t_cad[per] = (time.perf_counter() - tc) / steps
        c.close()
    t_coarse, rc_c = t_cad[True], rc_cad[True]
    f.step(steps, want_spectra=False)
    tf = time.perf_counter()
    rc_f = f.step(steps, want_spectra=True)[0]
    t_fine = (time.perf_counter() - tf) / steps
    f.close()
    # the network's kernels against the BF16 MFMA peak
    tr_ms, fw_ms = C.c_double(), C.c_double()
    if lib.bastr_prof_closure(net._h, ps._h, 20, 2048, C.byref(tr_ms), C.byref(fw_ms)) != 0:
        raise RuntimeError("bastr_prof_closure failed")
    fwd_flops, train_pts = CLOSURE_FWD_FLOPS, 2048 * 32
    train_tf = 3 * fwd_flops * train_pts / (tr_ms.value * 1e-3) / 1e12
    fwd_tf = fwd_flops * n * n / (fw_ms.value * 1e-3) / 1e12
    peak = MFMA_BF16_PEAK
    cpu = None
    if not args.no_cpu:
        # the oracle's coarse run with the same network in compute_ut
        from oracle import oracle
        threads = min(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1), 16)
        oracle.set_threads(threads)
        g0 = fine_run(fp, 21, 4.0)
        g0.step(800, want_spectra=False)
        cc = bastr_amd.solver(cp)
        filter_to(g0, cc)
        u = cc.get_fields()
        g0.close()
        cc.close()
        o = oracle.solver(cp)
        params, scales = net.params()
        oracle.set_closure(o, params, scales)
        oracle.set_closure_per_step(o, True)
        o.set_fields(*u)
        o.get_fields()
        its = 150
        tc = time.perf_counter()
        o.step(its)
        dt_cpu = time.perf_counter() - tc
        o.close()
        model, ncpu = cpu_model()
        cpu = {"value": n * n * its / dt_cpu, "unit": "grid-pt·steps/s", "cores": threads, "kind": "port",
               "cpu_model": model, "host_logical_cpus": ncpu,
               "sample": f"{its} iterations of the 256^2 closure run ({dt_cpu:.1f} s), FP64 C restatement with "
                         f"the network's forward pass in compute_ut once per time step, OpenMP"}
    out = {
        "metric": METRIC,
        "value": n * n / t_coarse,
        "unit": "grid-pt·steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": steps,
        "ms_per_step": t_coarse * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: pure-law random-phase fields E(k) = k^4 exp(-2 (k/4)^2), u_rms=1; training seeds 1-20, "
                "held-out seeds 21-25",
        "autotune": tune or None,
        "config": {"workload": "learned closure: coarse 2-D 256^2 Burgers, dt=4e-4, Re=5000, band forcing k=5 "
                               "E_t=0.075, lproject=t, with a 36-64-64-2 GELU MLP (BF16 MFMA) trained on filtered "
                               "(k_c=85) 2048^2 runs of this solver",
                   "grid": [n, n], "fine_grid": [N, N], "precision": "fp64 solve, bf16 network"},
        "roofline": {"bound": "mfma", "achieved": round(train_tf, 2), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(train_tf / peak, 5), "traffic": None, "kernel": "mlp::train_kernel",
                     "flops_per_launch": 3 * fwd_flops * train_pts, "mean_launch_us": round(tr_ms.value * 1e3, 2),
                     "inference": {"kernel": "mlp::forward_kernel", "achieved": round(fwd_tf, 2),
                                   "frac": round(fwd_tf / peak, 5), "flops_per_launch": fwd_flops * n * n,
                                   "mean_launch_us": round(fw_ms.value * 1e3, 2)}},
        "cpu_baseline": cpu,
        "closure": {
            "pairs": len(ps), "pairs_seconds": round(t_pairs, 2), "train_steps": iters,
            "train_seconds": round(t_train, 3), "batch_loss_first": float(losses[0]),
            "batch_loss_last50": float(np.mean(losses[-50:])),
            "held_out": summ, "held_out_per_seed": res,
            "cadence": "value, coarse_ms_per_step and simulated_time_ratio: the force evaluated once per "
                       "time step (bastr_closure_per_step); per_stage: evaluated at each RK3 stage",
            "coarse_ms_per_step": t_coarse * 1e3, "fine_ms_per_step": t_fine * 1e3,
            "fine_grid_pt_steps_per_s": N * N / t_fine,
            "simulated_time_per_s": {"closure_run": cp.deltat / t_coarse, "fine_run": fp.deltat / t_fine},
            "simulated_time_ratio": (cp.deltat / t_coarse) / (fp.deltat / t_fine),
            "per_stage": {"coarse_ms_per_step": t_cad[False] * 1e3,
                          "simulated_time_ratio": (cp.deltat / t_cad[False]) / (fp.deltat / t_fine),
                          "status": rc_cad[False]},
            "status": {"coarse": rc_c, "fine": rc_f},
        },
    }
    print(json.dumps(out))


def cfg_params(name, precision=None):
    from bastr_amd.solver import make_params
    cfg = CONFIGS[name]
    grid = cfg["grid"]
    return make_params(grid[0], grid[1], grid[2] if len(grid) == 3 else 0, reynolds=5000.0, deltat=1e-4,
                       timemethod=1, forcemethod=cfg["force"], forcek=5, target_energy=0.075,
                       lproject=cfg["proj"], kmax=cfg["kmax"],
                       precision=cfg["prec"] if precision is None else precision)


def new_field(bastr_amd, params):
    """A handle holding the workload's initial field: the pure-law random-phase
    field (k0 = 4, seed 1, u_rms = 1) built on the device."""
    from bastr_amd import abi
    s = bastr_amd.solver(params)
    s.set_spectrum_law(abi.LAW_PURE)
    s.init_random_phase(k0=4.0, seed=1, energy=1.0)
    return s


def measure(lib, bastr_amd, name, steps, warmup, chunk, copy=False):
    """Time `steps` iterations of workload `name` after `warmup` untimed ones,
    as a caller runs them: each bastr_step call (a chunk of iterations) replays
    its captured HIP graph on a small grid, or launches directly on a large one.
    Profiling events force direct launches, so
    the dominant kernel class (the largest share of a warm-up with every class
    timed) is timed in a second run of as many iterations from the same initial
    field: HIP events on the library's stream around each of its launches.
    That run's wall time is reported too (direct launches).  Returns the record
    and the last timed iteration's spectrum."""
    cfg = CONFIGS[name]
    grid, prec = cfg["grid"], cfg["prec"]
    # the timed run: unprofiled, so every call runs as a caller's does (where the
    # library replays graphs, the warm-up's last call captures the graph of the
    # chunk length the timed calls replay)
    g = new_field(bastr_amd, cfg_params(name))
    replay = lib.bastr_prof_graph(g._h) == 1
    run_steps(g, max(warmup, min(chunk, steps)), min(chunk, steps))
    t1 = time.perf_counter()
    last = run_steps(g, steps, chunk)
    wall = time.perf_counter() - t1
    g.close()
    # the profiled run (direct launches) from the same initial field (a fresh
    # field: the workload's field blows up after about 1860 iterations at 2048^2)
    s = new_field(bastr_amd, cfg_params(name))
    lib.bastr_prof_enable(s._h, (1 << NCLASS) - 1)
    run_steps(s, max(warmup, 1), chunk)
    warm = read_prof(lib, s._h)
    dom_name, (dom_cls, _, _, _) = max(warm.items(), key=lambda kv: kv[1][2])
    for cname, (c, n, ms, by) in sorted(warm.items(), key=lambda kv: -kv[1][2]):
        if n:
            print(f"  [{name}] warmup {cname:12s} launches={n:6d} mean={ms / n * 1e3:9.2f} us "
                  f"bytes/launch={by / 1e6:8.2f} MB  {by / (ms / n * 1e-3) / 1e9:8.1f} GB/s",
                  file=sys.stderr, flush=True)
    lib.bastr_prof_enable(s._h, 1 << dom_cls)
    t0 = time.perf_counter()
    run_steps(s, steps, chunk)
    wall_direct = time.perf_counter() - t0
    prof = read_prof(lib, s._h)
    _, n_l, ms_l, by_l = prof[dom_name]
    mean_s = ms_l / n_l * 1e-3
    achieved = by_l / mean_s / 1e9
    copy_gbps = float("nan")
    if copy:  # measured device copy bandwidth (1 GiB, read + write), beside the nominal peak
        cg = C.c_double()
        if lib.bastr_prof_copy_gbps(s._h, 1 << 30, 10, C.byref(cg)) == 0:
            copy_gbps = cg.value
    s.close()
    print(f"  [{name}] timed {steps} steps in {wall:.3f} s ({'graph replay' if replay else 'direct launches'}); "
          f"profiled run {wall_direct:.3f} s",
          file=sys.stderr, flush=True)
    npts = 1
    for n in grid:
        npts *= n
    T = reference_transforms(len(grid), 1, cfg["force"])
    b_step = T * 2 * (4 if prec == 1 else 8) * npts
    rec = {
        "value": npts * steps / wall, "unit": "grid-pt·steps/s", "steps": steps, "warmup": warmup,
        "ms_per_step": wall / steps * 1e3, "dtype": "f32" if prec == 1 else "f64",
        "launch": ("graph replay (each bastr_step call of up to 200 iterations replays its captured HIP graph)"
                   if replay else "direct launches (the library's default above 2^21 grid points: "
                   "faster there than replaying a captured graph)"),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profile(dom_name, prec, grid), "kernel": dom_name,
                     "bytes_per_launch": by_l, "mean_launch_us": round(mean_s * 1e6, 2),
                     "timing": f"HIP events around each of the class's {n_l} launches in a {steps}-iteration "
                               "run of the same workload from the same field (direct launches)"},
        "step_roofline": {"reference_transforms": T, "bytes_per_step": b_step,
                          "achieved": round(b_step * steps / wall / 1e9, 1), "unit": "GB/s",
                          "frac": round(b_step * steps / wall / 1e9 / HBM_PEAK_GBS, 4)},
        "direct_launch": {"ms_per_step": wall_direct / steps * 1e3, "value": npts * steps / wall_direct},
    }
    # every class of the warm-up (all classes timed with events): per launch
    # and per iteration, against its compulsory bytes
    nw = max(warmup, 1)
    rec["classes"] = {
        cname: {"us_per_launch": round(ms / n * 1e3, 2), "launches_per_iter": round(n / nw, 3),
                "us_per_iter": round(ms / nw * 1e3, 2), "bytes_per_launch": by,
                "frac": round(by / (ms / n * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        for cname, (c, n, ms, by) in sorted(warm.items(), key=lambda kv: -kv[1][2]) if n}
    if copy:
        rec["roofline"]["copy_measured"] = round(copy_gbps, 1)
        rec["roofline"]["frac_of_copy"] = round(achieved / copy_gbps, 4)
        rec["step_roofline"]["frac_of_copy"] = round(b_step * steps / wall / 1e9 / copy_gbps, 4)
    return rec, last


def workload_desc(name):
    cfg = CONFIGS[name]
    grid = cfg["grid"]
    dims = "x".join(map(str, grid))
    return (f"{len(grid)}-D viscous Burgers {dims}, Re=5000, dt=1e-4, RK3, "
            f"{'band forcing k=5 E_t=0.075' if cfg['force'] == 2 else 'free decay'}, "
            f"lproject={'t' if cfg['proj'] else 'f'}, spectra kmax={cfg['kmax']}")


# the other workloads of the metric, timed in the default run beside the
# headline 2048^2 FP64 one: (name, its FP64 counterpart for the accuracy check)
LEGS = [("2d-2048-fp32", "2d-2048"), ("3d-512", None), ("3d-512-fp32", "3d-512")]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", default="2d-2048", choices=sorted(CONFIGS) + ["2d-ensemble96", "2d-closure"])
    ap.add_argument("--chunk", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="2d-2048: time the headline workload only")
    ap.add_argument("--no-tune", action="store_true", help="keep the default radix plans (no autotuning)")
    ap.add_argument("--plans", default=None, metavar="N:PxW,...",
                    help="fix the column plans instead of tuning (e.g. 256:8x1,2048:16x1, applied to both "
                         "precisions): the profiles' runs use the plans the tuned run picked")
    ap.add_argument("--serial", action="store_true", help="2d-ensemble96: step the size groups one after another")
    ap.add_argument("--closure-kernels", action="store_true",
                    help="2d-closure: time only the network's kernels at the workload's shapes")
    ap.add_argument("--copy-only", type=int, default=None, metavar="MIB",
                    help="only run the library's device copy (bastr_prof_copy_gbps) of MIB MiB, 5 times: the "
                         "known-size kernel the HBM counter summaries are calibrated against")
    args = ap.parse_args()
    if args.copy_only is not None:
        import bastr_amd
        from bastr_amd.solver import make_params
        lib = bastr_amd.load()
        prof_api(lib)
        s = bastr_amd.solver(make_params(16, 16))
        g = C.c_double()
        rc = lib.bastr_prof_copy_gbps(s._h, args.copy_only << 20, 5, C.byref(g))
        s.close()
        print(f"copy of {args.copy_only} MiB x 5: {g.value:.1f} GB/s (rc {rc})")
        sys.exit(0 if rc == 0 else 1)
    if args.plans:
        # a fixed plan table, as a tuned run left it: no tuning iterations
        import bastr_amd
        for item in args.plans.split(","):
            n, pw = item.split(":")
            pts, wid = pw.split("x")
            for prec in (0, 1):
                try:
                    bastr_amd.tune_set(int(n), int(pts), prec, int(wid))
                except Exception:
                    pass  # no such plan in this precision: the default stays
        args.no_tune = True
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the whole periodic domain is resident in one "
              "MI355X's HBM); --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if args.config == "2d-ensemble96":
        return bench_ensemble(args)
    if args.config == "2d-closure":
        return bench_closure(args)

    import numpy as np

    import bastr_amd

    name = args.config
    cfg = CONFIGS[name]
    grid, prec = cfg["grid"], cfg["prec"]
    steps = args.steps if args.steps is not None else cfg["steps"]
    cpu_steps = args.cpu_steps if args.cpu_steps is not None else cfg["cpu"]
    dims = "x".join(map(str, grid))
    lib = bastr_amd.load()
    prof_api(lib)
    # radix plans of the column passes picked by the autotuner for every column
    # length this run uses (the table then serves all the timed workloads)
    tune = {}
    if not args.no_tune:
        names = [name] + ([leg for leg, _ in LEGS] if name == "2d-2048" and not args.no_legs else [])
        for nm in names:
            cg = CONFIGS[nm]
            n_col = cg["grid"][1]
            key = f"{n_col} {'fp32' if cg['prec'] == 1 else 'fp64'}"
            if key not in tune:
                # timed on the workload's own step (bastr_autotune_step): a handle
                # holding the workload's field, a few iterations in, every plan run
                # from that state (restored after each run)
                t_h = new_field(bastr_amd, cfg_params(nm))
                nd3 = len(cg["grid"]) == 3
                run_steps(t_h, 2 if nd3 else 5, 5)
                (pts, wid), ms = bastr_amd.autotune_step(t_h, 3 if nd3 else 20)
                t_h.close()
                tune[key] = {"points_per_thread": pts, "width": wid,
                             "ms_per_iteration": {f"{p}x{w}": (None if m != m else round(m, 4))
                                                  for (p, w), m in zip([(16, 1), (8, 1), (16, 2), (8, 2)], ms)},
                             "timed_on": f"{nm}: {3 if nd3 else 20} iterations per plan and round, two rounds"}
        print(f"  autotune: {tune}", file=sys.stderr, flush=True)

    rec, last = measure(lib, bastr_amd, name, steps, args.warmup, args.chunk, copy=True)
    finals = {name: last}
    accuracy = {}
    legs = {}
    if name == "2d-2048" and not args.no_legs:
        # the rest of the metric: the FP32 variant and the 512^3 workloads in
        # both precisions, each from the same initial field as its FP64 twin
        # and as many iterations, so their last spectra compare directly
        for leg, twin in LEGS:
            nd3 = len(CONFIGS[leg]["grid"]) == 3
            lsteps = min(steps, CONFIGS[leg]["steps"]) if nd3 else steps
            lwarm = min(args.warmup, 10) if nd3 else args.warmup
            lrec, llast = measure(lib, bastr_amd, leg, lsteps, lwarm, args.chunk)
            finals[leg] = llast
            lrec["config"] = {"workload": workload_desc(leg), "grid": list(CONFIGS[leg]["grid"])}
            if twin is not None and twin in finals and (twin != name or (lsteps, lwarm) == (steps, args.warmup)):
                e_abs, e_shell = ek_err(llast, finals[twin])
                lrec["accuracy_vs_fp64_device"] = {"iterations": lwarm + lsteps, "ek_rel_l2": e_abs,
                                                   "ek_err_max_per_shell": e_shell}
            legs[leg] = lrec
        ratios = {}
        for a32, a64 in (("2d-2048-fp32", None), ("3d-512-fp32", "3d-512")):
            v64 = rec["value"] if a64 is None else legs[a64]["value"]
            ratios["2048^2" if a64 is None else "512^3"] = round(legs[a32]["value"] / v64, 3)
        accuracy["fp32_vs_fp64_device"] = {k: v.get("accuracy_vs_fp64_device") for k, v in legs.items()
                                           if "accuracy_vs_fp64_device" in v}
    elif prec == 1:
        # an FP32 workload alone: its E(k) against the FP64 device path, same iterations
        nchk = min(steps, 200 if len(grid) == 2 else 20)
        res = []
        for pr in (1, 0):
            g = new_field(bastr_amd, cfg_params(name, pr))
            _, _, sp = g.step(nchk)
            res.append((sp[-1], g.get_fields()))
            g.close()
        e_abs, e_shell = ek_err(res[0][0], res[1][0])
        accuracy["fp32_vs_fp64_device"] = {"iterations": nchk, "ek_rel_l2": e_abs, "ek_err_max_per_shell": e_shell,
                                           "fields_rel_l2": rel_l2(res[0][1], res[1][1])}
    cpu = None
    if not args.no_cpu:
        # the same initial field, from the library's FP64 generator, on the CPU restatement
        g0 = new_field(bastr_amd, cfg_params(name, 0))
        fields = g0.get_fields()
        _, _, sp0 = g0.step(cpu_steps)
        dev = (sp0[-1], g0.get_fields())
        g0.close()
        v, thr, dt, osp, ofields = cpu_baseline(cfg_params(name, 0), fields, cpu_steps)
        model, ncpu = cpu_model()
        cpu = {"value": v, "unit": "grid-pt·steps/s", "cores": thr, "kind": "port",
               "cpu_model": model, "host_logical_cpus": ncpu,
               "sample": f"{cpu_steps} loop iteration(s) of the same {dims} FP64 workload ({dt:.1f} s), "
                         f"FP64 C restatement with its own radix-2 FFT, OpenMP"}
        e_abs, e_shell = ek_err(dev[0], osp)
        accuracy["fp64_device_vs_oracle"] = {"iterations": cpu_steps, "ek_rel_l2": e_abs,
                                             "ek_err_max_per_shell": e_shell, "fields_rel_l2": rel_l2(dev[1], ofields)}
    out = {
        "metric": METRIC,
        "value": rec["value"],
        "unit": "grid-pt·steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": rec["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": rec["dtype"],
        "data": "synthetic: seeded random-phase field, E(k) = k^4 exp(-2 (k/4)^2) (pure law, k0=4), seed 1, "
                "u_rms=1, built on the device",
        "config": {"workload": workload_desc(name), "grid": list(grid), "precision": "fp32" if prec == 1 else "fp64"},
        "roofline": rec["roofline"],
        "cpu_baseline": cpu,
        "step_roofline": rec["step_roofline"],
        "launch": rec["launch"],
        "direct_launch": rec["direct_launch"],
        "classes": rec["classes"],
        "autotune": tune or None,
        "accuracy": accuracy,
    }
    if legs:
        out["legs"] = legs
        out["fp32_over_fp64"] = ratios
    print(json.dumps(out))


if __name__ == "__main__":
    main()
