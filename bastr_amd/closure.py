"""Host side of the learned subgrid closure (include/bastr_closure.h).

A coarse run of the sharp-filtered field u_bar = P_c u (|k| < k_c = floor(n/3),
the reference's own spherical dealiasing, tool.F90:149-173) misses the
subgrid force tau = P_c N(u) - D_c N(u_bar), N(u) = -u_j d_j u_i.  Pairs
(coarse fields, tau) come from fine runs of this library; a per-point MLP
trained on them by the library's BF16-MFMA kernels is attached to a coarse
handle, whose every RK3 stage then adds the network's force to its nonlinear
term (where compute_ut2D forms the tendency, solution.F90:841-846).
"""
import ctypes as C

import numpy as np

from . import abi, load

NPARAM = abi.CLOSURE_NPARAM
NSCALE = abi.CLOSURE_NSCALE
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)


def _f(a):
    return a.ctypes.data_as(_F)


def _check(lib, rc):
    if rc < 0:
        from .solver import BastrError
        msg = lib.bastr_last_error()
        raise BastrError(rc, msg.decode() if msg else "")
    return rc


def filter_to(fine, coarse):
    """coarse's state := the sharp filter of fine's (bastr_filter_to)."""
    lib = load()
    _check(lib, lib.bastr_filter_to(fine._h, coarse._h))


class Pairs:
    """Training pairs of one n x n coarse grid, kept on the device."""

    def __init__(self, n, capacity):
        self._lib = load()
        self.n = int(n)
        h = C.c_void_p()
        _check(self._lib, self._lib.bastr_pairs_create(self.n, int(capacity), C.byref(h)))
        self._h = h

    def add(self, fine):
        _check(self._lib, self._lib.bastr_pairs_add(self._h, fine._h))

    def __len__(self):
        return self._lib.bastr_pairs_count(self._h)

    def get(self, i):
        """(fields[4, n, n], target[2, n, n]) of snapshot i, Fortran order per field."""
        n = self.n
        f = np.empty(4 * n * n, dtype=np.float32)
        t = np.empty(2 * n * n, dtype=np.float32)
        _check(self._lib, self._lib.bastr_pairs_get(self._h, int(i), _f(f), _f(t)))
        return (f.reshape(4, n, n).transpose(0, 2, 1), t.reshape(2, n, n).transpose(0, 2, 1))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bastr_pairs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Closure:
    """The per-point MLP (36 -> 64 -> 64 -> 2) of the closure."""

    def __init__(self, seed=0):
        self._lib = load()
        h = C.c_void_p()
        _check(self._lib, self._lib.bastr_closure_create(int(seed), C.byref(h)))
        self._h = h

    def params(self):
        p = np.empty(NPARAM, dtype=np.float32)
        s = np.empty(NSCALE, dtype=np.float32)
        _check(self._lib, self._lib.bastr_closure_get_params(self._h, _f(p), _f(s)))
        return p, s

    def set_params(self, params, scales):
        p = np.ascontiguousarray(params, dtype=np.float32)
        s = np.ascontiguousarray(scales, dtype=np.float32)
        if p.shape != (NPARAM,) or s.shape != (NSCALE,):
            raise ValueError("params / scales of the wrong length")
        _check(self._lib, self._lib.bastr_closure_set_params(self._h, _f(p), _f(s)))

    def fit_scales(self, pairs):
        _check(self._lib, self._lib.bastr_closure_fit_scales(self._h, pairs._h))

    def loss_grad(self, pairs, first=0, count=None):
        count = len(pairs) - first if count is None else count
        loss = C.c_double()
        g = np.empty(NPARAM, dtype=np.float32)
        _check(self._lib, self._lib.bastr_closure_loss_grad(self._h, pairs._h, int(first), int(count),
                                                             C.byref(loss), _f(g)))
        return loss.value, g

    def train(self, pairs, iters, batch=256, lr=1e-3, seed=1):
        """iters Adam steps on batches of `batch` random 32-point tiles; returns the losses."""
        losses = np.empty(int(iters), dtype=np.float64)
        _check(self._lib, self._lib.bastr_closure_train(self._h, pairs._h, int(iters), int(batch), float(lr),
                                                         int(seed), losses.ctypes.data_as(_D)))
        return losses

    def apply(self, pairs, i):
        """The network's force (scaled) on snapshot i: array [2, n, n]."""
        n = pairs.n
        f = np.empty(2 * n * n, dtype=np.float32)
        _check(self._lib, self._lib.bastr_closure_apply(self._h, pairs._h, int(i), _f(f)))
        return f.reshape(2, n, n).transpose(0, 2, 1)

    def attach(self, solver):
        _check(self._lib, self._lib.bastr_closure_attach(solver._h, self._h))

    @staticmethod
    def per_step(solver, on=True):
        """Cadence of the attached force: True evaluates the network once per
        time step (on the step's input field) and adds that force at the three
        RK3 stages; False (the default) evaluates it at every stage."""
        lib = load()
        _check(lib, lib.bastr_closure_per_step(solver._h, 1 if on else 0))

    @staticmethod
    def detach(solver):
        lib = load()
        _check(lib, lib.bastr_closure_attach(solver._h, None))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bastr_closure_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fine_run(params, seed, k0=4.0, law=1):
    """A fine (DNS) handle holding the seeded random-phase field (pure law by default)."""
    from . import solver as _solver
    s = _solver(params)
    s.set_spectrum_law(law)
    s.init_random_phase(k0=k0, seed=seed, energy=1.0)
    return s


def make_pairs(fine_params, n_coarse, seeds, snap_steps, k0=4.0, law=1, chunk=200):
    """Training pairs from fine runs: one run per seed, a snapshot after each
    iteration count in snap_steps (ascending)."""
    ps = Pairs(n_coarse, len(seeds) * len(snap_steps))
    for seed in seeds:
        f = fine_run(fine_params, seed, k0, law)
        done = 0
        for target in snap_steps:
            while done < target:
                k = min(chunk, target - done)
                rc, _, _ = f.step(k, want_spectra=False)
                if rc != 0:
                    raise RuntimeError(f"fine run (seed {seed}) stopped at iteration {done}: status {rc}")
                done += k
            ps.add(f)
        f.close()
    return ps


def shell_energy(spec_row):
    """E(k) = Es + Ed of one iteration's (kn, Es, Ed)."""
    return spec_row[1] + spec_row[2]


def evaluate(fine_params, coarse_params, seed, t0_steps, fine_steps, net=None, k0=4.0, law=1, per_step=False):
    """E(k) of a coarse run against the fine run of the same seed.  The coarse
    run starts from the fine field filtered after t0_steps fine iterations and
    advances (fine_steps - t0_steps) * dt_fine / dt_coarse iterations, with the
    network attached when given (per_step: its force evaluated once per time
    step, Closure.per_step).  Returns (E_fine[1..K], E_coarse[1..K]) after
    fine_steps fine iterations, K = the coarse cutoff - 1 (shells inside it)."""
    from . import solver as _solver
    ratio = coarse_params.deltat / fine_params.deltat
    ncoarse = (fine_steps - t0_steps) / ratio
    if abs(ncoarse - round(ncoarse)) > 1e-9:
        raise ValueError("the coarse steps must span the fine interval exactly")
    f = fine_run(fine_params, seed, k0, law)
    rc = f.step(t0_steps, want_spectra=False)[0] if t0_steps else 0
    c = _solver(coarse_params)
    filter_to(f, c)
    if net is not None:
        net.attach(c)
        Closure.per_step(c, per_step)
    rc2 = f.step(fine_steps - t0_steps, want_spectra=False)[0]
    rc3 = c.step(int(round(ncoarse)), want_spectra=False)[0]
    # the diagnostics of the next iteration are those of the field reached
    _, _, fsp = f.step(1, want_spectra=True)
    _, _, csp = c.step(1, want_spectra=True)
    K = coarse_params.ia // 3 - 1
    Ef = shell_energy(fsp[0])[:K]
    Ec = shell_energy(csp[0])[:K]
    f.close()
    c.close()
    if rc or rc2 or rc3:
        raise RuntimeError(f"a run stopped on a NaN (fine {rc}/{rc2}, coarse {rc3})")
    return Ef, Ec
