"""Host-side handle over the C-ABI of include/bastr.h.

`Solver` drives any library that exports the bastr entry points (the HIP
library under the prefix "bastr"); fields are numpy arrays in the reference's
Fortran order, shape (ia, ja) or (ia, ja, ka), so that u[i-1, j-1] is the
reference's u1(i, j, 0).
"""
import ctypes as C

import numpy as np

from . import abi


class BastrError(RuntimeError):
    """A nonzero status from the C-ABI (the reference's `stop '...'`)."""

    def __init__(self, code, msg):
        super().__init__(f"bastr error {code}: {msg}")
        self.code = code


def make_params(ia, ja, ka=0, reynolds=5000.0, deltat=1e-4, timemethod=1, forcemethod=0,
                forcek=5, target_energy=0.075, lproject=False, kmax=128, precision=abi.PREC_F64):
    return abi.Params(ia=ia, ja=ja, ka=ka, reynolds=reynolds, deltat=deltat, timemethod=timemethod,
                      forcemethod=forcemethod, forcek=forcek, target_energy=target_energy,
                      lproject=int(bool(lproject)), kmax=kmax, precision=precision)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    """One solver instance (bastr_create ... bastr_destroy)."""

    def __init__(self, lib, prefix, params):
        self._lib = lib
        self._pre = prefix
        self.params = params
        self.nd = 2 if params.ka == 0 else 3
        self.shape = (params.ia, params.ja) if self.nd == 2 else (params.ia, params.ja, params.ka)
        h = C.c_void_p()
        self._check(self._fn("create")(C.byref(params), C.byref(h)))
        self._h = h
        self.allkmax = self._fn("allkmax")(self._h)

    def _fn(self, name):
        return getattr(self._lib, f"{self._pre}_{name}")

    def _check(self, rc):
        if rc < 0:
            msg = self._fn("last_error")()
            raise BastrError(rc, msg.decode() if msg else "")
        return rc

    def close(self):
        if getattr(self, "_h", None):
            self._fn("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _field(self, u):
        a = np.asfortranarray(np.asarray(u, dtype=np.float64))
        if a.shape != self.shape:
            raise ValueError(f"field shape {a.shape} != grid {self.shape}")
        return a

    def set_fields(self, *u):
        fs = [self._field(x) for x in u[: self.nd]]
        ptrs = [_dptr(f) for f in fs] + [None] * (3 - self.nd)
        self._check(self._fn("set_fields")(self._h, *ptrs))

    def init_random_phase(self, k0=4.0, seed=1, energy=1.0):
        """Seeded random-phase field of Bastrpp's hitgen (bastr_init_random_phase):
        spectrum peak k0 (the reference's forcek), <sum u_i^2> = energy."""
        self._check(self._fn("init_random_phase")(self._h, float(k0), int(seed), float(energy)))

    def set_spectrum_law(self, law):
        """Law of later init_random_phase calls: abi.LAW_HITGEN (default) or abi.LAW_PURE."""
        self._check(self._fn("set_spectrum_law")(self._h, int(law)))

    def get_fields(self):
        fs = [np.empty(self.shape, dtype=np.float64, order="F") for _ in range(self.nd)]
        ptrs = [_dptr(f) for f in fs] + [None] * (3 - self.nd)
        self._check(self._fn("get_fields")(self._h, *ptrs))
        return fs

    def set_time(self, nstep, time):
        self._check(self._fn("set_time")(self._h, int(nstep), float(time)))

    def get_time(self):
        n = C.c_int64()
        t = C.c_double()
        self._check(self._fn("get_time")(self._h, C.byref(n), C.byref(t)))
        return n.value, t.value

    def step(self, nsteps, want_spectra=True):
        """Run nsteps loop iterations; returns (status, diags, spectra[n,3,allkmax])."""
        diags = (abi.Diag * nsteps)()
        spec = np.zeros((nsteps, 3, self.allkmax), dtype=np.float64) if want_spectra else None
        rc = self._check(self._fn("step")(self._h, nsteps, diags, _dptr(spec) if spec is not None else None))
        return rc, diags, spec

    def pp_spectra(self, *u):
        """Bastrpp spectra2d of the given fields: (Es[1..K], Ed[1..K]),
        (Esspe, Edspe, k2Edspe, kM1Edspe)."""
        K = self.allkmax
        fs = [self._field(x) for x in u[: self.nd]]
        ptrs = [_dptr(f) for f in fs] + [None] * (3 - self.nd)
        shells = np.zeros(2 * K)
        totals = np.zeros(4)
        self._check(self._fn("pp_spectra")(self._h, *ptrs, _dptr(shells), _dptr(totals)))
        return shells.reshape(2, K), totals

    def pp_transfer(self, *u):
        """Bastrpp transfer2d/3d of the given fields: array (3, K) of T, T1, T2."""
        K = self.allkmax
        fs = [self._field(x) for x in u[: self.nd]]
        ptrs = [_dptr(f) for f in fs] + [None] * (3 - self.nd)
        out = np.zeros(3 * K)
        self._check(self._fn("pp_transfer")(self._h, *ptrs, _dptr(out)))
        return out.reshape(3, K)

    def shell_counts(self):
        c = np.zeros(self.allkmax, dtype=np.int64)
        self._check(self._fn("shell_counts")(self._h, c.ctypes.data_as(C.POINTER(C.c_int64))))
        return c

    def _fft(self, name, a):
        z = np.asfortranarray(np.asarray(a, dtype=np.complex128)).copy(order="F")
        if z.shape != self.shape:
            raise ValueError(f"array shape {z.shape} != grid {self.shape}")
        self._check(self._fn(name)(self._h, _dptr(z)))
        return z

    def fft_forward(self, a):
        """fft2d/fft3d of fftwlink.F90:324-348 (normalised forward c2c)."""
        return self._fft("fft_forward", a)

    def fft_backward(self, a):
        """ifft2d/ifft3d of fftwlink.F90:350-368 (unnormalised backward c2c)."""
        return self._fft("fft_backward", a)


class Ensemble(Solver):
    """nmembers independent 2-D runs of one grid in one handle
    (bastr_create_ensemble): every launch of the step covers all members."""

    def __init__(self, lib, params, nmembers, forced=None):
        self._lib = lib
        self._pre = "bastr"
        self.params = params
        self.nd = 2
        self.shape = (params.ia, params.ja)
        self.nmembers = int(nmembers)
        flags = None
        if forced is not None:
            f = np.ascontiguousarray(np.asarray(forced, dtype=np.int32))
            if f.shape != (self.nmembers,):
                raise ValueError("forced needs one flag per member")
            flags = f.ctypes.data_as(C.POINTER(C.c_int32))
        h = C.c_void_p()
        self._check(self._fn("create_ensemble")(C.byref(params), self.nmembers, flags, C.byref(h)))
        self._h = h
        self.allkmax = self._fn("allkmax")(self._h)

    def set_member_fields(self, m, *u):
        fs = [self._field(x) for x in u[:2]]
        self._check(self._fn("set_member_fields")(self._h, int(m), _dptr(fs[0]), _dptr(fs[1]), None))

    def init_random_phase_member(self, m, k0=4.0, seed=1, energy=1.0):
        self._check(self._fn("init_random_phase_member")(self._h, int(m), float(k0), int(seed), float(energy)))

    def get_member_fields(self, m):
        fs = [np.empty(self.shape, dtype=np.float64, order="F") for _ in range(2)]
        self._check(self._fn("get_member_fields")(self._h, int(m), _dptr(fs[0]), _dptr(fs[1]), None))
        return fs

    def get_member_time(self, m):
        n = C.c_int64()
        t = C.c_double()
        self._check(self._fn("get_member_time")(self._h, int(m), C.byref(n), C.byref(t)))
        return n.value, t.value

    def step(self, nsteps, want_spectra=True):
        """Run nsteps iterations of every member; returns (status,
        diags[nmembers][nsteps], spectra[nmembers, nsteps, 3, allkmax])."""
        M = self.nmembers
        flat = (abi.Diag * (M * nsteps))()
        spec = np.zeros((M, nsteps, 3, self.allkmax), dtype=np.float64) if want_spectra else None
        rc = self._check(self._fn("step")(self._h, nsteps, flat, _dptr(spec) if spec is not None else None))
        diags = [flat[m * nsteps:(m + 1) * nsteps] for m in range(M)]
        return rc, diags, spec


def step_group(handles, nsteps, want_spectra=True):
    """bastr_step_group over Solver / Ensemble objects of one library: every
    handle's iterations are launched before any is waited for.  Returns
    (status, [(diags, spectra) per handle]) in each object's own step layout."""
    lib = handles[0]._lib
    n = len(handles)
    hs = (C.c_void_p * n)(*[h._h.value for h in handles])
    flats, specs = [], []
    for h in handles:
        M = getattr(h, "nmembers", 1)
        flats.append((abi.Diag * (M * nsteps))())
        shape = (M, nsteps, 3, h.allkmax) if isinstance(h, Ensemble) else (nsteps, 3, h.allkmax)
        specs.append(np.zeros(shape, dtype=np.float64) if want_spectra else None)
    dp = (C.POINTER(abi.Diag) * n)(*[C.cast(f, C.POINTER(abi.Diag)) for f in flats])
    sp = (C.POINTER(C.c_double) * n)(*[_dptr(x) if x is not None else None for x in specs])
    rc = handles[0]._check(lib.bastr_step_group(hs, n, nsteps, dp, sp))
    out = []
    for h, f, x in zip(handles, flats, specs):
        if isinstance(h, Ensemble):
            out.append(([f[m * nsteps:(m + 1) * nsteps] for m in range(h.nmembers)], x))
        else:
            out.append((f[:], x))
    return rc, out


def diag_table(diags, names):
    return np.array([[getattr(d, n) for n in names] for d in diags], dtype=np.float64)
