"""bastr_amd: MI355X-native pseudo-spectral time step of the Bastr Burgers solver.

The compute path is the HIP library bastr_amd/libbastr_hip.so (hand-written
gfx950 kernels behind the C-ABI of include/bastr.h).  There is no CPU
fallback: `load()` raises if the library is missing or lacks an entry point.
"""
import ctypes as C
import os

from . import abi
from .solver import BastrError, Ensemble, Solver, diag_table, make_params, step_group  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbastr_hip.so")
_LIB = None


def load():
    """Load the HIP library (fails loudly; never substitutes another path)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} missing (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        abi.bind(L, "bastr", ensemble=True)
        _LIB = L
    return _LIB


def solver(params):
    """A HIP-backed solver handle."""
    return Solver(load(), "bastr", params)


def autotune(n, precision=abi.PREC_F64, ncols=None, nplanes=1):
    """Time the plans of the length-n column passes (ncols columns of nplanes
    planes; default: the kept columns of an n x n 2-D step) and keep the fastest
    (bastr_autotune): returns ((points per thread, width multiplier),
    [ms of the plans (16, 1), (8, 1), (16, 2), (8, 2)], NaN where absent)."""
    L = load()
    if ncols is None:
        ncols = max(16, (n // 3 + 15) // 16 * 16)
    pts, wid = C.c_int32(), C.c_int32()
    ms = (C.c_double * 4)()
    rc = L.bastr_autotune(int(n), int(precision), int(ncols), int(nplanes), C.byref(pts), C.byref(wid), ms)
    if rc:
        raise BastrError(rc, L.bastr_last_error().decode())
    return (pts.value, wid.value), list(ms)


def autotune_step(solver, iters=20):
    """Time the column plans on the solver's (or ensemble's) own step, iters
    iterations per plan from its current state, which is restored, and keep
    the fastest (bastr_autotune_step): returns ((points per thread, width
    multiplier), [ms per iteration of the plans (16, 1), (8, 1), (16, 2),
    (8, 2)], NaN where absent)."""
    L = load()
    pts, wid = C.c_int32(), C.c_int32()
    ms = (C.c_double * 4)()
    rc = L.bastr_autotune_step(solver._h, int(iters), C.byref(pts), C.byref(wid), ms)
    if rc:
        raise BastrError(rc, L.bastr_last_error().decode())
    return (pts.value, wid.value), list(ms)


def tune_set(n, points, precision=abi.PREC_F64, width=1):
    L = load()
    rc = L.bastr_tune_set(int(n), int(precision), int(points), int(width))
    if rc:
        raise BastrError(rc, L.bastr_last_error().decode())


def tune_get(n, precision=abi.PREC_F64):
    """(points per thread, width multiplier) of the length-n column plan in use."""
    pts, wid = C.c_int32(), C.c_int32()
    load().bastr_tune_get(int(n), int(precision), C.byref(pts), C.byref(wid))
    return pts.value, wid.value


def ensemble(params, nmembers, forced=None):
    """A HIP-backed ensemble of nmembers independent 2-D runs on one grid;
    forced: None or nmembers flags (0 = free decay, 1 = forced as params say)."""
    return Ensemble(load(), params, nmembers, forced)
