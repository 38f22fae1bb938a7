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


def autotune(n, precision=abi.PREC_F64):
    """Time both radix plans of the length-n column passes and keep the faster
    (bastr_autotune): returns (points per thread, [ms of the 16-point plan,
    ms of the 8-point plan])."""
    L = load()
    pts = C.c_int32()
    ms = (C.c_double * 2)()
    rc = L.bastr_autotune(int(n), int(precision), C.byref(pts), ms)
    if rc:
        raise BastrError(rc, L.bastr_last_error().decode())
    return pts.value, [ms[0], ms[1]]


def tune_set(n, points, precision=abi.PREC_F64):
    L = load()
    rc = L.bastr_tune_set(int(n), int(precision), int(points))
    if rc:
        raise BastrError(rc, L.bastr_last_error().decode())


def tune_get(n, precision=abi.PREC_F64):
    return load().bastr_tune_get(int(n), int(precision))


def ensemble(params, nmembers, forced=None):
    """A HIP-backed ensemble of nmembers independent 2-D runs on one grid;
    forced: None or nmembers flags (0 = free decay, 1 = forced as params say)."""
    return Ensemble(load(), params, nmembers, forced)
