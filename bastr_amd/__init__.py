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


def ensemble(params, nmembers, forced=None):
    """A HIP-backed ensemble of nmembers independent 2-D runs on one grid;
    forced: None or nmembers flags (0 = free decay, 1 = forced as params say)."""
    return Ensemble(load(), params, nmembers, forced)
