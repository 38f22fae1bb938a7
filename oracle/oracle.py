"""TEST INFRASTRUCTURE ONLY: Python handle on the C oracle (oracle/liboracle.so).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg.  The oracle exports the same entry points as the HIP library under the
prefix "bo", so it is driven through the same bastr_amd.solver.Solver class.
"""
import ctypes as C
import os

import numpy as np

from bastr_amd import abi
from bastr_amd.solver import Solver

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} missing (run `make -C oracle`)")
        L = C.CDLL(path)
        abi.bind(L, "bo")
        L.bo_fft_c2c.restype = C.c_int
        L.bo_fft_c2c.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int]
        L.bo_set_closure.restype = C.c_int
        L.bo_set_closure.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.bo_set_closure_per_step.restype = C.c_int
        L.bo_set_closure_per_step.argtypes = [C.c_void_p, C.c_int32]
        L.bo_set_threads.argtypes = [C.c_int]
        L.bo_get_threads.restype = C.c_int
        _LIB = L
    return _LIB


def solver(params):
    return Solver(lib(), "bo", params)


def set_closure(solver, params, scales):
    """Attach the closure network (params, scales as float32 arrays; None: detach)
    to an oracle solver: every compute_ut then adds its force."""
    if params is None:
        rc = lib().bo_set_closure(solver._h, None, None)
    else:
        p = np.ascontiguousarray(params, dtype=np.float32)
        s = np.ascontiguousarray(scales, dtype=np.float32)
        rc = lib().bo_set_closure(solver._h, p.ctypes.data_as(C.POINTER(C.c_float)),
                                  s.ctypes.data_as(C.POINTER(C.c_float)))
    if rc:
        raise RuntimeError("bo_set_closure failed")


def set_closure_per_step(solver, on=True):
    """Cadence of the oracle's closure force: once per time step (stage A's
    force added at the three stages) or, on=False, at every stage."""
    if lib().bo_set_closure_per_step(solver._h, 1 if on else 0):
        raise RuntimeError("bo_set_closure_per_step failed")


def set_threads(n):
    lib().bo_set_threads(int(n))


def fft_c2c(a, direction):
    """Fortran-order c2c with fftwlink conventions: direction -1 forward (normalised), +1 backward."""
    z = np.asfortranarray(np.asarray(a, dtype=np.complex128)).copy(order="F")
    shp = z.shape
    n2 = shp[2] if len(shp) == 3 else 0
    rc = lib().bo_fft_c2c(shp[0], shp[1], n2, z.ctypes.data_as(C.POINTER(C.c_double)), int(direction))
    if rc:
        raise RuntimeError("bo_fft_c2c failed")
    return z
