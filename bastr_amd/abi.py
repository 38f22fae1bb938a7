"""ctypes mirror of include/bastr.h.

The structures here are the only Python-side description of the C-ABI; both
the HIP library (bastr_amd/libbastr_hip.so) and the CPU oracle expose
functions taking exactly these structures.
"""
import ctypes as C

BASTR_OK = 0
BASTR_STOP_NAN = 1
BASTR_EINVAL = -1
BASTR_EDEVICE = -2
BASTR_ENOMEM = -3

PREC_F64 = 0
PREC_F32 = 1

LAW_HITGEN = 0   # Bastrpp's piecewise hitgen law (Bastrpp.F90:654-658)
LAW_PURE = 1     # E(k) = k^4 exp(-2 (k/k0)^2) at every k


class Params(C.Structure):
    """bastr_params: run parameters of datin/input.dat (readwrite.F90:35-44)."""
    _fields_ = [
        ("ia", C.c_int32), ("ja", C.c_int32), ("ka", C.c_int32),
        ("reynolds", C.c_double),
        ("deltat", C.c_double),
        ("timemethod", C.c_int32),
        ("forcemethod", C.c_int32),
        ("forcek", C.c_int32),
        ("target_energy", C.c_double),
        ("lproject", C.c_int32),
        ("kmax", C.c_int32),
        ("precision", C.c_int32),
    ]


class Diag(C.Structure):
    """bastr_diag: one iteration's log values (specall, velgrad, factor_output)."""
    _fields_ = [
        ("nstep", C.c_int64),
        ("time", C.c_double),
        ("Esspe", C.c_double), ("Edspe", C.c_double), ("k2Edspe", C.c_double), ("kMEdspe", C.c_double),
        ("umumtheta2", C.c_double), ("umumijji", C.c_double), ("u2theta", C.c_double),
        ("dissp", C.c_double), ("eta_min", C.c_double), ("Taylength", C.c_double),
        ("ReTay", C.c_double), ("Kollength", C.c_double),
        ("force_factor", C.c_double), ("force_E", C.c_double),
        ("force_energy", C.c_double), ("force_Fen", C.c_double),
        ("nan_flag", C.c_int32), ("pad_", C.c_int32),
    ]

    SPECALL = ("Esspe", "Edspe", "k2Edspe", "kMEdspe")
    VELGRAD = ("umumtheta2", "umumijji", "u2theta", "dissp", "eta_min", "Taylength", "ReTay", "Kollength")


# Every entry point of include/bastr.h: (name, restype, argtypes) with the
# library prefix ("bastr" for the HIP library, "bo" for the oracle).
_P = C.c_void_p
_D = C.POINTER(C.c_double)
SIGNATURES = [
    ("last_error", C.c_char_p, []),
    ("create", C.c_int, [C.POINTER(Params), C.POINTER(_P)]),
    ("destroy", C.c_int, [_P]),
    ("allkmax", C.c_int, [_P]),
    ("set_fields", C.c_int, [_P, _D, _D, _D]),
    ("get_fields", C.c_int, [_P, _D, _D, _D]),
    ("set_time", C.c_int, [_P, C.c_int64, C.c_double]),
    ("get_time", C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    ("step", C.c_int, [_P, C.c_int32, C.POINTER(Diag), _D]),
    ("shell_counts", C.c_int, [_P, C.POINTER(C.c_int64)]),
    ("init_random_phase", C.c_int, [_P, C.c_double, C.c_uint64, C.c_double]),
    ("set_spectrum_law", C.c_int, [_P, C.c_int32]),
    ("pp_spectra", C.c_int, [_P, _D, _D, _D, _D, _D]),
    ("pp_transfer", C.c_int, [_P, _D, _D, _D, _D]),
    ("fft_forward", C.c_int, [_P, _D]),
    ("fft_backward", C.c_int, [_P, _D]),
]


# Ensemble entry points (HIP library only: the oracle is a one-field restatement).
ENSEMBLE_SIGNATURES = [
    ("create_ensemble", C.c_int, [C.POINTER(Params), C.c_int32, C.POINTER(C.c_int32), C.POINTER(_P)]),
    ("nmembers", C.c_int, [_P]),
    ("set_member_fields", C.c_int, [_P, C.c_int32, _D, _D, _D]),
    ("get_member_fields", C.c_int, [_P, C.c_int32, _D, _D, _D]),
    ("init_random_phase_member", C.c_int, [_P, C.c_int32, C.c_double, C.c_uint64, C.c_double]),
    ("get_member_time", C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    ("step_group", C.c_int, [C.POINTER(_P), C.c_int32, C.c_int32, C.POINTER(C.POINTER(Diag)), C.POINTER(_D)]),
]


# The step without device allocation and the column radix plans (HIP library
# only; no counterpart in the reference or the oracle).
NOALLOC_SIGNATURES = [
    ("step_noalloc", C.c_int, [_P, C.c_int32, C.POINTER(Diag), _D]),
    ("autotune", C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                           C.POINTER(C.c_int32), _D]),
    ("autotune_step", C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _D]),
    ("tune_set", C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    ("tune_get", C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
]


# The learned subgrid closure (include/bastr_closure.h; HIP library only).
_F = C.POINTER(C.c_float)
CLOSURE_NPARAM = 6658
CLOSURE_NSCALE = 6
CLOSURE_SIGNATURES = [
    ("filter_to", C.c_int, [_P, _P]),
    ("pairs_create", C.c_int, [C.c_int32, C.c_int32, C.POINTER(_P)]),
    ("pairs_add", C.c_int, [_P, _P]),
    ("pairs_count", C.c_int, [_P]),
    ("pairs_get", C.c_int, [_P, C.c_int32, _F, _F]),
    ("pairs_destroy", C.c_int, [_P]),
    ("closure_create", C.c_int, [C.c_uint64, C.POINTER(_P)]),
    ("closure_destroy", C.c_int, [_P]),
    ("closure_set_params", C.c_int, [_P, _F, _F]),
    ("closure_get_params", C.c_int, [_P, _F, _F]),
    ("closure_fit_scales", C.c_int, [_P, _P]),
    ("closure_loss_grad", C.c_int, [_P, _P, C.c_int32, C.c_int32, _D, _F]),
    ("closure_train", C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _D]),
    ("closure_apply", C.c_int, [_P, _P, C.c_int32, _F]),
    ("closure_attach", C.c_int, [_P, _P]),
    ("closure_per_step", C.c_int, [_P, C.c_int32]),
]


def bind(lib, prefix, ensemble=False):
    """Attach argtypes/restype for every entry point; raise if one is missing.
    ensemble=True also binds the HIP library's own entry points (ensembles,
    bastr_step_group, bastr_step_noalloc, the closure)."""
    for name, res, args in SIGNATURES + (ENSEMBLE_SIGNATURES + NOALLOC_SIGNATURES + CLOSURE_SIGNATURES
                                         if ensemble else []):
        fn = getattr(lib, f"{prefix}_{name}")  # AttributeError = missing export
        fn.restype = res
        fn.argtypes = args
    return lib
