"""ctypes binding of libbaorec_hip.so (declarations in include/baorec.h).

This is the Python mirror of the Julia ``ccall`` binding: one function per
C entry point, same argument order.  Loading fails loudly when the library is
missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.environ.get("BAOREC_HIP_LIB", os.path.join(_PKG_ROOT, "libbaorec_hip.so"))

OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_NO_DEVICE = -2
ERR_HIP = -3
ERR_OUT_OF_MEMORY = -4
ERR_OUT_OF_BOX = -5
ERR_NO_RESULT = -6
ERR_UNSUPPORTED = -7

MEM_HOST = 0
MEM_DEVICE = 1
ITERATIVE = 0
MULTIGRID = 1
FIELD = {"disp": 0, "rsd": 1, "sum": 2}
MAS = {"cic": 0, "tsc": 1}
MG_SCHEDULE = {"reference": 0, "rbgs": 1}


class Params(ctypes.Structure):
    """struct baorec_params (include/baorec.h)."""
    _fields_ = [
        ("bias", ctypes.c_float),
        ("f", ctypes.c_float),
        ("smoothing_radius", ctypes.c_float),
        ("box_size", ctypes.c_float * 3),
        ("box_min", ctypes.c_float * 3),
        ("has_los", ctypes.c_int32),
        ("los", ctypes.c_float * 3),
        ("n_iter", ctypes.c_int32),
        ("beta", ctypes.c_float),
        ("jacobi_damping_factor", ctypes.c_float),
        ("jacobi_niterations", ctypes.c_int32),
        ("vcycle_niterations", ctypes.c_int32),
        ("mas", ctypes.c_int32),
        ("mg_schedule", ctypes.c_int32),
        ("tolerance", ctypes.c_double),
        ("max_vcycles", ctypes.c_int32),
        ("rb_sweeps", ctypes.c_int32),
        ("fp64_coarse_max", ctypes.c_int32),
        ("box_pad", ctypes.c_float),
        ("ran_min", ctypes.c_float),
        ("mutate_inputs", ctypes.c_int32),
        ("mg_warm_start", ctypes.c_int32),
        ("mg_seed_steps", ctypes.c_int32),
        ("vcycles_done", ctypes.c_int32),
        ("final_residual", ctypes.c_double),
        ("final_change", ctypes.c_double),
        ("seed_residual", ctypes.c_double),
        ("rhs_shift", ctypes.c_double),
    ]


class Cosmology(ctypes.Structure):
    """struct baorec_cosmology (include/baorec.h)."""
    _fields_ = [
        ("omega_b", ctypes.c_float),
        ("omega_c", ctypes.c_float),
        ("h", ctypes.c_float),
        ("Neff", ctypes.c_float),
        ("Omega_k0", ctypes.c_float),
        ("T_cmb", ctypes.c_float),
        ("w0", ctypes.c_float),
        ("wa", ctypes.c_float),
        ("z_tab_min", ctypes.c_double),
        ("z_tab_max", ctypes.c_double),
        ("z_tab_num", ctypes.c_int32),
    ]


class PkParams(ctypes.Structure):
    """struct baorec_pk_params (include/baorec.h)."""
    _fields_ = [
        ("box_size", ctypes.c_float * 3),
        ("box_min", ctypes.c_float * 3),
        ("los", ctypes.c_float * 3),
        ("mas", ctypes.c_int32),
        ("interlace", ctypes.c_int32),
        ("kmin", ctypes.c_double),
        ("dk", ctypes.c_double),
        ("nbins", ctypes.c_int32),
    ]


P = ctypes.c_void_p
I32P = ctypes.POINTER(ctypes.c_int32)
FP = ctypes.c_void_p  # float* (host or device address)
I64 = ctypes.c_int64
INT = ctypes.c_int
PP = ctypes.POINTER(Params)

# name -> (restype, argtypes); mirrors include/baorec.h
SIGNATURES = {
    "baorec_params_default": (None, [PP, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "baorec_create": (INT, [ctypes.POINTER(P), INT]),
    "baorec_destroy": (INT, [P]),
    "baorec_last_error": (ctypes.c_char_p, [P]),
    "baorec_version": (ctypes.c_char_p, []),
    "baorec_run_iterative": (INT, [P, PP, I32P, INT, FP, FP, FP, FP, I64, FP, FP, FP, FP, I64]),
    "baorec_run_multigrid": (INT, [P, PP, I32P, INT, FP, FP, FP, FP, I64, FP, FP, FP, FP, I64]),
    "baorec_result_info": (INT, [P, I32P, I32P, ctypes.POINTER(ctypes.c_uint64)]),
    "baorec_result_copy": (INT, [P, INT, FP, ctypes.c_uint64]),
    "baorec_setup_overdensity": (INT, [P, PP, I32P, INT, FP, FP, FP, FP, FP, I64, FP, FP, FP, FP, I64]),
    "baorec_reconstructed_overdensity": (INT, [P, PP, I32P, INT, FP, FP, FP, FP, FP, I64, FP, FP, FP, FP, I64]),
    "baorec_reconstructed_potential": (INT, [P, PP, I32P, INT, FP, FP, FP, FP, FP, I64, FP, FP, FP, FP, I64]),
    "baorec_read_shifts": (INT, [P, PP, INT, FP, FP, FP, I64, FP, I32P, INT, INT, FP, FP, FP]),
    "baorec_reconstructed_positions": (INT, [P, PP, INT, FP, FP, FP, I64, FP, I32P, INT, INT, FP, FP, FP]),
    "baorec_setup_box": (INT, [P, INT, FP, FP, FP, I64, ctypes.c_float, FP, FP]),
    "baorec_cic": (INT, [P, PP, I32P, INT, FP, FP, FP, FP, FP, I64, INT]),
    "baorec_read_cic": (INT, [P, PP, I32P, INT, FP, FP, FP, FP, FP, I64]),
    "baorec_smooth": (INT, [P, PP, I32P, INT, FP]),
    "baorec_rfft3d": (INT, [P, I32P, INT, FP, FP]),
    "baorec_irfft3d": (INT, [P, I32P, INT, FP, FP]),
    "baorec_k_vec": (INT, [I32P, FP, FP, FP, FP]),
    "baorec_x_vec": (INT, [I32P, FP, FP, FP, FP, FP]),
    "baorec_mg_jacobi": (INT, [P, PP, I32P, INT, FP, FP]),
    "baorec_mg_gauss_seidel": (INT, [P, PP, I32P, INT, FP, FP, INT, INT]),
    "baorec_run_positions_graph": (INT, [P, PP, I32P, FP, FP, FP, FP, I64, FP, FP, FP]),
    "baorec_mg_residual": (INT, [P, PP, I32P, INT, FP, FP, FP]),
    "baorec_mg_reduce": (INT, [P, I32P, INT, FP, FP]),
    "baorec_mg_residual_restrict": (INT, [P, PP, I32P, INT, FP, FP, FP, INT]),
    "baorec_mg_prolong": (INT, [P, I32P, INT, FP, FP, INT]),
    "baorec_mg_vcycle": (INT, [P, PP, I32P, INT, FP, FP]),
    "baorec_mg_fmg": (INT, [P, PP, I32P, INT, FP, FP]),
    "baorec_cosmology_default": (None, [ctypes.POINTER(Cosmology)]),
    "baorec_cosmology_desi": (None, [ctypes.POINTER(Cosmology)]),
    "baorec_comoving_distance": (INT, [P, ctypes.POINTER(Cosmology), INT, FP, I64, FP]),
    "baorec_sky_to_cartesian": (INT, [P, ctypes.POINTER(Cosmology), INT, FP, FP, FP, I64, FP, FP, FP]),
    "baorec_cartesian_to_sky": (INT, [P, ctypes.POINTER(Cosmology), INT, FP, FP, FP, I64, FP, FP, FP]),
    "baorec_warmstart_nparams": (I64, []),
    "baorec_warmstart_init": (INT, [P, ctypes.c_uint64, ctypes.c_float]),
    "baorec_warmstart_save": (INT, [P, ctypes.c_char_p]),
    "baorec_warmstart_load": (INT, [P, ctypes.c_char_p]),
    "baorec_warmstart_params": (INT, [P, INT, FP, I64]),
    "baorec_warmstart_step": (INT, [P, PP, I32P, INT, FP, FP, ctypes.c_int32, ctypes.c_float,
                                    ctypes.POINTER(ctypes.c_double)]),
    "baorec_warmstart_eval": (INT, [P, I32P, INT, FP, FP, ctypes.c_int32, FP, ctypes.POINTER(ctypes.c_double)]),
    "baorec_warmstart_predict": (INT, [P, PP, I32P, INT, FP, FP]),
    "baorec_warmstart_tensor": (INT, [P, INT, INT, FP, I64]),
    "baorec_sky_to_cartesian_noalloc": (INT, [P, ctypes.POINTER(Cosmology), FP, FP, FP, I64, FP, FP, FP]),
    "baorec_pk_params_default": (None, [ctypes.POINTER(PkParams)]),
    "baorec_power_multipoles": (INT, [P, ctypes.POINTER(PkParams), I32P, INT, FP, FP, FP, FP, I64,
                                      FP, FP, FP, FP, FP]),
    "baorec_catalog_rows": (INT, [ctypes.c_char_p, ctypes.POINTER(I64), I32P]),
    "baorec_catalog_read": (INT, [ctypes.c_char_p, ctypes.c_int32, I32P, ctypes.POINTER(FP), I64,
                                  ctypes.POINTER(I64)]),
    "baorec_write_npy": (INT, [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(FP), I64]),
    "baorec_io_last_error": (ctypes.c_char_p, []),
    "baorec_run_batch": (INT, [P, PP, I32P, INT, INT, FP, FP, FP, FP, FP, FP, FP, FP]),
    "baorec_profile": (INT, [P, INT]),
    "baorec_profile_report": (I64, [P, ctypes.c_char_p, I64]),
    "baorec_autotune": (INT, [P, PP, I32P, INT, FP, FP, FP, FP, I64]),
    "baorec_tune_report": (I64, [P, ctypes.c_char_p, I64]),
    "baorec_tune_save": (INT, [P, ctypes.c_char_p]),
    "baorec_tune_load": (INT, [P, ctypes.c_char_p]),
}

_lib = None


def load():
    """Load libbaorec_hip.so (raises OSError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"libbaorec_hip.so not found at {LIB_PATH}: build it first "
                          "(python -c 'import __graft_entry__ as g; g.build()')")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
