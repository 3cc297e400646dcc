"""ctypes binding of libmadsdp_hip.so (declared in include/madsdp_hip.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared object is missing, and creating a solver fails loudly when
no HIP device is present.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

# MADSDP_HIP_LIB selects another build of the same library (the host-ASan
# one, libmadsdp_hip_asan.so)
LIB_PATH = Path(os.environ.get("MADSDP_HIP_LIB") or
                Path(__file__).resolve().parent / "libmadsdp_hip.so")

c_double_p = C.POINTER(C.c_double)
c_int64_p = C.POINTER(C.c_int64)
c_int32_p = C.POINTER(C.c_int32)


class Options(C.Structure):
    _fields_ = [
        ("tol", C.c_double), ("max_iter", C.c_int32), ("max_wall_time", C.c_double),
        ("mu_min", C.c_double), ("regularize_schur", C.c_double), ("max_schur_reg", C.c_int32),
        ("algorithm", C.c_int32), ("direction", C.c_int32), ("step", C.c_int32),
        ("centering", C.c_int32), ("initial_point", C.c_int32), ("schur", C.c_int32),
        ("tau", C.c_double), ("shrink", C.c_double), ("boundary", C.c_double),
        ("sigma_exponent", C.c_double), ("initial_ratio", C.c_double), ("ratio", C.c_double),
        ("identity_scale", C.c_double), ("initial_shift", C.c_double), ("max_shift", C.c_double),
        ("dc_beta", C.c_double), ("dc_delta", C.c_double), ("dc_A", C.c_double),
        ("dc_centering_max_iter", C.c_int32),
        ("precision", C.c_int32), ("mu_switch", C.c_double), ("refine_max_iter", C.c_int32),
        ("lowp_min_gflop", C.c_double), ("shared_device", C.c_int32),
    ]


class Info(C.Structure):
    _fields_ = [
        ("mu", C.c_double), ("sigma", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
        ("inf_pr", C.c_double), ("inf_du", C.c_double), ("inf_compl", C.c_double),
        ("obj_val", C.c_double), ("pobj", C.c_double), ("dobj", C.c_double),
        ("iter", C.c_int32), ("status", C.c_int32), ("n_newton", C.c_int32),
        ("n_predictor", C.c_int32), ("dc_t", C.c_double), ("dc_lambda", C.c_double),
        ("lowp_factorizations", C.c_int32), ("fp64_fallbacks", C.c_int32),
        ("refine_steps", C.c_int32), ("pad", C.c_int32), ("device_bytes", C.c_int64),
        ("schur_reg", C.c_double),
    ]


class RequestInfo(C.Structure):
    _fields_ = [("arrival", C.c_double), ("admitted", C.c_double), ("uploaded", C.c_double),
                ("done", C.c_double),
                ("footprint", C.c_int64), ("rc", C.c_int32), ("slot", C.c_int32),
                ("finished", C.c_int32), ("pad", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("peak_bytes", C.c_int64),
                ("slots", C.c_int32), ("peak_active", C.c_int32), ("elapsed", C.c_double)]


# name -> (restype, argtypes)
P = C.c_void_p
SIGNATURES = {
    "msdp_last_error": (C.c_char_p, []),
    "msdp_default_options": (None, [C.POINTER(Options)]),
    "msdp_solver_create": (C.c_int, [C.POINTER(P), C.c_int, C.c_int32, c_int64_p, C.c_int64, C.c_int64]),
    "msdp_solver_destroy": (C.c_int, [P]),
    "msdp_solver_set_cost_block": (C.c_int, [P, C.c_int32, c_double_p]),
    "msdp_solver_set_constraints_coo": (C.c_int, [P, C.c_int32, C.c_int64, c_int64_p, c_int64_p, c_int64_p, c_double_p]),
    "msdp_solver_set_constraints_dense": (C.c_int, [P, C.c_int32, C.c_int64, c_int64_p, c_double_p]),
    "msdp_solver_set_constraints_lowrank": (C.c_int, [P, C.c_int32, C.c_int64, c_int64_p, c_int64_p,
                                                      c_double_p, c_double_p]),
    "msdp_solver_set_rhs": (C.c_int, [P, c_double_p]),
    "msdp_solver_set_scalar_data": (C.c_int, [P, c_double_p, c_double_p]),
    "msdp_solver_configure": (C.c_int, [P, C.POINTER(Options)]),
    "msdp_solver_set_initial_point": (C.c_int, [P, c_double_p, c_double_p, c_double_p]),
    "msdp_solver_initialize": (C.c_int, [P]),
    "msdp_solver_iterate": (C.c_int, [P]),
    "msdp_solver_update_termination": (C.c_int, [P]),
    "msdp_solver_set_mu": (C.c_int, [P, C.c_double]),
    "msdp_solver_solve": (C.c_int, [P]),
    "msdp_solver_get_info": (C.c_int, [P, C.POINTER(Info)]),
    "msdp_solver_trace": (C.c_int, [P, C.c_int32, c_double_p, c_int32_p]),
    "msdp_solver_update_residuals": (C.c_int, [P]),
    "msdp_solver_prepare": (C.c_int, [P, c_int32_p]),
    "msdp_solver_factorize_schur": (C.c_int, [P, c_int32_p]),
    "msdp_solver_solve_directions": (C.c_int, [P, C.c_double, C.c_int32]),
    "msdp_solver_correction": (C.c_int, [P]),
    "msdp_solver_step_lengths": (C.c_int, [P, c_double_p, c_double_p]),
    "msdp_solver_affine_complementarity": (C.c_int, [P, C.c_double, C.c_double, c_double_p]),
    "msdp_solver_take_step": (C.c_int, [P, C.c_double, C.c_double]),
    "msdp_solver_schur_solve": (C.c_int, [P, c_double_p]),
    "msdp_solver_get": (C.c_int, [P, C.c_int32, C.c_int32, c_double_p]),
    "msdp_solver_set": (C.c_int, [P, C.c_int32, C.c_int32, c_double_p]),
    "msdp_solver_synchronize": (C.c_int, [P]),
    "msdp_solver_profile_begin": (C.c_int, [P, C.c_int32]),
    "msdp_solver_profile_end": (C.c_int, [P, c_double_p, c_int64_p, c_double_p]),
    "msdp_set_tuning": (C.c_int, [C.c_char_p, C.c_double]),
    "msdp_get_tuning": (C.c_int, [C.c_char_p, c_double_p]),
    "msdp_profile_begin": (C.c_int, [C.c_int32]),
    "msdp_profile_end": (C.c_int, [c_double_p, c_int64_p, c_double_p]),
    "msdp_solver_set_stream": (C.c_int, [P, P]),
    "msdp_solver_footprint": (C.c_int, [P, c_int64_p]),
    "msdp_solver_release": (C.c_int, [P]),
    "msdp_solver_upload": (C.c_int, [P]),
    "msdp_server_create": (C.c_int, [C.POINTER(P), C.c_int, C.c_int32, C.c_int64]),
    "msdp_server_start": (C.c_int, [P]),
    "msdp_server_submit": (C.c_int, [P, P, C.c_double, c_int64_p]),
    "msdp_server_wait_all": (C.c_int, [P]),
    "msdp_server_request": (C.c_int, [P, C.c_int64, C.POINTER(RequestInfo)]),
    "msdp_server_get_stats": (C.c_int, [P, C.POINTER(ServerStats)]),
    "msdp_server_destroy": (C.c_int, [P]),
    "msdp_k_gemm": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                              P, C.c_int64, P, C.c_int64, C.c_double, P, C.c_int64]),
    "msdp_k_gemm_tri": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_double, P, C.c_int64, P, C.c_int64, C.c_double, P, C.c_int64,
                                  C.c_int32, C.c_int32]),
    "msdp_k_gemm_noalloc": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_double, P, C.c_int64, P, C.c_int64, C.c_double, P,
                                      C.c_int64]),
    "msdp_k_gemm_noalloc_scratch_bytes": (C.c_size_t, []),
    "msdp_k_syrk_lower": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_double, P, C.c_int64,
                                    C.c_double, P, C.c_int64]),
    "msdp_k_potrf": (C.c_int, [P, C.c_int32, P, C.c_int64, P, P]),
    "msdp_k_potrf_work": (C.c_size_t, [C.c_int32]),
    "msdp_k_cholesky_factor": (C.c_int, [P, C.c_int32, P, C.c_int64, P, C.c_int64, P, P]),
    "msdp_k_cholesky_factor_scratch": (C.c_size_t, [C.c_int32]),
    "msdp_k_trtri": (C.c_int, [P, C.c_int32, P, C.c_int64, P, C.c_int64, P]),
    "msdp_k_trtri_work": (C.c_size_t, [C.c_int32]),
    "msdp_k_sytrf": (C.c_int, [P, C.c_int32, P, C.c_int64, P, P, P, P]),
    "msdp_k_sytrf_ctl_bytes": (C.c_size_t, []),
    "msdp_k_sytrs": (C.c_int, [P, C.c_int32, P, C.c_int64, P, P]),
    "msdp_k_syevd": (C.c_int, [P, C.c_int32, P, C.c_int64, P, P, C.c_int64, C.c_int32, P, P]),
    "msdp_k_syevd_work": (C.c_size_t, [C.c_int32]),
    "msdp_k_syevd_iwork": (C.c_size_t, [C.c_int32]),
    "msdp_k_eig_min": (C.c_int, [P, C.c_int32, P, C.c_int64, P, P]),
    "msdp_k_eig_min2": (C.c_int, [P, C.c_int32, P, P, C.c_int64, P, P, P]),
    "msdp_k_gemm32": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                P, C.c_int64, P, C.c_int64, C.c_double, P, C.c_int64]),
    "msdp_k_syrk_bf16x3": (C.c_int, [P, C.c_int32, C.c_int32, C.c_double, P, C.c_int64, C.c_double,
                                     P, C.c_int64, P]),
    "msdp_k_syrk_bf16x3_work_bytes": (C.c_size_t, [C.c_int32, C.c_int32]),
    "msdp_k_panel_apply": (C.c_int, [P, C.c_int32, C.c_int32, P, C.c_int64, P, P, C.c_int32, P]),
    "msdp_k_panel_apply_work": (C.c_size_t, [C.c_int32, C.c_int32]),
    "msdp_k_schur_onehot": (C.c_int, [P, C.c_int32, P, C.c_int64, P, C.c_int64, P, C.c_int64]),
}

_lib = None


def load():
    """Loads the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(the HIP library is required; there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    apply_tuning(lib)
    return lib


TUNING_PATH = Path(__file__).resolve().parent / "tuning_gfx950.json"


def apply_tuning(lib, path=None):
    """Applies the autotuner's table (madsdp_amd/autotune.py) if present;
    MADSDP_HIP_TUNING=0 keeps the compiled-in defaults."""
    import json
    path = Path(path or TUNING_PATH)
    if os.environ.get("MADSDP_HIP_TUNING", "1") == "0" or not path.exists():
        return {}
    params = json.loads(path.read_text())["params"]
    for name, value in params.items():
        rc = lib.msdp_set_tuning(name.encode(), float(value))
        if rc < 0:
            raise RuntimeError(f"{path}: {lib.msdp_last_error().decode()}")
    return params


class HipError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> int:
    if rc < 0:
        msg = load().msdp_last_error().decode(errors="replace")
        raise HipError(f"{what}: {msg}" if what else msg)
    return rc


def dptr(a):
    return a.ctypes.data_as(c_double_p)


def iptr(a):
    return a.ctypes.data_as(c_int64_p)
