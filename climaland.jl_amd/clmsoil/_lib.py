"""ctypes binding of libclmsoil.so (include/clm_soil.h).

The HIP library is the only compute path: if it is missing or fails to load
this module raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("CLMSOIL_LIB", os.path.join(PKG_ROOT, "libclmsoil.so"))

# mirrors of the enums in include/clm_soil.h
CLM_OK, CLM_E_ARG, CLM_E_HIP, CLM_E_DOMAIN, CLM_E_STATE = 0, -1, -2, -3, -4
MODEL_RICHARDS, MODEL_ENERGY_HYDROLOGY = 0, 1
BC_FLUX, BC_FREE_DRAINAGE, BC_MOISTURE_STATE, BC_TOPMODEL, BC_SURFACE_RUNOFF = 0, 1, 2, 3, 4
BC_ATMOS, BC_ATMOS_TOPMODEL = 5, 6
SOLVER_AUTO, SOLVER_GROUP, SOLVER_LANE = 0, 1, 2
JACOBIAN_PICARD, JACOBIAN_NEWTON = 0, 1
(P_NU, P_THETA_R, P_VG_ALPHA, P_VG_N, P_KSAT, P_SS, P_NU_SS_OM, P_NU_SS_QUARTZ,
 P_NU_SS_GRAVEL, P_KAPPA_DRY, P_KAPPA_SAT_FROZEN, P_KAPPA_SAT_UNFROZEN, P_RHOC_DS) = range(13)
NPARAM_RRE, NPARAM_EH = 6, 13
P_F_MAX, P_F_OVER, P_R_SB = 6, 7, 8
NPARAM_RRE_TOPMODEL = 9
P_EH_F_MAX, P_EH_F_OVER, P_EH_R_SB = 13, 14, 15
NPARAM_EH_TOPMODEL = 16
SURFACE_FIELDS = ("P_liq", "T_air", "lhf", "shf", "vapor_flux_liq", "vapor_flux_ice", "R_n")

EXPORTED = (
    "clm_abi_version", "clm_create", "clm_destroy", "clm_last_error", "clm_set_state",
    "clm_get_state", "clm_set_forcing", "clm_set_forcing_table", "clm_update_cache",
    "clm_exp_tendency", "clm_imp_tendency", "clm_jacobian", "clm_solve",
    "clm_update_implicit_cache", "clm_get_aux", "clm_get_jacobian", "clm_step", "clm_run",
    "clm_step_noalloc", "clm_totals", "clm_last_kernel_ms", "clm_get_shapes", "clm_set_shape", "clm_autotune",
    "clm_set_pairing", "clm_get_pairing",
    "clm_ws_init", "clm_ws_dims", "clm_ws_get_params", "clm_ws_set_params", "clm_ws_enable",
    "clm_ws_collect", "clm_ws_train", "clm_ws_features", "clm_ws_forward", "clm_ws_gradient",
    "clm_ws_time", "clm_ws_guess", "clm_save_checkpoint", "clm_read_checkpoint",
    "clm_set_surface_fluxes", "clm_set_surface_flux_table", "clm_soil_conductance",
)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)


class Config(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("bc_bottom", C.c_int32), ("ncol", C.c_int64),
        ("col_ptr", _ip), ("z_face", _dp), ("params", _dp), ("nparam", C.c_int32),
        ("jac_precision", C.c_int32), ("newton_max_iters", C.c_int32),
        ("newton_tol", C.c_double), ("solver", C.c_int32), ("bc_top", C.c_int32),
        ("jacobian", C.c_int32), ("reserved_", C.c_int32),
    ]


class State(C.Structure):
    _fields_ = [("vartheta_l", _dp), ("theta_i", _dp), ("rho_e_int", _dp),
                ("intF_water", _dp), ("intF_energy", _dp)]


class Aux(C.Structure):
    _fields_ = [("K", _dp), ("psi", _dp), ("theta_l", _dp), ("T", _dp), ("kappa", _dp),
                ("top_water", _dp), ("top_heat", _dp), ("bottom_water", _dp),
                ("bottom_heat", _dp), ("sink", _dp), ("r_ss", _dp), ("sink_energy", _dp),
                ("r_ess", _dp)]


class SurfaceFluxes(C.Structure):
    _fields_ = [(k, _dp) for k in SURFACE_FIELDS]


class StepStats(C.Structure):
    _fields_ = [("max_iters_used", C.c_int32), ("n_stalled", C.c_int32), ("mean_iters", C.c_double),
                ("n_nonconverged", C.c_int64), ("n_nan", C.c_int64),
                ("n_domain_error", C.c_int64), ("ms", C.c_double)]


class ClmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"clm error {code}: {msg}")
        self.code = code


_LIB = None


def load(path: str | None = None) -> C.CDLL:
    """Load libclmsoil.so (once).  Raises if it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"libclmsoil.so not found at {p}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(p)
    vp = C.c_void_p
    sig = {
        "clm_abi_version": ([], C.c_int),
        "clm_create": ([C.POINTER(vp), C.POINTER(Config), vp], C.c_int),
        "clm_destroy": ([vp], C.c_int),
        "clm_last_error": ([vp], C.c_char_p),
        "clm_set_state": ([vp, C.POINTER(State)], C.c_int),
        "clm_get_state": ([vp, C.POINTER(State)], C.c_int),
        "clm_set_forcing": ([vp, _dp, _dp, _dp, _dp], C.c_int),
        "clm_set_forcing_table": ([vp, C.c_int32, _dp, _dp], C.c_int),
        "clm_update_cache": ([vp, C.POINTER(State), C.c_double], C.c_int),
        "clm_exp_tendency": ([vp, C.POINTER(State), C.c_double, C.POINTER(State)], C.c_int),
        "clm_imp_tendency": ([vp, C.POINTER(State), C.c_double, C.POINTER(State)], C.c_int),
        "clm_jacobian": ([vp, C.POINTER(State), C.c_double, C.c_double], C.c_int),
        "clm_solve": ([vp, C.POINTER(State), C.POINTER(State)], C.c_int),
        "clm_update_implicit_cache": ([vp, C.POINTER(State), C.c_double], C.c_int),
        "clm_get_aux": ([vp, C.POINTER(Aux)], C.c_int),
        "clm_get_jacobian": ([vp, _dp, _dp, _dp], C.c_int),
        "clm_step": ([vp, C.c_double, C.c_double, C.POINTER(StepStats)], C.c_int),
        "clm_step_noalloc": ([vp, C.c_double, C.c_double, C.POINTER(StepStats)], C.c_int),
        "clm_set_pairing": ([vp, C.c_int32], C.c_int),
        "clm_get_pairing": ([vp, _ip], C.c_int),
        "clm_run": ([vp, C.c_double, C.c_double, C.c_int32, C.c_int32, C.POINTER(StepStats),
                     _ip], C.c_int),
        "clm_totals": ([vp, _dp, _dp], C.c_int),
        "clm_soil_conductance": ([vp, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp],
                                 C.c_int),
        "clm_last_kernel_ms": ([vp, _dp, _ip], C.c_int),
        "clm_get_shapes": ([vp, _ip, C.c_int32, _ip], C.c_int),
        "clm_set_shape": ([vp, C.c_int32, C.c_int32, C.c_int32], C.c_int),
        "clm_autotune": ([vp, C.c_double, C.c_int32], C.c_int),
        "clm_ws_init": ([vp, C.c_uint64], C.c_int),
        "clm_ws_dims": ([vp, _ip, C.POINTER(C.c_int64)], C.c_int),
        "clm_ws_get_params": ([vp, _fp], C.c_int),
        "clm_ws_set_params": ([vp, _fp], C.c_int),
        "clm_ws_enable": ([vp, C.c_int32], C.c_int),
        "clm_ws_collect": ([vp, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                            C.POINTER(C.c_int64)], C.c_int),
        "clm_ws_train": ([vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _fp], C.c_int),
        "clm_ws_features": ([vp, C.c_double, _fp], C.c_int),
        "clm_ws_forward": ([vp, _fp, C.c_int64, _fp], C.c_int),
        "clm_ws_gradient": ([vp, _fp, _fp, _fp, C.c_int64, _fp, _fp], C.c_int),
        "clm_ws_time": ([vp, C.c_double, C.c_int32, C.POINTER(C.c_double)], C.c_int),
        "clm_ws_guess": ([vp, C.c_double, C.c_int32, _dp, _dp], C.c_int),
        "clm_set_surface_fluxes": ([vp, C.POINTER(SurfaceFluxes)], C.c_int),
        "clm_set_surface_flux_table": ([vp, C.c_int32, C.POINTER(SurfaceFluxes)], C.c_int),
        "clm_save_checkpoint": ([vp, C.c_char_p, C.c_double, C.c_char_p, C.c_int32], C.c_int),
        "clm_read_checkpoint": ([vp, C.c_char_p, C.POINTER(C.c_double), _ip], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _LIB = lib
    return lib


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "float64 C-contiguous required"
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "float32 C-contiguous required"
    return a.ctypes.data_as(_fp)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def check(lib, ctx, code: int):
    if code != CLM_OK:
        msg = lib.clm_last_error(ctx)
        raise ClmError(code, msg.decode() if msg else "")
