"""ctypes wrapper of the CPU FP64 oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke()
and by bench.py's cpu_baseline leg, as the parity reference or the CPU baseline;
never by the library's product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
LIB_XPREC = os.path.join(HERE, "liborc_xprec.so")  # closures in long double
LIB_FMA = os.path.join(HERE, "liborc_fma.so")      # a*b+c contracted to FMA
VARIANTS = {"": LIB, "xprec": LIB_XPREC, "fma": LIB_FMA}
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

CONSTANT_NAMES = ("RHO_L", "RHO_I", "CP_L", "CP_I", "T_0", "T_FREEZE", "LH_V0", "LH_S0", "GRAV",
                  "K_THERM", "KERSTEN_ALPHA", "KERSTEN_BETA", "OMEGA", "GAMMA_VISC",
                  "GAMMA_T_REF", "LH_F0")


class Model(C.Structure):
    _fields_ = [("model", C.c_int32), ("bc_bottom", C.c_int32), ("ncol", C.c_int64),
                ("col_ptr", _ip), ("z_face", _dp), ("params", _dp), ("bc_top", C.c_int32),
                ("jacobian", C.c_int32)]


JAC_PICARD, JAC_NEWTON = 0, 1  # EnergyHydrology Newton matrix (clm_oracle.h)


BC_FLUX, BC_FREE_DRAINAGE, BC_MOISTURE_STATE, BC_TOPMODEL, BC_SURFACE_RUNOFF = 0, 1, 2, 3, 4
BC_ATMOS, BC_ATMOS_TOPMODEL = 5, 6
SURFACE_FIELDS = ("P_liq", "T_air", "lhf", "shf", "vapor_flux_liq", "vapor_flux_ice", "R_n")


class State(C.Structure):
    _fields_ = [("vartheta_l", _dp), ("theta_i", _dp), ("rho_e_int", _dp),
                ("intF_water", _dp), ("intF_energy", _dp)]


class Aux(C.Structure):
    _fields_ = [("K", _dp), ("psi", _dp), ("theta_l", _dp), ("T", _dp), ("kappa", _dp),
                ("top_water", _dp), ("top_heat", _dp), ("bot_water", _dp), ("bot_heat", _dp),
                ("sink", _dp), ("r_ss", _dp), ("sink_e", _dp), ("r_ess", _dp),
                ("vapor_ice", _dp)]


class Forcing(C.Structure):
    _fields_ = [("top_water", _dp), ("top_heat", _dp), ("bot_water", _dp), ("bot_heat", _dp)] + \
        [(k, _dp) for k in SURFACE_FIELDS]


_libs = {}


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr)


def lib(variant: str = ""):
    """The oracle ("") or one of its rounding probes: "xprec" (closures in
    long double) or "fma" (fused multiply-add contraction)."""
    path = VARIANTS[variant]
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        d = C.c_double
        for name, args in {
            "orc_effective_saturation": [d, d, d],
            "orc_volumetric_liquid_fraction": [d, d, d],
            "orc_matric_potential": [d, d, d],
            "orc_pressure_head": [d, d, d, d, d, d],
            "orc_dpsidtheta": [d, d, d, d, d, d],
            "orc_hydraulic_conductivity": [d, d, d],
            "orc_impedance_factor": [d, d],
            "orc_viscosity_factor": [d, d, d],
            "orc_volumetric_heat_capacity": [d, d, d],
            "orc_temperature_from_rho_e_int": [d, d, d],
            "orc_volumetric_internal_energy": [d, d, d],
            "orc_volumetric_internal_energy_liq": [d],
            "orc_kappa_sat": [d, d, d, d],
            "orc_thermal_conductivity": [d, d, d],
            "orc_relative_saturation": [d, d, d],
            "orc_kersten_number": [d, d, d, d, d, d, d],
            "orc_kappa_solid": [d, d, d, d, d],
            "orc_kappa_sat_frozen": [d, d, d],
            "orc_kappa_sat_unfrozen": [d, d, d],
            "orc_kappa_dry": [d, d, d, d],
            "orc_soil_Tf_depressed": [d, d, d, d, d, d],
            "orc_thermal_time": [d, d, d],
            "orc_heaviside": [d],
            "orc_vg_sc": [d],
            "orc_dry_soil_layer_thickness": [d, d, d, d],
            "orc_soil_conductance": [d, d, d, d, d, d],
        }.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = d
        L.orc_inverse_matric_potential.argtypes = [d, d, d, C.POINTER(C.c_int)]
        L.orc_inverse_matric_potential.restype = d
        L.orc_phase_change_source.argtypes = [d] * 8 + [C.POINTER(C.c_int)]
        L.orc_phase_change_source.restype = d
        L.orc_constants.argtypes = [_dp, C.c_int]
        M, S, A, F = (C.POINTER(t) for t in (Model, State, Aux, Forcing))
        L.orc_update_cache.argtypes = [M, S, F, A]
        L.orc_update_implicit_cache.argtypes = [M, S, F, A]
        L.orc_exp_tendency.argtypes = [M, S, A, S]
        L.orc_imp_tendency.argtypes = [M, S, A, S]
        L.orc_jacobian.argtypes = [M, S, A, d, _dp, _dp, _dp]
        L.orc_solve.argtypes = [M, _dp, _dp, _dp, S, S]
        L.orc_step.argtypes = [M, S, F, d, C.c_int, d, _ip]
        L.orc_step.restype = C.c_int64
        L.orc_step_guess.argtypes = [M, S, F, d, C.c_int, d, _ip, _dp, _dp]
        L.orc_step_guess.restype = C.c_int64
        L.orc_run.argtypes = [M, S, _dp, _dp, C.c_int, C.c_int, d, C.c_int, C.c_int, d, _ip,
                              C.c_int64, C.c_int]
        L.orc_run.restype = C.c_int64
        L.orc_totals.argtypes = [M, S, _dp, _dp]
        L.orc_surface_soil_conductance.argtypes = [M, S, d, d, d, d, _dp, _dp]
        _libs[path] = L
    return _libs[path]


def constants() -> dict:
    out = (C.c_double * len(CONSTANT_NAMES))()
    lib().orc_constants(out, len(CONSTANT_NAMES))
    return dict(zip(CONSTANT_NAMES, list(out)))


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class OracleModel:
    """Oracle view of a CSR column batch (same arrays the library takes)."""

    def __init__(self, model: int, bc_bottom: int, col_ptr, z_face, params, variant="",
                 bc_top: int = BC_FLUX, jacobian: int = JAC_PICARD):
        self.L = lib(variant)
        self.col_ptr = np.ascontiguousarray(col_ptr, dtype=np.int32)
        self.z_face = np.ascontiguousarray(z_face, dtype=np.float64)
        self.params = np.ascontiguousarray(params, dtype=np.float64)
        self.model = model
        self.eh = model == 1
        self.ncol = self.col_ptr.size - 1
        self.ncell = int(self.col_ptr[-1])
        self._m = Model(model, bc_bottom, self.ncol, self.col_ptr.ctypes.data_as(_ip),
                        _p(self.z_face), _p(self.params), bc_top, jacobian)

    @classmethod
    def from_batch(cls, b, variant="", jacobian: int = JAC_PICARD):
        return cls(b.model, b.bc_bottom, b.col_ptr, b.z_face, b.params, variant=variant,
                   bc_top=getattr(b, "bc_top", BC_FLUX), jacobian=jacobian)

    def new_state(self, src: dict | None = None) -> dict:
        st = {"vartheta_l": np.zeros(self.ncell), "intF_water": np.zeros(self.ncol)}
        if self.eh:
            st.update(theta_i=np.zeros(self.ncell), rho_e_int=np.zeros(self.ncell),
                      intF_energy=np.zeros(self.ncol))
        if src:
            for k in st:
                st[k][:] = src[k]
        return st

    def _s(self, st: dict) -> State:
        return State(*(_p(st.get(k)) for k in ("vartheta_l", "theta_i", "rho_e_int",
                                               "intF_water", "intF_energy")))

    def new_aux(self) -> dict:
        a = {k: np.zeros(self.ncell) for k in ("K", "psi", "theta_l", "T", "kappa", "sink",
                                               "sink_e")}
        a.update({k: np.zeros(self.ncol) for k in ("top_water", "top_heat", "bot_water",
                                                   "bot_heat", "r_ss", "r_ess", "vapor_ice")})
        return a

    def _a(self, a: dict) -> Aux:
        return Aux(*(_p(a.get(k)) for k in ("K", "psi", "theta_l", "T", "kappa", "top_water",
                                            "top_heat", "bot_water", "bot_heat", "sink",
                                            "r_ss", "sink_e", "r_ess", "vapor_ice")))

    @staticmethod
    def _f(tw, th=None, bw=None, bh=None, sfc=None) -> Forcing:
        """sfc: dict of SURFACE_FIELDS arrays (AtmosDrivenFluxBC inputs) or None."""
        s = sfc or {}
        return Forcing(_p(tw), _p(th), _p(bw), _p(bh),
                       *(_p(None if s.get(k) is None else
                            np.ascontiguousarray(s[k], dtype=np.float64)) for k in SURFACE_FIELDS))

    def update_cache(self, Y: dict, aux: dict, tw, th=None, bw=None, bh=None, sfc=None):
        # contiguous float64 copies held here, so that the pointers stay valid
        sfc = None if sfc is None else {k: np.ascontiguousarray(v, dtype=np.float64)
                                        for k, v in sfc.items()}
        f = self._f(tw, th, bw, bh, sfc)
        self.L.orc_update_cache(C.byref(self._m), C.byref(self._s(Y)), C.byref(f),
                               C.byref(self._a(aux)))

    def update_implicit_cache(self, Y: dict, aux: dict, tw=None, bw=None):
        """cache_imp!; under a top MoistureStateBC also the boundary fluxes,
        from the boundary states tw (top) and bw (bottom)."""
        f = self._f(tw, None, bw, None)
        self.L.orc_update_implicit_cache(C.byref(self._m), C.byref(self._s(Y)),
                                        C.byref(f) if tw is not None else None,
                                        C.byref(self._a(aux)))

    def exp_tendency(self, Y: dict, aux: dict) -> dict:
        dY = self.new_state()
        r = self.L.orc_exp_tendency(C.byref(self._m), C.byref(self._s(Y)), C.byref(self._a(aux)),
                                   C.byref(self._s(dY)))
        if r:
            raise ValueError("Matric potential is positive")
        return dY

    def imp_tendency(self, Y: dict, aux: dict) -> dict:
        dY = self.new_state()
        self.L.orc_imp_tendency(C.byref(self._m), C.byref(self._s(Y)), C.byref(self._a(aux)),
                               C.byref(self._s(dY)))
        return dY

    def jacobian(self, Y: dict, aux: dict, dtg: float):
        ww = np.zeros(3 * self.ncell)
        ee = np.zeros(3 * self.ncell) if self.eh else None
        ew = np.zeros(3 * self.ncell) if self.eh else None
        self.L.orc_jacobian(C.byref(self._m), C.byref(self._s(Y)), C.byref(self._a(aux)), dtg,
                           _p(ww), _p(ee), _p(ew))
        return ww, ee, ew

    def solve(self, W, R: dict) -> dict:
        ww, ee, ew = W
        dx = self.new_state()
        self.L.orc_solve(C.byref(self._m), _p(ww), _p(ee), _p(ew), C.byref(self._s(R)),
                        C.byref(self._s(dx)))
        return dx

    def step(self, Y: dict, dt: float, max_iters: int, tol: float, tw, th=None, bw=None,
             bh=None, guess_v=None, guess_e=None, sfc=None):
        """One ARS111 step in place; guess_v / guess_e (per cell) warm-start
        Newton at x0 = U0 + guess; sfc: AtmosDrivenFluxBC inputs."""
        it = np.zeros(self.ncol, dtype=np.int32)
        sfc = None if sfc is None else {k: np.ascontiguousarray(v, dtype=np.float64)
                                        for k, v in sfc.items()}
        f = self._f(tw, th, bw, bh, sfc)
        nc = self.L.orc_step_guess(C.byref(self._m), C.byref(self._s(Y)), C.byref(f), dt,
                                  max_iters, tol, it.ctypes.data_as(_ip), _p(guess_v),
                                  _p(guess_e))
        if nc < 0:
            raise ValueError("Matric potential is positive")
        return it, int(nc)

    def run(self, Y: dict, top_water, top_heat, steps_per_slot: int, dt: float, nsteps: int,
            max_iters: int, tol: float, ncol_limit: int = 0, nthreads: int = 0,
            want_iters: bool = True):
        tw = np.ascontiguousarray(top_water, dtype=np.float64)
        th = None if top_heat is None else np.ascontiguousarray(top_heat, dtype=np.float64)
        it = np.zeros((nsteps, self.ncol), dtype=np.int32) if want_iters else None
        nc = self.L.orc_run(C.byref(self._m), C.byref(self._s(Y)), _p(tw), _p(th), tw.shape[0],
                           steps_per_slot, dt, nsteps, max_iters, tol,
                           it.ctypes.data_as(_ip) if it is not None else None, ncol_limit,
                           nthreads)
        if nc < 0:
            raise ValueError("Matric potential is positive")
        return it, int(nc)

    def surface_soil_conductance(self, Y: dict, d_ds: float, evap_p: float, evap_alpha: float,
                                 D_vapor: float):
        g = np.zeros(self.ncol)
        dsl = np.zeros(self.ncol)
        r = self.L.orc_surface_soil_conductance(C.byref(self._m), C.byref(self._s(Y)), d_ds,
                                                evap_p, evap_alpha, D_vapor, _p(g), _p(dsl))
        if r:
            raise ValueError("surface soil conductance is defined for EnergyHydrology")
        return g, dsl

    def totals(self, Y: dict):
        w = np.zeros(self.ncol)
        e = np.zeros(self.ncol) if self.eh else None
        self.L.orc_totals(C.byref(self._m), C.byref(self._s(Y)), _p(w), _p(e))
        return w, e
