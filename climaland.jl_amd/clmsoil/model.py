"""Host-side mirror of ClimaLand's soil model interface over libclmsoil.so.

Names, argument meaning and call order follow the reference (ClimaLand.jl
v1.11.2) so code written against it reads the same:

  vanGenuchten(α=, n=)                      retention_models.jl:30-44
  RichardsParameters(ν=, hydrology_cm=, K_sat=, S_s=, θ_r=)   rre.jl:7-48
  EnergyHydrologyParameters(...)            energy_hydrology.jl:21-207
  WaterFluxBC, FreeDrainage, HeatFluxBC, WaterHeatBC, EnergyWaterFreeDrainage
                                            boundary_conditions.jl:45-60,452,568,588
  RichardsModel(parameters=, domain=, boundary_conditions=)   rre.jl:92-109
  EnergyHydrology(parameters=, domain=, boundary_conditions=) energy_hydrology.jl:251-271
  make_update_cache, make_exp_tendency, make_compute_imp_tendency,
  make_compute_jacobian, make_update_implicit_cache   models.jl:187-287; rre.jl; energy_hydrology.jl
  initialize_jacobian(model) -> jacobian with .ldiv(dx, R)  implicit_timestepping.jl:63-172
  LandSimulation(model, Y0, Δt, ...).step() / .solve()       Simulations.jl:115-332

Prognostic state Y is a dict of NumPy arrays in CSR layout (level 0 = the
bottom cell of a column): vartheta_l, intF_water and, for EnergyHydrology,
theta_i, rho_e_int, intF_energy.  The cache p lives on the device inside the
model's context; `model.aux()` copies it out.  Every computation runs in the
HIP library; errors come back as ClmError (the reference raises).
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib as L
from .inputs import KAPPA_ICE, KAPPA_LIQ, kappa_solid, KAPPA_OM, KAPPA_QUARTZ, KAPPA_MINERALS

# ---- parameters -------------------------------------------------------------


@dataclasses.dataclass
class vanGenuchten:  # noqa: N801 (reference name)
    α: np.ndarray
    n: np.ndarray

    @property
    def m(self):
        return 1.0 - 1.0 / np.asarray(self.n)

    @property
    def S_c(self):
        n = np.asarray(self.n)
        return (1.0 + ((n - 1.0) / n) ** (1.0 - 2.0 * n)) ** (-self.m)


@dataclasses.dataclass
class RichardsParameters:
    ν: np.ndarray
    hydrology_cm: vanGenuchten
    K_sat: np.ndarray
    S_s: np.ndarray
    θ_r: np.ndarray


@dataclasses.dataclass
class EnergyHydrologyParameters:
    ν: np.ndarray
    hydrology_cm: vanGenuchten
    K_sat: np.ndarray
    S_s: np.ndarray
    θ_r: np.ndarray
    ν_ss_om: np.ndarray
    ν_ss_quartz: np.ndarray
    ν_ss_gravel: np.ndarray
    ρc_ds: np.ndarray
    κ_dry: np.ndarray
    κ_sat_frozen: np.ndarray
    κ_sat_unfrozen: np.ndarray

    @classmethod
    def derived(cls, *, ν, hydrology_cm, K_sat, S_s, θ_r, ν_ss_om, ν_ss_quartz, ν_ss_gravel,
                ρc_ds, κ_dry):
        """κ_sat_{frozen,unfrozen} from the composition (energy_hydrology.jl:168-172)."""
        ks = kappa_solid(np.asarray(ν_ss_om), np.asarray(ν_ss_quartz))
        ν = np.asarray(ν)
        return cls(ν, hydrology_cm, K_sat, S_s, θ_r, ν_ss_om, ν_ss_quartz, ν_ss_gravel, ρc_ds,
                   κ_dry, ks ** (1.0 - ν) * KAPPA_ICE ** ν, ks ** (1.0 - ν) * KAPPA_LIQ ** ν)


# ---- domain and boundary conditions ------------------------------------------


@dataclasses.dataclass
class ColumnBatchDomain:
    """Many independent Columns (Domains.jl:137-247), CSR-packed."""
    col_ptr: np.ndarray  # int32 [ncol+1]
    z_face: np.ndarray   # float64 [ncell + ncol], bottom -> top per column

    @property
    def ncol(self):
        return self.col_ptr.size - 1

    @property
    def ncell(self):
        return int(self.col_ptr[-1])


@dataclasses.dataclass
class WaterFluxBC:
    """Prescribed water flux bc(p, t) per column (upward positive)."""
    bc: np.ndarray


@dataclasses.dataclass
class HeatFluxBC:
    bc: np.ndarray


@dataclasses.dataclass
class MoistureStateBC:
    """Prescribed ϑ_l at the boundary, bc(p, t) per column
    (boundary_conditions.jl:35-37, fluxes :227-325).  RichardsModel only."""
    bc: np.ndarray


class FreeDrainage:
    pass


class NoRunoff:
    """Runoff.jl:52-75: infiltration = precipitation."""


class SurfaceRunoff:
    """Runoff.jl:84-148: no infiltration into a saturated top layer, else
    max(-K_sat, P)."""


@dataclasses.dataclass
class TOPMODELRunoff:
    """Runoff.jl:199-271 (Niu et al. 2005): surface infiltration from the
    saturated fraction and the TOPMODELSubsurfaceRunoff sink (R_sb, f_over).
    Each field is a scalar or one value per column."""
    f_over: float | np.ndarray
    f_max: float | np.ndarray
    R_sb: float | np.ndarray


@dataclasses.dataclass
class SoilSurfaceFluxes:
    """What AtmosDrivenFluxBC's drivers give the soil-only boundary per column
    (boundary_conditions.jl:880-915): p.drivers.P_liq (m/s, negative down),
    p.drivers.T (K), and the outputs of turbulent_fluxes! (lhf, shf in W/m2;
    vapor_flux_liq, vapor_flux_ice in m/s) and net_radiation! (R_n, W/m2).
    SurfaceFluxes.jl's Monin-Obukhov solve that produces them stays with the
    caller.  Each field is a scalar or one value per column."""
    P_liq: float | np.ndarray
    T_air: float | np.ndarray
    lhf: float | np.ndarray = 0.0
    shf: float | np.ndarray = 0.0
    vapor_flux_liq: float | np.ndarray = 0.0
    vapor_flux_ice: float | np.ndarray = 0.0
    R_n: float | np.ndarray = 0.0


@dataclasses.dataclass
class AtmosDrivenFluxBC:
    """EnergyHydrology's top boundary driven by the atmosphere
    (boundary_conditions.jl:646-1003) in soil-only form, with NoRunoff or
    TOPMODELRunoff; adds the SoilSublimation source (energy_hydrology.jl:909-
    945).  `surface` holds the driver and turbulent-flux values at t_n."""
    surface: SoilSurfaceFluxes
    runoff: object = dataclasses.field(default_factory=NoRunoff)


@dataclasses.dataclass
class RichardsAtmosDrivenFluxBC:
    """boundary_conditions.jl:79-212: the top flux of a RichardsModel driven
    by the liquid precipitation rate `precip` (per column, m/s, negative
    downward) through a runoff model."""
    precip: np.ndarray
    runoff: object = dataclasses.field(default_factory=NoRunoff)


class EnergyWaterFreeDrainage:
    pass


@dataclasses.dataclass
class WaterHeatBC:
    water: WaterFluxBC
    heat: HeatFluxBC


# ---- models ------------------------------------------------------------------


class _SoilModel:
    model_id = -1

    def __init__(self, *, parameters, domain: ColumnBatchDomain, boundary_conditions: dict,
                 sources: tuple = (), jac_precision: int = 64, newton_max_iters: int = 3,
                 newton_tol: float = 0.0, solver: int = L.SOLVER_AUTO,
                 jacobian: int = L.JACOBIAN_PICARD, stream: int | None = None):
        """stream: a HIP stream handle (an integer, e.g. torch's
        `Stream.cuda_stream`) for the context's work, which the caller keeps
        alive until close(); None: the library's own stream."""
        self.parameters = parameters
        self.domain = domain
        self.boundary_conditions = boundary_conditions
        self.sources = sources
        self._lib = L.load()
        self._ctx = C.c_void_p()
        self._col_ptr = np.ascontiguousarray(domain.col_ptr, dtype=np.int32)
        self._z_face = np.ascontiguousarray(domain.z_face, dtype=np.float64)
        self._params = np.ascontiguousarray(self._param_rows(), dtype=np.float64)
        # the library reads ncol+1 offsets, ncell+ncol faces and nparam x ncol
        # parameters through plain pointers
        ncol = int(domain.ncol)
        if (self._col_ptr.size != ncol + 1 or self._col_ptr[0] != 0
                or np.any(np.diff(self._col_ptr) < 1)):
            raise ValueError("col_ptr must be ncol+1 increasing offsets starting at 0")
        if self._z_face.size != int(self._col_ptr[-1]) + ncol:
            raise ValueError(f"z_face must hold ncell + ncol = {int(self._col_ptr[-1]) + ncol} "
                             f"values, got {self._z_face.size}")
        if self._params.ndim != 2 or self._params.shape[1] != ncol:
            raise ValueError(f"parameter rows must be [nparam][{ncol}], got {self._params.shape}")
        cfg = L.Config(self.model_id, self._bottom_kind(), domain.ncol,
                       L.iptr(self._col_ptr), L.dptr(self._z_face), L.dptr(self._params),
                       self._params.shape[0], jac_precision, newton_max_iters, newton_tol,
                       solver, self._top_kind(), jacobian, 0)
        L.check(self._lib, None, self._lib.clm_create(C.byref(self._ctx), C.byref(cfg),
                                                      C.c_void_p(stream) if stream else None))
        self.jac_precision = jac_precision
        self.jacobian = jacobian
        self.set_forcing()

    # -- helpers
    def _check(self, code):
        L.check(self._lib, self._ctx, code)

    def close(self):
        if self._ctx:
            self._lib.clm_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def is_eh(self):
        return self.model_id == L.MODEL_ENERGY_HYDROLOGY

    def new_state(self) -> dict:
        d = self.domain
        st = {"vartheta_l": np.zeros(d.ncell), "intF_water": np.zeros(d.ncol)}
        if self.is_eh:
            st.update(theta_i=np.zeros(d.ncell), rho_e_int=np.zeros(d.ncell),
                      intF_energy=np.zeros(d.ncol))
        return st

    def _state(self, Y: dict) -> L.State:
        # the library reads / writes ncell or ncol values through each pointer
        d = self.domain
        for k in ("vartheta_l", "theta_i", "rho_e_int", "intF_water", "intF_energy"):
            a = Y.get(k)
            n = d.ncol if k.startswith("intF") else d.ncell
            if a is not None and np.asarray(a).size != n:
                raise ValueError(f"state field {k} must hold {n} values, got {np.asarray(a).size}")
        return L.State(*(L.dptr(Y.get(k)) for k in ("vartheta_l", "theta_i", "rho_e_int",
                                                    "intF_water", "intF_energy")))

    def _top_arrays(self):
        top = self.boundary_conditions["top"]
        if isinstance(top, WaterHeatBC):
            return np.asarray(top.water.bc, float), np.asarray(top.heat.bc, float)
        if isinstance(top, RichardsAtmosDrivenFluxBC):
            return np.asarray(top.precip, float), None
        return np.asarray(top.bc, float), None

    def _bottom_kind(self):
        b = self.boundary_conditions["bottom"]
        if isinstance(b, MoistureStateBC):
            return L.BC_MOISTURE_STATE
        return L.BC_FREE_DRAINAGE if isinstance(b, (FreeDrainage, EnergyWaterFreeDrainage)) \
            else L.BC_FLUX

    def _top_kind(self):
        top = self.boundary_conditions["top"]
        if isinstance(top, AtmosDrivenFluxBC):
            return L.BC_ATMOS_TOPMODEL if isinstance(top.runoff, TOPMODELRunoff) else L.BC_ATMOS
        if isinstance(top, RichardsAtmosDrivenFluxBC):
            if isinstance(top.runoff, TOPMODELRunoff):
                return L.BC_TOPMODEL
            if isinstance(top.runoff, SurfaceRunoff):
                return L.BC_SURFACE_RUNOFF
            return L.BC_FLUX  # NoRunoff: the precipitation is the flux
        return L.BC_MOISTURE_STATE if isinstance(top, MoistureStateBC) else L.BC_FLUX

    def _surface_struct(self, sfc: SoilSurfaceFluxes, shape):
        arrs = [np.ascontiguousarray(np.broadcast_to(getattr(sfc, k), shape), dtype=np.float64)
                for k in L.SURFACE_FIELDS]
        return L.SurfaceFluxes(*(L.dptr(a) for a in arrs)), arrs

    def set_forcing(self):
        """Push the current bc(p, t) values of the boundary conditions."""
        top = self.boundary_conditions["top"]
        if isinstance(top, AtmosDrivenFluxBC):
            st, keep = self._surface_struct(top.surface, (self.domain.ncol,))
            self._check(self._lib.clm_set_surface_fluxes(self._ctx, C.byref(st)))
            bw = bh = None
            b = self.boundary_conditions["bottom"]
            if isinstance(b, WaterHeatBC):
                bw = np.ascontiguousarray(np.broadcast_to(b.water.bc, (self.domain.ncol,)), float)
                bh = np.ascontiguousarray(np.broadcast_to(b.heat.bc, (self.domain.ncol,)), float)
            self._check(self._lib.clm_set_forcing(self._ctx, None, None, L.dptr(bw),
                                                  L.dptr(bh)))
            return
        tw, th = self._top_arrays()
        n = self.domain.ncol
        tw = np.ascontiguousarray(np.broadcast_to(tw, (n,)), dtype=np.float64)
        th = None if th is None else np.ascontiguousarray(np.broadcast_to(th, (n,)), float)
        bw = bh = None
        b = self.boundary_conditions["bottom"]
        if isinstance(b, WaterHeatBC):
            bw = np.ascontiguousarray(np.broadcast_to(b.water.bc, (n,)), float)
            bh = np.ascontiguousarray(np.broadcast_to(b.heat.bc, (n,)), float)
        elif isinstance(b, (WaterFluxBC, MoistureStateBC)):
            bw = np.ascontiguousarray(np.broadcast_to(b.bc, (n,)), float)
        self._check(self._lib.clm_set_forcing(self._ctx, L.dptr(tw), L.dptr(th), L.dptr(bw),
                                              L.dptr(bh)))

    def aux(self) -> dict:
        """Copy of the cache p.soil (K, ψ, [θ_l, T, κ], boundary fluxes,
        [TOPMODEL: subsurface sink per cell and R_ss])."""
        d = self.domain
        a = {"K": np.zeros(d.ncell), "psi": np.zeros(d.ncell), "top_water": np.zeros(d.ncol),
             "bottom_water": np.zeros(d.ncol)}
        if self.is_eh:
            a.update(theta_l=np.zeros(d.ncell), T=np.zeros(d.ncell), kappa=np.zeros(d.ncell),
                     top_heat=np.zeros(d.ncol), bottom_heat=np.zeros(d.ncol))
        if self._top_kind() in (L.BC_TOPMODEL, L.BC_ATMOS_TOPMODEL):
            a.update(sink=np.zeros(d.ncell), r_ss=np.zeros(d.ncol))
        if self._top_kind() == L.BC_ATMOS_TOPMODEL:
            a.update(sink_energy=np.zeros(d.ncell), r_ess=np.zeros(d.ncol))
        s = L.Aux(*(L.dptr(a.get(k)) for k in ("K", "psi", "theta_l", "T", "kappa", "top_water",
                                               "top_heat", "bottom_water", "bottom_heat",
                                               "sink", "r_ss", "sink_energy", "r_ess")))
        self._check(self._lib.clm_get_aux(self._ctx, C.byref(s)))
        return a

    # -- group-solver shapes and autotuner
    def shapes(self):
        """[(bucket max layers, G lanes/column, R rows/lane, columns)]."""
        n = C.c_int32(0)
        self._check(self._lib.clm_get_shapes(self._ctx, None, 0, C.byref(n)))
        rec = np.zeros(4 * max(n.value, 1), dtype=np.int32)
        self._check(self._lib.clm_get_shapes(self._ctx, L.iptr(rec), n.value, C.byref(n)))
        return [tuple(int(v) for v in rec[4 * i:4 * i + 4]) for i in range(n.value)]

    def set_shape(self, bucket_max_layers: int, G: int, R: int):
        self._check(self._lib.clm_set_shape(self._ctx, bucket_max_layers, G, R))

    def set_pairing(self, on: bool):
        """Paired waves for the <= 40 and <= 24 layer buckets (on by default)."""
        self._check(self._lib.clm_set_pairing(self._ctx, 1 if on else 0))

    def pairing_active(self) -> bool:
        a = C.c_int32(0)
        self._check(self._lib.clm_get_pairing(self._ctx, C.byref(a)))
        return bool(a.value)

    def autotune(self, dt: float, reps: int = 2):
        """Time every fitting shape per depth bucket on one step from the
        current device state (pushed by LandSimulation) and keep the fastest."""
        self._check(self._lib.clm_autotune(self._ctx, dt, reps))

    def get_jacobian(self):
        n = 3 * self.domain.ncell
        ww = np.zeros(n)
        ee = np.zeros(n) if self.is_eh else None
        ew = np.zeros(n) if self.is_eh else None
        self._check(self._lib.clm_get_jacobian(self._ctx, L.dptr(ww), L.dptr(ee), L.dptr(ew)))
        return ww, ee, ew


class RichardsModel(_SoilModel):
    model_id = L.MODEL_RICHARDS

    def _param_rows(self):
        p = self.parameters
        n = self.domain.ncol
        rows = [p.ν, p.θ_r, p.hydrology_cm.α, p.hydrology_cm.n, p.K_sat, p.S_s]
        top = self.boundary_conditions["top"]
        if isinstance(top, RichardsAtmosDrivenFluxBC) and isinstance(top.runoff, TOPMODELRunoff):
            rows += [top.runoff.f_max, top.runoff.f_over, top.runoff.R_sb]
        return np.stack([np.broadcast_to(np.asarray(r, float), (n,)) for r in rows])


class EnergyHydrology(_SoilModel):
    model_id = L.MODEL_ENERGY_HYDROLOGY

    def _param_rows(self):
        p = self.parameters
        n = self.domain.ncol
        rows = [p.ν, p.θ_r, p.hydrology_cm.α, p.hydrology_cm.n, p.K_sat, p.S_s, p.ν_ss_om,
                p.ν_ss_quartz, p.ν_ss_gravel, p.κ_dry, p.κ_sat_frozen, p.κ_sat_unfrozen,
                p.ρc_ds]
        top = self.boundary_conditions["top"]
        if isinstance(top, AtmosDrivenFluxBC) and isinstance(top.runoff, TOPMODELRunoff):
            rows += [top.runoff.f_max, top.runoff.f_over, top.runoff.R_sb]
        return np.stack([np.broadcast_to(np.asarray(r, float), (n,)) for r in rows])


# ---- hook factories (models.jl) ----------------------------------------------


def make_update_cache(model: _SoilModel):
    def update_cache(p, Y, t):
        model._check(model._lib.clm_update_cache(model._ctx, C.byref(model._state(Y)), t))
    return update_cache


def make_update_implicit_cache(model: _SoilModel):
    def update_implicit_cache(p, Y, t):
        model._check(model._lib.clm_update_implicit_cache(model._ctx,
                                                          C.byref(model._state(Y)), t))
    return update_implicit_cache


def make_exp_tendency(model: _SoilModel):
    """update_cache! + compute_exp_tendency! (models.jl:279-287)."""
    def exp_tendency(dY, Y, p, t):
        model._check(model._lib.clm_exp_tendency(model._ctx, C.byref(model._state(Y)), t,
                                                 C.byref(model._state(dY))))
    return exp_tendency


def make_compute_imp_tendency(model: _SoilModel):
    def compute_imp_tendency(dY, Y, p, t):
        model._check(model._lib.clm_imp_tendency(model._ctx, C.byref(model._state(Y)), t,
                                                 C.byref(model._state(dY))))
    return compute_imp_tendency


def total_liq_water_vol_per_area(model: _SoilModel, Y: dict) -> np.ndarray:
    """ClimaLand.total_liq_water_vol_per_area!(surface_field, model, Y, p, t)
    (rre.jl:502-511, energy_hydrology.jl:1282-1304): Σ ϑ_l Δz per column,
    plus θ_i ρ_i/ρ_l Δz for EnergyHydrology, evaluated on the device for the
    state Y (which becomes the model's device state)."""
    return _totals(model, Y)[0]


def total_energy_per_area(model: _SoilModel, Y: dict) -> np.ndarray:
    """ClimaLand.total_energy_per_area!(surface_field, model, Y, p, t)
    (energy_hydrology.jl:1306-1320): Σ ρe_int Δz per column (EnergyHydrology)."""
    if not model.is_eh:
        raise TypeError("total_energy_per_area is defined for EnergyHydrology")
    return _totals(model, Y)[1]


def soil_conductance(model: _SoilModel, Y: dict, d_ds: float, evap_p: float,
                     evap_alpha: float, D_vapor: float):
    """The surface soil conductance to water vapour that
    update_soil_evaporation's closure computes (energy_hydrology.jl:1103-1131):
    soil_conductance(S_l_sfc, S_c, d_ds, evap_p, evap_α, _D_vapor) with
    S_l_sfc the effective saturation of θ_l extrapolated to the surface
    (:1322-1368), per column, on the device for the state Y.  Returns
    (g_soil in m/s, dry soil layer thickness in m)."""
    if not model.is_eh:
        raise TypeError("soil_conductance is defined for EnergyHydrology")
    model._check(model._lib.clm_set_state(model._ctx, C.byref(model._state(Y))))
    n = model.domain.ncol
    g, dsl = np.zeros(n), np.zeros(n)
    model._check(model._lib.clm_soil_conductance(model._ctx, d_ds, evap_p, evap_alpha, D_vapor,
                                                 L.dptr(g), L.dptr(dsl)))
    return g, dsl


def _totals(model: _SoilModel, Y: dict):
    model._check(model._lib.clm_set_state(model._ctx, C.byref(model._state(Y))))
    n = model.domain.ncol
    w = np.zeros(n)
    e = np.zeros(n) if model.is_eh else None
    model._check(model._lib.clm_totals(model._ctx, L.dptr(w), L.dptr(e)))
    return w, e


class HIPJacobian:
    """jac_prototype: the block matrix lives on the device; ldiv solves with it."""

    def __init__(self, model: _SoilModel):
        self.model = model

    def ldiv(self, dx: dict, R: dict):
        m = self.model
        m._check(m._lib.clm_solve(m._ctx, C.byref(m._state(dx)), C.byref(m._state(R))))
        return dx


def initialize_jacobian(model: _SoilModel) -> HIPJacobian:
    return HIPJacobian(model)


def make_compute_jacobian(model: _SoilModel):
    def compute_jacobian(jacobian: HIPJacobian, Y, p, dtγ, t):
        model._check(model._lib.clm_jacobian(model._ctx, C.byref(model._state(Y)), dtγ, t))
    return compute_jacobian


# ---- simulation driver (Simulations.jl) ----------------------------------------


class NewtonWarmStart:
    """Learned Newton warm start of a soil model (clm_ws_* in clm_soil.h).

    A per-column MLP predicts the implicit part of the step's increment;
    Newton starts from U0 + guess and still iterates to the model's
    tolerance.  The network is trained on the device from (state, converged
    step) pairs that the plain solver produces (`collect`, then `train`).
    It has no counterpart in the reference, where Newton starts at U0.
    """

    def __init__(self, model: "_SoilModel", seed: int = 0):
        self.model = model
        model._check(model._lib.clm_ws_init(model._ctx, seed))
        d = np.zeros(4, dtype=np.int32)
        n = C.c_int64(0)
        model._check(model._lib.clm_ws_dims(model._ctx, L.iptr(d), C.byref(n)))
        self.F, self.Fp, self.H, self.Nout = (int(x) for x in d)
        self.n_params = int(n.value)

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        self.model._check(self.model._lib.clm_ws_get_params(self.model._ctx, L.fptr(p)))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self.model._check(self.model._lib.clm_ws_set_params(self.model._ctx, L.fptr(p)))

    def enable(self, on: bool = True):
        self.model._check(self.model._lib.clm_ws_enable(self.model._ctx, 1 if on else 0))

    def collect(self, sim: "LandSimulation", top_water, top_heat, nsteps: int,
                steps_per_slot: int, stride: int = 1) -> int:
        """Advance `sim` nsteps plain steps on the device, keeping the pairs
        of every stride-th step."""
        sim.set_forcing_table(top_water, top_heat)
        sim._push()
        rows = C.c_int64(0)
        m = self.model
        m._check(m._lib.clm_ws_collect(m._ctx, sim.Δt, nsteps, steps_per_slot, stride,
                                       C.byref(rows)))
        sim.t += nsteps * sim.Δt
        return int(rows.value)

    def train(self, epochs: int, batch: int = 8192, lr: float = 1e-3, seed: int = 1):
        losses = np.zeros(epochs, dtype=np.float32)
        m = self.model
        m._check(m._lib.clm_ws_train(m._ctx, epochs, batch, lr, seed, L.fptr(losses)))
        return losses

    def features(self, dt: float) -> np.ndarray:
        X = np.zeros((self.model.domain.ncol, self.Fp), dtype=np.float32)
        self.model._check(self.model._lib.clm_ws_features(self.model._ctx, dt, L.fptr(X)))
        return X

    def _rows(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.ndim != 2 or X.shape[1] != self.Fp:
            raise ValueError(f"feature rows must be [M][{self.Fp}], got shape {X.shape}")
        return X

    def forward(self, X: np.ndarray) -> np.ndarray:
        X = self._rows(X)
        Y = np.zeros((X.shape[0], self.Nout), dtype=np.float32)
        self.model._check(self.model._lib.clm_ws_forward(self.model._ctx, L.fptr(X), X.shape[0],
                                                         L.fptr(Y)))
        return Y

    def guess(self, dt: float, fused: bool = True):
        """Per-cell guesses (ϑ_l, ρe_int or None) for the device state."""
        m = self.model
        gv = np.zeros(m.domain.ncell)
        ge = np.zeros(m.domain.ncell) if m.is_eh else None
        m._check(m._lib.clm_ws_guess(m._ctx, dt, int(fused), L.dptr(gv), L.dptr(ge)))
        return gv, ge

    def inference_ms(self, dt: float, reps: int = 5) -> float:
        """Device time of one inference (features, network, guess), in ms."""
        ms = C.c_double(0.0)
        self.model._check(self.model._lib.clm_ws_time(self.model._ctx, dt, reps, C.byref(ms)))
        return ms.value

    def gradient(self, X: np.ndarray, T: np.ndarray, mask: np.ndarray):
        X = self._rows(X)
        T = np.ascontiguousarray(T, dtype=np.float32)
        mask = np.ascontiguousarray(mask, dtype=np.float32)
        if T.shape != (X.shape[0], self.Nout) or mask.size != X.shape[0]:
            raise ValueError(f"targets must be [{X.shape[0]}][{self.Nout}] and the mask "
                             f"{X.shape[0]} long, got {T.shape} and {mask.shape}")
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self.model._check(self.model._lib.clm_ws_gradient(
            self.model._ctx, L.fptr(X), L.fptr(T), L.fptr(mask), X.shape[0], L.fptr(loss),
            L.fptr(g)))
        return float(loss[0]), g


class LandSimulation:
    """ARS111 + NewtonsMethod stepping of a soil model.

    `step()` / `solve()` run whole steps on the device (clm_step / clm_run).
    `step_hooks()` drives the same step through the fine-grained hooks in
    the ClimaTimeSteppers order, which is what a ccall drop-in under the
    reference's own integrator executes.
    """

    def __init__(self, model: _SoilModel, Y0: dict, Δt: float, t0: float = 0.0,
                 max_iters: int = 3, tol: float = 0.0):
        self.model = model
        self.Y = {k: np.array(v, dtype=np.float64, copy=True) for k, v in Y0.items()}
        self.Δt = float(Δt)
        self.t = float(t0)
        self.max_iters = max_iters
        self.tol = tol
        self._on_device = False

    def _push(self):
        if not self._on_device:
            m = self.model
            m._check(m._lib.clm_set_state(m._ctx, C.byref(m._state(self.Y))))
            self._on_device = True

    def _pull(self):
        if self._on_device:
            m = self.model
            m._check(m._lib.clm_get_state(m._ctx, C.byref(m._state(self.Y))))
            self._on_device = False

    def state(self) -> dict:
        self._pull()
        return self.Y

    def totals(self):
        """(water, energy) column totals of the current state, computed on the
        device (energy is None for RichardsModel)."""
        m = self.model
        self._push()
        n = m.domain.ncol
        w = np.zeros(n)
        e = np.zeros(n) if m.is_eh else None
        m._check(m._lib.clm_totals(m._ctx, L.dptr(w), L.dptr(e)))
        return w, e

    def step(self) -> L.StepStats:
        m = self.model
        self._push()
        st = L.StepStats()
        m._check(m._lib.clm_step(m._ctx, self.t, self.Δt, C.byref(st)))
        self.t += self.Δt
        return st

    def step_noalloc(self) -> L.StepStats:
        """step!(sim) through clm_step_noalloc: the same step with its Newton
        flags and statistics in the library's static device scratch, so the
        call allocates nothing (state and forcing go up first if the host
        copy changed, as in step)."""
        m = self.model
        self._push()
        st = L.StepStats()
        m._check(m._lib.clm_step_noalloc(m._ctx, self.t, self.Δt, C.byref(st)))
        self.t += self.Δt
        return st

    def save_checkpoint(self, output_dir: str) -> str:
        """ClimaLand.save_checkpoint(Y, t, output_dir): the device state and
        time to <output_dir>/day<D>.<S>.clmchk; returns the file's path."""
        m = self.model
        self._push()
        buf = C.create_string_buffer(4096)
        m._check(m._lib.clm_save_checkpoint(m._ctx, output_dir.encode(), self.t, buf, 4096))
        return buf.value.decode()

    def restore_checkpoint(self, path: str) -> bool:
        """ClimaLand.read_checkpoint: load state and time; returns whether the
        file came from a model with the same parameters and grid."""
        m = self.model
        t = C.c_double(0.0)
        same = C.c_int32(0)
        m._check(m._lib.clm_read_checkpoint(m._ctx, path.encode(), C.byref(t), C.byref(same)))
        self._on_device = True
        self.t = t.value
        return bool(same.value)

    def set_forcing_table(self, top_water, top_heat=None):
        """Upload a [slot][col] forcing schedule (DriverUpdateCallback values).
        The library reads nslots x ncol values through plain pointers, so the
        shapes are checked here."""
        m = self.model
        ncol = m.domain.ncol
        tw = np.ascontiguousarray(top_water, dtype=np.float64)
        if tw.ndim != 2 or tw.shape[1] != ncol or tw.shape[0] < 1:
            raise ValueError(f"top_water must be [nslots][{ncol}], got shape {tw.shape}")
        th = None
        if top_heat is not None:
            th = np.ascontiguousarray(top_heat, dtype=np.float64)
            if th.shape != tw.shape:
                raise ValueError(f"top_heat must have top_water's shape {tw.shape}, got {th.shape}")
        m._check(m._lib.clm_set_forcing_table(m._ctx, tw.shape[0], L.dptr(tw), L.dptr(th)))

    def reset(self, Y0: dict, t0: float = 0.0):
        """Replace the state (host copy; pushed to the device on the next step)."""
        self.Y = {k: np.array(v, dtype=np.float64, copy=True) for k, v in Y0.items()}
        self._on_device = False
        self.t = float(t0)

    def set_surface_flux_table(self, table: dict):
        """AtmosDrivenFluxBC: {field: [slot][col] array} of SoilSurfaceFluxes fields."""
        m = self.model
        nslots = next(np.asarray(v).shape[0] for v in table.values())
        arrs = [np.ascontiguousarray(np.broadcast_to(table.get(k, 0.0), (nslots, m.domain.ncol)),
                                     dtype=np.float64) for k in L.SURFACE_FIELDS]
        st = L.SurfaceFluxes(*(L.dptr(a) for a in arrs))
        m._check(m._lib.clm_set_surface_flux_table(m._ctx, nslots, C.byref(st)))

    def solve(self, top_water, top_heat, nsteps: int, steps_per_slot: int,
              want_iters: bool = False, upload: bool = True, surface: dict | None = None):
        """nsteps steps with a [slot][col] forcing table (or, under
        AtmosDrivenFluxBC, the surface-flux table `surface`), as one graph replay."""
        m = self.model
        if upload and surface is not None:
            self.set_surface_flux_table(surface)
        elif upload and top_water is not None:
            self.set_forcing_table(top_water, top_heat)
        self._push()
        st = L.StepStats()
        it = np.zeros((nsteps, m.domain.ncol), dtype=np.int32) if want_iters else None
        m._check(m._lib.clm_run(m._ctx, self.t, self.Δt, nsteps, steps_per_slot, C.byref(st),
                                L.iptr(it) if it is not None else None))
        self.t += nsteps * self.Δt
        return st, it

    def step_hooks(self):
        """One ARS111 step through the fine-grained hooks (CTS call order)."""
        self._pull()
        m = self.model
        exp_tendency = make_exp_tendency(m)
        imp_tendency = make_compute_imp_tendency(m)
        jacobian = make_compute_jacobian(m)
        cache_imp = make_update_implicit_cache(m)
        W = initialize_jacobian(m)
        dt, t = self.Δt, self.t
        u = self.Y
        Texp = m.new_state()
        exp_tendency(Texp, u, None, t)
        U0 = {k: u[k] + dt * Texp[k] for k in u}
        x = {k: v.copy() for k, v in U0.items()}
        Timp, R, dx = m.new_state(), m.new_state(), m.new_state()
        iters = self.max_iters
        for k in range(1, self.max_iters + 1):
            jacobian(W, x, None, dt, t + dt)
            imp_tendency(Timp, x, None, t + dt)
            for key in x:
                R[key][:] = U0[key] + dt * Timp[key] - x[key]
            W.ldiv(dx, R)
            for key in x:
                x[key] -= dx[key]
            conv = False
            if self.tol > 0:
                err = _column_norm(m, dx, x)
                conv = bool(np.all(err < self.tol))
            cache_imp(None, x, t + dt)
            if conv:
                iters = k
                break
        self.Y = x
        self.t += dt
        return iters


def _column_norm(m: _SoilModel, dx: dict, x: dict) -> np.ndarray:
    """Per-column max|Δx| over the cell unknowns, ρe_int scaled by 1/ρc_s
    with the lagged θ_l and the iterate's θ_i."""
    from .inputs import CP_I, CP_L, RHO_I, RHO_L
    d = np.abs(dx["vartheta_l"])
    if m.is_eh:
        a = m.aux()
        rc = np.repeat(np.asarray(m.parameters.ρc_ds, float), np.diff(m.domain.col_ptr))
        rcs = rc + a["theta_l"] * (CP_L * RHO_L) + x["theta_i"] * (CP_I * RHO_I)
        d = np.maximum(np.maximum(d, np.abs(dx["theta_i"])), np.abs(dx["rho_e_int"]) / rcs)
    return np.maximum.reduceat(d, m.domain.col_ptr[:-1])


def model_from_batch(batch, jac_precision: int | None = None, max_iters: int | None = None,
                     tol: float | None = None, slot: int = 0, solver: int = L.SOLVER_AUTO,
                     jacobian: int = L.JACOBIAN_PICARD, stream: int | None = None):
    """Build the model a synthetic ColumnBatch describes.  jacobian: the
    Newton matrix of the EnergyHydrology group step (L.JACOBIAN_*); stream:
    a HIP stream handle for the context (None: the library's own)."""
    w = batch.workload
    P = batch.params
    hcm = vanGenuchten(P[L.P_VG_ALPHA], P[L.P_VG_N])
    dom = ColumnBatchDomain(batch.col_ptr, batch.z_face)
    kw = dict(jac_precision=jac_precision or w.jac_precision,
              newton_max_iters=max_iters or w.newton_max_iters,
              newton_tol=w.newton_tol if tol is None else tol, solver=solver)
    if jacobian != L.JACOBIAN_PICARD:
        kw["jacobian"] = jacobian
    if stream:
        kw["stream"] = stream
    if batch.model == L.MODEL_RICHARDS:
        par = RichardsParameters(P[L.P_NU], hcm, P[L.P_KSAT], P[L.P_SS], P[L.P_THETA_R])
        tw = batch.top_water[slot]
        if batch.bc_top == L.BC_MOISTURE_STATE:
            top = MoistureStateBC(tw)
        elif batch.bc_top == L.BC_TOPMODEL:
            top = RichardsAtmosDrivenFluxBC(tw, TOPMODELRunoff(
                f_over=P[L.P_F_OVER], f_max=P[L.P_F_MAX], R_sb=P[L.P_R_SB]))
        elif batch.bc_top == L.BC_SURFACE_RUNOFF:
            top = RichardsAtmosDrivenFluxBC(tw, SurfaceRunoff())
        else:
            top = WaterFluxBC(tw)
        bottom = {L.BC_FREE_DRAINAGE: FreeDrainage(),
                  L.BC_MOISTURE_STATE: MoistureStateBC(batch.bottom_water),
                  L.BC_FLUX: WaterFluxBC(batch.bottom_water)}[batch.bc_bottom]
        bcs = {"top": top, "bottom": bottom}
        return RichardsModel(parameters=par, domain=dom, boundary_conditions=bcs, **kw)
    par = EnergyHydrologyParameters(P[L.P_NU], hcm, P[L.P_KSAT], P[L.P_SS], P[L.P_THETA_R],
                                    P[L.P_NU_SS_OM], P[L.P_NU_SS_QUARTZ], P[L.P_NU_SS_GRAVEL],
                                    P[L.P_RHOC_DS], P[L.P_KAPPA_DRY], P[L.P_KAPPA_SAT_FROZEN],
                                    P[L.P_KAPPA_SAT_UNFROZEN])
    if batch.bc_top in (L.BC_ATMOS, L.BC_ATMOS_TOPMODEL):
        runoff = TOPMODELRunoff(f_over=P[L.P_EH_F_OVER], f_max=P[L.P_EH_F_MAX],
                                R_sb=P[L.P_EH_R_SB]) if batch.bc_top == L.BC_ATMOS_TOPMODEL \
            else NoRunoff()
        top = AtmosDrivenFluxBC(SoilSurfaceFluxes(**{k: v[slot] for k, v in
                                                     batch.surface.items()}), runoff)
    else:
        top = WaterHeatBC(WaterFluxBC(batch.top_water[slot]), HeatFluxBC(batch.top_heat[slot]))
    bcs = {"top": top, "bottom": EnergyWaterFreeDrainage()}
    return EnergyHydrology(parameters=par, domain=dom, boundary_conditions=bcs, **kw)


__all__ = [
    "vanGenuchten", "RichardsParameters", "EnergyHydrologyParameters", "ColumnBatchDomain",
    "WaterFluxBC", "HeatFluxBC", "FreeDrainage", "EnergyWaterFreeDrainage", "WaterHeatBC",
    "MoistureStateBC", "NewtonWarmStart", "RichardsAtmosDrivenFluxBC", "NoRunoff",
    "SurfaceRunoff", "TOPMODELRunoff", "AtmosDrivenFluxBC", "SoilSurfaceFluxes",
    "RichardsModel", "EnergyHydrology", "make_update_cache", "make_update_implicit_cache",
    "make_exp_tendency", "make_compute_imp_tendency", "make_compute_jacobian",
    "initialize_jacobian", "LandSimulation", "model_from_batch",
    "KAPPA_OM", "KAPPA_QUARTZ", "KAPPA_MINERALS",
]
