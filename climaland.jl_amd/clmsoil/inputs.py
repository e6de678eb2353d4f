"""Deterministic synthetic column batches (the benchmark and test workloads).

Every column draws from its own counter-based splitmix64 stream, seeded with
(config seed) ^ (column index * golden ratio), in a fixed order:
parameters -> geometry -> initial state -> forcing slots.  A batch of the
first n columns of a workload is therefore exactly a prefix of the full
workload, which is how the tests run the large workloads at small sizes.

Workloads (names are the ones bench.py and the tests use):
  rre_4096x50        RichardsModel, 4,096 columns x 50 uniform layers, z in
                     [-2, 0] m, top flux -1e-5 m/s, free drainage, dt 900 s,
                     96 steps, 3 fixed Newton iterations.
  rre_131072_uneven  RichardsModel, 131,072 columns, U{15..120} layers,
                     depth U[1,10] m stretched toward the surface (ratio
                     1.05), top flux U[-5e-5, 2e-6] m/s resampled every 6
                     steps, dt 1800 s, 48 steps, Newton to max|dx| < 1e-8.
  eh_65536x20        EnergyHydrology, 65,536 columns x 20 stretched layers to
                     15 m, T0 U[260,300] K (freeze-thaw active), dt 900 s, 96
                     steps, heat flux N(0, 80) W/m2, infiltration as above.
  eh_262144_uneven   EnergyHydrology, 262,144 columns, U{10..40} layers,
                     depth U[3,15] m, forcing as eh_65536x20, dt 1800 s,
                     240 steps.
Hydraulic parameters (all workloads): nu U[0.35,0.55], theta_r U[0.02,0.10],
alpha logU[0.5,10] 1/m, n U[1.2,3.0], K_sat logU[1e-7,1e-4] m/s, S_s 1e-3.
EnergyHydrology adds rho c_ds U[1.5e6,2.5e6], kappa_dry U[0.2,0.4], quartz
fraction U[0.1,0.9] (organic and gravel fractions 0).  Initial liquid water
theta_r + s (nu - theta_r) with s linear in depth between s_top, s_bot
U[0.3,0.9]; ice 0.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import _lib as L

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
BASE_SEED = 0xC1A4A1D0

# thermal constants used to derive EnergyHydrology parameters once per column
# (EnergyHydrologyParameters constructor, energy_hydrology.jl:149-172; values
# pinned by test/standalone/Soil/soil_parameterizations.jl:38-56)
KAPPA_QUARTZ, KAPPA_MINERALS, KAPPA_OM = 8.0, 2.5, 0.25
KAPPA_ICE, KAPPA_LIQ = 2.21, 0.57
RHO_L, RHO_I, CP_L, CP_I, T_0 = 1000.0, 916.7, 4181.0, 2100.0, 273.16
LH_F0 = 2834400.0 - 2500800.0
LH_V0, LH_S0 = 2500800.0, 2834400.0


class ColumnRNG:
    """splitmix64, one independent stream per column, vectorised over columns."""

    def __init__(self, seed: int, ncol: int):
        with np.errstate(over="ignore"):
            c = np.arange(ncol, dtype=np.uint64)
            self.state = np.uint64(seed) ^ (c * GOLDEN)

    def next_u64(self) -> np.ndarray:
        with np.errstate(over="ignore"):
            self.state = self.state + GOLDEN
            z = self.state.copy()
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))

    def uniform(self, a: float = 0.0, b: float = 1.0) -> np.ndarray:
        u = (self.next_u64() >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return a + (b - a) * u

    def log_uniform(self, a: float, b: float) -> np.ndarray:
        return np.exp(self.uniform(math.log(a), math.log(b)))

    def normal(self, sigma: float) -> np.ndarray:
        u1 = 1.0 - self.uniform()  # (0, 1]
        u2 = self.uniform()
        return sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)

    def int_range(self, lo: int, hi: int) -> np.ndarray:
        return np.minimum(lo + np.floor(self.uniform() * (hi - lo + 1)).astype(np.int64), hi)


@dataclasses.dataclass
class Workload:
    name: str
    config_id: int
    model: int
    ncol: int
    layers: tuple[int, int]           # inclusive range (equal ends = fixed)
    depth: tuple[float, float]
    stretch: float | None             # None = uniform layers
    dt: float
    nsteps: int
    steps_per_slot: int
    top_flux: tuple[float, float] | None  # None = constant -1e-5
    heat_sigma: float | None
    newton_max_iters: int
    newton_tol: float
    jac_precision: int = 64


WORKLOADS = {
    "rre_4096x50": Workload("rre_4096x50", 0, L.MODEL_RICHARDS, 4096, (50, 50), (2.0, 2.0),
                            None, 900.0, 96, 96, None, None, 3, 0.0),
    "rre_131072_uneven": Workload("rre_131072_uneven", 1, L.MODEL_RICHARDS, 131072, (15, 120),
                                  (1.0, 10.0), 1.05, 1800.0, 48, 6, (-5e-5, 2e-6), None, 50,
                                  1e-8),
    "eh_65536x20": Workload("eh_65536x20", 2, L.MODEL_ENERGY_HYDROLOGY, 65536, (20, 20),
                            (15.0, 15.0), 1.05, 900.0, 96, 6, (-5e-5, 2e-6), 80.0, 50, 1e-8,
                            32),
    "eh_262144_uneven": Workload("eh_262144_uneven", 4, L.MODEL_ENERGY_HYDROLOGY, 262144,
                                 (10, 40), (3.0, 15.0), 1.05, 1800.0, 240, 6, (-5e-5, 2e-6),
                                 80.0, 50, 1e-8, 32),
}


@dataclasses.dataclass
class ColumnBatch:
    """A CSR column batch with parameters, initial state and forcing table."""
    workload: Workload
    model: int
    bc_bottom: int
    col_ptr: np.ndarray   # int32 [ncol+1]
    z_face: np.ndarray    # float64 [ncell + ncol]
    params: np.ndarray    # float64 [nparam, ncol]
    vartheta_l: np.ndarray
    theta_i: np.ndarray | None
    rho_e_int: np.ndarray | None
    intF_water: np.ndarray
    intF_energy: np.ndarray | None
    top_water: np.ndarray  # [nslots, ncol]; ϑ_bc under a top MoistureStateBC
    top_heat: np.ndarray | None
    bc_top: int = L.BC_FLUX
    bottom_water: np.ndarray | None = None  # [ncol]: bottom flux or ϑ_bc
    surface: dict | None = None  # AtmosDrivenFluxBC inputs {field: [nslots, ncol]}

    @property
    def ncol(self) -> int:
        return self.col_ptr.size - 1

    @property
    def ncell(self) -> int:
        return int(self.col_ptr[-1])

    @property
    def nlayers(self) -> np.ndarray:
        return np.diff(self.col_ptr)

    @property
    def nslots(self) -> int:
        return self.top_water.shape[0]

    def state_arrays(self) -> dict:
        d = {"vartheta_l": self.vartheta_l, "intF_water": self.intF_water}
        if self.model == L.MODEL_ENERGY_HYDROLOGY:
            d.update(theta_i=self.theta_i, rho_e_int=self.rho_e_int, intF_energy=self.intF_energy)
        return d

    def copy_state(self) -> dict:
        return {k: v.copy() for k, v in self.state_arrays().items()}

    def forcing_slot(self, step: int) -> int:
        return (step // self.workload.steps_per_slot) % self.nslots

    def column_subset(self, cols: np.ndarray) -> "ColumnBatch":
        """A batch made of the given columns (in that order)."""
        cols = np.asarray(cols, dtype=np.int64)
        n = self.nlayers[cols]
        cp = np.zeros(cols.size + 1, dtype=np.int32)
        cp[1:] = np.cumsum(n)
        cell_idx = np.concatenate([np.arange(self.col_ptr[c], self.col_ptr[c + 1]) for c in cols])
        face_idx = np.concatenate([np.arange(self.col_ptr[c] + c, self.col_ptr[c + 1] + c + 1)
                                   for c in cols])
        eh = self.model == L.MODEL_ENERGY_HYDROLOGY
        return ColumnBatch(
            self.workload, self.model, self.bc_bottom, cp, self.z_face[face_idx].copy(),
            np.ascontiguousarray(self.params[:, cols]), self.vartheta_l[cell_idx].copy(),
            self.theta_i[cell_idx].copy() if eh else None,
            self.rho_e_int[cell_idx].copy() if eh else None, self.intF_water[cols].copy(),
            self.intF_energy[cols].copy() if eh else None,
            np.ascontiguousarray(self.top_water[:, cols]),
            np.ascontiguousarray(self.top_heat[:, cols]) if eh else None, self.bc_top,
            None if self.bottom_water is None else self.bottom_water[cols].copy(),
            None if self.surface is None else
            {k: np.ascontiguousarray(v[:, cols]) for k, v in self.surface.items()})

    def with_moisture_state(self, top_frac: float, bottom_frac: float | None = None,
                            seed: int = 0) -> "ColumnBatch":
        """The same RichardsModel columns under MoistureStateBC: ϑ_bc = θ_r +
        s (ν - θ_r) with s ~ U[top_frac - 0.05, top_frac + 0.05] per column
        at the top, and (if bottom_frac is given) likewise at the bottom;
        the bottom stays free drainage otherwise."""
        assert self.model == L.MODEL_RICHARDS
        rng = np.random.default_rng(seed)
        nu, thr = self.params[L.P_NU], self.params[L.P_THETA_R]
        s_top = top_frac + rng.uniform(-0.05, 0.05, self.ncol)
        top = np.broadcast_to(thr + s_top * (nu - thr), self.top_water.shape).copy()
        b = dataclasses.replace(self, top_water=top, bc_top=L.BC_MOISTURE_STATE)
        if bottom_frac is not None:
            s_bot = bottom_frac + rng.uniform(-0.05, 0.05, self.ncol)
            b = dataclasses.replace(b, bc_bottom=L.BC_MOISTURE_STATE,
                                    bottom_water=thr + s_bot * (nu - thr))
        return b


    def with_runoff(self, kind: int, seed: int = 0) -> "ColumnBatch":
        """The same RichardsModel columns under RichardsAtmosDrivenFluxBC with
        TOPMODELRunoff (kind L.BC_TOPMODEL) or SurfaceRunoff (L.BC_SURFACE_RUNOFF).

        The top forcing becomes the precipitation P = -K_sat U[0.1, 3] per
        column and slot, so that both branches of max(-K_sat, P) occur.  A
        water table at height f·depth above the bottom, f ~ U[0.1, 0.6], makes
        the cells below it saturated and hydrostatic, ϑ = ν + S_s (z_wt - z);
        a fifth of the columns have no water table (f = 0) and, under
        SurfaceRunoff, another fifth are saturated to the top (f = 1).
        TOPMODEL rows: f_max ~ U[0.2, 0.6], f_over = 3.28 1/m and
        R_sb = 1.484e-7 U[0.5, 2] m/s (the values of the reference's runoff
        test, test/standalone/Soil/runoff.jl:101-103)."""
        assert self.model == L.MODEL_RICHARDS
        assert kind in (L.BC_TOPMODEL, L.BC_SURFACE_RUNOFF)
        rng = np.random.default_rng(seed)
        P = self.params
        nc = self.ncol
        precip = -P[L.P_KSAT][None, :] * rng.uniform(0.1, 3.0, self.top_water.shape)
        f = rng.uniform(0.1, 0.6, nc)
        pick = rng.random(nc)
        f[pick < 0.2] = 0.0
        if kind == L.BC_SURFACE_RUNOFF:
            f[pick > 0.8] = 1.0
        nl = self.nlayers
        ccol = np.repeat(np.arange(nc), nl)
        fb = np.arange(self.ncell) + ccol
        zb, zt = self.z_face[self.col_ptr[:-1] + np.arange(nc)], self.z_face[self.col_ptr[1:] +
                                                                           np.arange(nc)]
        zwt = zb + f * (zt - zb)
        zc = 0.5 * (self.z_face[fb] + self.z_face[fb + 1])
        nu, ss = P[L.P_NU][ccol], P[L.P_SS][ccol]
        below = zc < zwt[ccol]
        vl = np.where(below, nu + ss * (zwt[ccol] - zc), np.minimum(self.vartheta_l, 0.98 * nu))
        if kind == L.BC_TOPMODEL:
            extra = np.stack([rng.uniform(0.2, 0.6, nc), np.full(nc, 3.28),
                              1.484e-7 * rng.uniform(0.5, 2.0, nc)])
            params = np.ascontiguousarray(np.concatenate([P[:L.NPARAM_RRE], extra]))
        else:
            params = P
        return dataclasses.replace(self, params=params, vartheta_l=vl,
                                   top_water=np.ascontiguousarray(precip), bc_top=kind)


    def with_atmos(self, topmodel: bool, seed: int = 0) -> "ColumnBatch":
        """The same EnergyHydrology columns under AtmosDrivenFluxBC (soil-only
        form) with synthetic driver and turbulent-flux values per column and
        slot: P_liq = -K_sat U[0.1, 3], T_air U[255, 300] K, lhf U[0, 150],
        shf U[-50, 100], R_n U[-400, 100] W/m2, vapor_flux_liq = lhf /
        (ρ_l L_v) on columns warmer than freezing, vapor_flux_ice likewise with
        L_s on the others.  With TOPMODEL the columns also get the water table
        of `with_runoff` and the TOPMODEL rows (f_max U[0.2, 0.6], f_over 3.28,
        R_sb 1.484e-7 U[0.5, 2])."""
        assert self.model == L.MODEL_ENERGY_HYDROLOGY
        rng = np.random.default_rng(seed)
        P = self.params
        shp = self.top_water.shape
        lhf = rng.uniform(0.0, 150.0, shp)
        T_air = rng.uniform(255.0, 300.0, shp)
        warm = T_air > 273.15
        sfc = {"P_liq": -P[L.P_KSAT][None, :] * rng.uniform(0.1, 3.0, shp), "T_air": T_air,
               "lhf": lhf, "shf": rng.uniform(-50.0, 100.0, shp),
               "vapor_flux_liq": np.where(warm, lhf / (RHO_L * LH_V0), 0.0),
               "vapor_flux_ice": np.where(warm, 0.0, lhf / (RHO_L * LH_S0)),
               "R_n": rng.uniform(-400.0, 100.0, shp)}
        b = dataclasses.replace(self, surface=sfc,
                                bc_top=L.BC_ATMOS_TOPMODEL if topmodel else L.BC_ATMOS)
        if not topmodel:
            return b
        nc = self.ncol
        f = rng.uniform(0.1, 0.6, nc)
        f[rng.random(nc) < 0.2] = 0.0
        ccol = np.repeat(np.arange(nc), self.nlayers)
        fb = np.arange(self.ncell) + ccol
        zb = self.z_face[self.col_ptr[:-1] + np.arange(nc)]
        zt = self.z_face[self.col_ptr[1:] + np.arange(nc)]
        zwt = zb + f * (zt - zb)
        zc = 0.5 * (self.z_face[fb] + self.z_face[fb + 1])
        nu, ss = P[L.P_NU][ccol], P[L.P_SS][ccol]
        vl = np.where(zc < zwt[ccol], nu + ss * (zwt[ccol] - zc),
                      np.minimum(self.vartheta_l, 0.98 * nu))
        # ρe_int keeps the column's temperature with the new liquid content
        thl = np.minimum(self.vartheta_l, nu - self.theta_i)
        rc_old = P[L.P_RHOC_DS][ccol] + thl * RHO_L * CP_L + self.theta_i * RHO_I * CP_I
        T = T_0 + (self.rho_e_int + self.theta_i * RHO_I * LH_F0) / rc_old
        thl_new = np.minimum(vl, nu - self.theta_i)
        rc = P[L.P_RHOC_DS][ccol] + thl_new * RHO_L * CP_L + self.theta_i * RHO_I * CP_I
        re = rc * (T - T_0) - self.theta_i * RHO_I * LH_F0
        extra = np.stack([rng.uniform(0.2, 0.6, nc), np.full(nc, 3.28),
                          1.484e-7 * rng.uniform(0.5, 2.0, nc)])
        return dataclasses.replace(b, params=np.ascontiguousarray(np.concatenate([P[:13], extra])),
                                   vartheta_l=vl, rho_e_int=re)


def kappa_solid(nu_om, nu_q):
    return KAPPA_OM ** nu_om * KAPPA_QUARTZ ** nu_q * KAPPA_MINERALS ** (1.0 - nu_om - nu_q)


def training_workload(name: str) -> str:
    """Name of a workload drawn like `name` from other seeds (config_id + 1000):
    columns for training the warm start that the evaluation never sees."""
    tname = name + "_train"
    if tname not in WORKLOADS:
        w = WORKLOADS[name]
        WORKLOADS[tname] = dataclasses.replace(w, name=tname, config_id=w.config_id + 1000)
    return tname


def make_batch(name: str, ncol: int | None = None, nsteps: int | None = None) -> ColumnBatch:
    """The first `ncol` columns of workload `name` (default: all of them)."""
    w = WORKLOADS[name]
    if nsteps is not None:
        w = dataclasses.replace(w, nsteps=nsteps)
    nc = w.ncol if ncol is None else int(ncol)
    rng = ColumnRNG(BASE_SEED + w.config_id, nc)
    eh = w.model == L.MODEL_ENERGY_HYDROLOGY
    # ---- parameters
    nu = rng.uniform(0.35, 0.55)
    theta_r = rng.uniform(0.02, 0.10)
    alpha = rng.log_uniform(0.5, 10.0)
    n = rng.uniform(1.2, 3.0)
    ksat = rng.log_uniform(1e-7, 1e-4)
    ss = np.full(nc, 1e-3)
    nparam = L.NPARAM_EH if eh else L.NPARAM_RRE
    params = np.zeros((nparam, nc))
    params[L.P_NU], params[L.P_THETA_R], params[L.P_VG_ALPHA] = nu, theta_r, alpha
    params[L.P_VG_N], params[L.P_KSAT], params[L.P_SS] = n, ksat, ss
    if eh:
        rhoc_ds = rng.uniform(1.5e6, 2.5e6)
        k_dry = rng.uniform(0.2, 0.4)
        quartz = rng.uniform(0.1, 0.9)
        ks = kappa_solid(0.0, quartz)
        params[L.P_NU_SS_OM] = 0.0
        params[L.P_NU_SS_QUARTZ] = quartz
        params[L.P_NU_SS_GRAVEL] = 0.0
        params[L.P_KAPPA_DRY] = k_dry
        params[L.P_KAPPA_SAT_FROZEN] = ks ** (1.0 - nu) * KAPPA_ICE ** nu
        params[L.P_KAPPA_SAT_UNFROZEN] = ks ** (1.0 - nu) * KAPPA_LIQ ** nu
        params[L.P_RHOC_DS] = rhoc_ds
    # ---- geometry
    if w.layers[0] == w.layers[1]:
        nl = np.full(nc, w.layers[0], dtype=np.int64)
    else:
        nl = rng.int_range(w.layers[0], w.layers[1])
    depth = (np.full(nc, w.depth[0]) if w.depth[0] == w.depth[1]
             else rng.uniform(w.depth[0], w.depth[1]))
    col_ptr = np.zeros(nc + 1, dtype=np.int32)
    col_ptr[1:] = np.cumsum(nl)
    nface = nl + 1
    fcol = np.repeat(np.arange(nc), nface)
    fidx = np.arange(fcol.size) - np.repeat(col_ptr[:-1] + np.arange(nc), nface)
    N_f, D_f = nl[fcol], depth[fcol]
    k_top = (N_f - fidx).astype(np.float64)  # layers between the face and the surface
    if w.stretch is None:
        z_face = -D_f * k_top / N_f
    else:
        r = w.stretch
        z_face = -D_f * (r ** k_top - 1.0) / (r ** N_f.astype(np.float64) - 1.0)
    # ---- initial state
    s_top = rng.uniform(0.3, 0.9)
    s_bot = rng.uniform(0.3, 0.9)
    ccol = np.repeat(np.arange(nc), nl)
    ci = np.arange(ccol.size) - np.repeat(col_ptr[:-1], nl)
    fb = col_ptr[ccol] + ccol + ci
    zc = 0.5 * (z_face[fb] + z_face[fb + 1])
    s = s_top[ccol] + (s_bot[ccol] - s_top[ccol]) * (-zc) / depth[ccol]
    vl = theta_r[ccol] + s * (nu[ccol] - theta_r[ccol])
    if eh:
        T0 = rng.uniform(260.0, 300.0)
        ti = np.zeros_like(vl)
        rcs = params[L.P_RHOC_DS][ccol] + vl * (CP_L * RHO_L) + ti * (CP_I * RHO_I)
        # volumetric_internal_energy, soil_heat_parameterizations.jl:201-212
        re = rcs * (T0[ccol] - T_0) - ti * RHO_I * LH_F0
    # ---- forcing table
    nslots = max(1, -(-w.nsteps // w.steps_per_slot))
    if w.top_flux is None:
        tw = np.full((nslots, nc), -1e-5)
    else:
        tw = np.zeros((nslots, nc))
    th = np.zeros((nslots, nc)) if eh else None
    for k in range(nslots):
        if w.top_flux is not None:
            tw[k] = rng.uniform(*w.top_flux)
        if eh:
            th[k] = rng.normal(w.heat_sigma)
    return ColumnBatch(
        w, w.model, L.BC_FREE_DRAINAGE, col_ptr, np.ascontiguousarray(z_face),
        np.ascontiguousarray(params), vl, ti if eh else None, re if eh else None,
        np.zeros(nc), np.zeros(nc) if eh else None, tw, th)
