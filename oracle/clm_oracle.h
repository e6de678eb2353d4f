/*
 * clm_oracle.h -- CPU FP64 restatement of ClimaLand's implicit soil-column
 * step (RichardsModel and EnergyHydrology), used ONLY as test
 * infrastructure: the parity reference for the HIP library and the CPU
 * baseline timed beside it.  Nothing in the product path links this.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the ClimaLand.jl v1.11.2 source tree).
 *
 * Layout at this boundary: columns are CSR-packed, level 0 = bottom cell.
 *   cells of column c : [col_ptr[c], col_ptr[c+1])
 *   faces of column c : [col_ptr[c] + c, col_ptr[c+1] + c + 1)  (N+1 faces)
 *   per-column params : params[p * ncol + c]  (SoA, index p below)
 */
#ifndef CLM_ORACLE_H
#define CLM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* per-column parameter rows (SoA) */
enum {
    ORC_P_NU = 0,          /* porosity ν */
    ORC_P_THETA_R,         /* residual water θ_r */
    ORC_P_VG_ALPHA,        /* van Genuchten α (1/m) */
    ORC_P_VG_N,            /* van Genuchten n */
    ORC_P_KSAT,            /* K_sat (m/s) */
    ORC_P_SS,              /* specific storativity S_s (1/m) */
    ORC_P_NU_SS_OM,        /* EnergyHydrology only from here */
    ORC_P_NU_SS_QUARTZ,
    ORC_P_NU_SS_GRAVEL,
    ORC_P_KAPPA_DRY,
    ORC_P_KAPPA_SAT_FROZEN,
    ORC_P_KAPPA_SAT_UNFROZEN,
    ORC_P_RHOC_DS,
    ORC_NPARAM_EH
};
#define ORC_NPARAM_RRE 6
/* RichardsModel with a TOPMODEL top boundary takes three more rows */
enum { ORC_P_F_MAX = 6, ORC_P_F_OVER = 7, ORC_P_R_SB = 8 };
#define ORC_NPARAM_RRE_TOPMODEL 9
/* EnergyHydrology with AtmosDrivenFluxBC and TOPMODELRunoff: three more rows */
enum { ORC_P_EH_F_MAX = 13, ORC_P_EH_F_OVER = 14, ORC_P_EH_R_SB = 15 };
#define ORC_NPARAM_EH_TOPMODEL 16

enum { ORC_MODEL_RICHARDS = 0, ORC_MODEL_ENERGY_HYDROLOGY = 1 };
/* Water boundary conditions.  Bottom: prescribed flux (WaterFluxBC /
 * WaterHeatBC of fluxes), FreeDrainage (RRE) / EnergyWaterFreeDrainage (EH),
 * or MoistureStateBC (RRE).  Top: prescribed flux (WaterFluxBC, with
 * HeatFluxBC for EH) or MoistureStateBC (RRE).  For a MoistureStateBC the
 * forcing value is the boundary ϑ_l, not a flux. */
enum { ORC_BC_FLUX = 0, ORC_BC_FREE_DRAINAGE = 1, ORC_BC_MOISTURE_STATE = 2,
       /* top only, RRE: RichardsAtmosDrivenFluxBC(precipitation, TOPMODELRunoff)
        * (boundary_conditions.jl:79-212, Runoff/Runoff.jl:174-486); the forcing
        * value is the liquid precipitation rate; parameter rows
        * ORC_P_F_MAX, ORC_P_F_OVER, ORC_P_R_SB */
       ORC_BC_TOPMODEL = 3,
       /* top only, RRE: RichardsAtmosDrivenFluxBC(precipitation, SurfaceRunoff)
        * (Runoff.jl:84-148); the forcing value is the precipitation rate */
       ORC_BC_SURFACE_RUNOFF = 4,
       /* top only, EH: AtmosDrivenFluxBC in soil-only form
        * (boundary_conditions.jl:646-1003, energy_hydrology.jl:909-945) with
        * the surface fluxes supplied per column (orc_forcing's surface
        * members), NoRunoff (5) or TOPMODELRunoff (6, rows ORC_P_EH_*) */
       ORC_BC_ATMOS = 5, ORC_BC_ATMOS_TOPMODEL = 6 };

typedef struct {
    int32_t model;
    int32_t bc_bottom;
    int64_t ncol;
    const int32_t* col_ptr;
    const double* z_face;
    const double* params;
    int32_t bc_top; /* ORC_BC_FLUX or ORC_BC_MOISTURE_STATE (RRE) */
    /* Newton matrix of EnergyHydrology: ORC_JAC_PICARD, the reference's
     * modified-Picard Jacobian (energy_hydrology.jl:466-576), or
     * ORC_JAC_NEWTON, the exact derivative of the same residual (no
     * counterpart in the reference; pinned by finite differences) */
    int32_t jacobian;
} orc_model;

enum { ORC_JAC_PICARD = 0, ORC_JAC_NEWTON = 1 };

/* prognostic state Y (EH uses all five; RRE uses vartheta_l, intF_water) */
typedef struct {
    double* vartheta_l;
    double* theta_i;
    double* rho_e_int;
    double* intF_water;
    double* intF_energy;
} orc_state;

/* cache p: per-cell closures and per-column boundary fluxes */
typedef struct {
    double* K;
    double* psi;
    double* theta_l;   /* EH */
    double* T;         /* EH */
    double* kappa;     /* EH */
    double* top_water;
    double* top_heat;
    double* bot_water;
    double* bot_heat;
    double* sink;      /* TOPMODEL subsurface runoff sink per cell (1/s), may be NULL */
    double* r_ss;      /* TOPMODEL subsurface runoff R_ss per column (m/s), may be NULL */
    double* sink_e;    /* EH TOPMODEL energy sink per cell (W/m3), may be NULL */
    double* r_ess;     /* EH TOPMODEL energy runoff R_ess per column (W/m2), may be NULL */
    double* vapor_ice; /* EH AtmosDrivenFluxBC: turbulent_fluxes.vapor_flux_ice per column,
                          read by the explicit sublimation source; may be NULL */
} orc_aux;

/* prescribed fluxes (per column), bottom ones used only for ORC_BC_FLUX */
typedef struct {
    const double* top_water;
    const double* top_heat;
    const double* bot_water;
    const double* bot_heat;
    /* ORC_BC_ATMOS*: p.drivers.P_liq (m/s), p.drivers.T (K), the turbulent
     * fluxes lhf, shf (W/m2), vapor_flux_liq, vapor_flux_ice (m/s) and R_n
     * (W/m2), per column */
    const double* P_liq;
    const double* T_air;
    const double* lhf;
    const double* shf;
    const double* vapor_liq;
    const double* vapor_ice;
    const double* R_n;
} orc_forcing;

/* ---- physical constants and pointwise closures (exported for KATs) ---- */
void orc_constants(double* out, int n);  /* see order in clm_oracle.c */
double orc_effective_saturation(double nu_eff, double vartheta, double theta_r);
double orc_volumetric_liquid_fraction(double vartheta, double nu_eff, double theta_r);
double orc_matric_potential(double alpha, double n, double S);
double orc_inverse_matric_potential(double alpha, double n, double psi, int* err);
double orc_pressure_head(double alpha, double n, double theta_r, double vartheta,
                         double nu_eff, double S_s);
double orc_dpsidtheta(double alpha, double n, double vartheta, double nu_eff,
                      double theta_r, double S_s);
double orc_hydraulic_conductivity(double n, double K_sat, double S);
double orc_impedance_factor(double f_i, double Omega);
double orc_viscosity_factor(double T, double gamma, double gamma_T_ref);
double orc_volumetric_heat_capacity(double theta_l, double theta_i, double rhoc_ds);
double orc_temperature_from_rho_e_int(double rho_e, double theta_i, double rhoc_s);
double orc_volumetric_internal_energy(double theta_i, double rhoc_s, double T);
double orc_volumetric_internal_energy_liq(double T);
double orc_kappa_sat(double theta_l, double theta_i, double k_unfrozen, double k_frozen);
double orc_thermal_conductivity(double kappa_dry, double K_e, double kappa_sat);
double orc_relative_saturation(double theta_l, double theta_i, double nu);
double orc_kersten_number(double theta_i, double S_r, double a, double b,
                          double nu_om, double nu_quartz, double nu_gravel);
double orc_kappa_solid(double nu_om, double nu_quartz, double k_om, double k_quartz,
                       double k_minerals);
double orc_kappa_sat_frozen(double k_solid, double nu, double k_ice);
double orc_kappa_sat_unfrozen(double k_solid, double nu, double k_liq);
double orc_kappa_dry(double rho_p, double nu, double k_solid, double k_air);
double orc_soil_Tf_depressed(double theta_l, double theta_i, double nu, double theta_r,
                             double alpha, double n);
double orc_thermal_time(double rhoc, double dz, double kappa);
double orc_phase_change_source(double theta_l, double theta_i, double T, double tau,
                               double nu, double theta_r, double alpha, double n,
                               int* err);
double orc_heaviside(double x);

/* vanGenuchten S_c (retention_models.jl:30-44) and the surface vapour
 * conductance closures (energy_hydrology.jl:1322-1368) */
double orc_vg_sc(double n);
double orc_dry_soil_layer_thickness(double S_l, double aSc, double d_ds, double p);
double orc_soil_conductance(double S_l, double S_c, double d_ds, double p, double alpha,
                            double D_vapor);

/* ---- batch hooks (OpenMP over columns) ---- */
/* make_update_cache: update_aux + update_boundary_fluxes at (Y, t) */
int orc_update_cache(const orc_model* m, const orc_state* Y, const orc_forcing* f,
                     orc_aux* p);
/* make_update_implicit_cache: implicit aux (+ implicit BC refresh: none here) */
/* cache_imp!: implicit aux; with a top MoistureStateBC (RRE) also the
 * boundary fluxes (rre.jl:460-468), which need the forcing f (may be NULL
 * otherwise) */
int orc_update_implicit_cache(const orc_model* m, const orc_state* Y, const orc_forcing* f,
                              orc_aux* p);
/* compute_exp_tendency (p must be current) */
int orc_exp_tendency(const orc_model* m, const orc_state* Y, const orc_aux* p,
                     orc_state* dY);
/* compute_imp_tendency (p must be current) */
int orc_imp_tendency(const orc_model* m, const orc_state* Y, const orc_aux* p,
                     orc_state* dY);
/* compute_jacobian!: W = dtγ·∂T_imp/∂Y − I.  Per cell, tridiagonal rows
 * (lower, diag, upper).  RRE: W_ww[3*cell].  EH additionally W_ee (ρe,ρe)
 * and W_ew (ρe,ϑ). */
int orc_jacobian(const orc_model* m, const orc_state* Y, const orc_aux* p,
                 double dtgamma, double* W_ww, double* W_ee, double* W_ew);
/* ldiv!(Δx, W, R) with the block structure of initialize_jacobian */
int orc_solve(const orc_model* m, const double* W_ww, const double* W_ee,
              const double* W_ew, const orc_state* R, orc_state* dx);
/* One ARS111 IMEX step with Newton, in place.  tol <= 0: exactly max_iters
 * iterations; else per-column stop when max|Δx| < tol (ρe scaled by
 * 1/ρc_s).  iters_out (per column) may be NULL.  Returns count of
 * columns that did not converge (tol > 0), or <0 on error. */
int64_t orc_step(const orc_model* m, orc_state* Y, const orc_forcing* f, double dt,
                 int max_iters, double tol, int32_t* iters_out);
/* nsteps steps; forcing slot s = (step / steps_per_slot) % nslots, slot
 * arrays laid out [slot][ncol].  iters_out is [nsteps][ncol] or NULL.
 * ncol_limit: only columns [0, ncol_limit) are advanced (sampling). */
/* orc_step with a Newton warm start: x0 = U0 + guess (guess_v on ϑ_l,
 * guess_e on ρe_int, per cell; either may be NULL), the implicit cache is
 * refreshed at x0, boundary fluxes stay those of u_n.  With both NULL this
 * is orc_step. */
int64_t orc_step_guess(const orc_model* m, orc_state* Y, const orc_forcing* f, double dt,
                       int max_iters, double tol, int32_t* iters_out, const double* guess_v,
                       const double* guess_e);

int64_t orc_run(const orc_model* m, orc_state* Y, const double* top_water_tab,
                const double* top_heat_tab, int nslots, int steps_per_slot,
                double dt, int nsteps, int max_iters, double tol,
                int32_t* iters_out, int64_t ncol_limit, int nthreads);

/* totals: Σ ϑ Δz (+ρ_i/ρ_l θ_i Δz), Σ ρe Δz  (rre.jl:502-511,
 * energy_hydrology.jl:1282-1320) */
int orc_totals(const orc_model* m, const orc_state* Y, double* water, double* energy);
/* surface soil conductance per column of the state Y (EnergyHydrology):
 * energy_hydrology.jl:1103-1131 with Domains.jl:883-895; dsl may be NULL */
int orc_surface_soil_conductance(const orc_model* m, const orc_state* Y, double d_ds,
                                 double evap_p, double evap_alpha, double D_vapor, double* g,
                                 double* dsl);

#ifdef __cplusplus
}
#endif
#endif
