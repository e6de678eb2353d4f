/*
 * clm_oracle.c -- CPU FP64 restatement of ClimaLand.jl's implicit
 * soil-column step.  TEST INFRASTRUCTURE ONLY: imported by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
 * reference for comparison or the CPU baseline, never as the measured or
 * shipped path.
 *
 * Restated from (ClimaLand.jl v1.11.2, paths relative to its root):
 *   src/standalone/Soil/soil_hydrology_parameterizations.jl  (vG closures)
 *   src/standalone/Soil/retention_models.jl:30-44           (vG m)
 *   src/standalone/Soil/soil_heat_parameterizations.jl       (heat closures)
 *   src/standalone/Soil/rre.jl:111-468                       (RichardsModel)
 *   src/standalone/Soil/energy_hydrology.jl:273-576,722-906  (EnergyHydrology)
 *   src/standalone/Soil/boundary_conditions.jl:168-180,340-353,465-477,
 *                                              590-608,709-724
 *   src/shared_utilities/implicit_timestepping.jl:63-172     (block structure)
 *   src/shared_utilities/utils.jl:83-99                      (heaviside)
 *   src/shared_utilities/Domains.jl:935-949                  (Δz fields)
 *   src/simulations/Simulations.jl:127-199                   (ARS111+Newton wiring)
 * The discrete operators (ClimaCore 0.15.2 InterpolateC2F / GradientC2F /
 * DivergenceF2C, MatrixFields Thomas solve) and the IMEX/Newton loop
 * (ClimaTimeSteppers 0.10.6) are not in the reference tree; they are
 * restated from their published definitions (second-order finite volume on
 * a staggered column; textbook Thomas algorithm; ARS111 tableau
 * a_exp=[0 0;1 0], a_imp=[0 0;0 1]; Newton x -= W\R with cache_imp! after
 * each update).
 *
 * Pinning status (what checks each part): see tests/test_oracle_*.py.
 *   - pointwise closures: the reference's own KATs
 *     (test/standalone/Soil/soil_parameterizations.jl) + mpmath closed forms.
 *   - Jacobian: the reference's KATs
 *     (test/shared_utilities/implicit_timestepping/{richards,energy_hydrology}_model.jl) + finite differences
 *     of this file's residual.
 *   - tendency: hydrostatic KAT (test/standalone/Soil/soiltest.jl:15-90),
 *     water/energy conservation identities.
 *   - solve: dense LU.  - step: brute-force dense Newton on tiny columns.
 *   - ClimaParams physical constants (rho_l, rho_i, cp_l, cp_i, T_0,
 *     T_freeze, LH_v0, LH_s0, g, K_therm): parity unpinned (values are the
 *     ClimaParams 1.1.4 defaults as published; no file in the reference
 *     tree states them).
 */
#include "clm_oracle.h"
#include <float.h>
#include <math.h>

/* Transcendentals of the closures.  The default build evaluates them in
 * double exactly as the reference does; -DORC_XPREC evaluates them in long
 * double (rounded back to double) -- the same mathematics with different
 * rounding, used by the tests to measure how strongly a computation reacts
 * to closure rounding alone. */
#ifdef ORC_XPREC
#define POW(x, y) ((double)powl((long double)(x), (long double)(y)))
#define EXP(x) ((double)expl((long double)(x)))
#define LOG(x) ((double)logl((long double)(x)))
#define SQRT(x) ((double)sqrtl((long double)(x)))
#else
#define POW(x, y) pow((x), (y))
#define EXP(x) exp(x)
#define LOG(x) log(x)
#define SQRT(x) sqrt(x)
#endif
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants ------------------------------------------------------- */
/* ClimaParams 1.1.4 defaults (read through src/shared_utilities/Parameters.jl:86-105) */
static const double RHO_L = 1000.0;       /* density_liquid_water */
static const double RHO_I = 916.7;        /* density_ice_water */
static const double CP_L = 4181.0;        /* isobaric_specific_heat_liquid */
static const double CP_I = 2100.0;        /* isobaric_specific_heat_ice */
static const double T_0 = 273.16;         /* thermodynamics_temperature_reference */
static const double T_FREEZE = 273.15;    /* temperature_water_freeze */
static const double LH_V0 = 2500800.0;    /* latent_heat_vaporization_at_reference */
static const double LH_S0 = 2834400.0;    /* latent_heat_sublimation_at_reference */
static const double GRAV = 9.81;          /* gravitational_acceleration */
static const double K_THERM = 0.024;      /* thermal_conductivity_of_air */
/* soil constants pinned by test/standalone/Soil/soil_parameterizations.jl:38-56,71-75 */
static const double KERSTEN_ALPHA = 0.24;
static const double KERSTEN_BETA = 18.3;
static const double OMEGA = 7.0;
static const double GAMMA_VISC = 2.64e-2;
static const double GAMMA_T_REF = 288.0;

#define LH_F0 (LH_S0 - LH_V0) /* Parameters.jl:58 */

/* SQRT(eps(Float64)) as used by the clipping in effective_saturation */
#define SQRT_EPS 1.4901161193847656e-08

void orc_constants(double* out, int n) {
    const double v[] = {RHO_L, RHO_I, CP_L, CP_I, T_0, T_FREEZE, LH_V0, LH_S0,
                        GRAV, K_THERM, KERSTEN_ALPHA, KERSTEN_BETA, OMEGA,
                        GAMMA_VISC, GAMMA_T_REF, LH_F0};
    for (int i = 0; i < n && i < (int)(sizeof(v) / sizeof(v[0])); ++i) out[i] = v[i];
}

/* ---- hydrology closures (soil_hydrology_parameterizations.jl) -------- */

/* :45-50 */
double orc_effective_saturation(double nu_eff, double vartheta, double theta_r) {
    double vs = fmax(vartheta, theta_r + SQRT_EPS);
    double ns = fmax(nu_eff, theta_r + SQRT_EPS);
    return (vs - theta_r) / (ns - theta_r);
}

/* :22-31 */
double orc_volumetric_liquid_fraction(double vartheta, double nu_eff, double theta_r) {
    double vs = fmax(vartheta, theta_r + SQRT_EPS);
    double ns = fmax(nu_eff, theta_r + SQRT_EPS);
    return vs < ns ? vs : ns;
}

/* :59-63, m = 1 - 1/n from retention_models.jl:40-44 */
double orc_matric_potential(double alpha, double n, double S) {
    double m = 1.0 - 1.0 / n;
    return -POW((POW(S, -1.0 / m) - 1.0) * POW(alpha, -n), 1.0 / n);
}

/* :72-77 (the reference raises for psi > 0) */
double orc_inverse_matric_potential(double alpha, double n, double psi, int* err) {
    if (psi > 0.0) {
        if (err) *err = 1;
        return NAN;
    }
    double m = 1.0 - 1.0 / n;
    return POW(1.0 + POW(alpha * fabs(psi), n), -m);
}

/* :109-127 */
double orc_pressure_head(double alpha, double n, double theta_r, double vartheta,
                         double nu_eff, double S_s) {
    double S = orc_effective_saturation(nu_eff, vartheta, theta_r);
    double vs = fmax(vartheta, theta_r + SQRT_EPS);
    double ns = fmax(nu_eff, theta_r + SQRT_EPS);
    if (S <= 1.0) return orc_matric_potential(alpha, n, S);
    return (vs - ns) / S_s;
}

/* :135-152 */
double orc_dpsidtheta(double alpha, double n, double vartheta, double nu_eff,
                      double theta_r, double S_s) {
    double ns = fmax(nu_eff, theta_r + SQRT_EPS);
    double S = orc_effective_saturation(ns, vartheta, theta_r);
    double m = 1.0 - 1.0 / n;
    if (S < 1.0) {
        return 1.0 / (alpha * m * n) / (ns - theta_r) *
               POW(POW(S, -1.0 / m) - 1.0, 1.0 / n - 1.0) * POW(S, -1.0 / m - 1.0);
    }
    return 1.0 / S_s;
}

/* :161-173 */
double orc_hydraulic_conductivity(double n, double K_sat, double S) {
    double m = 1.0 - 1.0 / n;
    double K;
    if (S < 1.0) {
        K = SQRT(S) * POW(1.0 - POW(1.0 - POW(S, 1.0 / m), m), 2.0);
    } else {
        K = 1.0;
    }
    return K * K_sat;
}

/* :302-305 */
double orc_impedance_factor(double f_i, double Omega) { return POW(10.0, -Omega * f_i); }

/* :320-324 */
double orc_viscosity_factor(double T, double gamma, double gamma_T_ref) {
    return EXP(gamma * (T - gamma_T_ref));
}

/* ---- heat closures (soil_heat_parameterizations.jl) ------------------- */

/* :157-171 */
double orc_volumetric_heat_capacity(double theta_l, double theta_i, double rhoc_ds) {
    double rcp_i = CP_I * RHO_I;
    double rcp_l = CP_L * RHO_L;
    return rhoc_ds + theta_l * rcp_l + theta_i * rcp_i;
}

/* :180-192 */
double orc_temperature_from_rho_e_int(double rho_e, double theta_i, double rhoc_s) {
    return T_0 + (rho_e + theta_i * RHO_I * LH_F0) / rhoc_s;
}

/* :201-212 */
double orc_volumetric_internal_energy(double theta_i, double rhoc_s, double T) {
    return rhoc_s * (T - T_0) - theta_i * RHO_I * LH_F0;
}

/* :221-231 */
double orc_volumetric_internal_energy_liq(double T) {
    double rcp_l = CP_L * RHO_L;
    return rcp_l * (T - T_0);
}

/* :243-256 */
double orc_kappa_sat(double theta_l, double theta_i, double k_unfrozen, double k_frozen) {
    double tw = theta_l + theta_i;
    if (tw < DBL_EPSILON) return (k_unfrozen + k_frozen) / 2.0;
    return POW(k_unfrozen, theta_l / tw) * POW(k_frozen, theta_i / tw);
}

/* :267-270 */
double orc_thermal_conductivity(double kappa_dry, double K_e, double kappa_sat) {
    return K_e * kappa_sat + (1.0 - K_e) * kappa_dry;
}

/* :283-285 */
double orc_relative_saturation(double theta_l, double theta_i, double nu) {
    return (theta_l + theta_i) / nu;
}

/* :301-320 (Balland & Arp) */
double orc_kersten_number(double theta_i, double S_r, double a, double b,
                          double nu_om, double nu_quartz, double nu_gravel) {
    if (theta_i < DBL_EPSILON) {
        return POW(S_r, (1.0 + nu_om - a * nu_quartz - nu_gravel) / 2.0) *
               POW(POW(1.0 + EXP(-b * S_r), -3.0) - POW((1.0 - S_r) / 2.0, 3.0),
                   1.0 - nu_om);
    }
    return POW(S_r, 1.0 + nu_om);
}

/* :338-348 */
double orc_kappa_solid(double nu_om, double nu_quartz, double k_om, double k_quartz,
                       double k_minerals) {
    return POW(k_om, nu_om) * POW(k_quartz, nu_quartz) *
           POW(k_minerals, 1.0 - nu_om - nu_quartz);
}

/* :359-361 */
double orc_kappa_sat_frozen(double k_solid, double nu, double k_ice) {
    return POW(k_solid, 1.0 - nu) * POW(k_ice, nu);
}

/* :371-373 */
double orc_kappa_sat_unfrozen(double k_solid, double nu, double k_liq) {
    return POW(k_solid, 1.0 - nu) * POW(k_liq, nu);
}

/* :386-392 */
double orc_kappa_dry(double rho_p, double nu, double k_solid, double k_air) {
    double a = 0.053;
    double rb = (1.0 - nu) * rho_p;
    double num = (a * k_solid - k_air) * rb + k_air * rho_p;
    double den = rho_p - (1.0 - a) * rb;
    return num / den;
}

/* utils.jl:93-99: H(x) = 1 iff x > eps */
double orc_heaviside(double x) { return (x > DBL_EPSILON) ? 1.0 : 0.0; }

/* :35-52, called from phase_change_source with (ρ_i, ρ_l) in the
 * (_ρ_ice, _ρ_liq) slots (:103-114) */
double orc_soil_Tf_depressed(double theta_l, double theta_i, double nu, double theta_r,
                             double alpha, double n) {
    double ttot = fmin(RHO_I / RHO_L * theta_i + theta_l, nu);
    double psiw0 = orc_matric_potential(alpha, n, orc_effective_saturation(nu, ttot, theta_r));
    return fmax(T_FREEZE * EXP(GRAV * psiw0 / LH_F0), 1.0);
}

/* :60-62 */
double orc_thermal_time(double rhoc, double dz, double kappa) {
    return 3.0 * rhoc * dz * dz / kappa;
}

/* :89-122 */
double orc_phase_change_source(double theta_l, double theta_i, double T, double tau,
                               double nu, double theta_r, double alpha, double n,
                               int* err) {
    double Tf = orc_soil_Tf_depressed(theta_l, theta_i, nu, theta_r, alpha, n);
    double psiT = LH_F0 / GRAV * LOG(T / Tf) * orc_heaviside(Tf - T);
    double ttot = fmin(RHO_I / RHO_L * theta_i + theta_l, nu);
    double psiw0 = orc_matric_potential(alpha, n, orc_effective_saturation(nu, ttot, theta_r));
    double thstar = orc_inverse_matric_potential(alpha, n, psiw0 + psiT, err) *
                        (nu - theta_r) + theta_r;
    return (theta_l - thstar) / tau;
}

/* retention_models.jl:30-44: S_c = (1 + ((n-1)/n)^(1-2n))^(-m), m = 1 - 1/n */
double orc_vg_sc(double n) {
    double m = 1.0 - 1.0 / n;
    return POW(1.0 + POW((n - 1.0) / n, 1.0 - 2.0 * n), -m);
}

/* energy_hydrology.jl:1352-1368: αS_c clamped to [0, 1]; d_ds ((αS_c - S_l)/S_l)^p
 * below it, 0 otherwise */
double orc_dry_soil_layer_thickness(double S_l, double aSc, double d_ds, double p) {
    aSc = fmin(fmax(aSc, 0.0), 1.0);
    return S_l < aSc ? d_ds * POW((aSc - S_l) / S_l, p) : 0.0;
}

/* energy_hydrology.jl:1322-1350: D_vapor / max(dsl(S_l, α S_c, d_ds, p), eps) */
double orc_soil_conductance(double S_l, double S_c, double d_ds, double p, double alpha,
                            double D_vapor) {
    double dsl = orc_dry_soil_layer_thickness(S_l, alpha * S_c, d_ds, p);
    return D_vapor / fmax(dsl, DBL_EPSILON);
}

/* ---- column geometry (Domains.jl:935-949; ClimaCore center = face
 * midpoint, interior face J = center spacing, boundary face J = 2x the
 * center-to-boundary distance) -------------------------------------- */
typedef struct {
    int N;
    const double* zf;  /* N+1 faces */
} col_geom;

/* max for the convergence norm that lets a NaN through (so a non-finite
 * iterate never counts as converged: MaximumError compares the norm with
 * `<`, false for NaN) */
static inline double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

static inline double zc(const col_geom* g, int i) { return 0.5 * (g->zf[i] + g->zf[i + 1]); }
static inline double dzc(const col_geom* g, int i) { return g->zf[i + 1] - g->zf[i]; }
/* interior face i (1..N-1): center spacing */
static inline double dzf(const col_geom* g, int i) { return zc(g, i) - zc(g, i - 1); }

/* ---- per-column parameter bundle ------------------------------------- */
typedef struct {
    double nu, theta_r, alpha, n, Ksat, Ss;
    double nu_om, nu_q, nu_g, k_dry, k_satf, k_satu, rhoc_ds;
    double f_max, f_over, R_sb; /* TOPMODEL */
} col_par;

static void load_par(const orc_model* m, int64_t c, col_par* q) {
    const double* P = m->params;
    int64_t nc = m->ncol;
    q->nu = P[ORC_P_NU * nc + c];
    q->theta_r = P[ORC_P_THETA_R * nc + c];
    q->alpha = P[ORC_P_VG_ALPHA * nc + c];
    q->n = P[ORC_P_VG_N * nc + c];
    q->Ksat = P[ORC_P_KSAT * nc + c];
    q->Ss = P[ORC_P_SS * nc + c];
    if (m->model == ORC_MODEL_ENERGY_HYDROLOGY) {
        q->nu_om = P[ORC_P_NU_SS_OM * nc + c];
        q->nu_q = P[ORC_P_NU_SS_QUARTZ * nc + c];
        q->nu_g = P[ORC_P_NU_SS_GRAVEL * nc + c];
        q->k_dry = P[ORC_P_KAPPA_DRY * nc + c];
        q->k_satf = P[ORC_P_KAPPA_SAT_FROZEN * nc + c];
        q->k_satu = P[ORC_P_KAPPA_SAT_UNFROZEN * nc + c];
        q->rhoc_ds = P[ORC_P_RHOC_DS * nc + c];
    } else {
        q->nu_om = q->nu_q = q->nu_g = q->k_dry = q->k_satf = q->k_satu = q->rhoc_ds = 0.0;
    }
    q->f_max = q->f_over = q->R_sb = 0.0;
    if (m->bc_top == ORC_BC_TOPMODEL) {
        q->f_max = P[ORC_P_F_MAX * nc + c];
        q->f_over = P[ORC_P_F_OVER * nc + c];
        q->R_sb = P[ORC_P_R_SB * nc + c];
    } else if (m->bc_top == ORC_BC_ATMOS_TOPMODEL) {
        q->f_max = P[ORC_P_EH_F_MAX * nc + c];
        q->f_over = P[ORC_P_EH_F_OVER * nc + c];
        q->R_sb = P[ORC_P_EH_R_SB * nc + c];
    }
}

/* ---- per-column views ------------------------------------------------- */
typedef struct {
    double *vl, *ti, *re; /* cells */
    double *iw, *ie;      /* scalars */
} col_state;

typedef struct {
    double *K, *psi, *thl, *T, *kap, *sink, *sink_e; /* cells */
    double *tw, *th, *bw, *bh, *rss, *ress, *vice;   /* scalars */
} col_aux;

/* AtmosDrivenFluxBC inputs of one column */
typedef struct {
    double P_liq, T_air, lhf, shf, vliq, vice, R_n;
} sfc_in;

/* ---- RichardsModel ---------------------------------------------------- */

/* rre.jl:368-380 make_update_aux */
static void rre_aux(const col_geom* g, const col_par* q, const double* vl, double* K,
                    double* psi) {
    for (int i = 0; i < g->N; ++i) {
        K[i] = orc_hydraulic_conductivity(q->n, q->Ksat,
                                          orc_effective_saturation(q->nu, vl[i], q->theta_r));
        psi[i] = orc_pressure_head(q->alpha, q->n, q->theta_r, vl[i], q->nu, q->Ss);
    }
}

/* distance from the top / bottom cell centre to the boundary face
 * (Domains.jl:935-949 Δz_top, Δz_bottom) */
static inline double dztop(const col_geom* g) { return g->zf[g->N] - zc(g, g->N - 1); }
static inline double dzbot(const col_geom* g) { return zc(g, 0) - g->zf[0]; }

/* TOPMODEL runoff for RichardsModel (Runoff/Runoff.jl:234-271, 321-344,
 * 373-440; Niu et al. 2005 eqs 8, 11, 12) at the state vl:
 *   is_sat_i = H((ϑ_i - θ_r) - (ν - θ_r)) (ϑ_i - θ_r)/(ν - θ_r),  H(x) = 1 iff x > eps
 *   h∇ = Σ is_sat_i Δz_i (column_integral_definite!), z∇ = depth - h∇
 *   f_sat = min(f_max exp(-f_over z∇ / 2), 1),  i_c = -K_sat
 *   infiltration = (1 - f_sat) max(i_c, P),  R_ss = R_sb exp(-f_over z∇)
 *   sink_i = R_ss / max(h∇, eps) is_sat_i   (dϑ_i -= sink_i, d∫F -= R_ss)
 * With no ice the surface and subsurface passes see the same saturation. */
static void rre_topmodel(const col_geom* g, const col_par* q, const double* vl, double precip,
                         double* infil, double* r_ss, double* sink) {
    int N = g->N;
    double nsat = q->nu - q->theta_r, h = 0.0;
    for (int i = 0; i < N; ++i) {
        double twc = vl[i] - q->theta_r;
        double s = orc_heaviside(twc - nsat) * twc / nsat;
        if (sink) sink[i] = s;
        h += s * dzc(g, i);
    }
    double zw = (g->zf[N] - g->zf[0]) - h;
    double fsat = fmin(q->f_max * EXP(-q->f_over / 2.0 * zw), 1.0);
    double ic = -1.0 * q->Ksat;
    *infil = (1.0 - fsat) * fmax(ic, precip);
    *r_ss = q->R_sb * EXP(-q->f_over * zw);
    if (sink) {
        double scale = *r_ss / fmax(h, DBL_EPSILON);
        for (int i = 0; i < N; ++i) sink[i] *= scale;
    }
}

/* rre.jl:111-149 boundary fluxes: WaterFluxBC (boundary_conditions.jl:168-180),
 * FreeDrainage (:340-353, -K at level 1) and MoistureStateBC (:227-325):
 * diffusive_flux(K_eff, x_2, x_1, Δz) = -K_eff (x_2 - x_1) / Δz with K_eff the
 * boundary cell's K, ψ_bc = pressure_head(ϑ_bc) with that cell's parameters,
 * and gravity entering as ψ + Δz on the upper point. */
static void rre_bc(const orc_model* m, const col_geom* g, const col_par* q, const double* vl,
                   const double* K, const double* psi, double f_top, double f_bot, double* tw,
                   double* bw, double* sink, double* r_ss) {
    int N = g->N;
    if (m->bc_top == ORC_BC_TOPMODEL) {
        double rs = 0.0;
        rre_topmodel(g, q, vl, f_top, tw, &rs, sink);
        if (r_ss) *r_ss = rs;
    } else if (m->bc_top == ORC_BC_SURFACE_RUNOFF) {
        /* Runoff.jl:107-148: (1 - H(ϑ_N - ν)) max(-K_sat, P) */
        *tw = (1.0 - orc_heaviside(vl[N - 1] - q->nu)) * fmax(-1.0 * q->Ksat, f_top);
    } else if (m->bc_top == ORC_BC_MOISTURE_STATE) {
        double dz = dztop(g);
        double psib = orc_pressure_head(q->alpha, q->n, q->theta_r, f_top, q->nu, q->Ss);
        *tw = -K[N - 1] * ((psib + dz) - psi[N - 1]) / dz;
    } else {
        *tw = f_top;
    }
    if (m->bc_bottom == ORC_BC_FREE_DRAINAGE) {
        *bw = -1.0 * K[0];
    } else if (m->bc_bottom == ORC_BC_MOISTURE_STATE) {
        double dz = dzbot(g);
        double psib = orc_pressure_head(q->alpha, q->n, q->theta_r, f_bot, q->nu, q->Ss);
        *bw = -K[0] * ((psi[0] + dz) - psib) / dz;
    } else {
        *bw = f_bot;
    }
}

/* set_dfluxBCdY! for a top MoistureStateBC (boundary_conditions.jl:375-411,
 * added in rre.jl:434-447): the top row of W gains
 * -dtγ K_N dψ/dϑ(ϑ_N) / (Δz_c,N Δz_top) on its diagonal. */
static void rre_top_bc_jacobian(const orc_model* m, const col_geom* g, const double* K,
                                const double* D, double dtg, double* W) {
    if (m->bc_top != ORC_BC_MOISTURE_STATE) return;
    int N = g->N;
    W[3 * (N - 1) + 1] -= dtg * K[N - 1] * D[N - 1] / (dzc(g, N - 1) * dztop(g));
}

/* Darcy face flux of the implicit tendency, rre.jl:191-192:
 * F = -interpc2f(K) * gradc2f(ψ + z), with SetValue fluxes at the boundary
 * faces (divf2c SetValue, :185-188). */
static inline double darcy_face(const col_geom* g, const double* K, const double* psi, int i) {
    double Kf = 0.5 * (K[i - 1] + K[i]);
    double h1 = psi[i] + zc(g, i);
    double h0 = psi[i - 1] + zc(g, i - 1);
    return -Kf * (h1 - h0) / dzf(g, i);
}

/* rre.jl:161-203 compute_imp_tendency */
static void rre_imp_tendency(const col_geom* g, const double* K, const double* psi,
                             double tw, double bw, const double* sink, double r_ss, double* dvl,
                             double* diw) {
    int N = g->N;
    *diw = -(tw - bw);
    double Flo = bw;
    for (int i = 0; i < N; ++i) {
        double Fhi = (i == N - 1) ? tw : darcy_face(g, K, psi, i + 1);
        dvl[i] = -(Fhi - Flo) / dzc(g, i);
        Flo = Fhi;
    }
    /* implicit sources (rre.jl:195-200): TOPMODEL subsurface runoff */
    if (sink) {
        for (int i = 0; i < N; ++i) dvl[i] -= sink[i];
        *diw -= r_ss;
    }
}

/* Modified-Picard tridiagonal block (rre.jl:403-454, EH :503-541 with
 * ν_eff): W = -dtγ · div_F2C · diag(-interp(Kf)) · grad_C2F(SetGradient 0) ·
 * diag(D) - I.  Boundary faces carry zero rows. */
static void picard_tridiag(const col_geom* g, const double* Kc, const double* D,
                           double dtg, double* W /* 3 per cell */) {
    int N = g->N;
    for (int i = 0; i < N; ++i) {
        double lo = 0.0, di = 0.0, up = 0.0;
        double inv_dzc = 1.0 / dzc(g, i);
        if (i > 0) {
            double Kf = 0.5 * (Kc[i - 1] + Kc[i]);
            double w = Kf / dzf(g, i);
            lo += w * D[i - 1];
            di -= w * D[i];
        }
        if (i < N - 1) {
            double Kf = 0.5 * (Kc[i] + Kc[i + 1]);
            double w = Kf / dzf(g, i + 1);
            up += w * D[i + 1];
            di -= w * D[i];
        }
        W[3 * i + 0] = dtg * lo * inv_dzc;
        W[3 * i + 1] = dtg * di * inv_dzc - 1.0;
        W[3 * i + 2] = dtg * up * inv_dzc;
    }
}

/* Thomas algorithm without pivoting (ClimaCore per-column band solve):
 * row i: lo x[i-1] + di x[i] + up x[i+1] = b[i].  scratch: 2N. */
static void thomas(int N, const double* W, const double* b, double* x, double* cp) {
    double* dp = cp + N;
    double beta = W[1];
    cp[0] = W[2] / beta;
    dp[0] = b[0] / beta;
    for (int i = 1; i < N; ++i) {
        double lo = W[3 * i], di = W[3 * i + 1], up = W[3 * i + 2];
        double den = di - lo * cp[i - 1];
        cp[i] = up / den;
        dp[i] = (b[i] - lo * dp[i - 1]) / den;
    }
    x[N - 1] = dp[N - 1];
    for (int i = N - 2; i >= 0; --i) x[i] = dp[i] - cp[i] * x[i + 1];
}

/* ---- EnergyHydrology -------------------------------------------------- */

/* energy_hydrology.jl:722-814 make_update_aux (explicit-stage cache) */
static void eh_aux(const col_geom* g, const col_par* q, const double* vl, const double* ti,
                   const double* re, double* thl, double* kap, double* T, double* K,
                   double* psi) {
    for (int i = 0; i < g->N; ++i) {
        double th = orc_volumetric_liquid_fraction(vl[i], q->nu - ti[i], q->theta_r);
        thl[i] = th;
        kap[i] = orc_thermal_conductivity(
            q->k_dry,
            orc_kersten_number(ti[i], orc_relative_saturation(th, ti[i], q->nu), KERSTEN_ALPHA,
                               KERSTEN_BETA, q->nu_om, q->nu_q, q->nu_g),
            orc_kappa_sat(th, ti[i], q->k_satu, q->k_satf));
        double t = orc_temperature_from_rho_e_int(
            re[i], ti[i], orc_volumetric_heat_capacity(th, ti[i], q->rhoc_ds));
        T[i] = t;
        K[i] = orc_impedance_factor(ti[i] / (th + ti[i] - q->theta_r), OMEGA) *
               orc_viscosity_factor(t, GAMMA_VISC, GAMMA_T_REF) *
               orc_hydraulic_conductivity(q->n, q->Ksat,
                                          orc_effective_saturation(q->nu, vl[i], q->theta_r));
        psi[i] = orc_pressure_head(q->alpha, q->n, q->theta_r, vl[i], q->nu - ti[i], q->Ss);
    }
}

/* energy_hydrology.jl:427-445 make_update_implicit_aux: T with
 * θ_l := min(ν-θ_i, ϑ_l) (no θ_r floor), ψ with ν_eff = ν-θ_i */
static void eh_imp_aux(const col_geom* g, const col_par* q, const double* vl,
                       const double* ti, const double* re, double* T, double* psi) {
    for (int i = 0; i < g->N; ++i) {
        T[i] = orc_temperature_from_rho_e_int(
            re[i], ti[i],
            orc_volumetric_heat_capacity(fmin(q->nu - ti[i], vl[i]), ti[i], q->rhoc_ds));
        psi[i] = orc_pressure_head(q->alpha, q->n, q->theta_r, vl[i], q->nu - ti[i], q->Ss);
    }
}

/* energy_hydrology.jl:273-298 + boundary_conditions.jl:709-724 (WaterHeatBC
 * of fluxes) and :590-608 (EnergyWaterFreeDrainage) */
static void eh_bc(int bc_bottom, const double* K, const double* T, double fw, double fh,
                  double bw_in, double bh_in, double* tw, double* th, double* bw,
                  double* bh) {
    *tw = fw;
    *th = fh;
    if (bc_bottom == ORC_BC_FREE_DRAINAGE) {
        *bw = -1.0 * K[0];
        *bh = -1.0 * K[0] * orc_volumetric_internal_energy_liq(T[0]);
    } else {
        *bw = bw_in;
        *bh = bh_in;
    }
}

/* TOPMODEL for EnergyHydrology (Runoff.jl:234-290, 373-416): the surface
 * pass weights layers by total water (ice included), with the infiltration
 * capacity -K_sat F(θ_i/(θ_l+θ_i-θ_r)) G(T) of the top layer; the subsurface
 * pass by liquid water only, and R_ess carries the liquid energy of the
 * saturated layers: R_ess = R_ss/max(h∇,eps) Σ w ρe_l(T) Δz.  The sources
 * (Runoff.jl:346-359) remove R_ss/max(h∇,eps) w from ϑ_l, R_ess/max(h∇,eps) w
 * from ρe_int, R_ss from ∫F_w and R_ess from ∫F_e. */
static void eh_topmodel(const col_geom* g, const col_par* q, const double* vl, const double* ti,
                        const double* thl, const double* T, double P, double* infil,
                        double* rss, double* ress, double* sink_v, double* sink_e) {
    const int N = g->N;
    const double nsat = q->nu - q->theta_r;
    double hs = 0.0;
    for (int i = 0; i < N; ++i) {
        double twc = vl[i] + ti[i] - q->theta_r;
        hs += orc_heaviside(twc - nsat) * twc / nsat * dzc(g, i);
    }
    const double depth = g->zf[N] - g->zf[0];
    const int t = N - 1;
    double ic = -q->Ksat * orc_impedance_factor(ti[t] / (thl[t] + ti[t] - q->theta_r), OMEGA) *
                orc_viscosity_factor(T[t], GAMMA_VISC, GAMMA_T_REF);
    double fsat = fmin(q->f_max * EXP(-q->f_over / 2.0 * (depth - hs)), 1.0);
    *infil = (1.0 - fsat) * fmax(ic, P);
    double h = 0.0, he = 0.0;
    for (int i = 0; i < N; ++i) {
        double twc = vl[i] - q->theta_r;
        double w = orc_heaviside(twc - nsat) * twc / nsat;
        sink_v[i] = w;
        h += w * dzc(g, i);
        he += w * orc_volumetric_internal_energy_liq(T[i]) * dzc(g, i);
    }
    *rss = q->R_sb * EXP(-q->f_over * (depth - h));
    *ress = he * (*rss / fmax(h, DBL_EPSILON));
    for (int i = 0; i < N; ++i) {
        sink_e[i] = (*ress / fmax(h, DBL_EPSILON)) * sink_v[i];
        sink_v[i] = (*rss / fmax(h, DBL_EPSILON)) * sink_v[i];
    }
}

/* AtmosDrivenFluxBC top fluxes in soil-only form (boundary_conditions.jl:880-
 * 915, 955-1003): infiltration (NoRunoff: the precipitation; TOPMODEL above)
 * plus the liquid vapour flux, and R_n + lhf + shf + infiltration ρe_l(T_air). */
static void eh_atmos_top(const orc_model* m, const col_geom* g, const col_par* q,
                         const double* vl, const double* ti, const double* thl, const double* T,
                         const sfc_in* s, double* tw, double* th, double* rss, double* ress,
                         double* sink_v, double* sink_e) {
    double infil = s->P_liq;
    *rss = *ress = 0.0;
    if (m->bc_top == ORC_BC_ATMOS_TOPMODEL)
        eh_topmodel(g, q, vl, ti, thl, T, s->P_liq, &infil, rss, ress, sink_v, sink_e);
    *tw = infil + s->vliq;
    *th = s->R_n + s->lhf + s->shf + infil * orc_volumetric_internal_energy_liq(s->T_air);
}

/* SoilSublimation source! (energy_hydrology.jl:931-945), explicit: the top
 * layer (H(z + 2 Δz_top), z the layer centre) loses ρ_l/ρ_i E_ice/(2 Δz_top)
 * of ice, ∫F_w the flux E_ice */
static void eh_sublimation(const col_geom* g, double vice, double* dti, double* diw) {
    const double dzt = dztop(g);
    for (int i = 0; i < g->N; ++i)
        dti[i] += -vice * RHO_L / RHO_I * orc_heaviside(zc(g, i) + 2.0 * dzt) / (2.0 * dzt);
    *diw += -vice;
}

/* energy_hydrology.jl:363-425 compute_imp_tendency */
static void eh_imp_tendency(const col_geom* g, const double* K, const double* psi,
                            const double* T, const double* kap, double tw, double th,
                            double bw, double bh, double* dvl, double* dti, double* dre,
                            double* diw, double* die) {
    int N = g->N;
    *diw = -(tw - bw);
    *die = -(th - bh);
    double Flo = bw, Glo = bh;
    for (int i = 0; i < N; ++i) {
        double Fhi, Ghi;
        if (i == N - 1) {
            Fhi = tw;
            Ghi = th;
        } else {
            int f = i + 1;
            double d = dzf(g, f);
            double Kf = 0.5 * (K[f - 1] + K[f]);
            double gh = ((psi[f] + zc(g, f)) - (psi[f - 1] + zc(g, f - 1))) / d;
            Fhi = -Kf * gh;
            double kf = 0.5 * (kap[f - 1] + kap[f]);
            double gT = (T[f] - T[f - 1]) / d;
            double eK = 0.5 * (orc_volumetric_internal_energy_liq(T[f - 1]) * K[f - 1] +
                               orc_volumetric_internal_energy_liq(T[f]) * K[f]);
            Ghi = -kf * gT - eK * gh;
        }
        dvl[i] = -(Fhi - Flo) / dzc(g, i);
        dre[i] = -(Ghi - Glo) / dzc(g, i);
        dti[i] = 0.0;
        Flo = Fhi;
        Glo = Ghi;
    }
}

/* implicit TOPMODELSubsurfaceRunoff source on EH (Runoff.jl:346-359) */
static void eh_runoff_sink(int N, const double* sink_v, const double* sink_e, double rss,
                           double ress, double* dvl, double* dre, double* diw, double* die) {
    for (int i = 0; i < N; ++i) {
        dvl[i] -= sink_v[i];
        dre[i] -= sink_e[i];
    }
    *diw -= rss;
    *die -= ress;
}

/* energy_hydrology.jl:838-906 PhaseChange source! on zeroed dY
 * (compute_exp_tendency :315-350).  τ uses the layer width Δz (:849). */
static int eh_exp_tendency(const col_geom* g, const col_par* q, const double* ti,
                           const double* thl, const double* T, const double* kap,
                           double* dvl, double* dti, double* dre, double* diw, double* die) {
    int err = 0;
    *diw = 0.0;
    *die = 0.0;
    for (int i = 0; i < g->N; ++i) {
        double rc = orc_volumetric_heat_capacity(thl[i], ti[i], q->rhoc_ds);
        double tau = orc_thermal_time(rc, dzc(g, i), kap[i]);
        double F = orc_phase_change_source(thl[i], ti[i], T[i], tau, q->nu, q->theta_r,
                                           q->alpha, q->n, &err);
        dvl[i] = 0.0 + -F;
        dti[i] = 0.0 + (RHO_L / RHO_I) * F;
        dre[i] = 0.0;
    }
    return err;
}

/* energy_hydrology.jl:466-576 compute_jacobian! */
static void eh_jacobian(const col_geom* g, const col_par* q, const double* vl,
                        const double* ti, const double* K, const double* T,
                        const double* kap, const double* thl, double dtg, double* Www,
                        double* Wee, double* Wew, double* tmpD, double* tmpK) {
    int N = g->N;
    /* ∂ϑres/∂ϑ: dψ/dϑ with ν_eff = ν - θ_i (:503-541) */
    for (int i = 0; i < N; ++i)
        tmpD[i] = orc_dpsidtheta(q->alpha, q->n, vl[i], q->nu - ti[i], q->theta_r, q->Ss);
    picard_tridiag(g, K, tmpD, dtg, Www);
    /* ∂ρeres/∂ϑ: diag(-interp(ρe_l(T) K)) ⋅ bidiag, then "- I" (:545-556) */
    for (int i = 0; i < N; ++i) tmpK[i] = orc_volumetric_internal_energy_liq(T[i]) * K[i];
    picard_tridiag(g, tmpK, tmpD, dtg, Wew);
    /* ∂ρeres/∂ρe: grad ⋅ diag(1/ρc_s(θ_l lagged, θ_i)), diag(interp(-κ)) (:559-573) */
    for (int i = 0; i < N; ++i)
        tmpD[i] = 1.0 / orc_volumetric_heat_capacity(thl[i], ti[i], q->rhoc_ds);
    picard_tridiag(g, kap, tmpD, dtg, Wee);
}

/* The exact Jacobian of the same EnergyHydrology residual, for the Newton
 * form of the inexact Newton (ORC_JAC_NEWTON; no counterpart in the
 * reference).  R(x) = U0 + Δt T_imp(x) - x with T_imp of
 * energy_hydrology.jl:363-425, whose x-dependence is the implicit cache
 * (:427-445): ψ(ϑ_l; ν - θ_i) and T = T_0 + (ρe + θ_i ρ_i L_f)/ρc_s(θ_l, θ_i)
 * with θ_l = min(ν - θ_i, ϑ_l).  K, κ and the boundary fluxes are lagged.
 * With h = ψ + z, E = ρe_l(T) K, E' = ρ_l c_l K and per interior face f
 * (cells f-1, f; gh = (h_f - h_{f-1})/Δz_f):
 *   G_f = -κ_f (T_f - T_{f-1})/Δz_f - ½(E_{f-1} + E_f) gh
 *   ∂G_f/∂T_{f-1} = κ_f/Δz_f - ½E'_{f-1} gh,  ∂G_f/∂T_f = -κ_f/Δz_f - ½E'_f gh
 *   ∂G_f/∂h_{f-1} = ½(E_{f-1}+E_f)/Δz_f = -∂G_f/∂h_f
 *   ∂T/∂ρe = 1/ρc_s,  ∂T/∂ϑ_l = -(T - T_0) ρ_l c_l/ρc_s while ϑ_l < ν - θ_i, else 0
 * The water block equals the reference's (K is lagged, so dψ/dϑ is its
 * whole derivative) and ∂R_ϑ/∂ρe = 0, so the system stays block lower
 * triangular; W_ew carries no "- I" (the reference's A_ρϑ does, :554-556). */
static void eh_jacobian_newton(const col_geom* g, const col_par* q, const double* vl,
                               const double* ti, const double* K, const double* T,
                               const double* kap, const double* psi, double dtg, double* Www,
                               double* Wee, double* Wew, double* tmp /* 5N */) {
    const int N = g->N;
    double *dpsi = tmp, *dTdv = tmp + N, *dTde = tmp + 2 * N, *E = tmp + 3 * N,
           *h = tmp + 4 * N;
    for (int j = 0; j < N; ++j) {
        const double nue = q->nu - ti[j];
        const double rc = orc_volumetric_heat_capacity(fmin(nue, vl[j]), ti[j], q->rhoc_ds);
        dTde[j] = 1.0 / rc;
        dTdv[j] = vl[j] < nue ? -(T[j] - T_0) * RHO_L * CP_L / rc : 0.0;
        dpsi[j] = orc_dpsidtheta(q->alpha, q->n, vl[j], nue, q->theta_r, q->Ss);
        E[j] = orc_volumetric_internal_energy_liq(T[j]) * K[j];
        h[j] = psi[j] + zc(g, j);
    }
    for (int i = 0; i < 3 * N; ++i) Www[i] = Wee[i] = Wew[i] = 0.0;
    for (int f = 1; f < N; ++f) {
        const double d = dzf(g, f);
        const double Kf = 0.5 * (K[f - 1] + K[f]), kf = 0.5 * (kap[f - 1] + kap[f]);
        const double gh = (h[f] - h[f - 1]) / d, Eb = 0.5 * (E[f - 1] + E[f]);
        const double dF_lo = Kf / d * dpsi[f - 1], dF_hi = -Kf / d * dpsi[f];
        const double dG_Tlo = kf / d - 0.5 * RHO_L * CP_L * K[f - 1] * gh;
        const double dG_Thi = -kf / d - 0.5 * RHO_L * CP_L * K[f] * gh;
        const double dG_vlo = dG_Tlo * dTdv[f - 1] + Eb / d * dpsi[f - 1];
        const double dG_vhi = dG_Thi * dTdv[f] - Eb / d * dpsi[f];
        const double dG_elo = dG_Tlo * dTde[f - 1], dG_ehi = dG_Thi * dTde[f];
        /* face f is the upper face of cell f-1 (T_imp gets -G_f/Δz) and the
         * lower face of cell f (+G_f/Δz) */
        double s = -dtg / dzc(g, f - 1);
        double* w = Www + 3 * (f - 1);
        w[1] += s * dF_lo, w[2] += s * dF_hi;
        w = Wew + 3 * (f - 1);
        w[1] += s * dG_vlo, w[2] += s * dG_vhi;
        w = Wee + 3 * (f - 1);
        w[1] += s * dG_elo, w[2] += s * dG_ehi;
        s = dtg / dzc(g, f);
        w = Www + 3 * f;
        w[0] += s * dF_lo, w[1] += s * dF_hi;
        w = Wew + 3 * f;
        w[0] += s * dG_vlo, w[1] += s * dG_vhi;
        w = Wee + 3 * f;
        w[0] += s * dG_elo, w[1] += s * dG_ehi;
    }
    for (int i = 0; i < N; ++i) {
        Www[3 * i + 1] -= 1.0;
        Wee[3 * i + 1] -= 1.0;
    }
}

/* BlockLowerTriangularSolve(@name(soil.ϑ_l)) (implicit_timestepping.jl:161-169):
 * x_ϑ = A_ϑϑ \ b_ϑ;  x_ρe = A_ρρ \ (b_ρe - A_ρϑ x_ϑ);  -I blocks: x = -b. */
static void eh_solve_col(int N, const double* Www, const double* Wee, const double* Wew,
                         const double* bvl, const double* bti, const double* bre,
                         double biw, double bie, double* xvl, double* xti, double* xre,
                         double* xiw, double* xie, double* scratch /* 3N */) {
    double* rhs = scratch + 2 * N;
    thomas(N, Www, bvl, xvl, scratch);
    for (int i = 0; i < N; ++i) {
        double s = Wew[3 * i + 1] * xvl[i];
        if (i > 0) s += Wew[3 * i] * xvl[i - 1];
        if (i < N - 1) s += Wew[3 * i + 2] * xvl[i + 1];
        rhs[i] = bre[i] - s;
    }
    thomas(N, Wee, rhs, xre, scratch);
    for (int i = 0; i < N; ++i) xti[i] = -bti[i];
    *xiw = -biw;
    *xie = -bie;
}

/* ---- per-column step: ARS111 + NewtonsMethod ------------------------------
 * ClimaTimeSteppers 0.10.6 semantics restated (Simulations.jl:127-135,
 * 177-199 wire it): stage 1: T_exp!(u_n) = update_cache!(p, u_n, t_n) +
 * compute_exp_tendency; U0 = u_n + Δt T_exp.  Stage 2: x = U0; for each
 * iteration: Wfact(x) with current p; R = U0 + Δt T_imp(x; p) - x;
 * Δ = W \ R; x -= Δ; stop if converged or last; cache_imp!(x).
 * u_{n+1} = x (b_imp = [0, 1]). */
typedef struct {
    double* buf;
    int cap;
} col_scratch;

static double* scratch_get(col_scratch* s, int need) {
    if (need > s->cap) {
        free(s->buf);
        s->buf = (double*)malloc(sizeof(double) * (size_t)need);
        s->cap = need;
    }
    return s->buf;
}

static int rre_step_col(const orc_model* m, int64_t c, col_state* Y, double f_top,
                        double f_bot, double dt, int max_iters, double tol, int* iters,
                        col_scratch* S, const double* gv) {
    int32_t a = m->col_ptr[c];
    int N = m->col_ptr[c + 1] - a;
    col_geom g = {N, m->z_face + a + c};
    col_par q;
    load_par(m, c, &q);
    double* w = scratch_get(S, 14 * N + 8);
    double *K = w, *psi = w + N, *U0 = w + 2 * N, *R = w + 3 * N, *dx = w + 4 * N,
           *W = w + 5 * N, *tmp = w + 8 * N, *D = w + 10 * N, *sink = w + 11 * N;
    double tw, bw, rss = 0.0;
    const int topm = m->bc_top == ORC_BC_TOPMODEL;
    /* stage 1 */
    rre_aux(&g, &q, Y->vl, K, psi);
    rre_bc(m, &g, &q, Y->vl, K, psi, f_top, f_bot, &tw, &bw, topm ? sink : NULL, &rss);
    const int imp_bc = m->bc_top == ORC_BC_MOISTURE_STATE;
    /* exp tendency is zero (rre.jl:214-243): U0 = u_n */
    memcpy(U0, Y->vl, sizeof(double) * N);
    double U0iw = *Y->iw;
    double* x = Y->vl;
    if (gv) { /* warm start x0 = U0 + guess; cache_imp!(x0) */
        for (int i = 0; i < N; ++i) x[i] += gv[i];
        rre_aux(&g, &q, x, K, psi);
        if (imp_bc) rre_bc(m, &g, &q, x, K, psi, f_top, f_bot, &tw, &bw, NULL, NULL);
    }
    int k;
    for (k = 1; k <= max_iters; ++k) {
        /* Wfact(x) */
        for (int i = 0; i < N; ++i)
            D[i] = orc_dpsidtheta(q.alpha, q.n, x[i], q.nu, q.theta_r, q.Ss);
        picard_tridiag(&g, K, D, dt, W);
        rre_top_bc_jacobian(m, &g, K, D, dt, W);
        /* residual */
        double diw;
        rre_imp_tendency(&g, K, psi, tw, bw, topm ? sink : NULL, rss, R, &diw);
        double err = 0.0;
        for (int i = 0; i < N; ++i) R[i] = U0[i] + dt * R[i] - x[i];
        double Riw = U0iw + dt * diw - *Y->iw;
        thomas(N, W, R, dx, tmp);
        for (int i = 0; i < N; ++i) {
            x[i] -= dx[i];
            err = nan_max(err, fabs(dx[i]));
        }
        /* ∫F block is -I (implicit_timestepping.jl:149-152): Δ = -R.  With
         * lagged boundary fluxes it is exact after the first iteration; the
         * convergence norm runs over the cell unknowns only. */
        double dxiw = -Riw;
        *Y->iw -= dxiw;
        int conv = (tol > 0.0) && (err < tol);
        /* cache_imp!(x): update_aux (models.jl:207-210); the boundary
         * fluxes are refreshed only under a top MoistureStateBC
         * (rre.jl:460-468), otherwise they stay lagged at u_n */
        rre_aux(&g, &q, x, K, psi);
        if (imp_bc) rre_bc(m, &g, &q, x, K, psi, f_top, f_bot, &tw, &bw, NULL, NULL);
        if (conv) break;
    }
    if (k > max_iters) {
        *iters = max_iters;
        return (tol > 0.0) ? 1 : 0;
    }
    *iters = k;
    return 0;
}

static int eh_step_col(const orc_model* m, int64_t c, col_state* Y, double f_top_w,
                       double f_top_h, double f_bot_w, double f_bot_h, const sfc_in* sf,
                       double dt, int max_iters, double tol, int* iters, col_scratch* S,
                       int* perr, const double* gv, const double* ge) {
    int32_t a = m->col_ptr[c];
    int N = m->col_ptr[c + 1] - a;
    col_geom g = {N, m->z_face + a + c};
    col_par q;
    load_par(m, c, &q);
    double* w = scratch_get(S, 44 * N + 8);
    double *K = w, *psi = w + N, *thl = w + 2 * N, *T = w + 3 * N, *kap = w + 4 * N;
    double *sv = w + 40 * N, *se = w + 41 * N, rss = 0.0, ress = 0.0;
    const int topm = m->bc_top == ORC_BC_ATMOS_TOPMODEL;
    double *U0v = w + 5 * N, *U0i = w + 6 * N, *U0e = w + 7 * N;
    double *Rv = w + 8 * N, *Ri = w + 9 * N, *Re = w + 10 * N;
    double *dv = w + 11 * N, *di = w + 12 * N, *de = w + 13 * N;
    double *Www = w + 14 * N, *Wee = w + 17 * N, *Wew = w + 20 * N;
    double *tmpD = w + 23 * N, *tmpK = w + 24 * N, *sc = w + 25 * N;
    double tw, th, bw, bh;
    /* stage 1: update_cache!(p, u_n) then exp tendency (phase change) */
    eh_aux(&g, &q, Y->vl, Y->ti, Y->re, thl, kap, T, K, psi);
    eh_bc(m->bc_bottom, K, T, f_top_w, f_top_h, f_bot_w, f_bot_h, &tw, &th, &bw, &bh);
    if (sf) eh_atmos_top(m, &g, &q, Y->vl, Y->ti, thl, T, sf, &tw, &th, &rss, &ress, sv, se);
    double dexw, dexe;
    *perr |= eh_exp_tendency(&g, &q, Y->ti, thl, T, kap, Rv, Ri, Re, &dexw, &dexe);
    if (sf) eh_sublimation(&g, sf->vice, Ri, &dexw);
    for (int i = 0; i < N; ++i) {
        U0v[i] = Y->vl[i] + dt * Rv[i];
        U0i[i] = Y->ti[i] + dt * Ri[i];
        U0e[i] = Y->re[i] + dt * Re[i];
    }
    double U0w = *Y->iw + dt * dexw, U0e_s = *Y->ie + dt * dexe;
    memcpy(Y->vl, U0v, sizeof(double) * N);
    memcpy(Y->ti, U0i, sizeof(double) * N);
    memcpy(Y->re, U0e, sizeof(double) * N);
    *Y->iw = U0w;
    *Y->ie = U0e_s;
    if (gv || ge) { /* warm start x0 = U0 + guess; cache_imp!(x0): T, ψ at x0 */
        for (int i = 0; i < N; ++i) {
            if (gv) Y->vl[i] += gv[i];
            if (ge) Y->re[i] += ge[i];
        }
        eh_imp_aux(&g, &q, Y->vl, Y->ti, Y->re, T, psi);
    }
    int k;
    for (k = 1; k <= max_iters; ++k) {
        if (m->jacobian == ORC_JAC_NEWTON)
            eh_jacobian_newton(&g, &q, Y->vl, Y->ti, K, T, kap, psi, dt, Www, Wee, Wew, tmpD);
        else
            eh_jacobian(&g, &q, Y->vl, Y->ti, K, T, kap, thl, dt, Www, Wee, Wew, tmpD, tmpK);
        double diw, die;
        eh_imp_tendency(&g, K, psi, T, kap, tw, th, bw, bh, Rv, Ri, Re, &diw, &die);
        if (topm) eh_runoff_sink(N, sv, se, rss, ress, Rv, Re, &diw, &die);
        for (int i = 0; i < N; ++i) {
            Rv[i] = U0v[i] + dt * Rv[i] - Y->vl[i];
            Ri[i] = U0i[i] + dt * Ri[i] - Y->ti[i];
            Re[i] = U0e[i] + dt * Re[i] - Y->re[i];
        }
        double Rw = U0w + dt * diw - *Y->iw, Rh = U0e_s + dt * die - *Y->ie;
        double dw, dh;
        eh_solve_col(N, Www, Wee, Wew, Rv, Ri, Re, Rw, Rh, dv, di, de, &dw, &dh, sc);
        double err = 0.0;
        for (int i = 0; i < N; ++i) {
            Y->vl[i] -= dv[i];
            Y->ti[i] -= di[i];
            Y->re[i] -= de[i];
            /* convergence norm: ρe scaled to a temperature by the lagged ρc_s */
            double rc = orc_volumetric_heat_capacity(thl[i], Y->ti[i], q.rhoc_ds);
            err = nan_max(err, nan_max(fabs(dv[i]), nan_max(fabs(di[i]), fabs(de[i]) / rc)));
        }
        *Y->iw -= dw;
        *Y->ie -= dh;
        int conv = (tol > 0.0) && (err < tol);
        eh_imp_aux(&g, &q, Y->vl, Y->ti, Y->re, T, psi);
        if (conv) break;
    }
    if (k > max_iters) {
        *iters = max_iters;
        return (tol > 0.0) ? 1 : 0;
    }
    *iters = k;
    return 0;
}

/* ---- batch drivers ---------------------------------------------------- */
static void view_state(const orc_model* m, const orc_state* Y, int64_t c, col_state* v) {
    int32_t a = m->col_ptr[c];
    v->vl = Y->vartheta_l + a;
    v->iw = Y->intF_water + c;
    if (m->model == ORC_MODEL_ENERGY_HYDROLOGY) {
        v->ti = Y->theta_i + a;
        v->re = Y->rho_e_int + a;
        v->ie = Y->intF_energy + c;
    } else {
        v->ti = v->re = v->ie = NULL;
    }
}

static void view_aux(const orc_model* m, const orc_aux* p, int64_t c, col_aux* v) {
    int32_t a = m->col_ptr[c];
    v->K = p->K + a;
    v->psi = p->psi + a;
    v->tw = p->top_water + c;
    v->bw = p->bot_water + c;
    const int topm = m->bc_top == ORC_BC_TOPMODEL || m->bc_top == ORC_BC_ATMOS_TOPMODEL;
    const int atm = m->bc_top == ORC_BC_ATMOS || m->bc_top == ORC_BC_ATMOS_TOPMODEL;
    v->sink = (topm && p->sink) ? p->sink + a : NULL;
    v->rss = (topm && p->r_ss) ? p->r_ss + c : NULL;
    v->sink_e = (m->bc_top == ORC_BC_ATMOS_TOPMODEL && p->sink_e) ? p->sink_e + a : NULL;
    v->ress = (m->bc_top == ORC_BC_ATMOS_TOPMODEL && p->r_ess) ? p->r_ess + c : NULL;
    v->vice = (atm && p->vapor_ice) ? p->vapor_ice + c : NULL;
    if (m->model == ORC_MODEL_ENERGY_HYDROLOGY) {
        v->thl = p->theta_l + a;
        v->T = p->T + a;
        v->kap = p->kappa + a;
        v->th = p->top_heat + c;
        v->bh = p->bot_heat + c;
    } else {
        v->thl = v->T = v->kap = v->th = v->bh = NULL;
    }
}

static inline double fval(const double* a, int64_t c) { return a ? a[c] : 0.0; }

static inline int is_atmos(const orc_model* m) {
    return m->bc_top == ORC_BC_ATMOS || m->bc_top == ORC_BC_ATMOS_TOPMODEL;
}
static sfc_in sfc_of(const orc_forcing* f, int64_t c) {
    sfc_in s = {fval(f->P_liq, c), fval(f->T_air, c),     fval(f->lhf, c), fval(f->shf, c),
                fval(f->vapor_liq, c), fval(f->vapor_ice, c), fval(f->R_n, c)};
    return s;
}

int orc_update_cache(const orc_model* m, const orc_state* Y, const orc_forcing* f,
                     orc_aux* p) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < m->ncol; ++c) {
        col_state y;
        col_aux a;
        view_state(m, Y, c, &y);
        view_aux(m, p, c, &a);
        int32_t o = m->col_ptr[c];
        col_geom g = {m->col_ptr[c + 1] - o, m->z_face + o + c};
        col_par q;
        load_par(m, c, &q);
        if (m->model == ORC_MODEL_RICHARDS) {
            rre_aux(&g, &q, y.vl, a.K, a.psi);
            rre_bc(m, &g, &q, y.vl, a.K, a.psi, fval(f->top_water, c), fval(f->bot_water, c),
                   a.tw, a.bw, a.sink, a.rss);
        } else {
            eh_aux(&g, &q, y.vl, y.ti, y.re, a.thl, a.kap, a.T, a.K, a.psi);
            eh_bc(m->bc_bottom, a.K, a.T, fval(f->top_water, c), fval(f->top_heat, c),
                  fval(f->bot_water, c), fval(f->bot_heat, c), a.tw, a.th, a.bw, a.bh);
            if (is_atmos(m)) {
                sfc_in s = sfc_of(f, c);
                double rss, ress;
                eh_atmos_top(m, &g, &q, y.vl, y.ti, a.thl, a.T, &s, a.tw, a.th, &rss, &ress,
                             a.sink, a.sink_e);
                if (a.rss) *a.rss = rss;
                if (a.ress) *a.ress = ress;
                if (a.vice) *a.vice = s.vice;
            }
        }
    }
    return 0;
}

int orc_update_implicit_cache(const orc_model* m, const orc_state* Y, const orc_forcing* f,
                              orc_aux* p) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < m->ncol; ++c) {
        col_state y;
        col_aux a;
        view_state(m, Y, c, &y);
        view_aux(m, p, c, &a);
        int32_t o = m->col_ptr[c];
        col_geom g = {m->col_ptr[c + 1] - o, m->z_face + o + c};
        col_par q;
        load_par(m, c, &q);
        if (m->model == ORC_MODEL_RICHARDS) {
            rre_aux(&g, &q, y.vl, a.K, a.psi);
            if (m->bc_top == ORC_BC_MOISTURE_STATE && f)
                rre_bc(m, &g, &q, y.vl, a.K, a.psi, fval(f->top_water, c),
                       fval(f->bot_water, c), a.tw, a.bw, NULL, NULL);
        } else {
            eh_imp_aux(&g, &q, y.vl, y.ti, y.re, a.T, a.psi);
        }
    }
    return 0;
}

int orc_exp_tendency(const orc_model* m, const orc_state* Y, const orc_aux* p,
                     orc_state* dY) {
    int err = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : err)
    for (int64_t c = 0; c < m->ncol; ++c) {
        col_state y, d;
        col_aux a;
        view_state(m, Y, c, &y);
        view_state(m, dY, c, &d);
        view_aux(m, p, c, &a);
        int32_t o = m->col_ptr[c];
        col_geom g = {m->col_ptr[c + 1] - o, m->z_face + o + c};
        col_par q;
        load_par(m, c, &q);
        if (m->model == ORC_MODEL_RICHARDS) {
            for (int i = 0; i < g.N; ++i) d.vl[i] = 0.0;
            *d.iw = 0.0;
        } else {
            err |= eh_exp_tendency(&g, &q, y.ti, a.thl, a.T, a.kap, d.vl, d.ti, d.re, d.iw, d.ie);
            if (a.vice) eh_sublimation(&g, *a.vice, d.ti, d.iw);
        }
    }
    return err ? -3 : 0;
}

int orc_imp_tendency(const orc_model* m, const orc_state* Y, const orc_aux* p,
                     orc_state* dY) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < m->ncol; ++c) {
        col_state d;
        col_aux a;
        view_state(m, dY, c, &d);
        view_aux(m, p, c, &a);
        int32_t o = m->col_ptr[c];
        col_geom g = {m->col_ptr[c + 1] - o, m->z_face + o + c};
        if (m->model == ORC_MODEL_RICHARDS)
            rre_imp_tendency(&g, a.K, a.psi, *a.tw, *a.bw, a.sink, a.rss ? *a.rss : 0.0, d.vl,
                             d.iw);
        else {
            eh_imp_tendency(&g, a.K, a.psi, a.T, a.kap, *a.tw, *a.th, *a.bw, *a.bh, d.vl, d.ti,
                            d.re, d.iw, d.ie);
            if (a.sink && a.sink_e)
                eh_runoff_sink(g.N, a.sink, a.sink_e, *a.rss, *a.ress, d.vl, d.re, d.iw, d.ie);
        }
    }
    (void)Y;
    return 0;
}

int orc_jacobian(const orc_model* m, const orc_state* Y, const orc_aux* p, double dtgamma,
                 double* W_ww, double* W_ee, double* W_ew) {
#pragma omp parallel
    {
        col_scratch S = {NULL, 0};
#pragma omp for schedule(dynamic, 64)
        for (int64_t c = 0; c < m->ncol; ++c) {
            col_state y;
            col_aux a;
            view_state(m, Y, c, &y);
            view_aux(m, p, c, &a);
            int32_t o = m->col_ptr[c];
            col_geom g = {m->col_ptr[c + 1] - o, m->z_face + o + c};
            col_par q;
            load_par(m, c, &q);
            double* t = scratch_get(&S, 5 * g.N + 2);
            if (m->model == ORC_MODEL_ENERGY_HYDROLOGY && m->jacobian == ORC_JAC_NEWTON) {
                eh_jacobian_newton(&g, &q, y.vl, y.ti, a.K, a.T, a.kap, a.psi, dtgamma,
                                   W_ww + 3 * (int64_t)o, W_ee + 3 * (int64_t)o,
                                   W_ew + 3 * (int64_t)o, t);
            } else if (m->model == ORC_MODEL_RICHARDS) {
                for (int i = 0; i < g.N; ++i)
                    t[i] = orc_dpsidtheta(q.alpha, q.n, y.vl[i], q.nu, q.theta_r, q.Ss);
                picard_tridiag(&g, a.K, t, dtgamma, W_ww + 3 * (int64_t)o);
                rre_top_bc_jacobian(m, &g, a.K, t, dtgamma, W_ww + 3 * (int64_t)o);
            } else {
                eh_jacobian(&g, &q, y.vl, y.ti, a.K, a.T, a.kap, a.thl, dtgamma,
                            W_ww + 3 * (int64_t)o, W_ee + 3 * (int64_t)o,
                            W_ew + 3 * (int64_t)o, t, t + g.N);
            }
        }
        free(S.buf);
    }
    return 0;
}

int orc_solve(const orc_model* m, const double* W_ww, const double* W_ee,
              const double* W_ew, const orc_state* R, orc_state* dx) {
#pragma omp parallel
    {
        col_scratch S = {NULL, 0};
#pragma omp for schedule(dynamic, 64)
        for (int64_t c = 0; c < m->ncol; ++c) {
            col_state r, x;
            view_state(m, R, c, &r);
            view_state(m, dx, c, &x);
            int32_t o = m->col_ptr[c];
            int N = m->col_ptr[c + 1] - o;
            double* t = scratch_get(&S, 3 * N + 2);
            if (m->model == ORC_MODEL_RICHARDS) {
                thomas(N, W_ww + 3 * (int64_t)o, r.vl, x.vl, t);
                *x.iw = -*r.iw; /* -I block: BlockDiagonalSolve */
            } else {
                eh_solve_col(N, W_ww + 3 * (int64_t)o, W_ee + 3 * (int64_t)o,
                             W_ew + 3 * (int64_t)o, r.vl, r.ti, r.re, *r.iw, *r.ie, x.vl,
                             x.ti, x.re, x.iw, x.ie, t);
            }
        }
        free(S.buf);
    }
    return 0;
}

int64_t orc_step(const orc_model* m, orc_state* Y, const orc_forcing* f, double dt,
                 int max_iters, double tol, int32_t* iters_out) {
    return orc_step_guess(m, Y, f, dt, max_iters, tol, iters_out, NULL, NULL);
}

int64_t orc_step_guess(const orc_model* m, orc_state* Y, const orc_forcing* f, double dt,
                       int max_iters, double tol, int32_t* iters_out, const double* guess_v,
                       const double* guess_e) {
    int64_t nonconv = 0;
    int err = 0;
#pragma omp parallel reduction(+ : nonconv) reduction(| : err)
    {
        col_scratch S = {NULL, 0};
#pragma omp for schedule(dynamic, 64)
        for (int64_t c = 0; c < m->ncol; ++c) {
            col_state y;
            view_state(m, Y, c, &y);
            int it = 0;
            int32_t o = m->col_ptr[c];
            const double* gv = guess_v ? guess_v + o : NULL;
            const double* ge = guess_e ? guess_e + o : NULL;
            if (m->model == ORC_MODEL_RICHARDS)
                nonconv += rre_step_col(m, c, &y, fval(f->top_water, c), fval(f->bot_water, c),
                                        dt, max_iters, tol, &it, &S, gv);
            else {
                sfc_in sf = sfc_of(f, c);
                nonconv += eh_step_col(m, c, &y, fval(f->top_water, c), fval(f->top_heat, c),
                                       fval(f->bot_water, c), fval(f->bot_heat, c),
                                       is_atmos(m) ? &sf : NULL, dt, max_iters, tol, &it, &S,
                                       &err, gv, ge);
            }
            if (iters_out) iters_out[c] = it;
        }
        free(S.buf);
    }
    return err ? -3 : nonconv;
}

int64_t orc_run(const orc_model* m, orc_state* Y, const double* top_water_tab,
                const double* top_heat_tab, int nslots, int steps_per_slot, double dt,
                int nsteps, int max_iters, double tol, int32_t* iters_out,
                int64_t ncol_limit, int nthreads) {
    int64_t nonconv = 0;
    int err = 0;
    int64_t nc = (ncol_limit > 0 && ncol_limit < m->ncol) ? ncol_limit : m->ncol;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    /* columns are independent: each thread advances a column through all
     * steps (identical arithmetic to stepping all columns per step) */
#pragma omp parallel num_threads(nthreads) reduction(+ : nonconv) reduction(| : err)
    {
        col_scratch S = {NULL, 0};
#pragma omp for schedule(dynamic, 16)
        for (int64_t c = 0; c < nc; ++c) {
            col_state y;
            view_state(m, Y, c, &y);
            for (int s = 0; s < nsteps; ++s) {
                int slot = (s / steps_per_slot) % nslots;
                double fw = top_water_tab[(int64_t)slot * m->ncol + c];
                double fh = top_heat_tab ? top_heat_tab[(int64_t)slot * m->ncol + c] : 0.0;
                int it = 0;
                if (m->model == ORC_MODEL_RICHARDS)
                    nonconv += rre_step_col(m, c, &y, fw, 0.0, dt, max_iters, tol, &it, &S,
                                            NULL);
                else
                    nonconv += eh_step_col(m, c, &y, fw, fh, 0.0, 0.0, NULL, dt, max_iters, tol,
                                           &it, &S, &err, NULL, NULL);
                if (iters_out) iters_out[(int64_t)s * m->ncol + c] = it;
            }
        }
        free(S.buf);
    }
    return err ? -3 : nonconv;
}

int orc_surface_soil_conductance(const orc_model* m, const orc_state* Y, double d_ds,
                                 double evap_p, double evap_alpha, double D_vapor, double* g,
                                 double* dsl) {
    if (m->model != ORC_MODEL_ENERGY_HYDROLOGY) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < m->ncol; ++c) {
        int32_t o = m->col_ptr[c];
        col_geom gm = {m->col_ptr[c + 1] - o, m->z_face + o + c};
        col_par q;
        load_par(m, c, &q);
        const int N = gm.N;
        /* p.soil.θ_l of the top two cells (volumetric_liquid_fraction, ν - θ_i) */
        double f1 = orc_volumetric_liquid_fraction(Y->vartheta_l[o + N - 1],
                                                   q.nu - Y->theta_i[o + N - 1], q.theta_r);
        double f2 = orc_volumetric_liquid_fraction(Y->vartheta_l[o + N - 2],
                                                   q.nu - Y->theta_i[o + N - 2], q.theta_r);
        double z1 = zc(&gm, N - 1), z2 = zc(&gm, N - 2);
        /* linear_interpolation_to_surface! (Domains.jl:883-895) */
        double th = (f1 - f2) / (z1 - z2) * (dztop(&gm) + z1 - z2) + f2;
        th = fmax(th, q.theta_r + DBL_EPSILON);
        double S = orc_effective_saturation(q.nu, th, q.theta_r);
        double Sc = orc_vg_sc(q.n);
        g[c] = orc_soil_conductance(S, Sc, d_ds, evap_p, evap_alpha, D_vapor);
        if (dsl) dsl[c] = orc_dry_soil_layer_thickness(S, evap_alpha * Sc, d_ds, evap_p);
    }
    return 0;
}

int orc_totals(const orc_model* m, const orc_state* Y, double* water, double* energy) {
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < m->ncol; ++c) {
        int32_t o = m->col_ptr[c];
        col_geom g = {m->col_ptr[c + 1] - o, m->z_face + o + c};
        double w = 0.0, e = 0.0;
        for (int i = 0; i < g.N; ++i) {
            double vol = Y->vartheta_l[o + i];
            if (m->model == ORC_MODEL_ENERGY_HYDROLOGY) {
                vol += Y->theta_i[o + i] * RHO_I / RHO_L;
                e += Y->rho_e_int[o + i] * dzc(&g, i);
            }
            w += vol * dzc(&g, i);
        }
        water[c] = w;
        if (energy) energy[c] = e;
    }
    return 0;
}
