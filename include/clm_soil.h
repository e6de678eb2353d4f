/*
 * clm_soil.h -- C ABI of libclmsoil.so: the MI355X (gfx950) implicit
 * soil-column step of ClimaLand.jl's RichardsModel and EnergyHydrology.
 *
 * The entry points replace, one for one, the hooks ClimaTimeSteppers calls
 * on a soil model (ClimaLand.jl v1.11.2):
 *
 *   clm_update_cache          <- make_update_cache(model)(p, Y, t)
 *                                src/shared_utilities/models.jl:187-195
 *   clm_exp_tendency          <- T_exp! = make_exp_tendency(model)
 *                                models.jl:279-287; rre.jl:214-243;
 *                                energy_hydrology.jl:315-350, 838-906
 *   clm_imp_tendency          <- T_imp! = make_compute_imp_tendency(model)
 *                                rre.jl:161-203; energy_hydrology.jl:363-425
 *   clm_jacobian              <- Wfact = make_compute_jacobian(model)
 *                                rre.jl:391-458; energy_hydrology.jl:466-576
 *   clm_solve                 <- ldiv!(Δx, jac_prototype, R) on the
 *                                FieldMatrixWithSolver built by
 *                                initialize_jacobian(Y),
 *                                src/shared_utilities/implicit_timestepping.jl:63-172
 *   clm_update_implicit_cache <- cache_imp! = make_update_implicit_cache(model)
 *                                models.jl:238-246; energy_hydrology.jl:427-455
 *   clm_step / clm_run        <- step!(sim) / solve!(sim) of a LandSimulation
 *                                with IMEXAlgorithm(ARS111, NewtonsMethod),
 *                                src/simulations/Simulations.jl:127-135,312-332
 *
 * Conventions
 *   - Host buffers are CSR-packed per column, level 0 = bottom cell:
 *     cells of column c are [col_ptr[c], col_ptr[c+1]); its N+1 faces are
 *     z_face[col_ptr[c] + c ... col_ptr[c+1] + c].
 *   - Per-column parameters are SoA: params[p * ncol + c], rows CLM_P_*.
 *   - All values are FP64 at the boundary, whatever the internal precision.
 *   - The library owns device memory; host pointers are borrowed for the
 *     duration of the call and copied.  Nothing allocates after
 *     clm_create except clm_set_forcing_table (which may grow its table).
 *   - Every function returns CLM_OK (0) or a negative CLM_E_* code and
 *     never aborts; clm_last_error(ctx) gives the message.
 *   - A context is used from one host thread at a time; its work is
 *     ordered on its own HIP stream.  Calls that return data to the host
 *     synchronise that stream.  Within a step the independent depth bins
 *     run on four internal streams that fork from that stream and join back
 *     into it (with events, so this also holds under graph capture): work
 *     queued on the ctx stream after a call sees all of the call's results.
 */
#ifndef CLM_SOIL_H
#define CLM_SOIL_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLM_ABI_VERSION 3

enum {
    CLM_OK = 0,
    CLM_E_ARG = -1,      /* invalid argument / shape */
    CLM_E_HIP = -2,      /* HIP runtime error (device, memory, launch) */
    CLM_E_DOMAIN = -3,   /* reference would raise: e.g. inverse_matric_potential(ψ>0) */
    CLM_E_STATE = -4     /* call order (e.g. clm_solve before clm_jacobian) */
};

enum { CLM_MODEL_RICHARDS = 0, CLM_MODEL_ENERGY_HYDROLOGY = 1 };

/* Water boundary conditions (boundary_conditions.jl).
 * Top:    CLM_BC_FLUX (WaterFluxBC, :168-180; EH: WaterHeatBC(WaterFluxBC,
 *         HeatFluxBC), :465-477, 709-724), CLM_BC_MOISTURE_STATE (RRE), or,
 *         for RRE, RichardsAtmosDrivenFluxBC(precipitation, runoff) (:79-212)
 *         with TOPMODELRunoff (CLM_BC_TOPMODEL) or SurfaceRunoff
 *         (CLM_BC_SURFACE_RUNOFF); with NoRunoff it is CLM_BC_FLUX with the
 *         precipitation as the flux.
 * Bottom: CLM_BC_FLUX, CLM_BC_FREE_DRAINAGE or CLM_BC_MOISTURE_STATE (RRE).
 * For CLM_BC_MOISTURE_STATE (MoistureStateBC, :227-325) the forcing value is
 * the boundary ϑ_l instead of a flux; with a top MoistureStateBC both
 * boundary fluxes follow the Newton iterate (rre.jl:460-468) and the
 * Jacobian's top row carries set_dfluxBCdY!'s term (:375-411).
 * For the runoff kinds the forcing value is the liquid precipitation rate
 * P_liq (negative downward); infiltration, and TOPMODEL's subsurface sink
 * (TOPMODELSubsurfaceRunoff, an implicit source: Runoff/Runoff.jl:174-183,
 * 321-332), are evaluated at u_n and held through the step. */
enum {
    CLM_BC_FLUX = 0,           /* prescribed fluxes */
    CLM_BC_FREE_DRAINAGE = 1,  /* FreeDrainage (RRE, :340-353) or
                                  EnergyWaterFreeDrainage (EH, :590-608) */
    CLM_BC_MOISTURE_STATE = 2, /* MoistureStateBC (RRE) */
    CLM_BC_TOPMODEL = 3,       /* top, RRE: TOPMODELRunoff (Runoff.jl:199-271) */
    CLM_BC_SURFACE_RUNOFF = 4, /* top, RRE: SurfaceRunoff (Runoff.jl:84-148) */
    /* top, EH: AtmosDrivenFluxBC(atmos, radiation, runoff) in soil-only form
     * (boundary_conditions.jl:646-1003, energy_hydrology.jl:909-945) with the
     * turbulent fluxes and net radiation supplied per column
     * (clm_set_surface_fluxes): NoRunoff, or TOPMODELRunoff with its
     * subsurface sinks of water and energy; adds the SoilSublimation source.
     * Group solver only. */
    CLM_BC_ATMOS = 5,
    CLM_BC_ATMOS_TOPMODEL = 6
};

/* Per-column parameter rows.  RRE uses the first CLM_NPARAM_RRE rows
 * (RichardsParameters, rre.jl:7-48); EH all CLM_NPARAM_EH rows
 * (EnergyHydrologyParameters, energy_hydrology.jl:21-78, with κ_dry,
 * κ_sat_frozen, κ_sat_unfrozen and ρc_ds as already-derived values). */
enum {
    CLM_P_NU = 0,
    CLM_P_THETA_R,
    CLM_P_VG_ALPHA,
    CLM_P_VG_N,
    CLM_P_KSAT,
    CLM_P_SS,
    CLM_P_NU_SS_OM,
    CLM_P_NU_SS_QUARTZ,
    CLM_P_NU_SS_GRAVEL,
    CLM_P_KAPPA_DRY,
    CLM_P_KAPPA_SAT_FROZEN,
    CLM_P_KAPPA_SAT_UNFROZEN,
    CLM_P_RHOC_DS,
    CLM_NPARAM_EH
};
#define CLM_NPARAM_RRE 6
/* RRE under CLM_BC_TOPMODEL: three more rows, TOPMODELRunoff's f_max (per
 * column), f_over (1/m) and R_sb (m/s) */
enum { CLM_P_F_MAX = 6, CLM_P_F_OVER = 7, CLM_P_R_SB = 8 };
#define CLM_NPARAM_RRE_TOPMODEL 9
/* EH under CLM_BC_ATMOS_TOPMODEL: the same three rows after the EH rows */
enum { CLM_P_EH_F_MAX = 13, CLM_P_EH_F_OVER = 14, CLM_P_EH_R_SB = 15 };
#define CLM_NPARAM_EH_TOPMODEL 16

/* Coarse-path solvers (clm_step / clm_run).  Both solve the same
 * block-tridiagonal Newton systems; they differ in how a column maps onto
 * the GPU.  AUTO picks GROUP for columns of up to 128 layers, else LANE. */
enum {
    CLM_SOLVER_AUTO = 0,
    CLM_SOLVER_GROUP = 1, /* 8-64 lanes per column, registers + parallel cyclic reduction */
    CLM_SOLVER_LANE = 2   /* one column per lane, Thomas through HBM scratch */
};

/* The Newton matrix of the EnergyHydrology group step (clm_step / clm_run).
 * PICARD: the reference's modified-Picard Jacobian (energy_hydrology.jl:
 * 466-576) -- the fine hooks always use it.  NEWTON: the exact derivative of
 * the same residual (no counterpart in the reference): the water block is the
 * reference's; the energy block adds the advection of heat by the Darcy flux
 * (∂(ρe_l K)/∂T) and takes ρc_s at the iterate's θ_l, and W_ρϑ carries ∂T/∂ϑ_l
 * and no "- I".  Same residual, same stopping rule, same root: an inexact
 * Newton that converges in fewer iterations.  A column whose Newton iterate
 * turns non-finite repeats the step with the reference's matrices (flag bit
 * 0x80000).  EnergyHydrology, group solver, prescribed-flux top boundary. */
enum { CLM_JACOBIAN_PICARD = 0, CLM_JACOBIAN_NEWTON = 1 };

typedef struct clm_ctx clm_ctx;

typedef struct {
    int32_t model;            /* CLM_MODEL_* */
    int32_t bc_bottom;        /* CLM_BC_* (bottom) */
    int64_t ncol;
    const int32_t* col_ptr;   /* ncol+1, col_ptr[0] = 0 */
    const double* z_face;     /* col_ptr[ncol] + ncol face heights, increasing */
    const double* params;     /* nparam * ncol */
    int32_t nparam;
    int32_t jac_precision;    /* 64: FP64 Jacobian+solve; 32: FP32 Jacobian+solve,
                                 residual/state/closures stay FP64 */
    int32_t newton_max_iters; /* NewtonsMethod(max_iters) */
    double newton_tol;        /* <= 0: exactly max_iters iterations (no convergence test);
                                 > 0: per-column stop when max|Δx| < tol over
                                 the cell unknowns (ρe_int scaled by 1/ρc_s) */
    int32_t solver;           /* CLM_SOLVER_*: how clm_step / clm_run solve */
    int32_t bc_top;           /* CLM_BC_FLUX, CLM_BC_MOISTURE_STATE, CLM_BC_TOPMODEL or
                                 CLM_BC_SURFACE_RUNOFF */
    int32_t jacobian;         /* CLM_JACOBIAN_* (0: the reference's) */
    int32_t reserved_;
} clm_config;

/* Prognostic state Y.soil.  RRE: vartheta_l, intF_water.
 * EH: vartheta_l, theta_i, rho_e_int, intF_water, intF_energy
 * (prognostic_vars, rre.jl:282, energy_hydrology.jl:624-625).
 * Unused members may be NULL. */
typedef struct {
    double* vartheta_l;
    double* theta_i;
    double* rho_e_int;
    double* intF_water;
    double* intF_energy;
} clm_state;

/* Cache p.soil members that the hooks read and write (auxiliary_vars,
 * rre.jl:292-305, energy_hydrology.jl:644-660).  Unused members may be NULL. */
typedef struct {
    double* K;
    double* psi;
    double* theta_l;
    double* T;
    double* kappa;
    double* top_water;   /* p.soil.top_bc(.water) */
    double* top_heat;    /* p.soil.top_bc.heat */
    double* bottom_water;
    double* bottom_heat;
    double* sink;        /* CLM_BC_TOPMODEL / ATMOS_TOPMODEL: subsurface runoff sink per
                            cell (1/s), (R_ss / max(h∇, eps)) · is_saturated
                            (Runoff.jl:321-332, 346-359) */
    double* r_ss;        /* p.soil.R_ss per column (m/s) */
    double* sink_energy; /* CLM_BC_ATMOS_TOPMODEL: (R_ess / max(h∇, eps)) · is_saturated
                            per cell (W/m3) */
    double* r_ess;       /* CLM_BC_ATMOS_TOPMODEL: p.soil.R_ess per column (W/m2) */
} clm_aux;

/* AtmosDrivenFluxBC inputs per column (CLM_BC_ATMOS*): what turbulent_fluxes!,
 * net_radiation! and the drivers give the soil-only boundary in the reference
 * (boundary_conditions.jl:880-915). */
typedef struct {
    const double* P_liq;          /* p.drivers.P_liq (m/s, negative downward) */
    const double* T_air;          /* p.drivers.T (K): energy of the infiltrating water */
    const double* lhf;            /* p.soil.turbulent_fluxes.lhf (W/m2) */
    const double* shf;            /* p.soil.turbulent_fluxes.shf (W/m2) */
    const double* vapor_flux_liq; /* p.soil.turbulent_fluxes.vapor_flux_liq (m/s) */
    const double* vapor_flux_ice; /* p.soil.turbulent_fluxes.vapor_flux_ice (m/s) */
    const double* R_n;            /* p.soil.R_n (W/m2) */
} clm_surface_fluxes;

typedef struct {
    int32_t max_iters_used;   /* max Newton iterations over columns */
    int32_t n_stalled;        /* column-steps whose increments met tol while the last
                                 residual stayed above 1e-3 (in ϑ_l, and in ρe_int/ρc_s
                                 in kelvin): stopped away from a root; counted in
                                 n_nonconverged too (saturates at INT32_MAX) */
    double mean_iters;        /* mean Newton iterations per column-step */
    int64_t n_nonconverged;   /* column-steps that hit max_iters with tol > 0, or stalled */
    int64_t n_nan;            /* columns with a non-finite state after the step */
    int64_t n_domain_error;   /* phase change reached inverse_matric_potential(ψ>0) */
    double ms;                /* device time of the step(s), HIP events */
} clm_step_stats;

/* ---- lifetime ---------------------------------------------------------- */
int clm_create(clm_ctx** out, const clm_config* cfg, void* hip_stream /* NULL: own */);
int clm_destroy(clm_ctx* ctx);
const char* clm_last_error(const clm_ctx* ctx); /* ctx may be NULL: last create error */
int clm_abi_version(void);

/* ---- state / forcing (host <-> device) -------------------------------- */
int clm_set_state(clm_ctx* ctx, const clm_state* Y);
int clm_get_state(clm_ctx* ctx, clm_state* Y);
/* Prescribed boundary fluxes (the values bc(p, t) returns), per column.
 * bottom_* are read only for CLM_BC_FLUX.  NULL arrays mean zero. */
int clm_set_forcing(clm_ctx* ctx, const double* top_water, const double* top_heat,
                    const double* bottom_water, const double* bottom_heat);
/* Device-resident forcing schedule for clm_run: nslots rows of ncol values
 * ([slot][col]); step s of a run uses slot (s / steps_per_slot) % nslots. */
int clm_set_forcing_table(clm_ctx* ctx, int32_t nslots, const double* top_water,
                          const double* top_heat);
/* CLM_BC_ATMOS*: the surface inputs at t_n for clm_step and the hooks, and
 * an [slot][col] schedule of them for clm_run (step s uses slot
 * (s / steps_per_slot) % nslots).  NULL members mean zero. */
int clm_set_surface_fluxes(clm_ctx* ctx, const clm_surface_fluxes* fluxes);
int clm_set_surface_flux_table(clm_ctx* ctx, int32_t nslots, const clm_surface_fluxes* fluxes);

/* ---- fine-grained hooks (1:1 with the ClimaLand closures) ----------------
 * Y / dY are host CSR buffers.  The cache p lives on the device inside ctx;
 * clm_get_aux copies it out. */
int clm_update_cache(clm_ctx* ctx, const clm_state* Y, double t);
int clm_exp_tendency(clm_ctx* ctx, const clm_state* Y, double t, clm_state* dY);
int clm_imp_tendency(clm_ctx* ctx, const clm_state* Y, double t, clm_state* dY);
int clm_jacobian(clm_ctx* ctx, const clm_state* Y, double dtgamma, double t);
int clm_solve(clm_ctx* ctx, clm_state* dx, const clm_state* R);
int clm_update_implicit_cache(clm_ctx* ctx, const clm_state* Y, double t);
int clm_get_aux(clm_ctx* ctx, clm_aux* p);
/* Jacobian rows as assembled by clm_jacobian, CSR cells x 3 (lower, diag,
 * upper) per block: ϑϑ; EH also ρeρe and ρeϑ (NULL to skip).  FP64 even when
 * jac_precision = 32 (values widened). */
int clm_get_jacobian(clm_ctx* ctx, double* W_ww, double* W_ee, double* W_ew);

/* ---- coarse path: whole ARS111 steps on the device ---------------------- */
/* One step of size dt from the state set by clm_set_state, forcing from
 * clm_set_forcing.  stats may be NULL (then nothing synchronises). */
int clm_step(clm_ctx* ctx, double t, double dt, clm_step_stats* stats);
/* nsteps steps with the device forcing table; the whole run is one graph
 * replay after the first call with the same (nsteps, steps_per_slot).
 * iters_out: NULL or host [nsteps][ncol] Newton iteration counts in the low
 * 16 bits, with flags: 0x10000 not converged (max_iters reached, or stalled),
 * 0x20000 non-finite iterate, 0x40000 inverse_matric_potential domain error,
 * 0x80000 the step was repeated from U0 with the reference's iteration (a
 * warm-started or Newton-matrix attempt went non-finite or stalled; the count
 * includes both attempts), 0x100000 stalled: max|Δ| < tol with the last
 * residual above 1e-3 (in ϑ_l, and in ρe_int/ρc_s in kelvin). */
int clm_run(clm_ctx* ctx, double t0, double dt, int32_t nsteps, int32_t steps_per_slot,
            clm_step_stats* stats, int32_t* iters_out);

/* clm_step without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same kernels and
 * results as clm_step; the Newton flags and statistics live in a static
 * device scratch of the library (one per GPU, shared by every context on
 * it, kept until the process ends): at most 1,048,576 columns (group
 * solver) or column slots (lane solver), else CLM_E_ARG.  Calls that overlap
 * on one GPU race on that scratch. */
int clm_step_noalloc(clm_ctx* ctx, double t, double dt, clm_step_stats* stats);

/* Column totals Σ ϑ_l Δz (+ ρ_i/ρ_l θ_i Δz) and Σ ρe_int Δz of the current
 * device state (total_liq_water_vol_per_area!, rre.jl:502-511,
 * energy_hydrology.jl:1282-1320).  energy may be NULL. */
int clm_totals(clm_ctx* ctx, double* water, double* energy);

/* Surface soil conductance to water vapour of the current device state, per
 * column (EnergyHydrology, group solver): the θ_l of the top two cells
 * extrapolated to the surface (linear_interpolation_to_surface!,
 * Domains.jl:883-895) and floored at θ_r + eps, its effective saturation,
 * then soil_conductance(S_l, S_c, d_ds, evap_p, evap_α, D_vapor) =
 * D_vapor / max(dry_soil_layer_thickness(S_l, evap_α S_c, d_ds, evap_p), eps)
 * (energy_hydrology.jl:1103-1131, 1322-1368), in m/s; dry_layer (may be NULL)
 * receives the dry-layer thickness (m).  The turbulent fluxes that use it
 * (SurfaceFluxes.jl) stay with the caller. */
int clm_soil_conductance(clm_ctx* ctx, double d_ds, double evap_p, double evap_alpha,
                         double D_vapor, double* g_soil, double* dry_layer);

/* ---- group-solver shapes (CLM_SOLVER_GROUP) --------------------------------
 * Columns are bucketed by depth (max layers 8, 16, 24, 32, 40, 48, 64, 80,
 * 96, 128); each bucket runs with one shape: G lanes per column, R rows per
 * lane, G*R >= the bucket's deepest column.  Records are
 * (bucket max layers, G, R, columns in bucket). */
int clm_get_shapes(clm_ctx* ctx, int32_t* records /* 4 x max_records */,
                   int32_t max_records, int32_t* nrecords);
/* Force a bucket's shape; (G, R) must be one of the compiled shapes. */
int clm_set_shape(clm_ctx* ctx, int32_t bucket_max_layers, int32_t G, int32_t R);
/* Paired waves (no counterpart in the reference; a launch layout): with the
 * buckets of <= 40 and <= 24 layers on shapes (64, 1) and (32, 1), an
 * EnergyHydrology step with a prescribed-flux top boundary runs one column
 * of each in a wave (40 + 24 lanes) instead of leaving lanes idle.  Results
 * are those of the separate launches, bit for bit.  On by default;
 * clm_get_pairing reports whether the current shapes use it. */
int clm_set_pairing(clm_ctx* ctx, int32_t on);
int clm_get_pairing(const clm_ctx* ctx, int32_t* active);
/* Autotuner: for every non-empty bucket, time each compiled shape that fits
 * (best of reps trials, each four consecutive steps from the current state
 * and clm_set_forcing values, regrouped by Newton count between steps; the
 * state is restored afterwards, the bins keep the regrouped column order)
 * and keep the fastest.  Invalidates the captured run graph.  Allocates a
 * temporary state copy. */
int clm_autotune(clm_ctx* ctx, double dt, int32_t reps);

/* Device time of the dominant kernel in the last clm_step/clm_run, per
 * launch (ms), measured with HIP events on the ctx stream. */
int clm_last_kernel_ms(const clm_ctx* ctx, double* ms_per_launch, int32_t* launches);

/* ---- checkpoints (save_checkpoint / read_checkpoint, checkpoints.jl:125-195) ----
 * clm_save_checkpoint writes the device state and time t to
 * <dir>/day<D>.<S>.clmchk (D = floor(t / 86400 s), S = floor(t mod 86400 s))
 * and returns that path in `path` (may be NULL).  clm_read_checkpoint loads a
 * checkpoint of the same model kind and column geometry into the device
 * state and returns its time; *same_model = 0 where the file was written by
 * a model with other parameters or grid (the reference warns and goes on).
 * File layout: 64-byte header (magic "CLMCHK01", model, bc_bottom, bc_top,
 * pad, ncol, ncell, t, model hash), then float64 vartheta_l[ncell],
 * intF_water[ncol] and, for EnergyHydrology, theta_i[ncell],
 * rho_e_int[ncell], intF_energy[ncol]. */
int clm_save_checkpoint(clm_ctx* ctx, const char* dir, double t, char* path, int32_t path_len);
int clm_read_checkpoint(clm_ctx* ctx, const char* path, double* t, int32_t* same_model);

/* ---- learned Newton warm start (group solver) ---------------------------------
 * A per-column MLP predicts the implicit part of the step's increment,
 * x_{n+1} - U0, at 32 nodes (layer-index resampling); Newton then starts at
 * x0 = U0 + guess and iterates to the configured tolerance, falling back to
 * x0 = U0 when the warm-started iterate turns non-finite (flag bit 0x80000;
 * iteration counts then include both attempts).  The network is the
 * project's own: BF16 weights/activations with FP32 accumulation on MFMA,
 * GELU, layers F -> H -> H -> H -> Nout with (F, H, Nout) = (105, 384, 64)
 * for EnergyHydrology (inputs padded to 128) and (40, 256, 32) for
 * RichardsModel (padded to 64).  No counterpart in the reference: it replaces
 * Newton's initial guess x0 = U0 (the predictor of ARS111 in
 * ClimaTimeSteppers' NewtonsMethod) and nothing else.
 * Parameters are one FP32 vector: per layer Wt [N x K] row-major, then b [N]. */
int clm_ws_init(clm_ctx* ctx, uint64_t seed);            /* allocate + random init */
int clm_ws_dims(clm_ctx* ctx, int32_t* dims /* F, Fp, H, Nout */, int64_t* nparams);
int clm_ws_get_params(clm_ctx* ctx, float* params);
int clm_ws_set_params(clm_ctx* ctx, const float* params);
/* Use (1) or stop using (0) the warm start in clm_step / clm_run. */
int clm_ws_enable(clm_ctx* ctx, int32_t on);
/* Training pairs: advance the device state nsteps plain steps with the
 * clm_set_forcing_table slots and, for every stride-th step, append
 * (features of u_n, node targets of x_{n+1} - U0, converged mask) per
 * column; returns the total row count. */
int clm_ws_collect(clm_ctx* ctx, double dt, int32_t nsteps, int32_t steps_per_slot,
                   int32_t stride, int64_t* rows);
/* Adam on the collected rows (masked MSE), minibatch `batch` (multiple of
 * 512), step size lr decayed over the epochs as lr (1 + cos(π ep/epochs))/2,
 * one loss value per epoch. */
int clm_ws_train(clm_ctx* ctx, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                 float* loss_per_epoch);
/* Checking hooks: features of the device state ([ncol x Fp], with the
 * clm_set_forcing fluxes), a forward pass of host rows X [M x Fp] -> Y
 * [M x Nout], and the loss and parameter gradient for host rows (M % 32 == 0). */
int clm_ws_features(clm_ctx* ctx, double dt, float* X);
int clm_ws_forward(clm_ctx* ctx, const float* X, int64_t M, float* Y);
int clm_ws_gradient(clm_ctx* ctx, const float* X, const float* T, const float* mask, int64_t M,
                    float* loss, float* grad);
/* The warm start's guess x0 - U0 for the device state and the clm_set_forcing
 * fluxes, per cell in CSR order: ϑ_l and (EH) ρe_int.  fused = 1 runs the
 * inference that clm_step / clm_run use (the feature kernel, then the four
 * layers in one kernel); fused = 0 the features, layer-by-layer GEMMs and
 * guess kernels (the same arithmetic, for checking). */
int clm_ws_guess(clm_ctx* ctx, double dt, int32_t fused, double* guess_vartheta_l,
                 double* guess_rho_e_int);
/* Device time of one warm-start inference as a step runs it (the feature
 * rows, then the network's output rows; the step kernels form the per-cell
 * guesses from those rows) on the current state and clm_set_forcing fluxes,
 * averaged over reps (HIP events). */
int clm_ws_time(clm_ctx* ctx, double dt, int32_t reps, double* ms_per_call);

#ifdef __cplusplus
}
#endif
#endif
