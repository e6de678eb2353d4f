// clm_rre.h -- RichardsModel kernels (one column per lane, packed layout).
//
// Discretisation (ClimaCore 0.15.2 operators as used in rre.jl:161-203 and
// rre.jl:391-458), per column with faces z_f[0..N] (bottom..top):
//   z_c[i] = (z_f[i] + z_f[i+1]) / 2,  Δz_c[i] = z_f[i+1] - z_f[i],
//   Δz_f[i] = z_c[i] - z_c[i-1]   (interior faces i = 1..N-1)
//   F[i]  = -(K[i-1]+K[i])/2 · ((ψ+z)[i] - (ψ+z)[i-1]) / Δz_f[i],
//   F[0]  = bottom flux, F[N] = top flux          (DivergenceF2C SetValue)
//   T[i]  = -(F[i+1] - F[i]) / Δz_c[i]
//   W     = dtγ ∂T/∂ϑ|_{K frozen} - I  (modified Picard; boundary faces 0)
#pragma once
#include "clm_layout.h"
#include "clm_physics.h"

namespace clm {

// BC_TOPMODEL and BC_SURFACE_RUNOFF are top kinds of RichardsModel:
// RichardsAtmosDrivenFluxBC(precipitation, runoff) (boundary_conditions.jl:79-212)
// with TOPMODELRunoff or SurfaceRunoff; the forcing value is the precipitation
enum { BC_FLUX = 0, BC_FREE_DRAINAGE = 1, BC_MOISTURE_STATE = 2, BC_TOPMODEL = 3,
       BC_SURFACE_RUNOFF = 4,
       // EnergyHydrology top: AtmosDrivenFluxBC in soil-only form with the
       // surface fluxes supplied per column, NoRunoff or TOPMODELRunoff
       BC_ATMOS = 5, BC_ATMOS_TOPMODEL = 6 };
// RichardsModel kernels take both boundary kinds packed as bottom | top << 4
__host__ __device__ inline int bc_bot(int bc) { return bc & 15; }
__host__ __device__ inline int bc_top(int bc) { return bc >> 4; }

// MoistureStateBC (boundary_conditions.jl:227-325): diffusive_flux(K_eff,
// x_2, x_1, Δz) = -K_eff (x_2 - x_1) / Δz with K_eff the boundary cell's K,
// ψ_bc = pressure_head(ϑ_bc) with that cell's parameters, and gravity as
// ψ + Δz on the upper point; Δz is the centre-to-boundary distance.
__device__ __forceinline__ double moisture_flux_top(const VG& v, double nu, double theta_bc,
                                                    double K_top, double psi_top, double dz) {
    double psib = psi_dpsi(v, nu, theta_bc).psi;
    return -K_top * ((psib + dz) - psi_top) / dz;
}
__device__ __forceinline__ double moisture_flux_bot(const VG& v, double nu, double theta_bc,
                                                    double K_bot, double psi_bot, double dz) {
    double psib = psi_dpsi(v, nu, theta_bc).psi;
    return -K_bot * ((psi_bot + dz) - psib) / dz;
}

// ---- runoff (Runoff/Runoff.jl) -------------------------------------------------
// saturation weight of a layer, is_saturated(ϑ_l - θ_r, ν - θ_r) (ϑ_l - θ_r)/(ν - θ_r)
// (Runoff.jl:261-262, 438-440; heaviside(x) = x > eps, utils.jl:93-99)
__device__ __forceinline__ double sat_weight(double vl, double nu, double thr) {
    double twc = vl - thr, nsat = nu - thr;
    return (twc - nsat > DEPS) ? twc / nsat : 0.0;
}
struct Runoff {
    double infil;  // top boundary flux (infiltration)
    double rss;    // subsurface runoff R_ss (m/s)
    double scale;  // sink of a layer = scale · saturation weight
};
// TOPMODEL (Runoff.jl:234-271, 373-376, 386-389, 427-429; Niu et al. 2005
// eqs 8, 11, 12) from h∇ = Σ weight Δz_c: z∇ = depth - h∇,
// f_sat = min(f_max e^{-f_over z∇/2}, 1), infiltration = (1 - f_sat) max(-K_sat, P),
// R_ss = R_sb e^{-f_over z∇}; the source (Runoff.jl:321-332) removes
// R_ss / max(h∇, eps) · weight from each layer and R_ss from ∫F
__device__ __forceinline__ Runoff topmodel(double h, double depth, double f_max, double f_over,
                                           double R_sb, double Ksat, double precip) {
    Runoff o;
    double zw = depth - h;
    double fsat = fmin(f_max * exp(-f_over / 2.0 * zw), 1.0);
    o.infil = (1.0 - fsat) * fmax(-1.0 * Ksat, precip);
    o.rss = R_sb * exp(-f_over * zw);
    o.scale = o.rss / fmax(h, DEPS);
    return o;
}
// SurfaceRunoff (Runoff.jl:107-148): (1 - is_saturated(ϑ_top, ν)) max(-K_sat, P)
__device__ __forceinline__ double surface_infiltration(double vl_top, double nu, double Ksat,
                                                       double precip) {
    return (vl_top - nu > DEPS) ? 0.0 : fmax(-1.0 * Ksat, precip);
}

// per-slot parameter rows (same order as CLM_P_* in include/clm_soil.h)
enum {
    P_NU = 0, P_THETA_R, P_VG_ALPHA, P_VG_N, P_KSAT, P_SS,
    P_NU_SS_OM, P_NU_SS_QUARTZ, P_NU_SS_GRAVEL, P_KAPPA_DRY,
    P_KAPPA_SAT_FROZEN, P_KAPPA_SAT_UNFROZEN, P_RHOC_DS, NPARAM_EH
};
// RichardsModel under a TOPMODEL top boundary: rows 6-8
enum { P_F_MAX = 6, P_F_OVER = 7, P_R_SB = 8, NPARAM_RRE_TOPMODEL = 9 };
// EnergyHydrology under AtmosDrivenFluxBC + TOPMODEL: rows 13-15
enum { P_EH_F_MAX = 13, P_EH_F_OVER = 14, P_EH_R_SB = 15, NPARAM_MAX = 16 };

__device__ __forceinline__ bool lane_column(const Layout& L, ColView& cv) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= L.nslot) return false;
    int32_t N = L.nl[s];
    if (N <= 0) return false;
    int64_t p = s >> 6;
    int l = (int)(s & 63);
    cv.cb = L.cell_off[p] + l;
    cv.fb = L.face_off[p] + l;
    cv.N = N;
    cv.s = s;
    return true;
}

__device__ __forceinline__ VG load_vg(const double* __restrict__ par, int64_t ns, int64_t s) {
    return make_vg(par[P_VG_ALPHA * ns + s], par[P_VG_N * ns + s], par[P_THETA_R * ns + s],
                   par[P_SS * ns + s], par[P_KSAT * ns + s]);
}

// ---- fine-grained hooks -----------------------------------------------------

// make_update_cache (rre.jl:368-380 update_aux + :111-149 boundary fluxes);
// cache_imp! with do_bc = 0 (boundary fluxes lagged) or, under a top
// MoistureStateBC, do_bc = 1 (rre.jl:460-468)
__global__ __launch_bounds__(256) void k_rre_update_cache(
    Layout L, const double* __restrict__ par, const double* __restrict__ zf,
    const double* __restrict__ vl, double* __restrict__ K, double* __restrict__ psi,
    const double* __restrict__ f_top, const double* __restrict__ f_bot, double* __restrict__ tw,
    double* __restrict__ bw, int bc, int do_bc, double* __restrict__ sink,
    double* __restrict__ rss) {
    ColView cv;
    if (!lane_column(L, cv)) return;
    VG v = load_vg(par, L.nslot, cv.s);
    double nu = par[P_NU * L.nslot + cv.s];
    const double thr = par[P_THETA_R * L.nslot + cv.s];
    const bool topm = do_bc && bc_top(bc) == BC_TOPMODEL;
    double K0 = 0.0, p0 = 0.0, KN = 0.0, pN = 0.0, xN = 0.0, h = 0.0;
    for (int j = 0; j < cv.N; ++j) {
        int64_t c = cv.cb + (int64_t)kWave * j;
        double x = vl[c];
        PsiD r = psi_dpsi(v, nu, x);
        double k = hyd_cond_from_a(v, r.S, r.a);
        K[c] = k;
        psi[c] = r.psi;
        if (j == 0) {
            K0 = k;
            p0 = r.psi;
        }
        KN = k;
        pN = r.psi;
        xN = x;
        if (topm) {
            double wgt = sat_weight(x, nu, thr);
            sink[c] = wgt;
            h += wgt * (zf[cv.fb + (int64_t)kWave * (j + 1)] - zf[cv.fb + (int64_t)kWave * j]);
        }
    }
    if (do_bc) {
        const int N = cv.N;
        if (topm) {
            const int64_t ns = L.nslot;
            Runoff o = topmodel(h, zf[cv.fb + (int64_t)kWave * N] - zf[cv.fb],
                                par[P_F_MAX * ns + cv.s], par[P_F_OVER * ns + cv.s],
                                par[P_R_SB * ns + cv.s], v.Ksat, f_top[cv.s]);
            tw[cv.s] = o.infil;
            rss[cv.s] = o.rss;
            for (int j = 0; j < N; ++j) sink[cv.cb + (int64_t)kWave * j] *= o.scale;
        } else if (bc_top(bc) == BC_SURFACE_RUNOFF) {
            tw[cv.s] = surface_infiltration(xN, nu, v.Ksat, f_top[cv.s]);
        } else if (bc_top(bc) == BC_MOISTURE_STATE) {
            double dz = 0.5 * (zf[cv.fb + (int64_t)kWave * N] - zf[cv.fb + (int64_t)kWave * (N - 1)]);
            tw[cv.s] = moisture_flux_top(v, nu, f_top[cv.s], KN, pN, dz);
        } else {
            tw[cv.s] = f_top[cv.s];  // WaterFluxBC
        }
        if (bc_bot(bc) == BC_FREE_DRAINAGE) {
            bw[cv.s] = -1.0 * K0;
        } else if (bc_bot(bc) == BC_MOISTURE_STATE) {
            double dz = 0.5 * (zf[cv.fb + kWave] - zf[cv.fb]);
            bw[cv.s] = moisture_flux_bot(v, nu, f_bot[cv.s], K0, p0, dz);
        } else {
            bw[cv.s] = f_bot[cv.s];
        }
    }
}

// compute_imp_tendency (rre.jl:161-203)
__global__ __launch_bounds__(256) void k_rre_imp_tendency(
    Layout L, const double* __restrict__ zf, const double* __restrict__ K,
    const double* __restrict__ psi, const double* __restrict__ tw,
    const double* __restrict__ bw, double* __restrict__ dvl, double* __restrict__ diw,
    const double* __restrict__ sink, const double* __restrict__ rss) {
    // sink / rss: TOPMODEL subsurface runoff (implicit source, Runoff.jl:321-332), or null
    ColView cv;
    if (!lane_column(L, cv)) return;
    const int N = cv.N;
    double top = tw[cv.s], bot = bw[cv.s];
    diw[cv.s] = sink ? -(top - bot) - rss[cv.s] : -(top - bot);
    double f0 = zf[cv.fb], f1 = zf[cv.fb + kWave];
    double zc_prev = 0.5 * (f0 + f1);
    double Flo = bot;
    double K_prev = K[cv.cb], h_prev = psi[cv.cb] + zc_prev;
    double dzc_prev = f1 - f0;
    for (int i = 0; i < N; ++i) {
        double Fhi, dzc_i = dzc_prev;
        if (i == N - 1) {
            Fhi = top;
        } else {
            int64_t c = cv.cb + (int64_t)kWave * (i + 1);
            double f2 = zf[cv.fb + (int64_t)kWave * (i + 2)];
            double zc_n = 0.5 * (f1 + f2);
            double Kn = K[c], hn = psi[c] + zc_n;
            Fhi = -0.5 * (K_prev + Kn) * (hn - h_prev) / (zc_n - zc_prev);
            dzc_prev = f2 - f1;
            f1 = f2;
            zc_prev = zc_n;
            K_prev = Kn;
            h_prev = hn;
        }
        double t = -(Fhi - Flo) / dzc_i;
        if (sink) t -= sink[cv.cb + (int64_t)kWave * i];
        dvl[cv.cb + (int64_t)kWave * i] = t;
        Flo = Fhi;
    }
}

// compute_jacobian! (rre.jl:391-458): rows (lo, di, up) of W
// Under a top MoistureStateBC the top row's diagonal gains
// -dtγ K_N dψ/dϑ_N / (Δz_c,N Δz_top) (set_dfluxBCdY!, boundary_conditions.jl:375-411,
// added in rre.jl:434-447).
template <typename JT>
__global__ __launch_bounds__(256) void k_rre_jacobian(
    Layout L, const double* __restrict__ par, const double* __restrict__ zf,
    const double* __restrict__ vl, const double* __restrict__ K, double dtg,
    JT* __restrict__ Wlo, JT* __restrict__ Wdi, JT* __restrict__ Wup, int bc) {
    ColView cv;
    if (!lane_column(L, cv)) return;
    const int N = cv.N;
    VG v = load_vg(par, L.nslot, cv.s);
    double nu = par[P_NU * L.nslot + cv.s];
    double f0 = zf[cv.fb], f1 = zf[cv.fb + kWave];
    double zc_c = 0.5 * (f0 + f1), dzc_c = f1 - f0;
    double K_c = K[cv.cb], D_c = psi_dpsi(v, nu, vl[cv.cb]).D;
    double D_m = 0.0, wlo = 0.0;
    for (int i = 0; i < N; ++i) {
        double wup = 0.0, D_p = 0.0, zc_n = 0.0, dzc_n = 0.0, K_n = 0.0;
        if (i < N - 1) {
            int64_t c = cv.cb + (int64_t)kWave * (i + 1);
            double f2 = zf[cv.fb + (int64_t)kWave * (i + 2)];
            zc_n = 0.5 * (f1 + f2);
            dzc_n = f2 - f1;
            K_n = K[c];
            D_p = psi_dpsi(v, nu, vl[c]).D;
            wup = 0.5 * (K_c + K_n) / (zc_n - zc_c);
            f1 = f2;
        }
        double s = dtg / dzc_c;
        int64_t c = cv.cb + (int64_t)kWave * i;
        double wbc = (i == N - 1 && bc_top(bc) == BC_MOISTURE_STATE) ? K_c / (0.5 * dzc_c) : 0.0;
        Wlo[c] = (JT)(s * wlo * D_m);
        Wdi[c] = (JT)(-s * (wlo + wup + wbc) * D_c - 1.0);
        Wup[c] = (JT)(s * wup * D_p);
        D_m = D_c;
        D_c = D_p;
        K_c = K_n;
        zc_c = zc_n;
        dzc_c = dzc_n;
        wlo = wup;
    }
}

// ldiv! for one tridiagonal block: Thomas without pivoting, one column per
// lane.  cp/dp scratch in the packed cell layout.
template <typename JT>
__device__ __forceinline__ void thomas_lane(const ColView& cv, const JT* __restrict__ Wlo,
                                            const JT* __restrict__ Wdi,
                                            const JT* __restrict__ Wup,
                                            const double* __restrict__ b,
                                            double* __restrict__ x, JT* __restrict__ cp,
                                            JT* __restrict__ dp) {
    const int N = cv.N;
    JT cpp = 0, dpp = 0;
    for (int i = 0; i < N; ++i) {
        int64_t c = cv.cb + (int64_t)kWave * i;
        JT lo = Wlo[c], di = Wdi[c], up = Wup[c];
        JT inv = (JT)1 / (di - lo * cpp);
        cpp = up * inv;
        dpp = ((JT)b[c] - lo * dpp) * inv;
        cp[c] = cpp;
        dp[c] = dpp;
    }
    JT xn = 0;
    for (int i = N - 1; i >= 0; --i) {
        int64_t c = cv.cb + (int64_t)kWave * i;
        xn = dp[c] - cp[c] * xn;
        x[c] = (double)xn;
    }
}

template <typename JT>
__global__ __launch_bounds__(256) void k_rre_solve(Layout L, const JT* __restrict__ Wlo,
                                                   const JT* __restrict__ Wdi,
                                                   const JT* __restrict__ Wup,
                                                   const double* __restrict__ rv,
                                                   const double* __restrict__ riw,
                                                   double* __restrict__ xv,
                                                   double* __restrict__ xiw,
                                                   JT* __restrict__ cp, JT* __restrict__ dp) {
    ColView cv;
    if (!lane_column(L, cv)) return;
    thomas_lane<JT>(cv, Wlo, Wdi, Wup, rv, xv, cp, dp);
    xiw[cv.s] = -riw[cv.s];  // -I block (implicit_timestepping.jl:149-152)
}

// ---- fused ARS111 step (explicit stage + whole Newton loop) -----------------
//
// One lane owns one column for the whole step.  Per Newton iteration:
// an upward sweep evaluates ψ, K, dψ/dϑ of each level once (sliding
// three-level window), forms the residual R = U0 + Δt T_imp(x) - x and the
// Picard row, and runs the Thomas forward elimination; a downward sweep
// back-substitutes Δx, updates x and measures max|Δx|.  Aux refresh
// (cache_imp! = update_aux for RRE, models.jl:207-210) is implicit: the
// next sweep re-evaluates the closures at the new x.  Boundary fluxes are
// lagged at u_n (rre.jl:460-468).  Iteration 1 also saves U0 (= u_n: the
// RRE explicit tendency is zero, rre.jl:214-243).
//
// iters[s] = Newton iterations | 0x10000 if not converged | 0x20000 if non-finite
// Under a top MoistureStateBC both boundary fluxes follow the iterate
// (rre.jl:460-468) and the top row carries the boundary term.
template <typename JT>
__global__ __launch_bounds__(256) void k_rre_step(
    Layout L, const double* __restrict__ par, const double* __restrict__ zf,
    double* __restrict__ x, double* __restrict__ U0, double* __restrict__ iw,
    const double* __restrict__ f_top, const double* __restrict__ f_bot, int bc,
    double dt, int max_iters, double tol, JT* __restrict__ cp, JT* __restrict__ dp,
    int32_t* __restrict__ iters) {
    ColView cv;
    if (!lane_column(L, cv)) return;
    const int N = cv.N;
    const int64_t cb = cv.cb, fb = cv.fb;
    VG v = load_vg(par, L.nslot, cv.s);
    const double nu = par[P_NU * L.nslot + cv.s];
    const int bc_bottom = bc_bot(bc);
    const bool top_state = bc_top(bc) == BC_MOISTURE_STATE;
    const bool topm = bc_top(bc) == BC_TOPMODEL;
    const double ftop = f_top[cv.s];
    double top = ftop;
    double bot = (bc_bottom == BC_FLUX) ? f_bot[cv.s] : 0.0;
    const double iw0 = iw[cv.s];
    const double thr = par[P_THETA_R * L.nslot + cv.s];
    // runoff from update_cache!(u_n), held for the step (Runoff.jl:234-271)
    double sscale = 0.0, rss = 0.0;
    if (topm) {
        double h = 0.0;
        for (int i = 0; i < N; ++i)
            h += sat_weight(x[cb + (int64_t)kWave * i], nu, thr) *
                 (zf[fb + (int64_t)kWave * (i + 1)] - zf[fb + (int64_t)kWave * i]);
        const int64_t ns = L.nslot;
        Runoff o = topmodel(h, zf[fb + (int64_t)kWave * N] - zf[fb], par[P_F_MAX * ns + cv.s],
                            par[P_F_OVER * ns + cv.s], par[P_R_SB * ns + cv.s], v.Ksat, ftop);
        top = o.infil;
        rss = o.rss;
        sscale = o.scale;
    } else if (bc_top(bc) == BC_SURFACE_RUNOFF) {
        top = surface_infiltration(x[cb + (int64_t)kWave * (N - 1)], nu, v.Ksat, ftop);
    }

    int k = 0;
    bool conv = false;
    double err = 0.0;
    for (k = 1; k <= max_iters; ++k) {
        const double* __restrict__ u0 = (k == 1) ? x : U0;
        // window: c = level i, n = level i+1
        double f0 = zf[fb], f1 = zf[fb + kWave];
        double zc_c = 0.5 * (f0 + f1), dzc_c = f1 - f0;
        double x_c = x[cb];
        PsiD r = psi_dpsi(v, nu, x_c);
        double K_c = hyd_cond_from_a(v, r.S, r.a), D_c = r.D, h_c = r.psi + zc_c;
        if (k == 1 || top_state) {  // bottom flux at u_n, or at the iterate under a top state BC
            if (bc_bottom == BC_FREE_DRAINAGE)
                bot = -1.0 * K_c;
            else if (bc_bottom == BC_MOISTURE_STATE)
                bot = moisture_flux_bot(v, nu, f_bot[cv.s], K_c, r.psi, 0.5 * dzc_c);
        }
        double Flo = bot, wlo = 0.0, D_m = 0.0;
        JT cpp = 0, dpp = 0;
        for (int i = 0; i < N; ++i) {
            const int64_t c = cb + (int64_t)kWave * i;
            double Fhi, wup = 0.0, D_n = 0.0, K_n = 0.0, h_n = 0.0, x_n = 0.0;
            double zc_n = 0.0, dzc_n = 0.0;
            if (i < N - 1) {
                double f2 = zf[fb + (int64_t)kWave * (i + 2)];
                x_n = x[c + kWave];
                zc_n = 0.5 * (f1 + f2);
                dzc_n = f2 - f1;
                PsiD rn = psi_dpsi(v, nu, x_n);
                K_n = hyd_cond_from_a(v, rn.S, rn.a);
                D_n = rn.D;
                h_n = rn.psi + zc_n;
                double Kf = 0.5 * (K_c + K_n);
                double idzf = 1.0 / (zc_n - zc_c);
                wup = Kf * idzf;
                Fhi = -Kf * (h_n - h_c) * idzf;
                f1 = f2;
            } else {
                if (top_state) top = moisture_flux_top(v, nu, ftop, K_c, h_c - zc_c, 0.5 * dzc_c);
                Fhi = top;
            }
            double idzc = 1.0 / dzc_c;
            double u0c = u0[c];
            if (k == 1) U0[c] = u0c;
            double R = u0c + dt * (-(Fhi - Flo) * idzc) - x_c;
            if (topm) R -= dt * (sscale * sat_weight(u0c, nu, thr));
            double s = dt * idzc;
            double wbc = (i == N - 1 && top_state) ? K_c / (0.5 * dzc_c) : 0.0;
            JT lo = (JT)(s * wlo * D_m);
            JT di = (JT)(-s * (wlo + wup + wbc) * D_c - 1.0);
            JT up = (JT)(s * wup * D_n);
            JT inv = (JT)1 / (di - lo * cpp);
            cpp = up * inv;
            dpp = ((JT)R - lo * dpp) * inv;
            cp[c] = cpp;
            dp[c] = dpp;
            // shift window
            Flo = Fhi;
            wlo = wup;
            D_m = D_c;
            D_c = D_n;
            K_c = K_n;
            h_c = h_n;
            x_c = x_n;
            zc_c = zc_n;
            dzc_c = dzc_n;
        }
        // back substitution + update
        err = 0.0;
        JT dxn = 0;
        for (int i = N - 1; i >= 0; --i) {
            const int64_t c = cb + (int64_t)kWave * i;
            dxn = dp[c] - cp[c] * dxn;
            double d = (double)dxn;
            x[c] = x[c] - d;
            err = nan_max(err, fabs(d));
        }
        iw[cv.s] = iw0 + dt * (-(top - bot) - rss);  // ∫F: -I block, fluxes of this iteration
        conv = (tol > 0.0) && (err < tol);
        // a non-finite iterate stays non-finite: stop (the state is NaN either way)
        if (conv || !isfinite(err)) break;
    }
    if (k > max_iters) k = max_iters;
    int32_t flag = k;
    if (tol > 0.0 && !conv) flag |= 0x10000;
    if (!isfinite(err)) flag |= 0x20000;
    iters[cv.s] = flag;
}

}  // namespace clm
