// clm_eh.h -- EnergyHydrology kernels (one column per lane, packed layout).
//
// Prognostics ϑ_l, θ_i, ρe_int (+ ∫F_water, ∫F_energy).  Explicit stage:
// update_aux (energy_hydrology.jl:722-814) + PhaseChange source!
// (:838-906).  Implicit: Darcy + conduction + advected liquid energy
// (:363-425) with K, κ, θ_l lagged at u_n and T, ψ refreshed by cache_imp!
// (:427-445).  Jacobian (:466-576): A_ϑϑ (dψ/dϑ with ν_eff = ν-θ_i),
// A_ρϑ (ρe_l(T)K; its diagonal carries the reference's -I), A_ρρ (κ,
// 1/ρc_s(θ_l, θ_i)); solved block-lower (implicit_timestepping.jl:161-169).
#pragma once
#include "clm_rre.h"

namespace clm {

struct EHPar {
    VG v;
    double nu, om, qz, gr, k_dry, k_satf, k_satu, rhoc_ds;
    double lk_u, lk_f;  // log κ_sat_unfrozen, log κ_sat_frozen
};

__device__ __forceinline__ EHPar load_eh(const double* __restrict__ par, int64_t ns, int64_t s) {
    EHPar e;
    e.v = load_vg(par, ns, s);
    e.nu = par[P_NU * ns + s];
    e.om = par[P_NU_SS_OM * ns + s];
    e.qz = par[P_NU_SS_QUARTZ * ns + s];
    e.gr = par[P_NU_SS_GRAVEL * ns + s];
    e.k_dry = par[P_KAPPA_DRY * ns + s];
    e.k_satf = par[P_KAPPA_SAT_FROZEN * ns + s];
    e.k_satu = par[P_KAPPA_SAT_UNFROZEN * ns + s];
    e.rhoc_ds = par[P_RHOC_DS * ns + s];
    e.lk_u = clm_log(e.k_satu);
    e.lk_f = clm_log(e.k_satf);
    return e;
}

// explicit-stage cache of one cell (energy_hydrology.jl:745-791), and ψ_w0
// of its phase change (matric_total)
struct EHAux {
    double thl, kap, T, K, psi, psiw0;
};

// with_psi false: a.psi is not wanted (a caller that evaluates ψ at its own
// iterate); a.psi is then left 0 where it would cost a separate evaluation
__device__ __forceinline__ EHAux eh_cell_aux(const EHPar& e, double vl, double ti, double re,
                                             bool with_psi = true) {
    EHAux a;
    a.thl = liq_fraction(e.v, vl, e.nu - ti);
    a.kap = thermal_cond(a.thl, ti, e.nu, e.om, e.qz, e.gr, e.k_dry, e.k_satu, e.k_satf,
                         e.lk_u, e.lk_f);
    a.T = temperature(re, ti, heat_capacity(a.thl, ti, e.rhoc_ds));
    double fi = ti / (a.thl + ti - e.v.theta_r);
    // impedance_factor 10^(-Ω f_i) times viscosity_factor e^(γ(T - γT_ref)), one exp
    double impvisc = exp(-OMEGA * LN10 * fi + GAMMA_VISC * (a.T - GAMMA_T_REF));
    if (__all(ti == 0.0)) {
        // no ice on the wave: ν_eff = ν, and the saturation of K, ψ and ψ_w0 is
        // one S (for S < 1; ψ_w0 = ψ(1) = -0 otherwise), so they share S and
        // S^(-1/m) -- the same expressions as the general branch, bit for bit
        PsiD p = psi_dpsi(e.v, e.nu, vl);
        a.K = impvisc * hyd_cond_from_a(e.v, p.S, p.a);
        a.psi = p.psi;
        a.psiw0 = p.S < 1.0 ? p.psi : -0.0;
    } else {
        a.K = impvisc * hyd_cond(e.v, eff_sat(e.v, e.nu, vl));
        a.psi = with_psi ? psi_dpsi(e.v, e.nu - ti, vl).psi : 0.0;
        a.psiw0 = matric_total(e.v, e.nu, a.thl, ti);
    }
    return a;
}

// implicit-stage T (energy_hydrology.jl:430-440: θ_l := min(ν-θ_i, ϑ_l))
__device__ __forceinline__ double eh_imp_T(const EHPar& e, double vl, double ti, double re) {
    return temperature(re, ti, heat_capacity(fmin(e.nu - ti, vl), ti, e.rhoc_ds));
}

// Newton form (no counterpart in the reference): ρc_s at θ_l = min(ν-θ_i, ϑ_l)
// of the iterate and ∂T/∂ϑ_l = -(T - T_0) ρ_l c_l / ρc_s where ϑ_l < ν - θ_i
// (0 above, where θ_l is capped), for a given T ...  ∂T/∂ϑ_l enters only the
// matrix: in Q = float forms (FP32 matrices) through the hardware
// reciprocal of ρc_s, in the FP64 form through the division
template <typename Q>
__device__ __forceinline__ double newton_dTdv(const EHPar& e, double vl, double ti, double T,
                                              double rc) {
    if (!(vl < e.nu - ti)) return 0.0;
    if constexpr (std::is_same<Q, float>::value)
        return -(T - T_0) * RCP_L * __builtin_amdgcn_rcp(rc);
    else
        return -(T - T_0) * RCP_L / rc;
}
template <typename Q>
__device__ __forceinline__ void newton_dT(const EHPar& e, double vl, double ti, double T,
                                          double& rc, double& dTdv) {
    rc = heat_capacity(fmin(e.nu - ti, vl), ti, e.rhoc_ds);
    dTdv = newton_dTdv<Q>(e, vl, ti, T, rc);
}
// ... and with T of the implicit cache, eh_imp_T's arithmetic
template <typename Q>
__device__ __forceinline__ void newton_T(const EHPar& e, double vl, double ti, double re,
                                         double& T, double& rc, double& dTdv) {
    rc = heat_capacity(fmin(e.nu - ti, vl), ti, e.rhoc_ds);
    T = temperature(re, ti, rc);
    dTdv = newton_dTdv<Q>(e, vl, ti, T, rc);
}

// ---- AtmosDrivenFluxBC, soil-only form (boundary_conditions.jl:880-915) --------
// The turbulent fluxes and net radiation come from the caller (SurfaceFluxes /
// radiation drivers are outside this path), per column:
struct Surface {
    const double *P, *Tair, *lhf, *shf, *vliq, *vice, *Rn;
};
// TOPMODEL for EnergyHydrology (Runoff.jl:234-290, 373-416): the surface pass
// weights layers by total water (ice included) and takes the capacity
// -K_sat F(θ_i/(θ_l+θ_i-θ_r)) G(T) of the top layer; the subsurface pass uses
// liquid water only, and R_ess = R_ss/max(h∇,eps) Σ w ρe_l(T) Δz.
__device__ __forceinline__ double surface_weight(double vl, double ti, double nu, double thr) {
    double twc = vl + ti - thr, nsat = nu - thr;
    return (twc - nsat > DEPS) ? twc / nsat : 0.0;
}
__device__ __forceinline__ double eh_infil_capacity(double Ksat, double thr, double thl_top,
                                                    double ti_top, double T_top) {
    return -Ksat * exp(-OMEGA * LN10 * (ti_top / (thl_top + ti_top - thr)) +
                       GAMMA_VISC * (T_top - GAMMA_T_REF));
}
// top fluxes from the infiltration: (water, heat)
__device__ __forceinline__ void eh_atmos_fluxes(const Surface& s, int64_t i, double infil,
                                                double* tw, double* th) {
    *tw = infil + s.vliq[i];
    *th = s.Rn[i] + s.lhf[i] + s.shf[i] + infil * energy_liq(s.Tair[i]);
}
// SoilSublimation source! (energy_hydrology.jl:931-945): θ_i tendency of a
// layer centred at zc (H(z + 2 Δz_top): the top layer when the surface is at 0)
__device__ __forceinline__ double sublimation(double vice, double zc, double dzt) {
    return -vice * RHO_L / RHO_I * (zc + 2.0 * dzt > DEPS ? 1.0 : 0.0) / (2.0 * dzt);
}

// ---- fine-grained hooks -----------------------------------------------------

// Under AtmosDrivenFluxBC (bc_top = BC_ATMOS*) sf holds the per-slot surface
// inputs; the TOPMODEL sinks (per cell) and R_ss, R_ess (per slot) go to
// sink_v, sink_e, rss, ress, and the ice vapour flux to vice (the explicit
// sublimation reads it)
__global__ __launch_bounds__(256) void k_eh_update_cache(
    Layout L, const double* __restrict__ par, const double* __restrict__ zf,
    const double* __restrict__ vl, const double* __restrict__ ti, const double* __restrict__ re,
    double* __restrict__ thl, double* __restrict__ kap, double* __restrict__ T,
    double* __restrict__ K, double* __restrict__ psi, const double* __restrict__ ftw,
    const double* __restrict__ fth, const double* __restrict__ fbw,
    const double* __restrict__ fbh, double* __restrict__ tw, double* __restrict__ th,
    double* __restrict__ bw, double* __restrict__ bh, int bc_bottom, int bc_top, Surface sf,
    double* __restrict__ sink_v, double* __restrict__ sink_e, double* __restrict__ rss,
    double* __restrict__ ress, double* __restrict__ vice) {
    ColView cv;
    if (!lane_column(L, cv)) return;
    EHPar e = load_eh(par, L.nslot, cv.s);
    double K0 = 0.0, T0 = 0.0;
    for (int j = 0; j < cv.N; ++j) {
        int64_t c = cv.cb + (int64_t)kWave * j;
        EHAux a = eh_cell_aux(e, vl[c], ti[c], re[c]);
        thl[c] = a.thl;
        kap[c] = a.kap;
        T[c] = a.T;
        K[c] = a.K;
        psi[c] = a.psi;
        if (j == 0) {
            K0 = a.K;
            T0 = a.T;
        }
    }
    if (bc_top == BC_ATMOS || bc_top == BC_ATMOS_TOPMODEL) {
        const int N = cv.N;
        const int64_t s = cv.s, ns = L.nslot;
        double infil = sf.P[s];
        if (bc_top == BC_ATMOS_TOPMODEL) {
            const double thr = e.v.theta_r, nu = e.nu;
            double hs = 0.0, h = 0.0, he = 0.0;
            for (int j = 0; j < N; ++j) {
                int64_t c = cv.cb + (int64_t)kWave * j;
                double dz = zf[cv.fb + (int64_t)kWave * (j + 1)] - zf[cv.fb + (int64_t)kWave * j];
                hs += surface_weight(vl[c], ti[c], nu, thr) * dz;
                double w = sat_weight(vl[c], nu, thr);
                sink_v[c] = w;
                h += w * dz;
                he += w * energy_liq(T[c]) * dz;
            }
            const int64_t ct = cv.cb + (int64_t)kWave * (N - 1);
            double ic = eh_infil_capacity(e.v.Ksat, thr, thl[ct], ti[ct], T[ct]);
            double depth = zf[cv.fb + (int64_t)kWave * N] - zf[cv.fb];
            double fsat = fmin(par[P_EH_F_MAX * ns + s] *
                                   exp(-par[P_EH_F_OVER * ns + s] / 2.0 * (depth - hs)),
                               1.0);
            infil = (1.0 - fsat) * fmax(ic, sf.P[s]);
            double r = par[P_EH_R_SB * ns + s] * exp(-par[P_EH_F_OVER * ns + s] * (depth - h));
            double re_ = he * (r / fmax(h, DEPS));
            rss[s] = r;
            ress[s] = re_;
            for (int j = 0; j < N; ++j) {
                int64_t c = cv.cb + (int64_t)kWave * j;
                double w = sink_v[c];
                sink_e[c] = (re_ / fmax(h, DEPS)) * w;
                sink_v[c] = (r / fmax(h, DEPS)) * w;
            }
        }
        eh_atmos_fluxes(sf, s, infil, &tw[s], &th[s]);
        vice[s] = sf.vice[s];
    } else {
        tw[cv.s] = ftw[cv.s];
        th[cv.s] = fth[cv.s];
    }
    if (bc_bottom == BC_FREE_DRAINAGE) {  // EnergyWaterFreeDrainage, boundary_conditions.jl:590-608
        bw[cv.s] = -1.0 * K0;
        bh[cv.s] = -1.0 * K0 * energy_liq(T0);
    } else {
        bw[cv.s] = fbw[cv.s];
        bh[cv.s] = fbh[cv.s];
    }
}

// make_update_implicit_aux (energy_hydrology.jl:427-445)
__global__ __launch_bounds__(256) void k_eh_imp_cache(Layout L, const double* __restrict__ par,
                                                      const double* __restrict__ vl,
                                                      const double* __restrict__ ti,
                                                      const double* __restrict__ re,
                                                      double* __restrict__ T,
                                                      double* __restrict__ psi) {
    ColView cv;
    if (!lane_column(L, cv)) return;
    EHPar e = load_eh(par, L.nslot, cv.s);
    for (int j = 0; j < cv.N; ++j) {
        int64_t c = cv.cb + (int64_t)kWave * j;
        double v = vl[c], t = ti[c];
        T[c] = eh_imp_T(e, v, t, re[c]);
        psi[c] = psi_dpsi(e.v, e.nu - t, v).psi;
    }
}

// compute_exp_tendency (:315-350) + PhaseChange source! (:838-906)
__global__ __launch_bounds__(256) void k_eh_exp_tendency(
    Layout L, const double* __restrict__ par, const double* __restrict__ zf,
    const double* __restrict__ ti, const double* __restrict__ thl,
    const double* __restrict__ T, const double* __restrict__ kap, double* __restrict__ dvl,
    double* __restrict__ dti, double* __restrict__ dre, double* __restrict__ diw,
    double* __restrict__ die, int32_t* __restrict__ bad_out, const double* __restrict__ vice) {
    // vice: AtmosDrivenFluxBC's ice vapour flux per slot (sublimation), or null
    ColView cv;
    if (!lane_column(L, cv)) return;
    EHPar e = load_eh(par, L.nslot, cv.s);
    int bad = 0;
    const int N = cv.N;
    const double ztop = zf[cv.fb + (int64_t)kWave * N];
    const double dzt = ztop - 0.5 * (zf[cv.fb + (int64_t)kWave * (N - 1)] + ztop);
    const double vi = vice ? vice[cv.s] : 0.0;
    for (int j = 0; j < cv.N; ++j) {
        int64_t c = cv.cb + (int64_t)kWave * j;
        double dz = zf[cv.fb + (int64_t)kWave * (j + 1)] - zf[cv.fb + (int64_t)kWave * j];
        double t = ti[c], l = thl[c];
        double tau = 3.0 * heat_capacity(l, t, e.rhoc_ds) * dz * dz / kap[c];  // thermal_time
        double F = phase_change(e.v, e.nu, l, t, T[c], tau, &bad);
        dvl[c] = -F;
        double d = (RHO_L / RHO_I) * F;
        if (vice) {
            double zc = 0.5 * (zf[cv.fb + (int64_t)kWave * j] + zf[cv.fb + (int64_t)kWave * (j + 1)]);
            d += sublimation(vi, zc, dzt);
        }
        dti[c] = d;
        dre[c] = 0.0;
    }
    diw[cv.s] = vice ? 0.0 + -vi : 0.0;
    die[cv.s] = 0.0;
    if (bad) bad_out[cv.s] = 1;
}

// compute_imp_tendency (:363-425)
__global__ __launch_bounds__(256) void k_eh_imp_tendency(
    Layout L, const double* __restrict__ zf, const double* __restrict__ K,
    const double* __restrict__ psi, const double* __restrict__ T,
    const double* __restrict__ kap, const double* __restrict__ tw,
    const double* __restrict__ th, const double* __restrict__ bw,
    const double* __restrict__ bh, double* __restrict__ dvl, double* __restrict__ dti,
    double* __restrict__ dre, double* __restrict__ diw, double* __restrict__ die,
    const double* __restrict__ sink_v, const double* __restrict__ sink_e,
    const double* __restrict__ rss, const double* __restrict__ ress) {
    // sink_*, rss, ress: TOPMODEL subsurface runoff of EH (Runoff.jl:346-359), or null
    ColView cv;
    if (!lane_column(L, cv)) return;
    const int N = cv.N;
    double topw = tw[cv.s], toph = th[cv.s], botw = bw[cv.s], both = bh[cv.s];
    diw[cv.s] = sink_v ? -(topw - botw) - rss[cv.s] : -(topw - botw);
    die[cv.s] = sink_v ? -(toph - both) - ress[cv.s] : -(toph - both);
    double f0 = zf[cv.fb], f1 = zf[cv.fb + kWave];
    double zc_c = 0.5 * (f0 + f1), dzc_c = f1 - f0;
    double K_c = K[cv.cb], h_c = psi[cv.cb] + zc_c, T_c = T[cv.cb], k_c = kap[cv.cb];
    double Flo = botw, Glo = both;
    for (int i = 0; i < N; ++i) {
        double Fhi, Ghi, zc_n = 0, dzc_n = 0, K_n = 0, h_n = 0, T_n = 0, k_n = 0;
        if (i == N - 1) {
            Fhi = topw;
            Ghi = toph;
        } else {
            int64_t c = cv.cb + (int64_t)kWave * (i + 1);
            double f2 = zf[cv.fb + (int64_t)kWave * (i + 2)];
            zc_n = 0.5 * (f1 + f2);
            dzc_n = f2 - f1;
            K_n = K[c];
            h_n = psi[c] + zc_n;
            T_n = T[c];
            k_n = kap[c];
            double d = zc_n - zc_c;
            double gh = (h_n - h_c) / d;
            Fhi = -0.5 * (K_c + K_n) * gh;
            double eK = 0.5 * (energy_liq(T_c) * K_c + energy_liq(T_n) * K_n);
            Ghi = -0.5 * (k_c + k_n) * ((T_n - T_c) / d) - eK * gh;
            f1 = f2;
        }
        int64_t c = cv.cb + (int64_t)kWave * i;
        double tv = -(Fhi - Flo) / dzc_c, te = -(Ghi - Glo) / dzc_c;
        if (sink_v) {
            tv -= sink_v[c];
            te -= sink_e[c];
        }
        dvl[c] = tv;
        dre[c] = te;
        dti[c] = 0.0;
        Flo = Fhi;
        Glo = Ghi;
        zc_c = zc_n;
        dzc_c = dzc_n;
        K_c = K_n;
        h_c = h_n;
        T_c = T_n;
        k_c = k_n;
    }
}

// compute_jacobian! (:466-576).  Rows for the three blocks.
template <typename JT>
__global__ __launch_bounds__(256) void k_eh_jacobian(
    Layout L, const double* __restrict__ par, const double* __restrict__ zf,
    const double* __restrict__ vl, const double* __restrict__ ti,
    const double* __restrict__ K, const double* __restrict__ T,
    const double* __restrict__ kap, const double* __restrict__ thl, double dtg,
    JT* __restrict__ Www, JT* __restrict__ Wee, JT* __restrict__ Wew, int64_t stride) {
    // W?? hold three planes (lo, di, up) each `stride` apart
    ColView cv;
    if (!lane_column(L, cv)) return;
    const int N = cv.N;
    EHPar e = load_eh(par, L.nslot, cv.s);
    auto Dw = [&](int64_t c) { return psi_dpsi(e.v, e.nu - ti[c], vl[c]).D; };
    auto Ew = [&](int64_t c) { return energy_liq(T[c]) * K[c]; };
    auto Dr = [&](int64_t c) { return 1.0 / heat_capacity(thl[c], ti[c], e.rhoc_ds); };
    double f0 = zf[cv.fb], f1 = zf[cv.fb + kWave];
    double zc_c = 0.5 * (f0 + f1), dzc_c = f1 - f0;
    int64_t c0 = cv.cb;
    double K_c = K[c0], E_c = Ew(c0), k_c = kap[c0];
    double Dw_c = Dw(c0), Dr_c = Dr(c0), Dw_m = 0, Dr_m = 0;
    double wK_lo = 0, wE_lo = 0, wk_lo = 0;
    for (int i = 0; i < N; ++i) {
        double wK_up = 0, wE_up = 0, wk_up = 0, Dw_n = 0, Dr_n = 0;
        double K_n = 0, E_n = 0, k_n = 0, zc_n = 0, dzc_n = 0;
        if (i < N - 1) {
            int64_t c = cv.cb + (int64_t)kWave * (i + 1);
            double f2 = zf[cv.fb + (int64_t)kWave * (i + 2)];
            zc_n = 0.5 * (f1 + f2);
            dzc_n = f2 - f1;
            K_n = K[c];
            E_n = Ew(c);
            k_n = kap[c];
            Dw_n = Dw(c);
            Dr_n = Dr(c);
            double idzf = 1.0 / (zc_n - zc_c);
            wK_up = 0.5 * (K_c + K_n) * idzf;
            wE_up = 0.5 * (E_c + E_n) * idzf;
            wk_up = 0.5 * (k_c + k_n) * idzf;
            f1 = f2;
        }
        double s = dtg / dzc_c;
        int64_t c = cv.cb + (int64_t)kWave * i;
        Www[c] = (JT)(s * wK_lo * Dw_m);
        Www[c + stride] = (JT)(-s * (wK_lo + wK_up) * Dw_c - 1.0);
        Www[c + 2 * stride] = (JT)(s * wK_up * Dw_n);
        Wew[c] = (JT)(s * wE_lo * Dw_m);
        Wew[c + stride] = (JT)(-s * (wE_lo + wE_up) * Dw_c - 1.0);  // "- I" (:554-556)
        Wew[c + 2 * stride] = (JT)(s * wE_up * Dw_n);
        Wee[c] = (JT)(s * wk_lo * Dr_m);
        Wee[c + stride] = (JT)(-s * (wk_lo + wk_up) * Dr_c - 1.0);
        Wee[c + 2 * stride] = (JT)(s * wk_up * Dr_n);
        Dw_m = Dw_c;
        Dw_c = Dw_n;
        Dr_m = Dr_c;
        Dr_c = Dr_n;
        K_c = K_n;
        E_c = E_n;
        k_c = k_n;
        zc_c = zc_n;
        dzc_c = dzc_n;
        wK_lo = wK_up;
        wE_lo = wE_up;
        wk_lo = wk_up;
    }
}

// ldiv! with BlockLowerTriangularSolve(@name(soil.ϑ_l))
template <typename JT>
__global__ __launch_bounds__(256) void k_eh_solve(
    Layout L, const JT* __restrict__ Www, const JT* __restrict__ Wee,
    const JT* __restrict__ Wew, int64_t stride, const double* __restrict__ rv,
    const double* __restrict__ ri, const double* __restrict__ re,
    const double* __restrict__ riw, const double* __restrict__ rie, double* __restrict__ xv,
    double* __restrict__ xi, double* __restrict__ xe, double* __restrict__ xiw,
    double* __restrict__ xie, JT* __restrict__ cp, JT* __restrict__ dp,
    double* __restrict__ rhs) {
    ColView cv;
    if (!lane_column(L, cv)) return;
    const int N = cv.N;
    thomas_lane<JT>(cv, Www, Www + stride, Www + 2 * stride, rv, xv, cp, dp);
    for (int i = 0; i < N; ++i) {
        int64_t c = cv.cb + (int64_t)kWave * i;
        double s = (double)Wew[c + stride] * xv[c];
        if (i > 0) s += (double)Wew[c] * xv[c - kWave];
        if (i < N - 1) s += (double)Wew[c + 2 * stride] * xv[c + kWave];
        rhs[c] = re[c] - s;
        xi[c] = -ri[c];
    }
    thomas_lane<JT>(cv, Wee, Wee + stride, Wee + 2 * stride, rhs, xe, cp, dp);
    xiw[cv.s] = -riw[cv.s];
    xie[cv.s] = -rie[cv.s];
}

// ---- fused ARS111 step --------------------------------------------------------
//
// Buffers (packed cells): x{v,i,e} state in/out (u_n in, u_{n+1} out);
// U0v/U0e stage values; Kl, kl, tl lagged K, κ, θ_l; Tp, Pp the cache T, ψ;
// Dp dψ/dϑ at the current iterate; scratch cpv, dpv, cpe, dpe (JT), Re (R_ρe),
// dV (Δϑ).
// Sweeps per Newton iteration: A (up) residuals + ϑ-block forward
// elimination; B (down) ϑ back substitution; C (up) ρe right-hand side
// b_ρ - A_ρϑ Δϑ + ρe-block forward elimination; D (down) ρe back
// substitution, convergence norm, and cache_imp! (T, ψ) plus dψ/dϑ for the
// next Jacobian at the new iterate.
struct EHBufs {
    double *xv, *xi, *xe, *U0v, *U0e, *Kl, *kl, *tl, *Tp, *Pp, *Re;
    double *iw, *ie;
};

template <typename JT>
struct EHScratch {
    JT *cpv, *dpv, *cpe, *dpe, *Dp, *dV;
};

template <typename JT>
__global__ __launch_bounds__(256) void k_eh_step(
    Layout L, const double* __restrict__ par, const double* __restrict__ zf, EHBufs B,
    EHScratch<JT> S, const double* __restrict__ ftw, const double* __restrict__ fth,
    const double* __restrict__ fbw, const double* __restrict__ fbh, int bc_bottom, double dt,
    int max_iters, double tol, int32_t* __restrict__ iters) {
    ColView cv;
    if (!lane_column(L, cv)) return;
    const int N = cv.N;
    const int64_t cb = cv.cb, fb = cv.fb;
    const EHPar e = load_eh(par, L.nslot, cv.s);
    const double topw = ftw[cv.s], toph = fth[cv.s];
    double botw = 0.0, both = 0.0;
    int bad = 0;

    // ---- stage 1: update_cache!(u_n) + phase change -> U0; D at U0 --------
    for (int j = 0; j < N; ++j) {
        const int64_t c = cb + (int64_t)kWave * j;
        double v = B.xv[c], t = B.xi[c], r = B.xe[c];
        EHAux a = eh_cell_aux(e, v, t, r);
        if (j == 0) {
            if (bc_bottom == BC_FREE_DRAINAGE) {
                botw = -1.0 * a.K;
                both = -1.0 * a.K * energy_liq(a.T);
            } else {
                botw = fbw[cv.s];
                both = fbh[cv.s];
            }
        }
        double dz = zf[fb + (int64_t)kWave * (j + 1)] - zf[fb + (int64_t)kWave * j];
        double tau = 3.0 * heat_capacity(a.thl, t, e.rhoc_ds) * dz * dz / a.kap;
        double F = phase_change_psi(e.v, e.nu, a.thl, t, a.T, tau, a.psiw0, &bad);
        double u0v = v + dt * (-F);
        double u0i = t + dt * ((RHO_L / RHO_I) * F);
        B.xv[c] = u0v;
        B.xi[c] = u0i;
        B.U0v[c] = u0v;
        B.U0e[c] = r;
        B.Kl[c] = a.K;
        B.kl[c] = a.kap;
        B.tl[c] = a.thl;
        B.Tp[c] = a.T;
        B.Pp[c] = a.psi;
        S.Dp[c] = (JT)psi_dpsi(e.v, e.nu - u0i, u0v).D;
    }
    // ∫F: zero explicit tendency, -I blocks: exact after the first iteration
    B.iw[cv.s] = B.iw[cv.s] + dt * (-(topw - botw));
    B.ie[cv.s] = B.ie[cv.s] + dt * (-(toph - both));

    int k;
    bool conv = false;
    double err = 0.0;
    for (k = 1; k <= max_iters; ++k) {
        // ---- sweep A --------------------------------------------------------
        {
            double f0 = zf[fb], f1 = zf[fb + kWave];
            double zc_c = 0.5 * (f0 + f1), dzc_c = f1 - f0;
            double K_c = B.Kl[cb], k_c = B.kl[cb], T_c = B.Tp[cb];
            double h_c = B.Pp[cb] + zc_c, D_c = (double)S.Dp[cb], D_m = 0.0;
            double Flo = botw, Glo = both, wlo = 0.0;
            JT cpp = 0, dpp = 0;
            for (int i = 0; i < N; ++i) {
                const int64_t c = cb + (int64_t)kWave * i;
                double Fhi, Ghi, wup = 0.0, D_n = 0.0, K_n = 0.0, k_n = 0.0, T_n = 0.0;
                double h_n = 0.0, zc_n = 0.0, dzc_n = 0.0;
                if (i < N - 1) {
                    const int64_t cn = c + kWave;
                    double f2 = zf[fb + (int64_t)kWave * (i + 2)];
                    zc_n = 0.5 * (f1 + f2);
                    dzc_n = f2 - f1;
                    K_n = B.Kl[cn];
                    k_n = B.kl[cn];
                    T_n = B.Tp[cn];
                    h_n = B.Pp[cn] + zc_n;
                    D_n = (double)S.Dp[cn];
                    double idzf = 1.0 / (zc_n - zc_c);
                    double gh = (h_n - h_c) * idzf;
                    double Kf = 0.5 * (K_c + K_n);
                    wup = Kf * idzf;
                    Fhi = -Kf * gh;
                    double eK = 0.5 * (energy_liq(T_c) * K_c + energy_liq(T_n) * K_n);
                    Ghi = -0.5 * (k_c + k_n) * ((T_n - T_c) * idzf) - eK * gh;
                    f1 = f2;
                } else {
                    Fhi = topw;
                    Ghi = toph;
                }
                double idzc = 1.0 / dzc_c;
                double Rv = B.U0v[c] + dt * (-(Fhi - Flo) * idzc) - B.xv[c];
                B.Re[c] = B.U0e[c] + dt * (-(Ghi - Glo) * idzc) - B.xe[c];
                double s = dt * idzc;
                JT lo = (JT)(s * wlo * D_m);
                JT di = (JT)(-s * (wlo + wup) * D_c - 1.0);
                JT up = (JT)(s * wup * D_n);
                JT inv = (JT)1 / (di - lo * cpp);
                cpp = up * inv;
                dpp = ((JT)Rv - lo * dpp) * inv;
                S.cpv[c] = cpp;
                S.dpv[c] = dpp;
                Flo = Fhi;
                Glo = Ghi;
                wlo = wup;
                D_m = D_c;
                D_c = D_n;
                K_c = K_n;
                k_c = k_n;
                T_c = T_n;
                h_c = h_n;
                zc_c = zc_n;
                dzc_c = dzc_n;
            }
        }
        // ---- sweep B --------------------------------------------------------
        double err_v = 0.0;
        {
            JT dn = 0;
            for (int i = N - 1; i >= 0; --i) {
                const int64_t c = cb + (int64_t)kWave * i;
                dn = S.dpv[c] - S.cpv[c] * dn;
                S.dV[c] = dn;
                double d = (double)dn;
                B.xv[c] = B.xv[c] - d;
                err_v = nan_max(err_v, fabs(d));
            }
        }
        // ---- sweep C --------------------------------------------------------
        {
            double f0 = zf[fb], f1 = zf[fb + kWave];
            double zc_c = 0.5 * (f0 + f1), dzc_c = f1 - f0;
            double E_c = energy_liq(B.Tp[cb]) * B.Kl[cb], k_c = B.kl[cb];
            double Dw_c = (double)S.Dp[cb], Dw_m = 0.0;
            double Dr_c = 1.0 / heat_capacity(B.tl[cb], B.xi[cb], e.rhoc_ds), Dr_m = 0.0;
            double x_c = (double)S.dV[cb], x_m = 0.0;
            double wE_lo = 0.0, wk_lo = 0.0;
            JT cpp = 0, dpp = 0;
            for (int i = 0; i < N; ++i) {
                const int64_t c = cb + (int64_t)kWave * i;
                double wE_up = 0.0, wk_up = 0.0, Dw_n = 0.0, Dr_n = 0.0, x_n = 0.0;
                double E_n = 0.0, k_n = 0.0, zc_n = 0.0, dzc_n = 0.0;
                if (i < N - 1) {
                    const int64_t cn = c + kWave;
                    double f2 = zf[fb + (int64_t)kWave * (i + 2)];
                    zc_n = 0.5 * (f1 + f2);
                    dzc_n = f2 - f1;
                    E_n = energy_liq(B.Tp[cn]) * B.Kl[cn];
                    k_n = B.kl[cn];
                    Dw_n = (double)S.Dp[cn];
                    Dr_n = 1.0 / heat_capacity(B.tl[cn], B.xi[cn], e.rhoc_ds);
                    x_n = (double)S.dV[cn];
                    double idzf = 1.0 / (zc_n - zc_c);
                    wE_up = 0.5 * (E_c + E_n) * idzf;
                    wk_up = 0.5 * (k_c + k_n) * idzf;
                    f1 = f2;
                }
                double s = dt / dzc_c;
                // A_ρϑ row times Δϑ
                double ew_lo = s * wE_lo * Dw_m, ew_up = s * wE_up * Dw_n;
                double ew_di = -s * (wE_lo + wE_up) * Dw_c - 1.0;
                double rhs = B.Re[c] - (ew_lo * x_m + ew_di * x_c + ew_up * x_n);
                JT lo = (JT)(s * wk_lo * Dr_m);
                JT di = (JT)(-s * (wk_lo + wk_up) * Dr_c - 1.0);
                JT up = (JT)(s * wk_up * Dr_n);
                JT inv = (JT)1 / (di - lo * cpp);
                cpp = up * inv;
                dpp = ((JT)rhs - lo * dpp) * inv;
                S.cpe[c] = cpp;
                S.dpe[c] = dpp;
                wE_lo = wE_up;
                wk_lo = wk_up;
                Dw_m = Dw_c;
                Dw_c = Dw_n;
                Dr_m = Dr_c;
                Dr_c = Dr_n;
                x_m = x_c;
                x_c = x_n;
                E_c = E_n;
                k_c = k_n;
                zc_c = zc_n;
                dzc_c = dzc_n;
            }
        }
        // ---- sweep D --------------------------------------------------------
        err = err_v;
        {
            JT dn = 0;
            for (int i = N - 1; i >= 0; --i) {
                const int64_t c = cb + (int64_t)kWave * i;
                dn = S.dpe[c] - S.cpe[c] * dn;
                double d = (double)dn;
                double xe = B.xe[c] - d;
                B.xe[c] = xe;
                double t = B.xi[c], v = B.xv[c];
                err = nan_max(err, fabs(d) / heat_capacity(B.tl[c], t, e.rhoc_ds));
                // cache_imp!(x) (:427-445) and dψ/dϑ(x) for the next Jacobian
                B.Tp[c] = eh_imp_T(e, v, t, xe);
                PsiD r = psi_dpsi(e.v, e.nu - t, v);
                B.Pp[c] = r.psi;
                S.Dp[c] = (JT)r.D;
            }
        }
        conv = (tol > 0.0) && (err < tol);
        // a non-finite iterate stays non-finite: stop (the state is NaN either way)
        if (conv || !isfinite(err)) break;
    }
    if (k > max_iters) k = max_iters;
    int32_t flag = k;
    if (tol > 0.0 && !conv) flag |= 0x10000;
    if (!isfinite(err)) flag |= 0x20000;
    if (bad) flag |= 0x40000;
    iters[cv.s] = flag;
}

}  // namespace clm
