// clm_physics.h -- pointwise soil closures for the gfx950 kernels.
//
// FP64 throughout.  The van Genuchten closures evaluated inside the Newton
// loop share their transcendental work: with lS = log S and a = S^(-1/m),
//   ψ      = -(a-1)^(1/n) / α                      (matric_potential)
//   dψ/dϑ  = (a-1)^(1/n-1) · a/S / (α m n (ν_eff-θ_r))
//   K      = K_sat √S (1 - (1 - 1/a)^m)^2          (S^(1/m) = 1/a)
// i.e. 2-3 log/exp pairs per cell instead of 5-7 pow() calls.  These are the
// same functions as soil_hydrology_parameterizations.jl:59-63, 135-152,
// 161-173 written with the identities (x·α^-n)^(1/n) = x^(1/n)/α and
// S^(-1/m-1) = S^(-1/m)/S; branch structure (S <= 1 for ψ, S < 1 for dψ/dϑ
// and K) and the √eps clipping are kept exactly.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace clm {

// ClimaParams 1.1.4 defaults (src/shared_utilities/Parameters.jl:86-105)
constexpr double RHO_L = 1000.0;
constexpr double RHO_I = 916.7;
constexpr double CP_L = 4181.0;
constexpr double CP_I = 2100.0;
constexpr double T_0 = 273.16;
constexpr double T_FREEZE = 273.15;
constexpr double LH_V0 = 2500800.0;
constexpr double LH_S0 = 2834400.0;
constexpr double GRAV = 9.81;
constexpr double K_THERM = 0.024;
constexpr double LH_F0 = LH_S0 - LH_V0;  // Parameters.jl:58
constexpr double RCP_L = CP_L * RHO_L;
constexpr double RCP_I = CP_I * RHO_I;
// soil constants (energy_hydrology.jl:124-180 lookups; values pinned by
// test/standalone/Soil/soil_parameterizations.jl:71-75)
constexpr double KERSTEN_A = 0.24;
constexpr double KERSTEN_B = 18.3;
constexpr double OMEGA = 7.0;
constexpr double GAMMA_VISC = 2.64e-2;
constexpr double GAMMA_T_REF = 288.0;

constexpr double DEPS = 2.220446049250313e-16;       // eps(Float64)
constexpr double SQRT_EPS = 1.4901161193847656e-08;  // sqrt(eps(Float64))

// max for Newton convergence norms that propagates NaN (a non-finite
// iterate never passes `err < tol`)
__device__ __forceinline__ double nan_max(double a, double b) {
    return (b > a || b != b) ? b : a;
}

// per-column van Genuchten constants, derived once per column
struct VG {
    double alpha, n, m, inv_m, inv_n, inv_alpha, theta_r, thr_safe, Ss, inv_Ss, Ksat;
    double cD;  // 1 / (α m n)
};

__host__ __device__ inline VG make_vg(double alpha, double n, double theta_r, double Ss,
                                      double Ksat) {
    VG v;
    v.alpha = alpha;
    v.n = n;
    v.m = 1.0 - 1.0 / n;  // retention_models.jl:40-44
    v.inv_m = 1.0 / v.m;
    v.inv_n = 1.0 / n;
    v.inv_alpha = 1.0 / alpha;
    v.theta_r = theta_r;
    v.thr_safe = theta_r + SQRT_EPS;
    v.Ss = Ss;
    v.inv_Ss = 1.0 / Ss;
    v.Ksat = Ksat;
    v.cD = 1.0 / (alpha * v.m * n);
    return v;
}

// effective_saturation, soil_hydrology_parameterizations.jl:45-50
__device__ __forceinline__ double eff_sat(const VG& v, double nu_eff, double vl) {
    double vs = fmax(vl, v.thr_safe);
    double ns = fmax(nu_eff, v.thr_safe);
    return (vs - v.theta_r) / (ns - v.theta_r);
}

// volumetric_liquid_fraction, :22-31
__device__ __forceinline__ double liq_fraction(const VG& v, double vl, double nu_eff) {
    double vs = fmax(vl, v.thr_safe);
    double ns = fmax(nu_eff, v.thr_safe);
    return vs < ns ? vs : ns;
}

// pressure_head (:109-127) and dψdϑ (:135-152) for the same (ϑ, ν_eff),
// sharing log/exp.  Also returns a = S^(-1/m) (for K when ν_eff = ν).
struct PsiD {
    double psi, D, S, a;
};

__device__ __forceinline__ PsiD psi_dpsi(const VG& v, double nu_eff, double vl) {
    PsiD r;
    double vs = fmax(vl, v.thr_safe);
    double ns = fmax(nu_eff, v.thr_safe);
    double span = ns - v.theta_r;
    double S = (vs - v.theta_r) / span;
    r.S = S;
    if (S < 1.0) {
        double lS = log(S);
        double a = exp(-lS * v.inv_m);
        double X = a - 1.0;
        double lX = log(X);
        double X1n = exp(lX * v.inv_n);
        r.a = a;
        r.psi = -X1n * v.inv_alpha;
        // (a-1)^(1/n-1) = X1n / X; X == 0 only within an ulp of S = 1,
        // where the reference's pow(0, 1/n-1) gives +Inf
        double Xq = (X > 0.0) ? X1n / X : INFINITY;
        r.D = v.cD / span * Xq * (a / S);
    } else {
        r.a = 1.0;
        r.psi = (S == 1.0) ? -0.0 : (vs - ns) * v.inv_Ss;
        r.D = v.inv_Ss;
    }
    return r;
}

// The same closures for the Newton loops of the fused kernels: the
// saturation span is fixed per cell during a step (ν_eff moves only with θ_i,
// which is explicit), so its reciprocal comes precomputed, and the Jacobian
// needs dψ/dϑ only through its reciprocal 1/D (scaled unknowns and the
// update Δ = y / D), which one division gives: 1/D = span X S / (cD X^(1/n) a).
struct PsiI {
    double psi, invD, S, a;
};

struct Span {
    double ns, inv;  // max(ν_eff, θ_r + sqrt(eps)) and 1 / (that - θ_r)
};
__device__ __forceinline__ Span make_span(const VG& v, double nu_eff) {
    Span sp;
    sp.ns = fmax(nu_eff, v.thr_safe);
    sp.inv = 1.0 / (sp.ns - v.theta_r);
    return sp;
}

__device__ __forceinline__ PsiI psi_invd(const VG& v, const Span& sp, double vl) {
    PsiI r;
    double vs = fmax(vl, v.thr_safe);
    double S = (vs - v.theta_r) * sp.inv;
    r.S = S;
    if (S < 1.0) {
        double lS = log(S);
        double a = exp(-lS * v.inv_m);
        double X = a - 1.0;
        double lX = log(X);
        double X1n = exp(lX * v.inv_n);
        r.a = a;
        r.psi = -X1n * v.inv_alpha;
        // D is +Inf where X == 0 (within an ulp of S = 1): 1/D = 0
        r.invD = (X > 0.0) ? (X * S) / (v.cD * sp.inv * X1n * a) : 0.0;
    } else {
        r.a = 1.0;
        r.psi = (S == 1.0) ? -0.0 : (vs - sp.ns) * v.inv_Ss;
        r.invD = v.Ss;
    }
    return r;
}

// hydraulic_conductivity (:161-173) given S and a = S^(-1/m)
__device__ __forceinline__ double hyd_cond_from_a(const VG& v, double S, double a) {
    if (S < 1.0) {
        double Y = 1.0 - 1.0 / a;  // 1 - S^(1/m)
        double Ym = exp(v.m * log(Y));
        double q = 1.0 - Ym;
        return sqrt(S) * (q * q) * v.Ksat;
    }
    return v.Ksat;
}

// hydraulic_conductivity with its own saturation
__device__ __forceinline__ double hyd_cond(const VG& v, double S) {
    if (S < 1.0) {
        double a = exp(-log(S) * v.inv_m);
        return hyd_cond_from_a(v, S, a);
    }
    return v.Ksat;
}

// matric_potential (:59-63)
__device__ __forceinline__ double matric(const VG& v, double S) {
    return -pow((pow(S, -v.inv_m) - 1.0) * pow(v.alpha, -v.n), v.inv_n);
}

// inverse_matric_potential (:72-77); *bad = 1 where the reference raises
__device__ __forceinline__ double inv_matric(const VG& v, double psi, int* bad) {
    if (psi > 0.0) *bad = 1;
    return pow(1.0 + pow(v.alpha * fabs(psi), v.n), -v.m);
}

// ---- heat (soil_heat_parameterizations.jl) --------------------------------
__device__ __forceinline__ double heat_capacity(double thl, double thi, double rhoc_ds) {
    return rhoc_ds + thl * RCP_L + thi * RCP_I;  // :157-171
}
__device__ __forceinline__ double temperature(double rhoe, double thi, double rhoc_s) {
    return T_0 + (rhoe + thi * RHO_I * LH_F0) / rhoc_s;  // :180-192
}
__device__ __forceinline__ double energy_liq(double T) { return RCP_L * (T - T_0); }  // :221-231

// κ (:243-320): Balland–Arp Kersten number and saturated conductivity
__device__ __forceinline__ double thermal_cond(double thl, double thi, double nu, double om,
                                               double qz, double gr, double k_dry,
                                               double k_satu, double k_satf) {
    double Sr = (thl + thi) / nu;
    double Ke;
    if (thi < DEPS) {
        double t1 = pow(1.0 + exp(-KERSTEN_B * Sr), -3.0);
        double h = (1.0 - Sr) / 2.0;
        Ke = pow(Sr, (1.0 + om - KERSTEN_A * qz - gr) / 2.0) *
             pow(t1 - h * h * h, 1.0 - om);
    } else {
        Ke = pow(Sr, 1.0 + om);
    }
    double tw = thl + thi;
    double ks = (tw < DEPS) ? (k_satu + k_satf) / 2.0
                            : pow(k_satu, thl / tw) * pow(k_satf, thi / tw);
    return Ke * ks + (1.0 - Ke) * k_dry;
}

// phase_change_source (:89-122) with soil_Tf_depressed (:35-52); τ from
// thermal_time (:60-62) on the layer width
__device__ __forceinline__ double phase_change(const VG& v, double nu, double thl, double thi,
                                               double T, double tau, int* bad) {
    double ttot = fmin(RHO_I / RHO_L * thi + thl, nu);
    double vs = fmax(ttot, v.thr_safe);
    double ns = fmax(nu, v.thr_safe);
    double Sw = (vs - v.theta_r) / (ns - v.theta_r);
    double psiw0 = matric(v, Sw);
    double Tf = fmax(T_FREEZE * exp(GRAV * psiw0 / LH_F0), 1.0);
    double H = (Tf - T > DEPS) ? 1.0 : 0.0;  // heaviside, utils.jl:93-99
    double psiT = LH_F0 / GRAV * log(T / Tf) * H;
    double thstar = inv_matric(v, psiw0 + psiT, bad) * (nu - v.theta_r) + v.theta_r;
    return (thl - thstar) / tau;
}

}  // namespace clm
