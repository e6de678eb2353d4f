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

#include <type_traits>

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
constexpr double LN10 = 2.302585092994046;  // ln 10, impedance 10^x = e^(x ln 10)

constexpr double DEPS = 2.220446049250313e-16;       // eps(Float64)
constexpr double SQRT_EPS = 1.4901161193847656e-08;  // sqrt(eps(Float64))

// max for Newton convergence norms that propagates NaN (a non-finite
// iterate never passes `err < tol`)
__device__ __forceinline__ double nan_max(double a, double b) {
    return (b > a || b != b) ? b : a;
}

// ---- logarithm and powers --------------------------------------------------
// Natural logarithm within about one ulp in ~40 instructions (the library's
// double-double routine takes ~60).  With x = 2^e m, m in [√½, √2),
// f = m - 1 (exact) and s = f / (2 + f):
//   log(1+f) = 2 atanh(s) = f - (f²/2 - s (f²/2 + R)),
//   R = Σ_{k=1..10} 2 s^(2k) / (2k+1)
// (the atanh series; |s| <= 0.1716, so the omitted tail is below 2^-60 of
// the result), and e·ln2 is added in two parts: ln2_hi carries 42
// significant bits, so e·ln2_hi is exact for every binary64 exponent.
// s enters only the correction term, so a reciprocal refined by two Newton
// steps replaces the IEEE division.  log(0) = -inf, log(+inf) = +inf,
// log(x < 0) = log(NaN) = NaN.
__host__ __device__ __forceinline__ double clm_log(double x) {
    constexpr double LN2_HI = 0.6931471805598903;     // 0x3fe62e42fefa3800
    constexpr double LN2_LO = 5.497923018708371e-14;  // ln 2 - LN2_HI
    constexpr double SQRT_HALF = 0.7071067811865476;
#if defined(__HIP_DEVICE_COMPILE__)
    double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1)
    int e = __builtin_amdgcn_frexp_exp(x);
#else
    int e;
    double m = frexp(x, &e);
#endif
    if (m < SQRT_HALF) {
        m += m;
        e -= 1;
    }
    const double f = m - 1.0;
    const double d = 2.0 + f;
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    r = fma(r, fma(-d, r, 1.0), r);
    const double s = f * r;
#else
    const double s = f / d;
#endif
    const double z = s * s;
    double p = 2.0 / 21.0;
    p = fma(p, z, 2.0 / 19.0);
    p = fma(p, z, 2.0 / 17.0);
    p = fma(p, z, 2.0 / 15.0);
    p = fma(p, z, 2.0 / 13.0);
    p = fma(p, z, 2.0 / 11.0);
    p = fma(p, z, 2.0 / 9.0);
    p = fma(p, z, 2.0 / 7.0);
    p = fma(p, z, 2.0 / 5.0);
    p = fma(p, z, 2.0 / 3.0);
    const double R = p * z;
    const double hfsq = 0.5 * f * f;
    const double de = (double)e;
    const double y = de * LN2_HI - ((hfsq - (s * (hfsq + R) + de * LN2_LO)) - f);
    const double sp = (x == 0.0) ? -INFINITY : ((x == INFINITY) ? x : NAN);
    return (x > 0.0 && x < INFINITY) ? y : sp;
}

// y log x with 0 log x = 0 (so x^0 = 1 also for x = 0, as in the reference's
// ^): the exponent of a factor that is absent from a product of powers
__host__ __device__ __forceinline__ double mul_log(double y, double log_x) {
    return y == 0.0 ? 0.0 : y * log_x;
}

// x^y = exp(y log x) for x >= 0: relative error about |y log x| ulp, which
// the closures' bases and exponents keep below ~50 ulp (1e-14).
__host__ __device__ __forceinline__ double clm_pow(double x, double y) {
    return exp(mul_log(y, clm_log(x)));
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

// vanGenuchten's S_c, retention_models.jl:30-44: (1 + ((n-1)/n)^(1-2n))^(-m)
__host__ __device__ inline double vg_sc(double n) {
    const double m = 1.0 - 1.0 / n;
    return pow(1.0 + pow((n - 1.0) / n, 1.0 - 2.0 * n), -m);
}

// dry_soil_layer_thickness, energy_hydrology.jl:1352-1368 (S_l in (0, 1])
__device__ __forceinline__ double dry_soil_layer_thickness(double S_l, double aSc, double d_ds,
                                                           double p) {
    aSc = fmin(fmax(aSc, 0.0), 1.0);
    return S_l < aSc ? d_ds * pow((aSc - S_l) / S_l, p) : 0.0;
}

// soil_conductance, :1322-1350: vapour conductance D_vapor / max(dsl, eps) (m/s)
__device__ __forceinline__ double soil_conductance(double S_l, double S_c, double d_ds, double p,
                                                   double alpha, double D_vapor) {
    const double dsl = dry_soil_layer_thickness(S_l, alpha * S_c, d_ds, p);
    return D_vapor / fmax(dsl, DEPS);
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
        double lS = clm_log(S);
        double a = exp(-lS * v.inv_m);
        double X = a - 1.0;
        double lX = clm_log(X);
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

// Q: the precision of 1/D.  double: the division, as the FP64 form's
// Jacobian needs; float (FP32-Jacobian forms, whose matrices are FP32 and
// whose update only has to be an inexact-Newton step): the hardware FP64
// reciprocal times the numerator, well inside FP32 accuracy over the whole
// FP64 range, at a fraction of a division's cost.
template <typename Q = double>
__device__ __forceinline__ PsiI psi_invd(const VG& v, const Span& sp, double vl) {
    PsiI r;
    double vs = fmax(vl, v.thr_safe);
    double S = (vs - v.theta_r) * sp.inv;
    r.S = S;
    if (S < 1.0) {
        double lS = clm_log(S);
        double a = exp(-lS * v.inv_m);
        double X = a - 1.0;
        double lX = clm_log(X);
        double X1n = exp(lX * v.inv_n);
        r.a = a;
        r.psi = -X1n * v.inv_alpha;
        // D is +Inf where X == 0 (within an ulp of S = 1): 1/D = 0
        const double den = v.cD * sp.inv * X1n * a;
        if constexpr (std::is_same<Q, float>::value)
            r.invD = (X > 0.0) ? (X * S) * __builtin_amdgcn_rcp(den) : 0.0;
        else
            r.invD = (X > 0.0) ? (X * S) / den : 0.0;
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
        double Ym = exp(v.m * clm_log(Y));
        double q = 1.0 - Ym;
        return sqrt(S) * (q * q) * v.Ksat;
    }
    return v.Ksat;
}

// hydraulic_conductivity with its own saturation
__device__ __forceinline__ double hyd_cond(const VG& v, double S) {
    if (S < 1.0) {
        double a = exp(-clm_log(S) * v.inv_m);
        return hyd_cond_from_a(v, S, a);
    }
    return v.Ksat;
}

// matric_potential (:59-63), -((S^(-1/m) - 1) α^(-n))^(1/n), written as
// -(S^(-1/m) - 1)^(1/n) / α
__device__ __forceinline__ double matric(const VG& v, double S) {
    double X = exp(-clm_log(S) * v.inv_m) - 1.0;
    return -exp(clm_log(X) * v.inv_n) * v.inv_alpha;
}

// inverse_matric_potential (:72-77); *bad = 1 where the reference raises
__device__ __forceinline__ double inv_matric(const VG& v, double psi, int* bad) {
    if (psi > 0.0) *bad = 1;
    return clm_pow(1.0 + clm_pow(v.alpha * fabs(psi), v.n), -v.m);
}

// ---- heat (soil_heat_parameterizations.jl) --------------------------------
__device__ __forceinline__ double heat_capacity(double thl, double thi, double rhoc_ds) {
    return rhoc_ds + thl * RCP_L + thi * RCP_I;  // :157-171
}
__device__ __forceinline__ double temperature(double rhoe, double thi, double rhoc_s) {
    return T_0 + (rhoe + thi * RHO_I * LH_F0) / rhoc_s;  // :180-192
}
__device__ __forceinline__ double energy_liq(double T) { return RCP_L * (T - T_0); }  // :221-231

// κ (:243-320): Balland–Arp Kersten number and saturated conductivity.
// Both Kersten branches are one power product Sr^p g^q (frozen: p = 1+ν_om,
// g^q = 1), evaluated without divergence as exp(p log Sr + q log g);
// κ_u^(θ_l/θ_w) κ_f^(θ_i/θ_w) = exp((θ_l log κ_u + θ_i log κ_f) / θ_w) with
// the per-column logarithms lk_u, lk_f.
__device__ __forceinline__ double thermal_cond(double thl, double thi, double nu, double om,
                                               double qz, double gr, double k_dry,
                                               double k_satu, double k_satf, double lk_u,
                                               double lk_f) {
    double Sr = (thl + thi) / nu;
    const bool unfrozen = thi < DEPS;
    double xk = 1.0 + exp(-KERSTEN_B * Sr);
    double t1 = 1.0 / (xk * xk * xk);  // (1 + e^(-β Sr))^-3
    double h = (1.0 - Sr) / 2.0;
    double p = unfrozen ? (1.0 + om - KERSTEN_A * qz - gr) / 2.0 : 1.0 + om;
    double q = unfrozen ? 1.0 - om : 0.0;
    double g = unfrozen ? t1 - h * h * h : 1.0;
    double Ke = exp(mul_log(p, clm_log(Sr)) + mul_log(q, clm_log(g)));
    double tw = thl + thi;
    // κ_f^0 = 1 also for κ_f = 0 (an ice-free cell of a column with κ = 0)
    double ks = (tw < DEPS) ? (k_satu + k_satf) / 2.0
                            : exp((mul_log(thl, lk_u) + mul_log(thi, lk_f)) / tw);
    return Ke * ks + (1.0 - Ke) * k_dry;
}

// ψ_w0 of phase_change_source (:89-100): the matric potential of the total
// water θ_l + (ρ_i/ρ_l) θ_i at porosity ν
__device__ __forceinline__ double matric_total(const VG& v, double nu, double thl, double thi) {
    double ttot = fmin(RHO_I / RHO_L * thi + thl, nu);
    double vs = fmax(ttot, v.thr_safe);
    double ns = fmax(nu, v.thr_safe);
    double Sw = (vs - v.theta_r) / (ns - v.theta_r);
    return matric(v, Sw);
}

// phase_change_source (:89-122) with soil_Tf_depressed (:35-52) given ψ_w0; τ
// from thermal_time (:60-62) on the layer width.  Without ice and above the
// depressed freezing point, θ* = inverse_matric_potential(ψ_w0)(ν-θ_r)+θ_r is
// θ_l itself (ψ_w0 is the matric potential of θ_l), so F = 0 exactly; the
// θ* round trip, which would return θ_l up to rounding, is skipped when that
// holds on the whole wave.
__device__ __forceinline__ double phase_change_psi(const VG& v, double nu, double thl,
                                                   double thi, double T, double tau,
                                                   double psiw0, int* bad) {
    double Tf = fmax(T_FREEZE * exp(GRAV * psiw0 / LH_F0), 1.0);
    double H = (Tf - T > DEPS) ? 1.0 : 0.0;  // heaviside, utils.jl:93-99
    const bool still = thi == 0.0 && H == 0.0;
    if (__all(still)) return 0.0;
    double psiT = LH_F0 / GRAV * clm_log(T / Tf) * H;
    double thstar = inv_matric(v, psiw0 + psiT, bad) * (nu - v.theta_r) + v.theta_r;
    return still ? 0.0 : (thl - thstar) / tau;
}

__device__ __forceinline__ double phase_change(const VG& v, double nu, double thl, double thi,
                                               double T, double tau, int* bad) {
    return phase_change_psi(v, nu, thl, thi, T, tau, matric_total(v, nu, thl, thi), bad);
}

}  // namespace clm
