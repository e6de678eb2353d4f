// clm_nodes.h -- the warm-start network's node grid: each column is resampled
// to kNodes nodes in layer-index space (node j sits at fractional layer
// (j + 0.5) N / kNodes - 0.5, node 0 at the bottom) and back, the units of
// the network's outputs, and the arithmetic of a column's feature row.
// Shared by the network kernels (features, guesses) and the step kernels,
// which turn a column's output row into its per-cell Newton guess.
#pragma once
#include "clm_eh.h"

namespace clm {
namespace ws {

typedef __bf16 bf16;

constexpr int kNodes = 32;
constexpr double kScaleV = 1e-2;  // ϑ_l output unit
constexpr double kScaleT = 1.0;   // ρe/ρc_s output unit (K)

// ---- resampling between layers and nodes ------------------------------------
// node j of a column with N cells lies between cells i0 and i1, at weight t
// of i1
struct NodeAt {
    int i0, i1;
    double t;
};
__device__ __forceinline__ NodeAt node_at(int j, int N) {
    double q = ((double)j + 0.5) * (double)N / (double)kNodes - 0.5;
    q = fmin(fmax(q, 0.0), (double)(N - 1));
    NodeAt a;
    a.i0 = (int)q;
    a.i1 = a.i0 + 1 < N ? a.i0 + 1 : N - 1;
    a.t = q - (double)a.i0;
    return a;
}
__device__ __forceinline__ double lerp_node(const NodeAt& a, double v0, double v1) {
    return (1.0 - a.t) * v0 + a.t * v1;
}
// value at node j of a column with N cells (v: cell accessor)
template <typename F>
__device__ __forceinline__ double at_node(int j, int N, F v) {
    const NodeAt a = node_at(j, N);
    return lerp_node(a, v(a.i0), v(a.i1));
}
// value at cell i of node values o[0..kNodes)
__device__ __forceinline__ double at_cell(int i, int N, const float* __restrict__ o) {
    double q = ((double)i + 0.5) * (double)kNodes / (double)N - 0.5;
    q = fmin(fmax(q, 0.0), (double)(kNodes - 1));
    int j0 = (int)q;
    int j1 = j0 + 1 < kNodes ? j0 + 1 : kNodes - 1;
    double t = q - (double)j0;
    return (1.0 - t) * (double)o[j0] + t * (double)o[j1];
}

// ---- feature rows -------------------------------------------------------------
// A row holds the state at the nodes (ϑ_l; EnergyHydrology also T and θ_i:
// state_features(eh) entries), then the tail features (per-column constants, the
// top forcing and Δt) and zero padding to Fp entries.
constexpr int kFpEh = 128, kFpRre = 64;
constexpr int kTails = 8;  // per-column static tail features, precomputed
__host__ __device__ constexpr int state_features(bool eh) { return (eh ? 3 : 1) * kNodes; }

// Features are O(1) by construction.  A column whose iterate blew up enters
// the network as 0 (non-finite) or clamped to +-20 (finite but absurd), so
// that no activation overflows: such columns are masked out of training
// (a zero loss gradient times an overflowed activation would still be NaN)
// and their guess is discarded when Newton falls back to the plain start.
__device__ __forceinline__ bf16 feat(double x) {
    return (bf16)(float)(isfinite(x) ? fmin(fmax(x, -20.0), 20.0) : 0.0);
}
// The state features are formed in FP32 from FP32 copies of the cell values
// (every feature is rounded to BF16): node positions are exact in FP32
// ((j + 1/2) N / 32 with N <= 64), divisions are hardware reciprocals
struct NodeAtF {
    int i0, i1;
    float t;
};
__device__ __forceinline__ NodeAtF node_at_f(int j, int N) {
    float q = ((float)j + 0.5f) * (float)N * (1.0f / (float)kNodes) - 0.5f;
    q = fminf(fmaxf(q, 0.0f), (float)(N - 1));
    NodeAtF a;
    a.i0 = (int)q;
    a.i1 = a.i0 + 1 < N ? a.i0 + 1 : N - 1;
    a.t = q - (float)a.i0;
    return a;
}
// one explicit FMA: the compiler would contract (1 - t) v0 + t v1 either way
// round, differently in different kernels
__device__ __forceinline__ float lerp_f(const NodeAtF& a, float v0, float v1) {
    return fmaf(a.t, v1, (1.0f - a.t) * v0);
}
__device__ __forceinline__ bf16 featf(float x) {
    return (bf16)(isfinite(x) ? fminf(fmaxf(x, -20.0f), 20.0f) : 0.0f);
}
// T - T_0 of a cell (energy_hydrology.jl:430-440 with θ_l = min(ν - θ_i, ϑ_l)),
// from the FP64 state
__device__ __forceinline__ float feat_dT(double nu, double rcd, double v, double t, double e) {
    return (float)(e + t * RHO_I * LH_F0) *
           __builtin_amdgcn_rcpf((float)heat_capacity(fmin(nu - t, v), t, rcd));
}
// node a's state features from the cells around it (FP32 cell values ϑ_l,
// θ_i and T - T_0 at a.i0 and a.i1; thr = θ_r, nu = ν):
// (ϑ_l - θ_r)/(ν - θ_r), (T - 273.15)/10, θ_i/ν
__device__ __forceinline__ bf16 feat_v(const NodeAtF& a, float thr, float nu, float v0, float v1) {
    return featf((lerp_f(a, v0, v1) - thr) * __builtin_amdgcn_rcpf(nu - thr));
}
__device__ __forceinline__ bf16 feat_T(const NodeAtF& a, float d0, float d1) {
    return featf(((float)(T_0 - 273.15) + lerp_f(a, d0, d1)) * 0.1f);
}
__device__ __forceinline__ bf16 feat_i(const NodeAtF& a, float nu, float t0, float t1) {
    return featf(lerp_f(a, t0, t1) * __builtin_amdgcn_rcpf(nu));
}
// tail entry q (0-based after the state entries) of a column: its static
// tails tc[0..5] (log10 α, n - 2, K_sat, ν, θ_r, N / 40), the top water and
// heat fluxes fw, fh and Δt in the network's units; RichardsModel has no heat
// flux entry
__device__ __forceinline__ bf16 tail_entry(bool eh, int q, const bf16* __restrict__ tc,
                                           double fw, double fh, double dt) {
    int i = q;
    if (!eh && i >= 6) ++i;
    if (q >= (eh ? 9 : 8)) return (bf16)0.0f;
    if (i < 5) return tc[i];
    if (i == 8) return tc[5];
    return feat(i == 5 ? fw * 1e5 : (i == 6 ? fh * 0.01 : dt * (1.0 / 3600.0)));
}

}  // namespace ws
}  // namespace clm
