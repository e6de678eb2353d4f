// clm_nodes.h -- the warm-start network's node grid: each column is resampled
// to kNodes nodes in layer-index space (node j sits at fractional layer
// (j + 0.5) N / kNodes - 0.5, node 0 at the bottom) and back, and the units of
// the network's outputs.  Shared by the network kernels (features, guesses)
// and the step kernels, which turn a column's output row into its per-cell
// Newton guess.
#pragma once

namespace clm {
namespace ws {

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

}  // namespace ws
}  // namespace clm
