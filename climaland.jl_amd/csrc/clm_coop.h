// clm_coop.h -- group-cooperative step kernels: G lanes own one column.
//
// A wavefront holds 64/G columns; lane l of a group owns rows R*l .. R*l+R-1
// of its column (level 0 = bottom).  The whole ARS111 step -- explicit stage
// and every Newton iteration -- runs out of registers: the column's state
// is read from HBM once and written once per step.  Neighbour values move
// through the wave's LDS slice in the tridiagonal solves (Xch below) and by
// lane permutes elsewhere.  Newton iterations are group-uniform, so a
// column that needs many iterations holds back only its own wavefront.
//
// Linear solve.  The modified-Picard block W = dtγ J - I with
// J = diag(1/Δz_c) M diag(D) (M: the symmetric face-conductance Laplacian,
// D: dψ/dϑ or 1/ρc_s) is solved in the equivalent form
//     (dtγ M - diag(Δz_c / D)) y = Δz_c ∘ R,     Δ = y / D,
// exactly the same linear system with rows scaled by Δz_c and unknowns by
// D.  That matrix is symmetric and strictly diagonally dominant, so
// parallel cyclic reduction (log2 G exchange steps; one local cyclic-
// reduction level first when R = 2) is stable without pivoting.  For
// EnergyHydrology the ρe-block unknown y_ρ = Δρe / ρc_s is exactly the
// scaled quantity of the convergence norm.
#pragma once
#include <type_traits>

#include "clm_eh.h"
#include "clm_nodes.h"

namespace clm {

// Lane permutes.  Every lane of a group must execute each one: a lane
// masked off by divergent control flow reads back as zero to the lane that
// asks for it.
template <int G, typename T>
__device__ __forceinline__ T g_up(T v, int d) {  // value of lane l-d in the group
    return __shfl_up(v, (unsigned)d, G);
}
template <int G, typename T>
__device__ __forceinline__ T g_down(T v, int d) {  // value of lane l+d in the group
    return __shfl_down(v, (unsigned)d, G);
}
template <int G, typename T>
__device__ __forceinline__ T g_bcast(T v, int src) {
    return __shfl(v, src, G);
}
template <int G>
__device__ __forceinline__ double g_sum(double v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}
template <int G>
__device__ __forceinline__ double g_max(double v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v = nan_max(v, __shfl_xor(v, o, G));
    return v;
}
// group max of 32-bit values: DPP within rows of 16 lanes (quad swaps, then
// the half-row and row mirrors), then across rows the gfx950 lane swaps:
// v_permlane16_swap exchanges rows 0<->1 and 2<->3, v_permlane32_swap the
// two halves of the wave, so with v in both operands the pair holds the
// lane's own value and that of lane l^16 (l^32) -- a VALU operation where a
// lane permute is an LDS round trip
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}
template <int G>
__device__ __forceinline__ uint32_t g_umax(uint32_t v) {
    if constexpr (G >= 2) v = max(v, dpp_u32<0xB1>(v));   // quad_perm [1,0,3,2]
    if constexpr (G >= 4) v = max(v, dpp_u32<0x4E>(v));   // quad_perm [2,3,0,1]
    if constexpr (G >= 8) v = max(v, dpp_u32<0x141>(v));  // row_half_mirror
    if constexpr (G >= 16) v = max(v, dpp_u32<0x140>(v)); // row_mirror
    if constexpr (G >= 32) {
        auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        v = max((uint32_t)p[0], (uint32_t)p[1]);
    }
    if constexpr (G >= 64) {
        auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        v = max((uint32_t)p[0], (uint32_t)p[1]);
    }
    return v;
}

// Newton's convergence test on the group's max |Δ| without an FP64
// reduction: the high words of non-negative doubles are ordered like the
// values, so the max of the high words decides `max < tol` unless it equals
// tol's high word, where the exact FP64 max is taken (a group-uniform branch).
// Returns NaN when some |Δ| is not finite, the exact max in the tie, and
// otherwise 0 or tol: a finite value on the same side of tol as the max.
template <int G>
__device__ __forceinline__ double g_err(double e, double tol) {
    const uint32_t h = (uint32_t)((uint64_t)__double_as_longlong(e) >> 32) & 0x7fffffffu;
    const uint32_t th = (uint32_t)((uint64_t)__double_as_longlong(tol) >> 32) & 0x7fffffffu;
    const uint32_t H = g_umax<G>(h);
    if (H >= 0x7ff00000u) return NAN;
    if (H == th) return g_max<G>(e);
    return H < th ? 0.0 : tol;
}

// A paired wave holds two columns: lanes [0, kPairSplit) one of up to
// kPairSplit rows, lanes [kPairSplit, 64) one of up to 64 - kPairSplit rows.
constexpr int kPairSplit = 40;

// The reductions and broadcasts of one column's lanes: a group of G lanes,
// or (PAIR) the column's part of a paired wave.  The paired reductions work
// within aligned 8-lane half-rows (DPP), none of which holds lanes of both
// columns (kPairSplit is a multiple of 8), then take the half-rows' results
// with lane reads: each column reads only its own lanes, so they stay right
// while the other column's lanes have left a loop.
template <int G, bool PAIR>
struct Grp {
    bool inA = true;  // PAIR: a lane of the first column
    static_assert(!PAIR || (G == 64 && kPairSplit % 8 == 0), "paired wave layout");
    __device__ __forceinline__ uint32_t umax(uint32_t v) const {
        if constexpr (!PAIR) {
            return g_umax<G>(v);
        } else {
            v = g_umax<8>(v);
            uint32_t a = 0, b = 0;
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)v, 8 * h + 7);
                if (8 * h < kPairSplit)
                    a = max(a, r);
                else
                    b = max(b, r);
            }
            return inA ? a : b;
        }
    }
    __device__ __forceinline__ double dmax(double v) const {
        if constexpr (!PAIR) {
            return g_max<G>(v);
        } else {
#pragma unroll
            for (int o = 4; o > 0; o >>= 1) v = nan_max(v, __shfl_xor(v, o, 8));
            const long long q = __double_as_longlong(v);
            double a = 0.0, b = 0.0;
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const int lo = __builtin_amdgcn_readlane((int)(uint32_t)q, 8 * h + 7);
                const int hi = __builtin_amdgcn_readlane((int)(uint32_t)(q >> 32), 8 * h + 7);
                const double r = __longlong_as_double(
                    (long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
                if (8 * h < kPairSplit)
                    a = nan_max(a, r);
                else
                    b = nan_max(b, r);
            }
            return inA ? a : b;
        }
    }
    template <typename T>
    __device__ __forceinline__ T bcast(T v, int src) const {
        if constexpr (!PAIR)
            return g_bcast<G>(v, src);
        else
            return __shfl(v, (inA ? 0 : kPairSplit) + src, 64);
    }
    // g_err for the column's lanes
    __device__ __forceinline__ double err(double e, double tol) const {
        if constexpr (!PAIR) {
            return g_err<G>(e, tol);
        } else {
            const uint32_t h = (uint32_t)((uint64_t)__double_as_longlong(e) >> 32) & 0x7fffffffu;
            const uint32_t th =
                (uint32_t)((uint64_t)__double_as_longlong(tol) >> 32) & 0x7fffffffu;
            const uint32_t H = umax(h);
            if (H >= 0x7ff00000u) return NAN;
            if (H == th) return dmax(e);
            return H < th ? 0.0 : tol;
        }
    }
};

// Reciprocal of a pivot: exact division in FP64; in the FP32 solve of the
// mixed form (an inexact Newton whose FP64 residual decides convergence) the
// hardware reciprocal (1 ulp), one instruction instead of the ten of an IEEE
// division.  The pivots are those of a strictly diagonally dominant system:
// never zero, never denormal.
template <typename JT>
__device__ __forceinline__ JT recip(JT x) {
    if constexpr (std::is_same<JT, float>::value)
        return __builtin_amdgcn_rcpf(x);
    else
        return (JT)1 / x;
}

// Stall test.  The stopping rule is the reference's, max|Δ| < tol on the
// increment (MaximumError); where an iterate reaches the region in which
// effective_saturation clips ϑ_l (soil_hydrology_parameterizations.jl:45-50),
// ψ stops changing with ϑ_l while dψ/dϑ is huge, the steps shrink below tol
// and the iteration stops far from a root.  The residual of the last
// iteration tells the two apart: on 36,157 increment-converged RichardsModel
// and 23,498 EnergyHydrology column-steps of the oracle it stayed below 1e-4
// (in ϑ_l, and in ρe_int/ρc_s, kelvin) except where it was above 1e-2, most
// of those above 1e6.  A column-step that stops with a last residual above
// kStall is flagged (0x100000) and counted as not converged; a warm-started
// or Newton-matrix attempt that stalls is repeated from U0 with the
// reference's iteration, as one that turns non-finite is.
constexpr double kStall = 1e-3;

// Neighbour exchange of the tridiagonal solves through LDS.  Each group of
// G lanes owns 2G consecutive entries of its wave's LDS slice: its lanes at
// [G/2, 3G/2) and identity rows (a = c = d = 0, b = 1/b = 1) in the G/2
// entries on either side, so a read at distance s <= G/2 past either end of
// the column returns an identity row with no select.  One 4-wide entry per
// exchange replaces four lane permutes and their boundary selects.  Only the
// owning wave touches its slice, and a wave's LDS operations complete in
// issue order: the wave barriers keep the compiler from moving the accesses
// across each other, which is all the ordering needed.
template <typename JT>
struct alignas(4 * sizeof(JT)) Q4 {
    JT x, y, z, w;
};

// A paired wave (PAIR) holds two columns (see kPairSplit).  Its slice is
// [32 identity][first column][32 identity][second column][32 identity]:
// reads at distance <= 32 (the six levels of a 64-lane solve) never leave a
// column's own entries and padding.

template <int G, typename JT, bool PAIR = false>
struct Xch {
    static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "group of 2..64 lanes");
    static_assert(!PAIR || G == 64, "a paired layout spans the wave");
    // (64 / G) groups x 2G entries; paired: 3 x 32 padding + 64 lanes
    static constexpr int kWaveEntries = PAIR ? 160 : 128;
    Q4<JT>* e;  // this lane's entry
    int gl;
    __device__ __forceinline__ Xch(Q4<JT>* block_slices, int gl_) : gl(gl_) {
        const int lane = threadIdx.x & 63;
        const Q4<JT> id{(JT)0, (JT)0, (JT)0, (JT)1};
        if constexpr (PAIR) {
            Q4<JT>* seg = block_slices + (threadIdx.x >> 6) * kWaveEntries;
            e = seg + (lane < kPairSplit ? 32 + lane : 104 + (lane - kPairSplit));
            seg[lane < 32 ? lane : 72 + (lane - 32)] = id;  // [0, 32) and [72, 104)
            if (lane < 32) seg[128 + lane] = id;             // [128, 160)
        } else {
            Q4<JT>* seg =
                block_slices + (threadIdx.x >> 6) * kWaveEntries + (lane / G) * (2 * G);
            e = seg + G / 2 + gl;
            Q4<JT>* pad = (gl < G / 2) ? seg + gl : seg + G + gl;
            *pad = id;
        }
    }
    __device__ __forceinline__ void put(JT x, JT y, JT z, JT w) {
        __builtin_amdgcn_wave_barrier();
        *e = Q4<JT>{x, y, z, w};
        __builtin_amdgcn_wave_barrier();
    }
    // one component only (the readers of this exchange use that slot alone)
    template <int SLOT>
    __device__ __forceinline__ void put1(JT v) {
        __builtin_amdgcn_wave_barrier();
        (&e->x)[SLOT] = v;
        __builtin_amdgcn_wave_barrier();
    }
    __device__ __forceinline__ Q4<JT> at(int off) const { return e[off]; }
};

// Parallel cyclic reduction, one row per lane (rows >= N are identity rows
// a = c = d = 0, b = 1).  Returns x with a x_{l-1} + b x_l + c x_{l+1} = d.
template <int G, typename JT, typename XC>
__device__ __forceinline__ JT pcr1(XC& xc, JT a, JT b, JT c, JT d) {
#pragma unroll
    for (int s = 1; s < G; s <<= 1) {
        xc.put(a, c, d, recip(b));
        const Q4<JT> m = xc.at(-s), p = xc.at(s);
        JT al = -a * m.w, ga = -c * p.w;
        JT na = al * m.x, nc = ga * p.y;
        b = b + al * m.y + ga * p.x;
        d = d + al * m.z + ga * p.z;
        a = na;
        c = nc;
    }
    if constexpr (std::is_same<JT, float>::value)
        return d * recip(b);
    else
        return d / b;
}

// Two rows per lane (rows 2l, 2l+1): one cyclic-reduction level folds the
// even rows into the odd ones, PCR solves the odd rows across lanes, the
// even rows are recovered by back substitution.  The next lane's even row
// (a, c, d, b) comes through an exchange entry, whose identity padding
// (b in the fourth slot) is again the row past the column's top.
template <int G, typename JT, typename XC>
__device__ __forceinline__ void pcr2(XC& xc, const JT* a, const JT* b, const JT* c,
                                     const JT* d, JT* x) {
    xc.put(a[0], c[0], d[0], b[0]);
    const Q4<JT> n = xc.at(1);  // row 2l+2
    JT al, ga;
    if constexpr (std::is_same<JT, float>::value) {
        al = -a[1] * recip(b[0]);
        ga = -c[1] * recip(n.w);
    } else {
        al = -a[1] / b[0];
        ga = -c[1] / n.w;
    }
    JT oa = al * a[0];
    JT oc = ga * n.y;
    JT ob = b[1] + al * c[0] + ga * n.x;
    JT od = d[1] + al * d[0] + ga * n.z;
    JT xo = pcr1<G, JT>(xc, oa, ob, oc, od);
    xc.template put1<0>(xo);
    const JT xm = xc.at(-1).x;  // x_{2l-1}; 0 below the column
    x[1] = xo;
    if constexpr (std::is_same<JT, float>::value)
        x[0] = (d[0] - a[0] * xm - c[0] * xo) * recip(b[0]);
    else
        x[0] = (d[0] - a[0] * xm - c[0] * xo) / b[0];
}

// Factor once, solve many: the multipliers of pcr1's elimination depend on
// the matrix only, so a matrix that stays fixed across Newton iterations (the
// EnergyHydrology ρe block: lagged κ, lagged ρc_s) is reduced once per step
// and each iteration only carries its right-hand side through the levels.
template <int G>
struct Log2 {
    static constexpr int v = G <= 1 ? 0 : 1 + Log2<G / 2>::v;
};

template <int G, typename JT>
struct Pcr1Factor {
    static constexpr int L = Log2<G>::v;
    JT al[L], ga[L];
    JT ib;
    template <typename XC>
    __device__ __forceinline__ void factor(XC& xc, JT a, JT b, JT c) {
        int lev = 0;
#pragma unroll
        for (int s = 1; s < G; s <<= 1, ++lev) {
            xc.put(a, c, (JT)0, (JT)1 / b);
            const Q4<JT> m = xc.at(-s), p = xc.at(s);
            JT l = -a * m.w, g = -c * p.w;
            al[lev] = l;
            ga[lev] = g;
            b = b + l * m.y + g * p.x;
            a = l * m.x;
            c = g * p.y;
        }
        ib = (JT)1 / b;
    }
    template <typename XC>
    __device__ __forceinline__ JT solve(XC& xc, JT d) const {
        int lev = 0;
#pragma unroll
        for (int s = 1; s < G; s <<= 1, ++lev) {
            xc.template put1<2>(d);
            const JT dm = xc.at(-s).z, dp = xc.at(s).z;
            d = d + al[lev] * dm + ga[lev] * dp;
        }
        return d * ib;
    }
};

// pcr2 factored: the even-row elimination multipliers, the reduced odd
// system's factor and what the even-row recovery needs
template <int G, typename JT>
struct Pcr2Factor {
    JT al, ga, a0, c0, ib0;
    Pcr1Factor<G, JT> odd;
    template <typename XC>
    __device__ __forceinline__ void factor(XC& xc, const JT* a, const JT* b, const JT* c) {
        xc.put(a[0], c[0], (JT)0, b[0]);
        const Q4<JT> n = xc.at(1);
        ib0 = (JT)1 / b[0];
        al = -a[1] * ib0;
        ga = -c[1] / n.w;
        a0 = a[0];
        c0 = c[0];
        odd.factor(xc, al * a[0], b[1] + al * c[0] + ga * n.x, ga * n.y);
    }
    template <typename XC>
    __device__ __forceinline__ void solve(XC& xc, const JT* d, JT* x) const {
        xc.template put1<2>(d[0]);
        const JT d2 = xc.at(1).z;
        JT xo = odd.solve(xc, d[1] + al * d[0] + ga * d2);
        xc.template put1<0>(xo);
        const JT xm = xc.at(-1).x;
        x[1] = xo;
        x[0] = (d[0] - a0 * xm - c0 * xo) * ib0;
    }
};

// R >= 2 rows per lane: the partition method.  The lane's interior rows
// 1..R-2 form a tridiagonal system with its first and last unknowns as
// boundary values; three Thomas sweeps (right-hand sides d, -a_1 e_1,
// -c_{R-2} e_{R-2}) give x_k = α_k + β_k x_0 + γ_k x_{R-1}.  Substituting
// into rows 0 and R-1 leaves two rows per lane, tridiagonal in the order
// (.., x_{R-1} of lane l-1, x_0, x_{R-1}, x_0 of lane l+1, ..): the Schur
// complement of a diagonally dominant matrix, solved by pcr2.  Interior
// unknowns follow from the formula above.
template <int G, int R, typename JT, typename XC>
__device__ __forceinline__ void pcr_solve(XC& xc, JT (&a)[R], JT (&b)[R], JT (&c)[R],
                                          JT (&d)[R], JT (&x)[R]) {
    if constexpr (R == 1) {
        x[0] = pcr1<G, JT>(xc, a[0], b[0], c[0], d[0]);
    } else if constexpr (R == 2) {
        pcr2<G, JT>(xc, a, b, c, d, x);
    } else {
        constexpr int M = R - 2;  // interior rows 1..R-2
        JT cp[M], al[M], be[M], ga[M];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const int k = j + 1;
            JT den = b[k] - (j > 0 ? a[k] * cp[j - 1] : (JT)0);
            JT inv = recip(den);
            cp[j] = c[k] * inv;
            al[j] = (d[k] - (j > 0 ? a[k] * al[j - 1] : (JT)0)) * inv;
            be[j] = ((j == 0 ? -a[1] : (JT)0) - (j > 0 ? a[k] * be[j - 1] : (JT)0)) * inv;
            ga[j] = ((j == M - 1 ? -c[R - 2] : (JT)0) - (j > 0 ? a[k] * ga[j - 1] : (JT)0)) * inv;
        }
#pragma unroll
        for (int j = M - 2; j >= 0; --j) {
            al[j] -= cp[j] * al[j + 1];
            be[j] -= cp[j] * be[j + 1];
            ga[j] -= cp[j] * ga[j + 1];
        }
        JT ra[2], rb[2], rc[2], rd[2], rx[2];
        ra[0] = a[0];
        rb[0] = b[0] + c[0] * be[0];
        rc[0] = c[0] * ga[0];
        rd[0] = d[0] - c[0] * al[0];
        ra[1] = a[R - 1] * be[M - 1];
        rb[1] = b[R - 1] + a[R - 1] * ga[M - 1];
        rc[1] = c[R - 1];
        rd[1] = d[R - 1] - a[R - 1] * al[M - 1];
        pcr2<G, JT>(xc, ra, rb, rc, rd, rx);
        x[0] = rx[0];
        x[R - 1] = rx[1];
#pragma unroll
        for (int j = 0; j < M; ++j) x[j + 1] = al[j] + be[j] * rx[0] + ga[j] * rx[1];
    }
}

// neighbour values across the lane boundary: the row below my first row is
// the previous lane's last row, the row above my last row the next lane's
// first row
template <int G, int R>
__device__ __forceinline__ double below(const double (&v)[R], int k) {
    return k > 0 ? v[k - 1] : g_up<G>(v[R - 1], 1);
}

struct GroupId {
    int gl, grp;
    bool live;
};

template <int G>
__device__ __forceinline__ GroupId group_id(int ncols_bin) {
    constexpr int GPW = 64 / G;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int lane = threadIdx.x & 63;
    GroupId id;
    id.gl = lane % G;
    id.grp = (t >> 6) * GPW + lane / G;
    id.live = id.grp < ncols_bin;
    return id;
}

// ---- RichardsModel ----------------------------------------------------------
template <int G, int R, typename JT>
__global__ __launch_bounds__(256) void k_rre_coop(
    const int32_t* __restrict__ col_ptr, const int32_t* __restrict__ bin_cols, int ncols_bin,
    const double* __restrict__ par, int64_t ncol, const double* __restrict__ zf,
    double* __restrict__ xg, double* __restrict__ iw, const double* __restrict__ f_top,
    const double* __restrict__ f_bot, int bc, double dt, int max_iters, double tol,
    int32_t* __restrict__ iters, const float* __restrict__ wso, double* __restrict__ u0out) {
    GroupId id = group_id<G>(ncols_bin);
    if (!id.live) return;
    const int gl = id.gl;
    __shared__ Q4<JT> xmem[4 * Xch<G, JT>::kWaveEntries];  // 256 threads: 4 waves
    Xch<G, JT> xc(xmem, gl);
    const int col = bin_cols[id.grp];
    const int64_t base = col_ptr[col];
    const int N = (int)(col_ptr[col + 1] - base);
    const int64_t fbase = base + col;
    const bool warm = wso != nullptr;
    const VG v = make_vg(par[P_VG_ALPHA * ncol + col], par[P_VG_N * ncol + col],
                         par[P_THETA_R * ncol + col], par[P_SS * ncol + col],
                         par[P_KSAT * ncol + col]);
    const double nu = par[P_NU * ncol + col];
    double top = f_top[col];

    double x[R], u0[R], zc[R], dzc[R], dzf[R];
    bool on[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        int r = gl * R + k;
        on[k] = r < N;
        double z0 = 0.0, z1 = 1.0, xv = 0.0;
        if (on[k]) {
            xv = xg[base + r];
            z0 = zf[fbase + r];
            z1 = zf[fbase + r + 1];
        }
        x[k] = xv;
        u0[k] = xv;  // explicit tendency is zero (rre.jl:214-243): U0 = u_n
        zc[k] = 0.5 * (z0 + z1);
        dzc[k] = z1 - z0;
        if (u0out && on[k]) u0out[base + r] = xv;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) dzf[k] = zc[k] - below<G, R>(zc, k);  // face below row
    double idzf[R];  // reciprocal spacings (row 0's face is the boundary: unused)
#pragma unroll
    for (int k = 0; k < R; ++k) idzf[k] = dzf[k] != 0.0 ? 1.0 / dzf[k] : 0.0;
    const Span sp = make_span(v, nu);

    // runoff from update_cache!(u_n), held for the step (Runoff.jl:234-271):
    // TOPMODEL's infiltration and subsurface sink, or SurfaceRunoff's infiltration
    const bool topm = bc_top(bc) == BC_TOPMODEL;
    double sink[R], rss = 0.0;
#pragma unroll
    for (int k = 0; k < R; ++k) sink[k] = 0.0;
    if (topm) {
        const double thr = par[P_THETA_R * ncol + col];
        double hp = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            sink[k] = on[k] ? sat_weight(u0[k], nu, thr) : 0.0;
            hp += sink[k] * dzc[k];
        }
        Runoff o = topmodel(g_sum<G>(hp), zf[fbase + N] - zf[fbase], par[P_F_MAX * ncol + col],
                            par[P_F_OVER * ncol + col], par[P_R_SB * ncol + col], v.Ksat, top);
        top = o.infil;
        rss = o.rss;
#pragma unroll
        for (int k = 0; k < R; ++k) sink[k] *= o.scale;
    } else if (bc_top(bc) == BC_SURFACE_RUNOFF) {
        top = surface_infiltration(xg[base + N - 1], nu, v.Ksat, top);
    }

    // boundary fluxes from update_cache!(u_n); under a top MoistureStateBC
    // they follow the iterate (rre.jl:460-468) and the top row carries the
    // boundary term of set_dfluxBCdY! (boundary_conditions.jl:375-411)
    const int bc_bottom = bc_bot(bc);
    const bool top_state = bc_top(bc) == BC_MOISTURE_STATE;
    double bot;
    {
        PsiD p0 = psi_dpsi(v, nu, u0[0]);
        double K0 = on[0] ? hyd_cond_from_a(v, p0.S, p0.a) : 0.0;
        double Kb = g_bcast<G>(K0, 0), pb = g_bcast<G>(p0.psi, 0), db = g_bcast<G>(dzc[0], 0);
        if (bc_bottom == BC_FREE_DRAINAGE)
            bot = -1.0 * Kb;
        else if (bc_bottom == BC_MOISTURE_STATE)
            bot = moisture_flux_bot(v, nu, f_bot[col], Kb, pb, 0.5 * db);
        else
            bot = f_bot[col];
    }
    double top_used = top;
    int kit = 0, kit_total = 0;
    bool conv = false, fell_back = false, stalled = false;
    double err = 0.0;
    for (int attempt = 0; attempt < 2; ++attempt) {
    if (attempt == 0 && warm) {  // x0 = U0 + guess; cache_imp!(x0) happens at the loop top
        // the guess at each cell from the column's network output row
        const float* o = wso + (int64_t)col * ws::kNodes;
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (on[k]) x[k] = u0[k] + ws::at_cell(gl * R + k, N, o) * ws::kScaleV;
    }
    if (attempt == 1) {  // the warm start went non-finite or stalled: plain Newton from U0
        if (!(warm && (!isfinite(err) || stalled))) break;
        fell_back = true;
#pragma unroll
        for (int k = 0; k < R; ++k) x[k] = u0[k];
    }
    bool stl = false;  // this lane's rows of the last residual above kStall
    for (kit = 1; kit <= max_iters; ++kit) {
        double K[R], h[R], iD[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            PsiI p = psi_invd<JT>(v, sp, x[k]);
            K[k] = on[k] ? hyd_cond_from_a(v, p.S, p.a) : 0.0;
            h[k] = p.psi + zc[k];
            iD[k] = p.invD;
        }
        if (top_state && gl == 0 && bc_bottom != BC_FLUX) {  // bottom flux at the iterate
            double psi0 = h[0] - zc[0];
            bot = (bc_bottom == BC_FREE_DRAINAGE)
                      ? -1.0 * K[0]
                      : moisture_flux_bot(v, nu, f_bot[col], K[0], psi0, 0.5 * dzc[0]);
        }
        // face below each row: conductance w and flux F
        double w[R], F[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            int r = gl * R + k;
            double Km = below<G, R>(K, k), hm = below<G, R>(h, k);
            double ww = 0.5 * (Km + K[k]) * idzf[k];
            w[k] = (r == 0) ? 0.0 : ww;
            F[k] = (r == 0) ? bot : -ww * (h[k] - hm);
        }
        JT a[R], b[R], c[R], d[R], y[R];
        stl = false;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            int r = gl * R + k;
            double wu = (k + 1 < R) ? w[k + 1] : g_down<G>(w[0], 1);
            double Fu = (k + 1 < R) ? F[k + 1] : g_down<G>(F[0], 1);
            double wbc = 0.0;
            if (r >= N - 1) {
                wu = 0.0;
                Fu = top;
                if (top_state && r == N - 1) {
                    Fu = moisture_flux_top(v, nu, top, K[k], h[k] - zc[k], 0.5 * dzc[k]);
                    wbc = K[k] / (0.5 * dzc[k]);
                    top_used = Fu;
                }
            }
            // Δz_c R = Δz_c (U0 - x) - Δt (F_up - F_down) - Δt Δz_c sink
            double rhs = dzc[k] * (u0[k] - x[k]) - dt * (Fu - F[k]);
            if (topm) rhs -= dt * (dzc[k] * sink[k]);
            stl |= on[k] && fabs(rhs) > kStall * dzc[k];
            a[k] = (JT)(dt * w[k]);
            c[k] = (JT)(dt * wu);
            b[k] = (JT)(-dt * (w[k] + wu + wbc) - dzc[k] * iD[k]);
            d[k] = (JT)rhs;
            if (!on[k]) {
                a[k] = 0;
                b[k] = 1;
                c[k] = 0;
                d[k] = 0;
            }
        }
        pcr_solve<G, R, JT>(xc, a, b, c, d, y);
        double e = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            double dx = on[k] ? (double)y[k] * iD[k] : 0.0;
            x[k] -= dx;
            e = nan_max(e, fabs(dx));
        }
        err = g_err<G>(e, tol);
        conv = (tol > 0.0) && (err < tol);
        if (conv || !isfinite(err)) break;
    }
    if (kit > max_iters) kit = max_iters;
    kit_total += kit;
    {
        const bool any_stl = g_umax<G>((uint32_t)stl) != 0u;  // every lane of the group
        stalled = conv && any_stl;
    }
    }
#pragma unroll
    for (int k = 0; k < R; ++k)
        if (on[k]) xg[base + gl * R + k] = x[k];
    if (top_state) top_used = g_bcast<G>(top_used, (N - 1) / R);
    if (gl == 0) {
        iw[col] = iw[col] + dt * (-(top_used - bot) - rss);  // ∫F: -I block, fluxes of the last iteration
        int32_t flag = kit_total;
        if (tol > 0.0 && (!conv || stalled)) flag |= 0x10000;
        if (!isfinite(err)) flag |= 0x20000;
        if (fell_back) flag |= 0x80000;
        if (stalled) flag |= 0x100000;
        iters[col] = flag;
    }
}

// ---- EnergyHydrology ----------------------------------------------------------
// ATM: top boundary AtmosDrivenFluxBC in soil-only form (bc_top BC_ATMOS or
// BC_ATMOS_TOPMODEL, surface inputs sf per column; ftw / fth are not read).
// In update_cache!(u_n): the infiltration (TOPMODEL or the precipitation),
// the top fluxes, and the subsurface sinks held through the step; in the
// explicit stage the sublimation source on θ_i and ∫F_w.
// Blocks per CU the EnergyHydrology kernels are compiled for: the mixed form
// with one row per lane fits four waves per SIMD (<= 128 VGPRs, a few spills
// outside the Newton loop), which hides more of the latency of the dependent
// solve levels than the three its register need would give.  With several
// rows per lane (deeper columns) four waves would spill inside the loop, and
// the FP64 form spills at four: those keep their register need.
template <typename JT, int R>
constexpr int kEhMinBlocks = (std::is_same<JT, float>::value && R == 1) ? 4 : 1;

// NEWTON: the Newton form of the inexact Newton (no counterpart in the
// reference).  The same FP64 residual, the same water block (K is lagged, so
// dψ/dϑ is its whole derivative), and the exact derivative of the energy
// residual in place of the modified-Picard blocks: with ΔT = Δρe/ρc_s(x) +
// ∂T/∂ϑ Δϑ (ρc_s and ∂T/∂ϑ = -(T - T_0) ρ_l c_l/ρc_s at θ_l = min(ν - θ_i,
// ϑ_l) of the iterate) the heat flux of the face below a row changes by
//   ΔG = cTlo ΔT_below + cThi ΔT + wE (Δh_below - Δh),
//   cTlo = κ_f/Δz_f - ½ ρ_l c_l K_below gh,  cThi = -κ_f/Δz_f - ½ ρ_l c_l K gh,
// (gh the face's head gradient).  The energy block then carries the
// advection of heat by the Darcy flux, is no longer symmetric and changes
// with every iterate, so it is solved in full each iteration; W_ρϑ has no
// "- I".  A column whose Newton iterate turns non-finite is stepped again from
// U0 with the reference's matrices (flag bit 0x80000).
//
// PAIR (G = 64, R = 1, prescribed-flux top): wave w holds column bin_cols[w]
// (up to kPairSplit layers) in lanes [0, kPairSplit) and column bin_cols_b[w]
// (up to 64 - kPairSplit layers; none for w >= ncols_b) in the other lanes,
// so a column of 33-40 layers no longer leaves a third of its wave idle.
// Each column's rows see only its own rows and identity padding (Xch), its
// own lanes' reductions and broadcasts (Grp), and neighbour values from the
// other column only where a boundary select discards them, so its results
// are those of a wave of its own, bit for bit.
template <int G, int R, typename JT, bool ATM = false, bool NEWTON = false, bool PAIR = false>
__global__ __launch_bounds__(256, (kEhMinBlocks<JT, R>)) void k_eh_coop(
    const int32_t* __restrict__ col_ptr, const int32_t* __restrict__ bin_cols, int ncols_bin,
    const double* __restrict__ par, int64_t ncol, const double* __restrict__ zf,
    double* __restrict__ xvg, double* __restrict__ xig, double* __restrict__ xeg,
    double* __restrict__ iw, double* __restrict__ ie, const double* __restrict__ ftw,
    const double* __restrict__ fth, const double* __restrict__ fbw,
    const double* __restrict__ fbh, int bc_bottom, double dt, int max_iters, double tol,
    int32_t* __restrict__ iters, const float* __restrict__ wso,
    double* __restrict__ u0out, int bc_top, Surface sf,
    const int32_t* __restrict__ bin_cols_b, int ncols_b) {
    static_assert(!PAIR || (G == 64 && R == 1 && !ATM), "paired waves: 64 lanes, one row each");
    int gl, wave = 0;
    Grp<G, PAIR> grp;
    if constexpr (PAIR) {
        wave = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
        if (wave >= ncols_bin) return;  // wave-uniform
        const int lane = threadIdx.x & 63;
        grp.inA = lane < kPairSplit;
        gl = grp.inA ? lane : lane - kPairSplit;
    } else {
        GroupId id = group_id<G>(ncols_bin);
        if (!id.live) return;
        gl = id.gl;
        wave = id.grp;
    }
    __shared__ Q4<JT> xmem[4 * Xch<G, JT, PAIR>::kWaveEntries];  // 256 threads: 4 waves
    Xch<G, JT, PAIR> xc(xmem, gl);
    // PAIR: the lanes of a missing second column run on the first one's index
    // with no rows.  A 64-lane group is the whole wave: its column index,
    // offsets and parameters are wave-uniform (scalar registers and loads).
    const bool has_col = !PAIR || grp.inA || wave < ncols_b;
    const int col = PAIR ? ((grp.inA || !has_col) ? bin_cols[wave] : bin_cols_b[wave])
                         : (G == 64 ? __builtin_amdgcn_readfirstlane(bin_cols[wave])
                                    : bin_cols[wave]);
    const int64_t base = col_ptr[col];
    const int N = has_col ? (int)(col_ptr[col + 1] - base) : 0;
    const int64_t fbase = base + col;
    const EHPar e = load_eh(par, ncol, col);
    double topw = ATM ? 0.0 : ftw[col], toph = ATM ? 0.0 : fth[col];
    // AtmosDrivenFluxBC: subsurface runoff sinks per row, R_ss, R_ess, and the
    // explicit ∫F_w term of sublimation
    double snk_v[ATM ? R : 1], snk_e[ATM ? R : 1];
    double rss = 0.0, ress = 0.0, vice = 0.0, dzt = 1.0;
    if constexpr (ATM) {
        vice = sf.vice[col];
        const double ztop = zf[fbase + N];
        dzt = ztop - 0.5 * (zf[fbase + N - 1] + ztop);
    }

    const bool warm = wso != nullptr;
    double xv[R], xi[R], xe[R], u0v[R], u0e[R], zc[R], dzc[R], dzf[R];
    double Kl[R], kl[R], rcs[R], Tp[R], Pp[R], iD[R];
    // Newton form: ρc_s and ∂T/∂ϑ_l at the iterate
    double rcx[NEWTON ? R : 1], dTdv[NEWTON ? R : 1];
    Span sp[R];  // saturation span of each row (θ_i is fixed for the step after stage 1)
    bool on[R];
    int bad = 0;
    // ---- stage 1: update_cache!(u_n) + phase change -> U0 -------------------
#pragma unroll
    for (int k = 0; k < R; ++k) {
        int r = gl * R + k;
        on[k] = r < N;
        double z0 = 0.0, z1 = 1.0, v = 0.3, t = 0.0, re = 0.0;
        if (on[k]) {
            v = xvg[base + r];
            t = xig[base + r];
            re = xeg[base + r];
            z0 = zf[fbase + r];
            z1 = zf[fbase + r + 1];
        }
        zc[k] = 0.5 * (z0 + z1);
        dzc[k] = z1 - z0;
        EHAux a = eh_cell_aux(e, v, t, re, !warm);  // a warm start evaluates ψ at its x0
        double F = 0.0;
        if (on[k]) {
            double tau = 3.0 * heat_capacity(a.thl, t, e.rhoc_ds) * dzc[k] * dzc[k] / a.kap;
            F = phase_change_psi(e.v, e.nu, a.thl, t, a.T, tau, a.psiw0, &bad);
        }
        u0v[k] = v + dt * (-F);
        if (u0out && on[k]) u0out[base + r] = u0v[k];
        double dti = (RHO_L / RHO_I) * F;
        if constexpr (ATM) {
            if (on[k]) dti += sublimation(vice, zc[k], dzt);
            // TOPMODEL weights of u_n: surface (with ice) in snk_e, liquid in snk_v
            snk_e[k] = on[k] ? surface_weight(v, t, e.nu, e.v.theta_r) : 0.0;
            snk_v[k] = on[k] ? sat_weight(v, e.nu, e.v.theta_r) : 0.0;
            // the top layer's θ_l, θ_i, T for the infiltration capacity
            if (on[k] && r == N - 1) {
                rss = a.thl;
                ress = t;
                dzf[k] = a.T;  // scratch until the spacings are formed below
            }
        }
        xi[k] = t + dt * dti;
        u0e[k] = re;
        xv[k] = u0v[k];
        xe[k] = re;
        Kl[k] = on[k] ? a.K : 0.0;
        kl[k] = on[k] ? a.kap : 0.0;
        rcs[k] = heat_capacity(a.thl, xi[k], e.rhoc_ds);  // 1/D of the ρe block (lagged θ_l)
        Tp[k] = a.T;
        Pp[k] = a.psi;
        sp[k] = make_span(e.v, e.nu - xi[k]);
        if (!warm) {  // Wfact at x = U0 (a warm start evaluates it at its own x0)
            iD[k] = psi_invd<JT>(e.v, sp[k], xv[k]).invD;
            if constexpr (NEWTON) newton_dT<JT>(e, xv[k], xi[k], Tp[k], rcx[k], dTdv[k]);
        }
    }
    if constexpr (ATM) {
        // update_infiltration_water_flux! (Runoff.jl:234-290) and the top fluxes
        double Ttop = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (on[k] && gl * R + k == N - 1) Ttop = dzf[k];
        const int tl = (N - 1) / R;
        const double thl_t = g_bcast<G>(rss, tl), ti_t = g_bcast<G>(ress, tl);
        Ttop = g_bcast<G>(Ttop, tl);
        double infil = sf.P[col];
        rss = ress = 0.0;
        if (bc_top == BC_ATMOS_TOPMODEL) {
            double hs = 0.0, h = 0.0, he = 0.0;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                hs += snk_e[k] * dzc[k];
                h += snk_v[k] * dzc[k];
                he += snk_v[k] * energy_liq(Tp[k]) * dzc[k];
            }
            hs = g_sum<G>(hs);
            h = g_sum<G>(h);
            he = g_sum<G>(he);
            const double depth = zf[fbase + N] - zf[fbase];
            const double f_over = par[P_EH_F_OVER * ncol + col];
            double ic = eh_infil_capacity(e.v.Ksat, e.v.theta_r, thl_t, ti_t, Ttop);
            double fsat = fmin(par[P_EH_F_MAX * ncol + col] * exp(-f_over / 2.0 * (depth - hs)),
                               1.0);
            infil = (1.0 - fsat) * fmax(ic, sf.P[col]);
            rss = par[P_EH_R_SB * ncol + col] * exp(-f_over * (depth - h));
            ress = he * (rss / fmax(h, DEPS));
#pragma unroll
            for (int k = 0; k < R; ++k) {
                snk_e[k] = (ress / fmax(h, DEPS)) * snk_v[k];
                snk_v[k] = (rss / fmax(h, DEPS)) * snk_v[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < R; ++k) snk_v[k] = snk_e[k] = 0.0;
        }
        eh_atmos_fluxes(sf, col, infil, &topw, &toph);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) dzf[k] = zc[k] - below<G, R>(zc, k);
    double idzf[R];  // reciprocal spacings (row 0's face is the boundary: unused)
#pragma unroll
    for (int k = 0; k < R; ++k) idzf[k] = dzf[k] != 0.0 ? 1.0 / dzf[k] : 0.0;
    double botw, both;
    {
        double K0 = grp.bcast(Kl[0], 0), T0 = grp.bcast(Tp[0], 0);
        if (bc_bottom == BC_FREE_DRAINAGE) {  // boundary_conditions.jl:590-608
            botw = -1.0 * K0;
            both = -1.0 * K0 * energy_liq(T0);
        } else {
            botw = fbw[col];
            both = fbh[col];
        }
    }
    // face means and conductances of the face below each row (lagged K, κ)
    double Kf[R], kf[R], wK[R], wk[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        int r = gl * R + k;
        double Km = below<G, R>(Kl, k), km = below<G, R>(kl, k);
        Kf[k] = 0.5 * (Km + Kl[k]);
        kf[k] = 0.5 * (km + kl[k]);
        wK[k] = (r == 0) ? 0.0 : Kf[k] * idzf[k];
        wk[k] = (r == 0) ? 0.0 : kf[k] * idzf[k];
    }
    // the same conductances on the face above each row (zero above the top row)
    double wKa[R], wka[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const bool last = gl * R + k >= N - 1;
        double wKn = (k + 1 < R) ? wK[k + 1] : g_down<G>(wK[0], 1);
        double wkn = (k + 1 < R) ? wk[k + 1] : g_down<G>(wk[0], 1);
        wKa[k] = last ? 0.0 : wKn;
        wka[k] = last ? 0.0 : wkn;
    }

    // the ρe block of W is fixed for the whole step (lagged κ and ρc_s): with
    // one or two rows per lane it is reduced once here
    using EFac = typename std::conditional<R == 1, Pcr1Factor<G, JT>, Pcr2Factor<G, JT>>::type;
    EFac efac;
    if constexpr (R <= 2 && !NEWTON) {
        JT ea[R], eb[R], ec[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            ea[k] = (JT)(dt * wk[k]);
            ec[k] = (JT)(dt * wka[k]);
            eb[k] = (JT)(-dt * (wk[k] + wka[k]) - dzc[k] * rcs[k]);
            if (!on[k]) {
                ea[k] = 0;
                eb[k] = 1;
                ec[k] = 0;
            }
        }
        if constexpr (R == 1)
            efac.factor(xc, ea[0], eb[0], ec[0]);
        else
            efac.factor(xc, ea, eb, ec);
    }
    int kit = 0, kit_total = 0;
    bool conv = false, fell_back = false, stalled = false;
    double err = 0.0;
    for (int attempt = 0; attempt < 2; ++attempt) {
    // the Newton matrices on the first attempt, the reference's on a fallback
    const bool nw = NEWTON && attempt == 0;
    if (attempt == 0 && warm) {  // x0 = U0 + guess, cache_imp!(x0) and Wfact(x0)
        // the guess at each cell from the column's network output row: ϑ_l,
        // then ρe_int / ρc_s(u_n) in kelvin, ρc_s at θ_l = min(ν - θ_i, ϑ_l) of
        // u_n (re-read: HBM holds u_n until the end)
        const float* o = wso + (int64_t)col * (2 * ws::kNodes);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            int r = gl * R + k;
            if (on[k]) {
                const double v0 = xvg[base + r], t0 = xig[base + r];
                xv[k] = u0v[k] + ws::at_cell(r, N, o) * ws::kScaleV;
                xe[k] = u0e[k] + ws::at_cell(r, N, o + ws::kNodes) * ws::kScaleT *
                                     heat_capacity(fmin(e.nu - t0, v0), t0, e.rhoc_ds);
            }
            if constexpr (NEWTON)
                newton_T<JT>(e, xv[k], xi[k], xe[k], Tp[k], rcx[k], dTdv[k]);
            else
                Tp[k] = eh_imp_T(e, xv[k], xi[k], xe[k]);
            PsiI pd = psi_invd<JT>(e.v, sp[k], xv[k]);
            Pp[k] = pd.psi;
            iD[k] = pd.invD;
        }
    }
    if (attempt == 1) {  // the warm start or the Newton matrices went non-finite or
                         // stalled: Newton with the reference's matrices from U0
        if (!((warm || NEWTON) && (!isfinite(err) || stalled))) break;
        fell_back = true;
#pragma unroll
        for (int k = 0; k < R; ++k) {  // the cache of u_n, re-read (HBM holds u_n until the end)
            int r = gl * R + k;
            double v0 = 0.3, t0 = 0.0, e0 = 0.0;
            if (on[k]) {
                v0 = xvg[base + r];
                t0 = xig[base + r];
                e0 = xeg[base + r];
            }
            EHAux a = eh_cell_aux(e, v0, t0, e0);
            Tp[k] = a.T;
            Pp[k] = a.psi;
            xv[k] = u0v[k];
            xe[k] = u0e[k];
            iD[k] = psi_invd<JT>(e.v, sp[k], xv[k]).invD;
        }
    }
    bool stl = false;  // this lane's rows of the last residual above kStall
    for (kit = 1; kit <= max_iters; ++kit) {
        double h[R], E[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            h[k] = Pp[k] + zc[k];                          // p.ψ
            E[k] = energy_liq(Tp[k]) * Kl[k];              // p.T
        }
        // face below each row: Darcy flux F, heat flux G, advective conductance wE
        double F[R], Gf[R], wE[R];
        // Newton form: ∂G/∂T of the face below each row (zero on the boundary
        // faces and below rows past the column's top)
        double ctl[NEWTON ? R : 1], cth[NEWTON ? R : 1];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            int r = gl * R + k;
            double hm = below<G, R>(h, k), Tm = below<G, R>(Tp, k), Em = below<G, R>(E, k);
            double gh = (h[k] - hm) * idzf[k];
            double eK = 0.5 * (Em + E[k]);
            wE[k] = (r == 0) ? 0.0 : eK * idzf[k];
            F[k] = (r == 0) ? botw : -Kf[k] * gh;
            Gf[k] = (r == 0) ? both : -kf[k] * ((Tp[k] - Tm) * idzf[k]) - eK * gh;
            if constexpr (NEWTON) {
                const double Km = 2.0 * Kf[k] - Kl[k];  // lagged K of the row below
                const bool face = nw && r > 0 && on[k];
                ctl[k] = face ? wk[k] - 0.5 * RCP_L * Km * gh : 0.0;
                cth[k] = face ? -wk[k] - 0.5 * RCP_L * Kl[k] * gh : 0.0;
            }
        }
        JT a[R], b[R], c[R], d[R], y[R];
        double dRe[R];  // Δz_c R_ρe
        stl = false;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            int r = gl * R + k;
            bool last = r >= N - 1;
            // shuffles run on every lane of the group (a lane that skips one
            // would read as zero to its neighbour); boundary selects come after
            double Fu = (k + 1 < R) ? F[k + 1] : g_down<G>(F[0], 1);
            double Gu = (k + 1 < R) ? Gf[k + 1] : g_down<G>(Gf[0], 1);
            const double wu = wKa[k];
            if (last) {
                Fu = topw;
                Gu = toph;
            }
            double rv = dzc[k] * (u0v[k] - xv[k]) - dt * (Fu - F[k]);
            dRe[k] = dzc[k] * (u0e[k] - xe[k]) - dt * (Gu - Gf[k]);
            if constexpr (ATM) {  // implicit TOPMODELSubsurfaceRunoff (Runoff.jl:346-359)
                rv -= dt * (dzc[k] * snk_v[k]);
                dRe[k] -= dt * (dzc[k] * snk_e[k]);
            }
            stl |= on[k] && (fabs(rv) > kStall * dzc[k] ||
                             fabs(dRe[k]) > kStall * (dzc[k] * rcs[k]));
            a[k] = (JT)(dt * wK[k]);
            c[k] = (JT)(dt * wu);
            b[k] = (JT)(-dt * (wK[k] + wu) - dzc[k] * iD[k]);
            d[k] = (JT)rv;
            if (!on[k]) {
                a[k] = 0;
                b[k] = 1;
                c[k] = 0;
                d[k] = 0;
            }
        }
        pcr_solve<G, R, JT>(xc, a, b, c, d, y);  // y = D ∘ Δϑ
        // ρe block right-hand side: Δz_c (R_ρe - W_ρϑ Δϑ), with
        // Δz_c W_ρϑ Δϑ = dt M_E y - Δz_c Δϑ (the reference's "- I", :554-556)
        double yv[R], dv[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            yv[k] = on[k] ? (double)y[k] : 0.0;
            dv[k] = on[k] ? yv[k] * iD[k] : 0.0;
        }
        // y of the row below the lane's first row and above its last (0 past
        // the column's ends, where the conductances are 0); in the Newton form
        // also the temperature change ∂T/∂ϑ Δϑ of those rows
        double tau[NEWTON ? R : 1];
        if constexpr (NEWTON) {
#pragma unroll
            for (int k = 0; k < R; ++k) tau[k] = nw ? dTdv[k] * dv[k] : 0.0;
            xc.put(on[R - 1] ? y[R - 1] : (JT)0, on[0] ? y[0] : (JT)0, (JT)tau[R - 1],
                   (JT)tau[0]);
        } else {
            xc.put(on[R - 1] ? y[R - 1] : (JT)0, on[0] ? y[0] : (JT)0, (JT)0, (JT)0);
        }
        const Q4<JT> nb = xc.at(-1), na = xc.at(1);
        const double yb = (double)nb.x, ya = (double)na.y;
        // Newton form: the known part of each face-below heat-flux change,
        // cTlo τ_below + cThi τ, and the face-above values from the next row
        double kt[NEWTON ? R : 1], ctl_u[NEWTON ? R : 1], cth_u[NEWTON ? R : 1],
            kt_u[NEWTON ? R : 1];
        if constexpr (NEWTON) {
#pragma unroll
            for (int k = 0; k < R; ++k)
                kt[k] = ctl[k] * (k > 0 ? tau[k - 1] : (double)nb.z) + cth[k] * tau[k];
            xc.put((JT)ctl[0], (JT)cth[0], (JT)kt[0], (JT)0);
            const Q4<JT> nu = xc.at(1);
#pragma unroll
            for (int k = 0; k < R; ++k) {
                ctl_u[k] = (k + 1 < R) ? ctl[k + 1] : (double)nu.x;
                cth_u[k] = (k + 1 < R) ? cth[k + 1] : (double)nu.y;
                kt_u[k] = (k + 1 < R) ? kt[k + 1] : (double)nu.z;
            }
        }
        double e2 = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            int r = gl * R + k;
            bool last = r >= N - 1;
            double wEn = (k + 1 < R) ? wE[k + 1] : g_down<G>(wE[0], 1);
            double ym = k > 0 ? yv[k - 1] : yb;
            double yp = (k + 1 < R) ? yv[k + 1] : ya;
            double wEu = last ? 0.0 : wEn;
            double MEy = wE[k] * (ym - yv[k]) + wEu * (yp - yv[k]);
            if (NEWTON && nw) {
                double rhs = dRe[k] - dt * MEy + dt * (kt_u[k] - kt[k]);
                a[k] = (JT)(dt * ctl[k]);
                c[k] = (JT)(-dt * cth_u[k]);
                b[k] = (JT)(-dzc[k] * rcx[k] + dt * (cth[k] - ctl_u[k]));
                d[k] = (JT)rhs;
            } else {
                double rhs = dRe[k] - (dt * MEy - dzc[k] * dv[k]);
                a[k] = (JT)(dt * wk[k]);
                c[k] = (JT)(dt * wka[k]);
                b[k] = (JT)(-dt * (wk[k] + wka[k]) - dzc[k] * rcs[k]);
                d[k] = (JT)rhs;
            }
            if (!on[k]) {
                a[k] = 0;
                b[k] = 1;
                c[k] = 0;
                d[k] = 0;
            }
            xv[k] -= dv[k];
            e2 = nan_max(e2, fabs(dv[k]));
        }
        if constexpr (R == 1 && !NEWTON)
            y[0] = efac.solve(xc, d[0]);  // y = Δρe / ρc_s
        else if constexpr (R == 2 && !NEWTON)
            efac.solve(xc, d, y);
        else
            pcr_solve<G, R, JT>(xc, a, b, c, d, y);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            double yr = on[k] ? (double)y[k] : 0.0;
            // Newton form: y = Δρe / ρc_s(x); the norm is Δρe / ρc_s(θ_l lagged)
            if (NEWTON && nw) {
                xe[k] -= yr * rcx[k];
                e2 = nan_max(e2, fabs(yr * (rcx[k] / rcs[k])));
            } else {
                xe[k] -= yr * rcs[k];
                e2 = nan_max(e2, fabs(yr));
            }
            // cache_imp!(x): T and ψ at the new iterate (energy_hydrology.jl:427-445);
            // dψ/dϑ of the same evaluation is the next Wfact(x)
            if constexpr (NEWTON)
                newton_T<JT>(e, xv[k], xi[k], xe[k], Tp[k], rcx[k], dTdv[k]);
            else
                Tp[k] = eh_imp_T(e, xv[k], xi[k], xe[k]);
            PsiI pd = psi_invd<JT>(e.v, sp[k], xv[k]);
            Pp[k] = pd.psi;
            iD[k] = pd.invD;
        }
        err = grp.err(e2, tol);
        conv = (tol > 0.0) && (err < tol);
        if (conv || !isfinite(err)) break;
    }
    if (kit > max_iters) kit = max_iters;
    kit_total += kit;
    {
        const bool any_stl = grp.umax((uint32_t)stl) != 0u;  // every lane of the group
        stalled = conv && any_stl;
    }
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (on[k]) {
            int64_t i = base + gl * R + k;
            xvg[i] = xv[k];
            xig[i] = xi[k];
            xeg[i] = xe[k];
        }
    }
    const uint32_t bad_g = grp.umax((uint32_t)bad);
    if (gl == 0 && has_col) {
        // ∫F: U0 (explicit: sublimation) then the -I block with the step's fluxes
        double iw0 = ATM ? iw[col] + dt * (0.0 + -vice) : iw[col];
        iw[col] = iw0 + dt * (-(topw - botw) - rss);
        ie[col] = ie[col] + dt * (-(toph - both) - ress);
        int32_t flag = kit_total;
        if (tol > 0.0 && (!conv || stalled)) flag |= 0x10000;
        if (!isfinite(err)) flag |= 0x20000;
        if (bad_g) flag |= 0x40000;
        if (fell_back) flag |= 0x80000;
        if (stalled) flag |= 0x100000;
        iters[col] = flag;
    }
}

}  // namespace clm
