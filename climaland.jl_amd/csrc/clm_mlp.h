// clm_mlp.h -- learned Newton warm start: a per-column MLP that predicts the
// implicit part of the step's increment, x_{n+1} - U0, from the column state.
//
// Network: input features -> H -> H -> H -> 2x32 node outputs, GELU (tanh form) between
// layers, BF16 weights and activations, FP32 accumulation on MFMA
// (v_mfma_f32_32x32x16_bf16).  Columns are resampled to 32 nodes in layer-
// index space (node j sits at fractional layer (j + 0.5) N / 32 - 0.5, node 0
// at the bottom), so the stretched grid keeps its surface resolution.
//
// Features (one row per column, zero-padded to Fp columns):
//   [0, 32)   (ϑ_l - θ_r) / (ν - θ_r) at the nodes
//   EnergyHydrology only: [32, 64) (T - 273.15) / 10, [64, 96) θ_i / ν
//   then log10 α, n - 2, (log10 K_sat + 5.5) / 1.5, (ν - 0.45) / 0.1,
//   (θ_r - 0.06) / 0.04, top water flux / 1e-5, [top heat flux / 100],
//   Δt / 3600, N / 40.
// Outputs (Nout = 32 RRE, 64 EH): node values of (x_ϑ - U0_ϑ) / 1e-2 and,
// for EH, of (x_ρe - u_ρe) / ρc_s(u) in kelvin.
//
// The Newton loop starts from x0 = U0 + guess and still decides convergence: it
// iterates to the same tolerance, and falls back to the plain start U0 when
// the warm-started iterate turns non-finite.
#pragma once
#include "clm_eh.h"
#include "clm_nodes.h"

namespace clm {
namespace ws {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;
// 16 bytes as a native vector: copies through registers are plain loads and
// stores (a uint4 struct copies through memcpy, which can leave the staging
// registers of a tile in scratch)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

struct Dims {
    int F, Fp, H, Nout;
};
__host__ __device__ inline Dims dims_for(bool eh) {
    return eh ? Dims{105, kFpEh, 384, 64} : Dims{40, kFpRre, 256, 32};
}

// The per-column kernels below give each column a group of 32 lanes (lane j
// owns node j, or cells j, j+32, ...) so that rows of X, T and Y and the
// column's cells are read and written contiguously.
constexpr int kGroup = 32;

__device__ __forceinline__ double tail_feature(int i, bool eh, double al, double n, double ks,
                                               double nu, double thr, double fw, double fh,
                                               double dt, int N) {
    if (!eh && i >= 6) ++i;  // RichardsModel has no heat-flux feature
    switch (i) {
        case 0: return log10(al);
        case 1: return n - 2.0;
        case 2: return (log10(ks) + 5.5) / 1.5;
        case 3: return (nu - 0.45) / 0.1;
        case 4: return (thr - 0.06) / 0.04;
        case 5: return fw / 1e-5;
        case 6: return fh / 100.0;
        case 7: return dt / 3600.0;
        default: return (double)N / 40.0;
    }
}

// ---- features of the current state: 32 lanes per column -----------------------
// Tail features that depend on the column alone (log10 α, n - 2, K_sat, ν,
// θ_r, N / 40) are rounded once per context into tails[c][0..5]
__global__ __launch_bounds__(256) void k_static_tails(bool eh, const int32_t* __restrict__ col_ptr,
                                                      int64_t ncol, const double* __restrict__ par,
                                                      bf16* __restrict__ tails) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    const int N = (int)(col_ptr[c + 1] - col_ptr[c]);
    bf16* o = tails + c * kTails;
    for (int i = 0; i < 5; ++i)
        o[i] = feat(tail_feature(i, eh, par[P_VG_ALPHA * ncol + c], par[P_VG_N * ncol + c],
                                 par[P_KSAT * ncol + c], par[P_NU * ncol + c],
                                 par[P_THETA_R * ncol + c], 0.0, 0.0, 0.0, N));
    o[5] = feat((double)N / 40.0);
    o[6] = o[7] = (bf16)0.0f;
}

// k_features gives each column 16 lanes: lane j0 forms nodes j0 and j0 + 16
// (their loads independent, in flight together) and tail entries j0,
// j0 + 16 of the column's row (half the waves of one lane per node, each
// with twice the loads in flight)
constexpr int kFeatGroup = 16;

// lane j0 of column c's group: node j's state features from the two cells
// around it (each loaded once), for j = j0, j0 + 16; then the tail entries
__device__ __forceinline__ void feature_lane(bool eh, const int32_t* __restrict__ col_ptr,
                                             int64_t ncol, const double* __restrict__ par,
                                             const bf16* __restrict__ tails,
                                             const double* __restrict__ xv,
                                             const double* __restrict__ xi,
                                             const double* __restrict__ xe,
                                             const double* __restrict__ ftw,
                                             const double* __restrict__ fth, double dt, int Fp,
                                             int64_t c, int j0, bf16* __restrict__ row) {
    const int64_t b = col_ptr[c];
    const int N = (int)(col_ptr[c + 1] - b);
    const double nu = par[P_NU * ncol + c], thr = par[P_THETA_R * ncol + c];
    const double rcd = eh ? par[P_RHOC_DS * ncol + c] : 0.0;
#pragma unroll
    for (int u = 0; u < kNodes / kFeatGroup; ++u) {
        const int j = j0 + u * kFeatGroup;
        const NodeAtF a = node_at_f(j, N);
        const double v0 = xv[b + a.i0], v1 = xv[b + a.i1];
        row[j] = feat_v(a, (float)thr, (float)nu, (float)v0, (float)v1);
        if (eh) {
            const double t0 = xi[b + a.i0], t1 = xi[b + a.i1];
            row[kNodes + j] = feat_T(a, feat_dT(nu, rcd, v0, t0, xe[b + a.i0]),
                                     feat_dT(nu, rcd, v1, t1, xe[b + a.i1]));
            row[2 * kNodes + j] = feat_i(a, (float)nu, (float)t0, (float)t1);
        }
    }
    const int S = state_features(eh);
    for (int f = S + j0; f < Fp; f += kFeatGroup)
        row[f] = tail_entry(eh, f - S, tails + c * kTails, ftw[c], eh ? fth[c] : 0.0, dt);
}

__global__ __launch_bounds__(256) void k_features(
    bool eh, const int32_t* __restrict__ col_ptr, int64_t ncol, const double* __restrict__ par,
    const bf16* __restrict__ tails, const double* __restrict__ xv, const double* __restrict__ xi,
    const double* __restrict__ xe, const double* __restrict__ ftw, const double* __restrict__ fth,
    double dt, int Fp, bf16* __restrict__ X) {
    int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t c = gt / kFeatGroup;
    const int j0 = (int)(gt % kFeatGroup);
    if (c >= ncol) return;
    feature_lane(eh, col_ptr, ncol, par, tails, xv, xi, xe, ftw, fth, dt, Fp, c, j0,
                 X + c * (int64_t)Fp);
}

// ---- training targets from (u_n, U0_ϑ, x_{n+1}): 32 lanes per column ----------
// mask = 1 for a column-step that converged with finite iterates and a
// target within 1e3 output units
__global__ __launch_bounds__(256) void k_targets(
    bool eh, const int32_t* __restrict__ col_ptr, int64_t ncol, const double* __restrict__ par,
    const double* __restrict__ u0v, const double* __restrict__ unv, const double* __restrict__ uni,
    const double* __restrict__ une, const double* __restrict__ xv, const double* __restrict__ xe,
    const int32_t* __restrict__ flags, int Nout, float* __restrict__ T, float* __restrict__ mask) {
    int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t c = gt / kGroup;
    const int j = (int)(gt % kGroup);
    const bool live = c < ncol;
    bool ok = false;
    double dv = 0.0, de = 0.0;
    if (live) {
        const int64_t b = col_ptr[c];
        const int N = (int)(col_ptr[c + 1] - b);
        ok = (flags[c] & 0x70000) == 0;
        dv = at_node(j, N, [&](int i) { return xv[b + i] - u0v[b + i]; }) / kScaleV;
        ok = ok && isfinite(dv) && fabs(dv) <= 1e3;
        if (eh) {
            const double nu = par[P_NU * ncol + c], rcd = par[P_RHOC_DS * ncol + c];
            de = at_node(j, N, [&](int i) {
                double v = unv[b + i], t = uni[b + i];
                return (xe[b + i] - une[b + i]) / heat_capacity(fmin(nu - t, v), t, rcd);
            }) / kScaleT;
            ok = ok && isfinite(de) && fabs(de) <= 1e3;
        }
    }
    // the column is usable only if every node of its group is
    unsigned long long bal = __ballot(ok);
    unsigned long long gmask = 0xffffffffull << (threadIdx.x & 32);
    bool all_ok = (bal & gmask) == gmask;
    if (!live) return;
    float* row = T + c * (int64_t)Nout;
    row[j] = all_ok ? (float)dv : 0.0f;
    if (eh) row[kNodes + j] = all_ok ? (float)de : 0.0f;
    if (j == 0) mask[c] = all_ok ? 1.0f : 0.0f;
}

// ---- network output -> per-cell guess (FP64, CSR): 32 lanes per column ---------
__device__ __forceinline__ void guess_lane(bool eh, const int32_t* __restrict__ col_ptr,
                                           int64_t ncol, const double* __restrict__ par,
                                           const double* __restrict__ xv,
                                           const double* __restrict__ xi,
                                           const float* __restrict__ o, int64_t c, int j,
                                           double* __restrict__ gv, double* __restrict__ ge) {
    const int64_t b = col_ptr[c];
    const int N = (int)(col_ptr[c + 1] - b);
    const double nu = eh ? par[P_NU * ncol + c] : 0.0;
    const double rcd = eh ? par[P_RHOC_DS * ncol + c] : 0.0;
    for (int i = j; i < N; i += kGroup) {
        gv[b + i] = at_cell(i, N, o) * kScaleV;
        if (eh) {
            double v = xv[b + i], t = xi[b + i];
            ge[b + i] = at_cell(i, N, o + kNodes) * kScaleT * heat_capacity(fmin(nu - t, v), t, rcd);
        }
    }
}

__global__ __launch_bounds__(256) void k_guess(
    bool eh, const int32_t* __restrict__ col_ptr, int64_t ncol, const double* __restrict__ par,
    const double* __restrict__ xv, const double* __restrict__ xi, const float* __restrict__ Y,
    int Nout, double* __restrict__ gv, double* __restrict__ ge) {
    int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t c = gt / kGroup;
    const int j = (int)(gt % kGroup);
    if (c >= ncol) return;
    guess_lane(eh, col_ptr, ncol, par, xv, xi, Y + c * (int64_t)Nout, c, j, gv, ge);
}

// ---- GEMM: C[M x N] = A[M x K] . B[N x K]^T (both K-contiguous), BF16 in, FP32 acc
// 256 x 128 tile per 256-thread block: wave w owns rows 128 (w / 2) .. +128 and
// columns 64 (w % 2) .. +64 as 4 x 2 MFMA 32x32x16 tiles (128 accumulators per
// lane).  The MFMA computes each tile transposed (B fragment first), so a
// lane's accumulators r = 4q..4q+3 are four consecutive columns of one row
// and the epilogue stores them as one vector.  K advances in chunks of 32 staged through LDS; the next chunk's
// global loads are issued before the current chunk's MFMAs (register
// double buffering).  Requires K % 32 == 0 and lda, ldb % 8 == 0 (16-byte
// loads); rows beyond M / N read as zero and are not stored.  blockIdx.z
// splits K into slices of kslice (EPI_F32 only: slice z writes its partial
// product to Cf + z * M * ldcf).
enum Epi { EPI_F32 = 0, EPI_BIAS_F32 = 1, EPI_BIAS_GELU = 2, EPI_DGELU = 3 };

// GELU, tanh form: z Φ(z) ≈ z σ(u), u = 2 sqrt(2/π) (z + 0.044715 z^3), and
// its derivative.  σ(u) = 1 / (1 + 2^t) with t = -u log2(e) = z (A + B z^2)
// (the constants folded: two multiplies and an FMA before the hardware
// exp2 and reciprocal, a few FP32 ulp, far below the BF16 rounding of the
// result)
constexpr float kGeluA = -2.302208198144325f;   // -2 sqrt(2/π) / ln 2
constexpr float kGeluB = -0.1029432395800235f;  // kGeluA * 0.044715
__device__ __forceinline__ float gelu_sigma(float z, float z2) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * fmaf(z2, kGeluB, kGeluA)));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_sigma(z, z * z); }
// d/dz z σ(u) = σ + z σ (1 - σ) du/dz, du/dz = 2 sqrt(2/π) (1 + 3 0.044715 z^2)
__device__ __forceinline__ float dgelu(float z) {
    const float z2 = z * z;
    const float sg = gelu_sigma(z, z2);
    return sg + z * sg * (1.0f - sg) * fmaf(z2, 0.21406444881780076f, 1.5957691216057308f);
}

constexpr int kBM = 256, kBN = 128, kBK = 32, kLdsRow = kBK + 8;

template <int EPI>
__global__ __launch_bounds__(256, 2) void k_gemm_nt(
    int M, int N, int K, int kslice, const bf16* __restrict__ A, int lda,
    const bf16* __restrict__ B, int ldb, const float* __restrict__ bias, float* __restrict__ Cf,
    int ldcf, bf16* __restrict__ Cb, int ldcb, const float* __restrict__ Z, int ldz) {
    __shared__ __attribute__((aligned(16))) bf16 As[kBM * kLdsRow];
    __shared__ __attribute__((aligned(16))) bf16 Bs[kBN * kLdsRow];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.x * kBM, n0 = blockIdx.y * kBN;
    const int kbeg = blockIdx.z * kslice;
    const int kend = min(K, kbeg + kslice);
    floatx16 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    const u32x4 zero4 = {0u, 0u, 0u, 0u};
    u32x4 ra[4], rb[2];
    auto gload = [&](int k0) {
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            int cidx = t + 256 * h;
            int row = cidx >> 2, kq = (cidx & 3) * 8;
            ra[h] = (m0 + row < M) ? *(const u32x4*)(A + (int64_t)(m0 + row) * lda + k0 + kq)
                                   : zero4;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            int cidx = t + 256 * h;
            int row = cidx >> 2, kq = (cidx & 3) * 8;
            rb[h] = (n0 + row < N) ? *(const u32x4*)(B + (int64_t)(n0 + row) * ldb + k0 + kq)
                                   : zero4;
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            int cidx = t + 256 * h;
            *(u32x4*)(As + (cidx >> 2) * kLdsRow + (cidx & 3) * 8) = ra[h];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            int cidx = t + 256 * h;
            *(u32x4*)(Bs + (cidx >> 2) * kLdsRow + (cidx & 3) * 8) = rb[h];
        }
    };
    if (kbeg < kend) {
        gload(kbeg);
        lstore();
    }
    __syncthreads();
    for (int k0 = kbeg; k0 < kend; k0 += kBK) {
        const bool more = k0 + kBK < kend;
        if (more) gload(k0 + kBK);  // in flight during this chunk's MFMAs
#pragma unroll
        for (int ks = 0; ks < kBK; ks += 16) {
            bf16x8 a[4], b[2];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                a[i] = *(const bf16x8*)(As + (wm * 128 + i * 32 + (lane & 31)) * kLdsRow + ks +
                                        (lane >> 5) * 8);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *(const bf16x8*)(Bs + (wn * 64 + j * 32 + (lane & 31)) * kLdsRow + ks +
                                        (lane >> 5) * 8);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[j], a[i], acc[i][j], 0,
                                                                        0, 0);
        }
        __syncthreads();
        if (more) {
            lstore();
            __syncthreads();
        }
    }
    // tile (i, j), accumulator r = 4q + e: row m = 32 i + lane % 32, column
    // n = 32 j + 8 q + 4 (lane / 32) + e  (N % 4 == 0: a group of four is
    // wholly inside or outside the matrix)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + wm * 128 + i * 32 + (lane & 31);
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int n = n0 + wn * 64 + j * 32 + 8 * q + 4 * (lane >> 5);
                if (n >= N) continue;
                float4 v = {acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2],
                            acc[i][j][4 * q + 3]};
                if constexpr (EPI == EPI_F32) {
                    *(float4*)(Cf + (int64_t)blockIdx.z * M * ldcf + (int64_t)m * ldcf + n) = v;
                } else if constexpr (EPI == EPI_BIAS_F32) {
                    float4 bb = *(const float4*)(bias + n);
                    v.x += bb.x;
                    v.y += bb.y;
                    v.z += bb.z;
                    v.w += bb.w;
                    *(float4*)(Cf + (int64_t)m * ldcf + n) = v;
                } else if constexpr (EPI == EPI_BIAS_GELU) {
                    float4 bb = *(const float4*)(bias + n);
                    v.x += bb.x;
                    v.y += bb.y;
                    v.z += bb.z;
                    v.w += bb.w;
                    if (Cf) *(float4*)(Cf + (int64_t)m * ldcf + n) = v;
                    bf16x4 h = {(bf16)gelu(v.x), (bf16)gelu(v.y), (bf16)gelu(v.z),
                                (bf16)gelu(v.w)};
                    *(bf16x4*)(Cb + (int64_t)m * ldcb + n) = h;
                } else {
                    float4 z = *(const float4*)(Z + (int64_t)m * ldz + n);
                    bf16x4 h = {(bf16)(v.x * dgelu(z.x)), (bf16)(v.y * dgelu(z.y)),
                                (bf16)(v.z * dgelu(z.z)), (bf16)(v.w * dgelu(z.w))};
                    *(bf16x4*)(Cb + (int64_t)m * ldcb + n) = h;
                }
            }
    }
}

// ---- wide GEMM for the hidden layers: one block owns full output rows --------
// C[M x N] = A[M x K] . B[N x K]^T with N = 128 NTW (384 or 256 here): a
// 512-thread block computes 128 rows x N columns, so each row of A is read
// from HBM once; waves are 2 (rows) x 4 (columns), each 64 x 32 NTW as
// 2 x NTW MFMA 32x32x16 tiles.  K chunks of 32 alternate between two LDS
// buffers: the next chunk's global loads are issued before the current
// chunk's MFMAs and stored into the other buffer, one barrier per chunk.
// Same epilogues and operand orientation as k_gemm_nt; K % 32 == 0.
constexpr int kWBM = 128;

template <int EPI, int NTW>
__global__ __launch_bounds__(512) void k_gemm_wide(
    int M, int K, const bf16* __restrict__ A, int lda, const bf16* __restrict__ B, int ldb,
    const float* __restrict__ bias, float* __restrict__ Cf, int ldcf, bf16* __restrict__ Cb,
    int ldcb, const float* __restrict__ Z, int ldz) {
    constexpr int BN = 128 * NTW;
    constexpr int NA = kWBM * kBK / 8 / 512;  // 16-byte pieces of A per thread per chunk (1)
    constexpr int NB = BN * kBK / 8 / 512;    // of B per thread per chunk (NTW)
    __shared__ __attribute__((aligned(16))) bf16 As[2][kWBM * kLdsRow];
    __shared__ __attribute__((aligned(16))) bf16 Bs[2][BN * kLdsRow];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 2, wn = w & 3;
    const int m0 = blockIdx.x * kWBM;
    floatx16 acc[2][NTW];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < NTW; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    const u32x4 zero4 = {0u, 0u, 0u, 0u};
    u32x4 ra[NA], rb[NB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int h = 0; h < NA; ++h) {
            int cidx = t + 512 * h;
            int row = cidx >> 2, kq = (cidx & 3) * 8;
            ra[h] = (m0 + row < M) ? *(const u32x4*)(A + (int64_t)(m0 + row) * lda + k0 + kq)
                                   : zero4;
        }
#pragma unroll
        for (int h = 0; h < NB; ++h) {
            int cidx = t + 512 * h;
            int row = cidx >> 2, kq = (cidx & 3) * 8;
            rb[h] = *(const u32x4*)(B + (int64_t)row * ldb + k0 + kq);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int h = 0; h < NA; ++h) {
            int cidx = t + 512 * h;
            *(u32x4*)(&As[buf][(cidx >> 2) * kLdsRow + (cidx & 3) * 8]) = ra[h];
        }
#pragma unroll
        for (int h = 0; h < NB; ++h) {
            int cidx = t + 512 * h;
            *(u32x4*)(&Bs[buf][(cidx >> 2) * kLdsRow + (cidx & 3) * 8]) = rb[h];
        }
    };
    gload(0);
    lstore(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += kBK) {
        const bool more = k0 + kBK < K;
        if (more) gload(k0 + kBK);
#pragma unroll
        for (int ks = 0; ks < kBK; ks += 16) {
            bf16x8 a[2], b[NTW];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *(const bf16x8*)(&As[buf][(wm * 64 + i * 32 + (lane & 31)) * kLdsRow + ks +
                                                 (lane >> 5) * 8]);
#pragma unroll
            for (int j = 0; j < NTW; ++j)
                b[j] = *(const bf16x8*)(&Bs[buf][(wn * 32 * NTW + j * 32 + (lane & 31)) * kLdsRow +
                                                 ks + (lane >> 5) * 8]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < NTW; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[j], a[i], acc[i][j], 0,
                                                                        0, 0);
        }
        if (more) lstore(buf ^ 1);  // the other buffer was last read before the previous barrier
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int m = m0 + wm * 64 + i * 32 + (lane & 31);
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < NTW; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int n = wn * 32 * NTW + j * 32 + 8 * q + 4 * (lane >> 5);
                float4 v = {acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2],
                            acc[i][j][4 * q + 3]};
                if constexpr (EPI == EPI_BIAS_GELU) {
                    float4 bb = *(const float4*)(bias + n);
                    v.x += bb.x;
                    v.y += bb.y;
                    v.z += bb.z;
                    v.w += bb.w;
                    if (Cf) *(float4*)(Cf + (int64_t)m * ldcf + n) = v;
                    bf16x4 h = {(bf16)gelu(v.x), (bf16)gelu(v.y), (bf16)gelu(v.z),
                                (bf16)gelu(v.w)};
                    *(bf16x4*)(Cb + (int64_t)m * ldcb + n) = h;
                } else if constexpr (EPI == EPI_DGELU) {
                    float4 z = *(const float4*)(Z + (int64_t)m * ldz + n);
                    bf16x4 h = {(bf16)(v.x * dgelu(z.x)), (bf16)(v.y * dgelu(z.y)),
                                (bf16)(v.z * dgelu(z.z)), (bf16)(v.w * dgelu(z.w))};
                    *(bf16x4*)(Cb + (int64_t)m * ldcb + n) = h;
                } else {
                    *(float4*)(Cf + (int64_t)m * ldcf + n) = v;
                }
            }
    }
}

// ---- fused inference: the four layers in one kernel -----------------------------
// A 512-thread block takes 128 columns: their feature rows (from k_features,
// a launch of its own at full occupancy -- inside this kernel, whose LDS
// admits one block per CU, the features' dependent loads waited with eight
// waves per CU) are copied into LDS, pass through the four layers with the
// activations kept in LDS (one [128][H + 8] BF16 buffer, overwritten by each
// hidden layer's output once every wave has finished reading it), and the
// output layer writes the FP32 output rows [ncol][Nout]; the step kernels
// turn a column's row into its per-cell guesses as k_guess does.  Weights
// stream from L2 in chunk-major order ([K/32][N][32], written by
// k_cast_weights / k_adam) through a double-buffered [2][N][32 + 8] LDS
// buffer, one barrier per 32-wide K chunk.  Each layer keeps the
// accumulation order of k_gemm_wide / k_gemm_nt (K chunks of 32 ascending,
// two 16-deep MFMA steps per chunk, the same transposed operand order), so
// the outputs equal the unfused path's bit for bit.
constexpr int kIBM = 128;

template <bool EH>
struct NetDims;
template <>
struct NetDims<true> {
    static constexpr int Fp = kFpEh, H = 384, No = 64;
};
template <>
struct NetDims<false> {
    static constexpr int Fp = kFpRre, H = 256, No = 32;
};

// weight chunk ch ([N][32] of the chunk-major layout) -> registers -> LDS
template <int N>
struct Chunk {
    static constexpr int NU = N * kBK / 8;  // 16-byte pieces per chunk
    static constexpr int NB = (NU + 511) / 512;
    u32x4 r[NB];
    // threads past the chunk's end repeat its last piece (same value, same
    // LDS slot): every access is unconditional, so r stays in registers and
    // the loads stay in flight until the store
    static __device__ __forceinline__ int slot(int h) {
        int idx = (int)threadIdx.x + 512 * h;
        return NU % 512 == 0 ? idx : (idx < NU ? idx : NU - 1);
    }
    __device__ __forceinline__ void load(const bf16* __restrict__ Wpk, int ch) {
        const u32x4* src = (const u32x4*)(Wpk + (int64_t)ch * N * kBK);
#pragma unroll
        for (int h = 0; h < NB; ++h) r[h] = src[slot(h)];
    }
    __device__ __forceinline__ void store(bf16* __restrict__ Wb, int buf) const {
#pragma unroll
        for (int h = 0; h < NB; ++h) {
            int idx = slot(h);
            *(u32x4*)(Wb + buf * N * kLdsRow + (idx >> 2) * kLdsRow + (idx & 3) * 8) = r[h];
        }
    }
};

// one layer over the block's 128 rows: Act[:, 0:K] . W^T + bias; hidden layers
// write GELU(.) as BF16 into Act[:, 0:N], the output layer writes FP32 rows
// [rows][N] to Y in global memory (rows past `rows` are not stored).  On entry
// chunk 0 of W is in Wb buffer 0 (behind a barrier) and chunk 1 in s1; on
// exit, if NN > 0, chunk 0 of the next layer's weights Wnext ([NN][K']
// chunk-major) is in buffer 0 and its chunk 1 in n1.
template <int N, int K, bool HIDDEN, int NN>
__device__ __forceinline__ void infer_layer(bf16* __restrict__ Act, int lds_row,
                                            bf16* __restrict__ Wb,
                                            const bf16* __restrict__ Wpk,
                                            const float* __restrict__ bias,
                                            float* __restrict__ Y, int rows,
                                            const bf16* __restrict__ Wnext,
                                            Chunk<N>& s1,
                                            Chunk<NN ? NN : 32>& n0,
                                            Chunk<NN ? NN : 32>& n1) {
    constexpr int WN = N / 32 < 4 ? N / 32 : 4;          // waves across columns
    constexpr int WM = 8 / WN < 4 ? 8 / WN : 4;          // waves across rows
    constexpr int TM = 4 / WM, TN = N / (32 * WN);       // 32x32 tiles per wave
    constexpr int WS = kLdsRow;                           // weight chunk row stride
    static_assert(K % kBK == 0 && N % 32 == 0 && TN >= 1, "layer shape");
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w / WN, wn = w % WN;
    const bool active = w < WM * WN;
    floatx16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    // Chunk c of W sits in register set c % 2 (s0 even, s1 odd; s1 arrives
    // holding chunk 1) from two chunks before its use, so its load has two
    // chunks of MFMAs to arrive in, and goes into LDS buffer c % 2 behind
    // the barrier that ends the reads of that buffer.  The last two chunks'
    // slots load the next layer's chunks 0 and 1 (n0, n1): chunk 0 goes into
    // buffer 0 once this layer's reads are done, chunk 1 stays in n1.
    constexpr int NCH = K / kBK;
    static_assert(NCH >= 2, "two chunks at least");
    Chunk<N> s0;
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
        if (ch + 2 < NCH) {
            if (ch % 2 == 0)
                s0.load(Wpk, ch + 2);
            else
                s1.load(Wpk, ch + 2);
        } else if constexpr (NN > 0) {
            if (ch + 2 == NCH)
                n0.load(Wnext, 0);
            else
                n1.load(Wnext, 1);
        }
        if (active) {
#pragma unroll
            for (int ks = 0; ks < kBK; ks += 16) {
                bf16x8 a[TM], b[TN];
#pragma unroll
                for (int i = 0; i < TM; ++i)
                    a[i] = *(const bf16x8*)(Act + (wm * TM * 32 + i * 32 + (lane & 31)) * lds_row +
                                            ch * kBK + ks + (lane >> 5) * 8);
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    b[j] = *(const bf16x8*)(Wb + (ch % 2) * N * WS +
                                            (wn * TN * 32 + j * 32 + (lane & 31)) * WS + ks +
                                            (lane >> 5) * 8);
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[j], a[i], acc[i][j],
                                                                            0, 0, 0);
            }
        }
        // chunk ch + 1 into the other buffer, last read before the previous barrier
        if (ch + 1 < NCH) {
            if (ch % 2 == 0)
                s1.store(Wb, 1);
            else
                s0.store(Wb, 0);
        }
        __syncthreads();
    }
    // every wave is past its last read of Act and Wb
    if constexpr (NN > 0) {
        static_assert(HIDDEN, "only hidden layers are followed by a layer");
        n0.store(Wb, 0);
    }
    if (active) {
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            const int m = wm * TM * 32 + i * 32 + (lane & 31);
#pragma unroll
            for (int j = 0; j < TN; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int n = wn * TN * 32 + j * 32 + 8 * q + 4 * (lane >> 5);
                    float4 bb = *(const float4*)(bias + n);
                    float4 v = {acc[i][j][4 * q] + bb.x, acc[i][j][4 * q + 1] + bb.y,
                                acc[i][j][4 * q + 2] + bb.z, acc[i][j][4 * q + 3] + bb.w};
                    if constexpr (HIDDEN) {
                        bf16x4 hv = {(bf16)gelu(v.x), (bf16)gelu(v.y), (bf16)gelu(v.z),
                                     (bf16)gelu(v.w)};
                        *(bf16x4*)(Act + m * lds_row + n) = hv;
                    } else if (m < rows) {
                        *(float4*)(Y + (int64_t)m * N + n) = v;
                    }
                }
        }
    }
    __syncthreads();
}

// X: feature rows [ncol][Fp] (k_features); Wpk[l]: chunk-major BF16 weights
// of layer l; bias[l]: its FP32 bias
template <bool EH>
__global__ __launch_bounds__(512, 1) void k_ws_infer(
    int64_t ncol, const bf16* __restrict__ X, const bf16* __restrict__ W0,
    const bf16* __restrict__ W1, const bf16* __restrict__ W2, const bf16* __restrict__ W3,
    const float* __restrict__ b0, const float* __restrict__ b1, const float* __restrict__ b2,
    const float* __restrict__ b3, float* __restrict__ Yg) {
    using D = NetDims<EH>;
    constexpr int LR = D::H + 8;  // LDS row stride of the activations
    __shared__ __attribute__((aligned(16))) bf16 Act[kIBM * LR];
    __shared__ __attribute__((aligned(16))) bf16 Wb[2 * D::H * kLdsRow];
    const int64_t c0 = (int64_t)blockIdx.x * kIBM;
    Chunk<D::H> c1, n0a, n1a, n0b, n1b;
    Chunk<D::No> n0c, n1c;
    Chunk<32> d0, d1;  // (no layer after the output layer)
    c1.load(W0, 1);    // layer 1's chunk 1, in flight through the copy and chunk 0
    {
        Chunk<D::H> first;
        first.load(W0, 0);  // layer 1's first weight chunk, in flight during the copy
        // the block's feature rows; rows past the last column repeat it
        // (their outputs are never stored)
        constexpr int RU = D::Fp / 8;  // 16-byte pieces per row
        for (int q = threadIdx.x; q < kIBM * RU; q += 512) {
            const int lc = q / RU, k = q % RU;
            const int64_t c = c0 + lc < ncol ? c0 + lc : ncol - 1;
            *(u32x4*)(Act + lc * LR + 8 * k) = *(const u32x4*)(X + c * D::Fp + 8 * k);
        }
        first.store(Wb, 0);
    }
    __syncthreads();
    const int rows = (int)min<int64_t>(kIBM, ncol - c0);
    infer_layer<D::H, D::Fp, true, D::H>(Act, LR, Wb, W0, b0, nullptr, 0, W1, c1, n0a, n1a);
    infer_layer<D::H, D::H, true, D::H>(Act, LR, Wb, W1, b1, nullptr, 0, W2, n1a, n0b, n1b);
    infer_layer<D::H, D::H, true, D::No>(Act, LR, Wb, W2, b2, nullptr, 0, W3, n1b, n0c, n1c);
    infer_layer<D::No, D::H, false, 0>(Act, LR, Wb, W3, b3, Yg + c0 * D::No, rows, nullptr, n1c,
                                       d0, d1);
}

// ---- small training kernels ---------------------------------------------------
// out[c][r] = in[r][c] (bf16), 32 x 32 tiles through LDS
__global__ __launch_bounds__(256) void k_transpose(int R, int Cn, const bf16* __restrict__ in,
                                                   int ldi, bf16* __restrict__ out, int ldo) {
    __shared__ bf16 tile[32][33];
    int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows per pass
    for (int y = ty; y < 32; y += 8) {
        int r = r0 + y, c = c0 + tx;
        tile[y][tx] = (r < R && c < Cn) ? in[(int64_t)r * ldi + c] : (bf16)0.0f;
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        int c = c0 + y, r = r0 + tx;
        if (c < Cn && r < R) out[(int64_t)c * ldo + r] = tile[tx][y];
    }
}

// sum of S partial products: out[i] = sum_s part[s * n + i]
__global__ __launch_bounds__(256) void k_sum_slices(int64_t n, int S, const float* __restrict__ part,
                                                    float* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < S; ++k) s += part[(int64_t)k * n + i];
    out[i] = s;
}

// partial column sums of a bf16 [M x N] matrix (bias gradients): block
// (x, y) sums rows 256 y .. 256 y + 255 of columns 64 x .. 64 x + 63 into
// part[y * N + n]; k_sum_slices then adds the gridDim.y partials
__global__ __launch_bounds__(256) void k_colsum(int M, int N, const bf16* __restrict__ D,
                                                float* __restrict__ part) {
    __shared__ float red[4][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + tx;
    const int r0 = blockIdx.y * 256;
    float s = 0.0f;
    if (n < N)
        for (int m = r0 + ty; m < min(M, r0 + 256); m += 4) s += (float)D[(int64_t)m * N + n];
    red[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && n < N)
        part[(int64_t)blockIdx.y * N + n] = red[0][tx] + red[1][tx] + red[2][tx] + red[3][tx];
}

// MSE loss sum(mask (Y - T)^2) / (M Nout) and its gradient dY (bf16)
__global__ __launch_bounds__(256) void k_mse(int M, int Nout, const float* __restrict__ Y,
                                             const float* __restrict__ T,
                                             const float* __restrict__ mask,
                                             bf16* __restrict__ dY, float* __restrict__ loss) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < (int64_t)M * Nout) {
        float d = mask[i / Nout] > 0.0f ? Y[i] - T[i] : 0.0f;
        float inv = 1.0f / ((float)M * (float)Nout);
        l = d * d * inv;
        dY[i] = (bf16)(2.0f * d * inv);
    }
    // one atomic per block (per wave, the adds on one address serialised)
    __shared__ float part[4];
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (part[0] + part[1]) + (part[2] + part[3]));
}

// rows of the training set gathered into a minibatch
__global__ __launch_bounds__(256) void k_gather(int M, const int32_t* __restrict__ idx, int Fp,
                                                const bf16* __restrict__ Xd, int Nout,
                                                const float* __restrict__ Td,
                                                const float* __restrict__ Md, bf16* __restrict__ X,
                                                float* __restrict__ T, float* __restrict__ mk) {
    int m = blockIdx.x;
    if (m >= M) return;
    int64_t s = idx[m];
    for (int f = threadIdx.x; f < Fp; f += blockDim.x) X[(int64_t)m * Fp + f] = Xd[s * Fp + f];
    for (int f = threadIdx.x; f < Nout; f += blockDim.x)
        T[(int64_t)m * Nout + f] = Td[s * Nout + f];
    if (threadIdx.x == 0) mk[m] = Md[s];
}

// Adam step on one layer's FP32 master weights Wt [N x K] and bias [N]; writes
// the BF16 copies Wt (forward) and W = Wt^T (backward data gradient)
__global__ __launch_bounds__(256) void k_adam(int N, int K, float* __restrict__ W,
                                              float* __restrict__ bias,
                                              const float* __restrict__ gW,
                                              const float* __restrict__ gb, float* __restrict__ m1,
                                              float* __restrict__ m2, float lr, float b1t,
                                              float b2t, bf16* __restrict__ Wt_b,
                                              bf16* __restrict__ W_b, bf16* __restrict__ W_p) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int64_t nw = (int64_t)N * K;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nw + N) return;
    float g = i < nw ? gW[i] : gb[i - nw];
    float mm = b1 * m1[i] + (1.0f - b1) * g;
    float vv = b2 * m2[i] + (1.0f - b2) * g * g;
    m1[i] = mm;
    m2[i] = vv;
    float step = lr * (mm / (1.0f - b1t)) / (sqrtf(vv / (1.0f - b2t)) + eps);
    if (i < nw) {
        float p = W[i] - step;
        W[i] = p;
        int64_t n = i / K, k = i % K;
        Wt_b[i] = (bf16)p;
        W_b[k * N + n] = (bf16)p;
        W_p[(k / kBK) * N * kBK + n * kBK + k % kBK] = (bf16)p;
    } else {
        bias[i - nw] -= step;
    }
}

// BF16 copies of FP32 master weights (after init or a host upload)
__global__ __launch_bounds__(256) void k_cast_weights(int N, int K, const float* __restrict__ W,
                                                      bf16* __restrict__ Wt_b,
                                                      bf16* __restrict__ W_b,
                                                      bf16* __restrict__ W_p) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)N * K) return;
    int64_t n = i / K, k = i % K;
    Wt_b[i] = (bf16)W[i];
    W_b[k * N + n] = (bf16)W[i];
    W_p[(k / kBK) * N * kBK + n * kBK + k % kBK] = (bf16)W[i];
}

__global__ __launch_bounds__(256) void k_f32_to_bf16(int64_t n, const float* __restrict__ in,
                                                     bf16* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (bf16)in[i];
}

__global__ __launch_bounds__(256) void k_bf16_to_f32(int64_t n, const bf16* __restrict__ in,
                                                     float* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

__global__ void k_zero_f32(int64_t n, float* __restrict__ p) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace ws
}  // namespace clm
