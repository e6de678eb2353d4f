// clm_mlp.h -- learned Newton warm start: a per-column MLP that predicts the
// implicit part of the step's increment, x_{n+1} - U0, from the column state.
//
// Network: input features -> H -> H -> H -> 2x32 node outputs, GELU between
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
// The Newton loop starts from x0 = U0 + guess and remains the judge: it
// iterates to the same tolerance, and falls back to the plain start U0 when
// the warm-started iterate turns non-finite.
#pragma once
#include "clm_eh.h"

namespace clm {
namespace ws {

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

constexpr int kNodes = 32;
constexpr double kScaleV = 1e-2;  // ϑ_l output unit
constexpr double kScaleT = 1.0;   // ρe/ρc_s output unit (K)

struct Dims {
    int F, Fp, H, Nout;
};
__host__ __device__ inline Dims dims_for(bool eh) {
    return eh ? Dims{105, 128, 384, 64} : Dims{40, 64, 256, 32};
}

// ---- resampling between layers and nodes ------------------------------------
// value at node j of a column with N cells (v: cell accessor)
template <typename F>
__device__ __forceinline__ double at_node(int j, int N, F v) {
    double q = ((double)j + 0.5) * (double)N / (double)kNodes - 0.5;
    q = fmin(fmax(q, 0.0), (double)(N - 1));
    int i0 = (int)q;
    int i1 = i0 + 1 < N ? i0 + 1 : N - 1;
    double t = q - (double)i0;
    return (1.0 - t) * v(i0) + t * v(i1);
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

// non-finite inputs (a column whose iterate blew up) enter the network as 0;
// such columns are masked out of training and fall back to the plain start
__device__ __forceinline__ bf16 feat(double x) { return (bf16)(float)(isfinite(x) ? x : 0.0); }

// ---- features of the current state: one thread per column --------------------
__global__ __launch_bounds__(256) void k_features(
    bool eh, const int32_t* __restrict__ col_ptr, int64_t ncol, const double* __restrict__ par,
    const double* __restrict__ xv, const double* __restrict__ xi, const double* __restrict__ xe,
    const double* __restrict__ ftw, const double* __restrict__ fth, double dt, int Fp,
    bf16* __restrict__ X) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    const int64_t b = col_ptr[c];
    const int N = (int)(col_ptr[c + 1] - b);
    bf16* row = X + c * (int64_t)Fp;
    const double nu = par[P_NU * ncol + c], thr = par[P_THETA_R * ncol + c];
    const double al = par[P_VG_ALPHA * ncol + c], n = par[P_VG_N * ncol + c];
    const double ks = par[P_KSAT * ncol + c];
    int f = 0;
    for (int j = 0; j < kNodes; ++j)
        row[f++] = feat((at_node(j, N, [&](int i) { return xv[b + i]; }) - thr) /
                                 (nu - thr));
    if (eh) {
        const double rcd = par[P_RHOC_DS * ncol + c];
        for (int j = 0; j < kNodes; ++j) {
            double T = at_node(j, N, [&](int i) {
                double v = xv[b + i], t = xi[b + i];
                return temperature(xe[b + i], t, heat_capacity(fmin(nu - t, v), t, rcd));
            });
            row[f++] = feat((T - 273.15) / 10.0);
        }
        for (int j = 0; j < kNodes; ++j)
            row[f++] = feat(at_node(j, N, [&](int i) { return xi[b + i]; }) / nu);
    }
    row[f++] = feat(log10(al));
    row[f++] = feat(n - 2.0);
    row[f++] = feat((log10(ks) + 5.5) / 1.5);
    row[f++] = feat((nu - 0.45) / 0.1);
    row[f++] = feat((thr - 0.06) / 0.04);
    row[f++] = feat(ftw[c] / 1e-5);
    if (eh) row[f++] = feat(fth[c] / 100.0);
    row[f++] = feat(dt / 3600.0);
    row[f++] = feat((double)N / 40.0);
    for (; f < Fp; ++f) row[f] = (bf16)0.0f;
}

// ---- training targets from (u_n, U0_ϑ, x_{n+1}): one thread per column -------
// mask = 1 for a column-step that converged with finite iterates
__global__ __launch_bounds__(256) void k_targets(
    bool eh, const int32_t* __restrict__ col_ptr, int64_t ncol, const double* __restrict__ par,
    const double* __restrict__ u0v, const double* __restrict__ unv, const double* __restrict__ uni,
    const double* __restrict__ une, const double* __restrict__ xv, const double* __restrict__ xe,
    const int32_t* __restrict__ flags, int Nout, float* __restrict__ T, float* __restrict__ mask) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    const int64_t b = col_ptr[c];
    const int N = (int)(col_ptr[c + 1] - b);
    float* row = T + c * (int64_t)Nout;
    bool ok = (flags[c] & 0x70000) == 0;
    for (int j = 0; j < kNodes; ++j) {
        double d = at_node(j, N, [&](int i) { return xv[b + i] - u0v[b + i]; }) / kScaleV;
        ok = ok && isfinite(d);
        row[j] = (float)d;
    }
    if (eh) {
        const double nu = par[P_NU * ncol + c], rcd = par[P_RHOC_DS * ncol + c];
        for (int j = 0; j < kNodes; ++j) {
            double d = at_node(j, N, [&](int i) {
                double v = unv[b + i], t = uni[b + i];
                return (xe[b + i] - une[b + i]) / heat_capacity(fmin(nu - t, v), t, rcd);
            }) / kScaleT;
            ok = ok && isfinite(d);
            row[kNodes + j] = (float)d;
        }
    }
    if (!ok)
        for (int j = 0; j < Nout; ++j) row[j] = 0.0f;
    mask[c] = ok ? 1.0f : 0.0f;
}

// ---- network output -> per-cell guess (FP64, CSR): one thread per column ------
__global__ __launch_bounds__(256) void k_guess(
    bool eh, const int32_t* __restrict__ col_ptr, int64_t ncol, const double* __restrict__ par,
    const double* __restrict__ xv, const double* __restrict__ xi, const float* __restrict__ Y,
    int Nout, double* __restrict__ gv, double* __restrict__ ge) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    const int64_t b = col_ptr[c];
    const int N = (int)(col_ptr[c + 1] - b);
    const float* o = Y + c * (int64_t)Nout;
    for (int i = 0; i < N; ++i) gv[b + i] = at_cell(i, N, o) * kScaleV;
    if (eh) {
        const double nu = par[P_NU * ncol + c], rcd = par[P_RHOC_DS * ncol + c];
        for (int i = 0; i < N; ++i) {
            double v = xv[b + i], t = xi[b + i];
            ge[b + i] = at_cell(i, N, o + kNodes) * kScaleT *
                        heat_capacity(fmin(nu - t, v), t, rcd);
        }
    }
}

// ---- GEMM: C[M x N] = A[M x K] . B[N x K]^T (both K-contiguous), BF16 in, FP32 acc
// 128 x 128 tile per 256-thread block (4 waves of 64 x 64, 2 x 2 MFMA 32x32x16
// tiles each), K in chunks of 32 staged through LDS.  K % 32 == 0 and lda,
// ldb % 8 == 0 (16-byte loads).  blockIdx.z splits K into slices of kslice
// (EPI_F32 only: slice z writes its partial product to Cf + z * M * ldcf).
enum Epi { EPI_F32 = 0, EPI_BIAS_F32 = 1, EPI_BIAS_GELU = 2, EPI_DGELU = 3 };

__device__ __forceinline__ float gelu(float z) {
    return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float dgelu(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
           z * 0.39894228040143268f * expf(-0.5f * z * z);
}

constexpr int kBM = 128, kBN = 128, kBK = 32, kLdsRow = kBK + 8;

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_nt(
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
    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    const uint4 zero4 = {0u, 0u, 0u, 0u};
    for (int k0 = kbeg; k0 < kend; k0 += kBK) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            int cidx = t + 256 * h;
            int row = cidx >> 2, kq = (cidx & 3) * 8;
            uint4 va = zero4, vb = zero4;
            if (m0 + row < M) va = *(const uint4*)(A + (int64_t)(m0 + row) * lda + k0 + kq);
            if (n0 + row < N) vb = *(const uint4*)(B + (int64_t)(n0 + row) * ldb + k0 + kq);
            *(uint4*)(As + row * kLdsRow + kq) = va;
            *(uint4*)(Bs + row * kLdsRow + kq) = vb;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kBK; ks += 16) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *(const bf16x8*)(As + (wm * 64 + i * 32 + (lane & 31)) * kLdsRow + ks +
                                        (lane >> 5) * 8);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *(const bf16x8*)(Bs + (wn * 64 + j * 32 + (lane & 31)) * kLdsRow + ks +
                                        (lane >> 5) * 8);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0,
                                                                        0, 0);
        }
        __syncthreads();
    }
    // D[row][col] of a 32x32 tile: col = lane % 32, row = 8 (r / 4) + 4 (lane / 32) + r % 4
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                int row = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                int col = n0 + wn * 64 + j * 32 + (lane & 31);
                if (row >= M || col >= N) continue;
                float v = acc[i][j][r];
                if constexpr (EPI == EPI_F32) {
                    Cf[(int64_t)blockIdx.z * M * ldcf + (int64_t)row * ldcf + col] = v;
                } else if constexpr (EPI == EPI_BIAS_F32) {
                    Cf[(int64_t)row * ldcf + col] = v + bias[col];
                } else if constexpr (EPI == EPI_BIAS_GELU) {
                    float z = v + bias[col];
                    if (Cf) Cf[(int64_t)row * ldcf + col] = z;
                    Cb[(int64_t)row * ldcb + col] = (bf16)gelu(z);
                } else {
                    Cb[(int64_t)row * ldcb + col] = (bf16)(v * dgelu(Z[(int64_t)row * ldz + col]));
                }
            }
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

// column sums of a bf16 [M x N] matrix (bias gradients)
__global__ __launch_bounds__(256) void k_colsum(int M, int N, const bf16* __restrict__ D,
                                                float* __restrict__ out) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int m = 0; m < M; ++m) s += (float)D[(int64_t)m * N + n];
    out[n] = s;
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
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
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
                                              bf16* __restrict__ W_b) {
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
    } else {
        bias[i - nw] -= step;
    }
}

// BF16 copies of FP32 master weights (after init or a host upload)
__global__ __launch_bounds__(256) void k_cast_weights(int N, int K, const float* __restrict__ W,
                                                      bf16* __restrict__ Wt_b,
                                                      bf16* __restrict__ W_b) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)N * K) return;
    int64_t n = i / K, k = i % K;
    Wt_b[i] = (bf16)W[i];
    W_b[k * N + n] = (bf16)W[i];
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
