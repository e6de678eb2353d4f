// clm_soil.hip -- libclmsoil.so: C ABI (include/clm_soil.h) over the gfx950
// soil-column kernels in clm_rre.h / clm_eh.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/clm_soil.h"
#include "clm_coop.h"
#include "clm_eh.h"
#include "clm_layout.h"
#include "clm_mlp.h"
#include "clm_rre.h"

using namespace clm;

namespace {

thread_local std::string g_create_error;

struct DevStats {
    unsigned long long sum_iters, nonconv, nan, domain, nrec, stalled;
    unsigned int max_iters;
    unsigned int pad;
};

constexpr int64_t kScratchFlags = int64_t(1) << 20;
constexpr size_t kScratchHead = 256;  // DevStats, padded
static_assert(sizeof(DevStats) <= kScratchHead, "statistics record fits the head");

}  // namespace

// clm_step_noalloc's scratch (its layout and use: scratch_address below)
static __device__ __align__(256) unsigned char scratch[kScratchHead + 4 * kScratchFlags];

struct clm_ctx {
    int model = 0, bc_bottom = 0, bc_top = 0, jac_prec = 64, max_iters = 3;
    int jacobian = CLM_JACOBIAN_PICARD;
    double tol = 0.0;
    int64_t ncol = 0, ncell = 0;
    int nparam = 0;
    std::vector<int32_t> col_ptr;
    std::vector<double> h_zf, h_par;  // host copies (checkpoint model hash)
    // layout
    int32_t npack = 0;
    int64_t nslot = 0, npcell = 0, npface = 0;
    std::vector<int32_t> slot_col, slot_nl;
    std::vector<int64_t> cell_off, face_off;
    int64_t* d_cell_off = nullptr;
    int64_t* d_face_off = nullptr;
    int32_t* d_nl = nullptr;
    int32_t* d_slot_col = nullptr;
    double* d_par = nullptr;
    double* d_zf = nullptr;
    // state
    double *d_xv = nullptr, *d_xi = nullptr, *d_xe = nullptr, *d_iw = nullptr, *d_ie = nullptr;
    // forcing
    double *d_ftw = nullptr, *d_fth = nullptr, *d_fbw = nullptr, *d_fbh = nullptr;
    double *d_tab_w = nullptr, *d_tab_h = nullptr;
    int32_t tab_slots = 0;
    // aux (fine hooks)
    double *d_K = nullptr, *d_psi = nullptr, *d_thl = nullptr, *d_T = nullptr, *d_kap = nullptr;
    double *d_tw = nullptr, *d_th = nullptr, *d_bw = nullptr, *d_bh = nullptr;
    double *d_sink = nullptr, *d_rss = nullptr;  // TOPMODEL subsurface runoff
    // EnergyHydrology AtmosDrivenFluxBC: energy sink, R_ess, ice vapour flux
    double *d_sinke = nullptr, *d_ress = nullptr, *d_vice = nullptr;
    // its surface inputs (P_liq, T_air, lhf, shf, vapor_liq, vapor_ice, R_n):
    // slot order (hooks), column order (group kernels), [slot][col] table
    double* d_sfc[7] = {nullptr};
    double* d_csfc[7] = {nullptr};
    double* d_tab_sfc[7] = {nullptr};
    int32_t sfc_slots = 0;
    bool have_cache = false, have_jac = false;
    // fine-hook work buffers
    double* d_wy[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double* d_wd[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    void* d_W = nullptr;  // 9 planes of JT (ww lo/di/up, ee, ew)
    int32_t* d_bad = nullptr;
    // fused-step buffers
    double *d_U0v = nullptr, *d_U0e = nullptr, *d_Kl = nullptr, *d_kl = nullptr, *d_tl = nullptr;
    double *d_Tp = nullptr, *d_Pp = nullptr, *d_Re = nullptr;
    void* d_scr = nullptr;  // 6 planes of JT
    int32_t* d_iters = nullptr;
    int64_t iters_cap = 0;  // in steps
    DevStats* d_stats = nullptr;
    // group-cooperative path: CSR state in column order, columns binned by depth
    int solver = CLM_SOLVER_GROUP;
    int32_t* d_col_ptr = nullptr;
    double *d_cpar = nullptr, *d_czf = nullptr;
    double *d_cv = nullptr, *d_ci = nullptr, *d_ce = nullptr, *d_ciw = nullptr, *d_cie = nullptr;
    double *d_cftw = nullptr, *d_cfth = nullptr, *d_cfbw = nullptr, *d_cfbh = nullptr;
    struct Bin {
        int G, R;
        int32_t n;
        int32_t* d_cols;  // into d_bincols
    };
    // all bins' column lists back to back (bin b at bin_off[b]), a copy of
    // them, and the regrouping histogram / fill pointers [bin][bucket]
    int32_t *d_bincols = nullptr, *d_bintmp = nullptr, *d_rghist = nullptr, *d_rgfill = nullptr;
    int32_t bin_off[17] = {0};
    std::vector<Bin> bins;  // one per depth bucket (n may be 0)
    int bucket_G[16], bucket_R[16], bucket_nmax[16];
    struct WsState* ws = nullptr;  // learned warm start (clm_ws_*)
    // stream, graph, events
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // the depth bins of a step are independent: each launches on a stream of
    // its own (forked from and joined back into `stream`, also under graph
    // capture), so one bin's last waves share the GPU with the next bin's
    // first instead of draining alone
    static constexpr int kBinStreams = 4;
    hipStream_t bin_stream[kBinStreams] = {};
    hipEvent_t bin_fork = nullptr, bin_join[kBinStreams] = {};
    // inside a run, a step's regrouping runs on bin_stream[0] beside the next
    // step's warm-start inference; the next step's fork waits for rg_done
    hipEvent_t rg_start = nullptr, rg_done = nullptr;
    bool rg_pending = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int32_t g_nsteps = -1, g_sps = -1, g_tab = -1;
    double g_dt = 0.0;
    double last_ms = 0.0;
    int32_t last_launches = 0;
    // paired waves for the bins of <= 40 and <= 24 layers (launch_pair)
    bool pairing = true;
    // device address of this GPU's static scratch (clm_step_noalloc)
    unsigned char* d_scratch = nullptr;
    std::string err;
};

// ---------------------------------------------------------------------------
namespace {

int set_err(clm_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c)
        c->err = buf;
    else
        g_create_error = buf;
    return code;
}

#define HIPCHK(ctx, call)                                                                \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return set_err((ctx), CLM_E_HIP, "%s failed: %s (%s:%d)", #call,              \
                           hipGetErrorString(e_), __FILE__, __LINE__);                   \
    } while (0)

// Allocate and zero on the context stream (stream-ordered with all later
// work of the context; the null stream does not order against it).
template <typename T>
int dalloc(clm_ctx* c, T** p, int64_t n) {
    size_t bytes = sizeof(T) * (size_t)std::max<int64_t>(n, 1);
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess)
        return set_err(c, CLM_E_HIP, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    e = hipMemsetAsync(*p, 0, bytes, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
        return set_err(c, CLM_E_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
    return CLM_OK;
}

template <typename T>
void dfree(T*& p) {
    if (p) hipFree((void*)p);
    p = nullptr;
}

bool is_eh(const clm_ctx* c) { return c->model == CLM_MODEL_ENERGY_HYDROLOGY; }
// RichardsModel kernels take both boundary kinds as bottom | top << 4
int rre_bc(const clm_ctx* c) { return c->bc_bottom | (c->bc_top << 4); }
bool is_group(const clm_ctx* c) { return c->solver == CLM_SOLVER_GROUP; }
// length of per-column device vectors of the coarse path
int64_t ncolvec(const clm_ctx* c) { return is_group(c) ? c->ncol : c->nslot; }

// Group-solver shapes: G lanes per column, R rows per lane (capacity G*R).
struct Shape {
    int G, R;
};
constexpr Shape kMenu[] = {{8, 1},  {16, 1}, {32, 1}, {64, 1}, {8, 2},  {16, 2}, {32, 2},
                           {64, 2}, {8, 3},  {16, 3}, {32, 3}, {4, 4},  {8, 4},  {16, 4},
                           {32, 4}, {8, 5},  {16, 5}, {8, 6},  {8, 8},  {16, 8}};
constexpr int kNMenu = sizeof(kMenu) / sizeof(kMenu[0]);
// depth buckets: columns with kBucketMax[b-1] < N <= kBucketMax[b] share a shape
constexpr int kBucketMax[] = {8, 16, 24, 32, 40, 48, 64, 80, 96, 128};
constexpr int kNBucket = sizeof(kBucketMax) / sizeof(kBucketMax[0]);
constexpr int kRgBuckets = 64;  // iteration-count buckets of the bin regrouping
// shapes before autotuning: few lanes per column, several rows per lane
constexpr Shape kDefaultShape[kNBucket] = {{8, 1},  {8, 2},  {8, 3},  {8, 4},  {16, 3},
                                           {16, 3}, {16, 4}, {32, 3}, {32, 3}, {32, 4}};

int bucket_of(int N) {
    for (int b = 0; b < kNBucket; ++b)
        if (N <= kBucketMax[b]) return b;
    return -1;
}
int menu_index(int G, int R) {
    for (int i = 0; i < kNMenu; ++i)
        if (kMenu[i].G == G && kMenu[i].R == R) return i;
    return -1;
}
size_t jt_size(const clm_ctx* c) { return c->jac_prec == 32 ? sizeof(float) : sizeof(double); }

Layout layout(const clm_ctx* c) {
    Layout L;
    L.cell_off = c->d_cell_off;
    L.face_off = c->d_face_off;
    L.nl = c->d_nl;
    L.slot_col = c->d_slot_col;
    L.npack = c->npack;
    L.nslot = c->nslot;
    return L;
}

dim3 grid_slots(const clm_ctx* c) { return dim3((unsigned)((c->nslot + 255) / 256)); }

// host CSR <-> packed
void pack_cells_host(const clm_ctx* c, const double* csr, std::vector<double>& out) {
    out.assign((size_t)c->npcell, 0.0);
    for (int64_t s = 0; s < c->nslot; ++s) {
        int32_t col = c->slot_col[s];
        if (col < 0) continue;
        int64_t base = c->cell_off[s >> 6] + (s & 63);
        const double* src = csr + c->col_ptr[col];
        for (int32_t j = 0; j < c->slot_nl[s]; ++j) out[base + 64 * j] = src[j];
    }
}

void unpack_cells_host(const clm_ctx* c, const std::vector<double>& in, double* csr) {
    for (int64_t s = 0; s < c->nslot; ++s) {
        int32_t col = c->slot_col[s];
        if (col < 0) continue;
        int64_t base = c->cell_off[s >> 6] + (s & 63);
        double* dst = csr + c->col_ptr[col];
        for (int32_t j = 0; j < c->slot_nl[s]; ++j) dst[j] = in[base + 64 * j];
    }
}

void pack_cols_host(const clm_ctx* c, const double* col, std::vector<double>& out) {
    out.assign((size_t)c->nslot, 0.0);
    if (!col) return;
    for (int64_t s = 0; s < c->nslot; ++s)
        if (c->slot_col[s] >= 0) out[s] = col[c->slot_col[s]];
}

void unpack_cols_host(const clm_ctx* c, const std::vector<double>& in, double* col) {
    for (int64_t s = 0; s < c->nslot; ++s)
        if (c->slot_col[s] >= 0) col[c->slot_col[s]] = in[s];
}

int upload_cells(clm_ctx* c, const double* csr, double* d) {
    if (!csr) return set_err(c, CLM_E_ARG, "missing cell array");
    std::vector<double> h;
    pack_cells_host(c, csr, h);
    HIPCHK(c, hipMemcpyAsync(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int upload_cols(clm_ctx* c, const double* col, double* d) {
    std::vector<double> h;
    pack_cols_host(c, col, h);
    HIPCHK(c, hipMemcpyAsync(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int download_cells(clm_ctx* c, const double* d, double* csr) {
    if (!csr) return CLM_OK;
    std::vector<double> h((size_t)c->npcell);
    HIPCHK(c, hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unpack_cells_host(c, h, csr);
    return CLM_OK;
}

int download_cols(clm_ctx* c, const double* d, double* col) {
    if (!col) return CLM_OK;
    std::vector<double> h((size_t)c->nslot);
    HIPCHK(c, hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unpack_cols_host(c, h, col);
    return CLM_OK;
}

// upload a host state into the fine-hook work buffers
int upload_state(clm_ctx* c, const clm_state* Y, double* const* dst) {
    if (!Y || !Y->vartheta_l) return set_err(c, CLM_E_ARG, "state: vartheta_l required");
    int r;
    if ((r = upload_cells(c, Y->vartheta_l, dst[0]))) return r;
    if (is_eh(c)) {
        if (!Y->theta_i || !Y->rho_e_int)
            return set_err(c, CLM_E_ARG, "EnergyHydrology state needs theta_i and rho_e_int");
        if ((r = upload_cells(c, Y->theta_i, dst[1]))) return r;
        if ((r = upload_cells(c, Y->rho_e_int, dst[2]))) return r;
        if ((r = upload_cols(c, Y->intF_energy, dst[4]))) return r;
    }
    return upload_cols(c, Y->intF_water, dst[3]);
}

int download_state(clm_ctx* c, double* const* src, clm_state* Y) {
    int r;
    if ((r = download_cells(c, src[0], Y->vartheta_l))) return r;
    if ((r = download_cols(c, src[3], Y->intF_water))) return r;
    if (is_eh(c)) {
        if ((r = download_cells(c, src[1], Y->theta_i))) return r;
        if ((r = download_cells(c, src[2], Y->rho_e_int))) return r;
        if ((r = download_cols(c, src[4], Y->intF_energy))) return r;
    }
    return CLM_OK;
}

// Column totals (total_liq_water_vol_per_area!, rre.jl:502-511 and
// energy_hydrology.jl:1282-1304; total_energy_per_area!, :1306-1320):
// column_integral_definite of ϑ_l (+ θ_i ρ_i/ρ_l) and of ρe_int, Σ f Δz over
// the layers.  One column per lane: CSR order (group path) ...
__global__ __launch_bounds__(256) void k_totals_csr(const int32_t* __restrict__ col_ptr,
                                                    const double* __restrict__ zf,
                                                    const double* __restrict__ vl,
                                                    const double* __restrict__ ti,
                                                    const double* __restrict__ re, int64_t ncol,
                                                    double* __restrict__ water,
                                                    double* __restrict__ energy) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncol) return;
    const int64_t a = col_ptr[c], n = col_ptr[c + 1] - a;
    const double* z = zf + a + c;
    double w = 0.0, e = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double dz = z[i + 1] - z[i];
        double vol = vl[a + i];
        if (ti) {
            vol += ti[a + i] * RHO_I / RHO_L;
            e += re[a + i] * dz;
        }
        w += vol * dz;
    }
    water[c] = w;
    if (energy) energy[c] = e;
}
// ... and the packed lane layout (slot order; a level is one coalesced row)
__global__ __launch_bounds__(256) void k_totals_packed(Layout L, const double* __restrict__ zf,
                                                       const double* __restrict__ vl,
                                                       const double* __restrict__ ti,
                                                       const double* __restrict__ re,
                                                       double* __restrict__ water,
                                                       double* __restrict__ energy) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= L.nslot) return;
    const int64_t cb = L.cell_off[s >> 6] + (s & 63), fb = L.face_off[s >> 6] + (s & 63);
    const int n = L.nl[s];
    double w = 0.0, e = 0.0;
    for (int j = 0; j < n; ++j) {
        const double dz = zf[fb + 64 * (j + 1)] - zf[fb + 64 * j];
        double vol = vl[cb + 64 * j];
        if (ti) {
            vol += ti[cb + 64 * j] * RHO_I / RHO_L;
            e += re[cb + 64 * j] * dz;
        }
        w += vol * dz;
    }
    water[s] = w;
    if (energy) energy[s] = e;
}

// Surface soil conductance to water vapour (energy_hydrology.jl:1103-1131):
// θ_l of the top two cells (volumetric_liquid_fraction with ν - θ_i) extrapolated
// linearly to the surface (linear_interpolation_to_surface!, Domains.jl:883-895),
// floored at θ_r + eps, its effective saturation with ν, then
// soil_conductance(S_l, S_c, d_ds, evap_p, evap_α, D_vapor).  One column per
// lane, CSR state of the group path.
__global__ __launch_bounds__(256) void k_soil_conductance(
    const int32_t* __restrict__ col_ptr, const double* __restrict__ zf,
    const double* __restrict__ par, int64_t ncol, const double* __restrict__ vl,
    const double* __restrict__ ti, double d_ds, double evap_p, double evap_alpha, double D_vapor,
    double* __restrict__ g_out, double* __restrict__ dsl_out) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncol) return;
    const int64_t a = col_ptr[c], n = col_ptr[c + 1] - a;
    const double* z = zf + a + c;
    const VG v = make_vg(par[P_VG_ALPHA * ncol + c], par[P_VG_N * ncol + c],
                         par[P_THETA_R * ncol + c], par[P_SS * ncol + c], par[P_KSAT * ncol + c]);
    const double nu = par[P_NU * ncol + c];
    const int64_t t1 = a + n - 1, t2 = a + n - 2;
    const double f1 = liq_fraction(v, vl[t1], nu - ti[t1]);
    const double f2 = liq_fraction(v, vl[t2], nu - ti[t2]);
    const double z1 = 0.5 * (z[n - 1] + z[n]), z2 = 0.5 * (z[n - 2] + z[n - 1]);
    const double dz_top = z[n] - z1;
    double th = (f1 - f2) / (z1 - z2) * (dz_top + z1 - z2) + f2;
    th = fmax(th, v.theta_r + DEPS);
    const double S = eff_sat(v, nu, th);
    const double Sc = vg_sc(v.n);
    g_out[c] = soil_conductance(S, Sc, d_ds, evap_p, evap_alpha, D_vapor);
    if (dsl_out) dsl_out[c] = dry_soil_layer_thickness(S, evap_alpha * Sc, d_ds, evap_p);
}

__global__ void k_zero_stats(DevStats* s) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        s->sum_iters = s->nonconv = s->nan = s->domain = s->nrec = s->stalled = 0;
        s->max_iters = 0;
    }
}

__global__ __launch_bounds__(256) void k_reduce_stats(const int32_t* __restrict__ it,
                                                      const int32_t* __restrict__ nl,
                                                      int64_t nslot, int64_t nrec,
                                                      DevStats* __restrict__ st) {
    unsigned long long sum = 0, nc = 0, nn = 0, nd = 0, cnt = 0, ns = 0;
    unsigned int mx = 0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrec;
         r += (int64_t)gridDim.x * blockDim.x) {
        if (nl && nl[r % nslot] <= 0) continue;
        int32_t v = it[r];
        unsigned int k = (unsigned int)(v & 0xffff);
        sum += k;
        mx = k > mx ? k : mx;
        nc += (v & 0x10000) ? 1 : 0;
        nn += (v & 0x20000) ? 1 : 0;
        nd += (v & 0x40000) ? 1 : 0;
        ns += (v & 0x100000) ? 1 : 0;
        cnt += 1;
    }
    // wavefront reduction then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_down(sum, off);
        nc += __shfl_down(nc, off);
        nn += __shfl_down(nn, off);
        nd += __shfl_down(nd, off);
        cnt += __shfl_down(cnt, off);
        ns += __shfl_down(ns, off);
        unsigned int o = __shfl_down(mx, off);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&st->sum_iters, sum);
        atomicAdd(&st->nonconv, nc);
        atomicAdd(&st->nan, nn);
        atomicAdd(&st->domain, nd);
        atomicAdd(&st->nrec, cnt);
        atomicAdd(&st->stalled, ns);
        atomicMax(&st->max_iters, mx);
    }
}

int collect_stats(clm_ctx* c, int64_t nsteps, clm_step_stats* out,
                  const int32_t* iters = nullptr, DevStats* dst = nullptr) {
    if (!iters) iters = c->d_iters;
    if (!dst) dst = c->d_stats;
    hipLaunchKernelGGL(k_zero_stats, dim3(1), dim3(64), 0, c->stream, dst);
    int64_t nv = ncolvec(c);
    int64_t nrec = nsteps * nv;
    unsigned g = (unsigned)std::min<int64_t>((nrec + 255) / 256, 4096);
    hipLaunchKernelGGL(k_reduce_stats, dim3(g), dim3(256), 0, c->stream, iters,
                       is_group(c) ? nullptr : c->d_nl, nv, nrec, dst);
    HIPCHK(c, hipGetLastError());
    DevStats h;
    HIPCHK(c, hipMemcpyAsync(&h, dst, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out->max_iters_used = (int32_t)h.max_iters;
    out->n_stalled = (int32_t)std::min<unsigned long long>(h.stalled, 0x7fffffffull);
    out->mean_iters = h.nrec ? (double)h.sum_iters / (double)h.nrec : 0.0;
    out->n_nonconverged = (int64_t)h.nonconv;
    out->n_nan = (int64_t)h.nan;
    out->n_domain_error = (int64_t)h.domain;
    return CLM_OK;
}

void destroy_graph(clm_ctx* c);

int ensure_iters(clm_ctx* c, int64_t nsteps) {
    if (nsteps <= c->iters_cap) return CLM_OK;
    // a captured run graph writes the old table: drop it with the table, and
    // keep iters_cap at 0 until the new one exists
    destroy_graph(c);
    c->iters_cap = 0;
    dfree(c->d_iters);
    int r = dalloc(c, &c->d_iters, nsteps * ncolvec(c));
    if (r) return r;
    c->iters_cap = nsteps;
    return CLM_OK;
}

template <typename JT>
EHScratch<JT> eh_scratch(clm_ctx* c) {
    JT* b = (JT*)c->d_scr;
    int64_t n = c->npcell;
    return EHScratch<JT>{b, b + n, b + 2 * n, b + 3 * n, b + 4 * n, b + 5 * n};
}

EHBufs eh_bufs(clm_ctx* c) {
    return EHBufs{c->d_xv, c->d_xi, c->d_xe, c->d_U0v, c->d_U0e, c->d_Kl, c->d_kl,
                  c->d_tl, c->d_Tp, c->d_Pp, c->d_Re, c->d_iw, c->d_ie};
}

// launch one fused step on the ctx stream (capturable: kernels only)
template <typename JT>
void launch_step_t(clm_ctx* c, const double* ftw, const double* fth, double dt,
                   int32_t* iters) {
    Layout L = layout(c);
    dim3 g = grid_slots(c), b(256);
    if (!is_eh(c)) {
        JT* cp = (JT*)c->d_scr;
        JT* dp = cp + c->npcell;
        hipLaunchKernelGGL(k_rre_step<JT>, g, b, 0, c->stream, L, c->d_par, c->d_zf, c->d_xv,
                           c->d_U0v, c->d_iw, ftw, c->d_fbw, rre_bc(c), dt, c->max_iters,
                           c->tol, cp, dp, iters);
    } else {
        hipLaunchKernelGGL(k_eh_step<JT>, g, b, 0, c->stream, L, c->d_par, c->d_zf,
                           eh_bufs(c), eh_scratch<JT>(c), ftw, fth, c->d_fbw, c->d_fbh,
                           c->bc_bottom, dt, c->max_iters, c->tol, iters);
    }
}

// one launch per depth bin; every kernel is independent of the others
// optional per-step extras of the group kernels: a Newton warm start
// (x0 = U0 + guess) and an output copy of the predictor U0_ϑ
struct StepExtra {
    const float* wso = nullptr;  // warm-start network output rows [ncol][Nout]
    double* u0out = nullptr;
    Surface sf{};  // AtmosDrivenFluxBC inputs per column (EH)
    // state in other device memory (CSR order) instead of the context's
    double *xv = nullptr, *xi = nullptr, *xe = nullptr, *iw = nullptr, *ie = nullptr;
};

bool is_atmos(const clm_ctx* c) {
    return c->model == CLM_MODEL_ENERGY_HYDROLOGY &&
           (c->bc_top == CLM_BC_ATMOS || c->bc_top == CLM_BC_ATMOS_TOPMODEL);
}
Surface surface_of(double* const* a, int64_t off) {
    Surface s;
    s.P = a[0] + off;
    s.Tair = a[1] + off;
    s.lhf = a[2] + off;
    s.shf = a[3] + off;
    s.vliq = a[4] + off;
    s.vice = a[5] + off;
    s.Rn = a[6] + off;
    return s;
}
// extras of a plain step on the current forcing
StepExtra current_extra(const clm_ctx* c) {
    StepExtra x;
    if (is_atmos(c)) x.sf = surface_of(c->d_csfc, 0);
    return x;
}

template <typename JT, int G, int R>
void launch_bin(clm_ctx* c, const clm_ctx::Bin& bn, const double* ftw, const double* fth,
                double dt, int32_t* iters, const StepExtra& x) {
    constexpr int gpb = 256 / G;  // columns per 256-thread block
    dim3 g((unsigned)((bn.n + gpb - 1) / gpb)), b(256);
    double* xv = x.xv ? x.xv : c->d_cv;
    double* xi = x.xv ? x.xi : c->d_ci;
    double* xe = x.xv ? x.xe : c->d_ce;
    double* iw = x.xv ? x.iw : c->d_ciw;
    double* ie = x.xv ? x.ie : c->d_cie;
    if (!is_eh(c)) {
        hipLaunchKernelGGL((k_rre_coop<G, R, JT>), g, b, 0, c->stream, c->d_col_ptr, bn.d_cols,
                           bn.n, c->d_cpar, c->ncol, c->d_czf, xv, iw, ftw, c->d_cfbw,
                           rre_bc(c), dt, c->max_iters, c->tol, iters, x.wso, x.u0out);
    } else {
        if (is_atmos(c))
            hipLaunchKernelGGL((k_eh_coop<G, R, JT, true>), g, b, 0, c->stream, c->d_col_ptr,
                               bn.d_cols, bn.n, c->d_cpar, c->ncol, c->d_czf, xv, xi, xe, iw, ie,
                               ftw, fth, c->d_cfbw, c->d_cfbh, c->bc_bottom, dt, c->max_iters,
                               c->tol, iters, x.wso, x.u0out, c->bc_top, x.sf, nullptr, 0);
        else if (c->jacobian == CLM_JACOBIAN_NEWTON)
            hipLaunchKernelGGL((k_eh_coop<G, R, JT, false, true>), g, b, 0, c->stream,
                               c->d_col_ptr, bn.d_cols, bn.n, c->d_cpar, c->ncol, c->d_czf, xv,
                               xi, xe, iw, ie, ftw, fth, c->d_cfbw, c->d_cfbh, c->bc_bottom, dt,
                               c->max_iters, c->tol, iters, x.wso, x.u0out, c->bc_top,
                               x.sf, nullptr, 0);
        else
            hipLaunchKernelGGL((k_eh_coop<G, R, JT, false>), g, b, 0, c->stream, c->d_col_ptr,
                               bn.d_cols, bn.n, c->d_cpar, c->ncol, c->d_czf, xv, xi, xe, iw, ie,
                               ftw, fth, c->d_cfbw, c->d_cfbh, c->bc_bottom, dt, c->max_iters,
                               c->tol, iters, x.wso, x.u0out, c->bc_top, x.sf, nullptr, 0);
    }
}

template <typename JT, int I>
void launch_menu(clm_ctx* c, int idx, const clm_ctx::Bin& bn, const double* ftw,
                 const double* fth, double dt, int32_t* iters, const StepExtra& x) {
    if constexpr (I < kNMenu) {
        if (idx == I)
            launch_bin<JT, kMenu[I].G, kMenu[I].R>(c, bn, ftw, fth, dt, iters, x);
        else
            launch_menu<JT, I + 1>(c, idx, bn, ftw, fth, dt, iters, x);
    }
}

template <typename JT>
void launch_bin_any(clm_ctx* c, const clm_ctx::Bin& bn, const double* ftw, const double* fth,
                    double dt, int32_t* iters, const StepExtra& x = StepExtra()) {
    if (bn.n > 0) launch_menu<JT, 0>(c, menu_index(bn.G, bn.R), bn, ftw, fth, dt, iters, x);
}

// Paired waves (k_eh_coop<64, 1, JT, false, NEWTON, true>): the bin of 33-40
// layers on shape (64, 1) and the bin of 17-24 layers on shape (32, 1) share
// waves, the i-th column of each list in wave i (both lists are ordered by
// the last step's Newton counts, so a wave's columns iterate alike); the
// shallow bin's columns beyond the deep bin's count run on their own shape.
// EnergyHydrology with a prescribed-flux top boundary.
constexpr int kPairDeepBin = 4, kPairShallowBin = 2;  // buckets of <= 40 and <= 24 layers

bool pairing_applies(const clm_ctx* c) {
    if (!c->pairing || !is_eh(c) || is_atmos(c) || !is_group(c)) return false;
    const clm_ctx::Bin &d = c->bins[kPairDeepBin], &h = c->bins[kPairShallowBin];
    static_assert(kBucketMax[kPairDeepBin] == kPairSplit &&
                      kBucketMax[kPairShallowBin] <= 64 - kPairSplit,
                  "paired bins fit the wave's two parts");
    return d.n > 0 && h.n > 0 && d.G == 64 && d.R == 1 && h.G == 32 && h.R == 1;
}

template <typename JT>
void launch_pair(clm_ctx* c, const double* ftw, const double* fth, double dt, int32_t* iters,
                 const StepExtra& x) {
    const clm_ctx::Bin &d = c->bins[kPairDeepBin], &h = c->bins[kPairShallowBin];
    const int nb = std::min(d.n, h.n);
    dim3 g((unsigned)((d.n + 3) / 4)), b(256);  // one wave per deep column
    double* xv = x.xv ? x.xv : c->d_cv;
    double* xi = x.xv ? x.xi : c->d_ci;
    double* xe = x.xv ? x.xe : c->d_ce;
    double* iw = x.xv ? x.iw : c->d_ciw;
    double* ie = x.xv ? x.ie : c->d_cie;
    if (c->jacobian == CLM_JACOBIAN_NEWTON)
        hipLaunchKernelGGL((k_eh_coop<64, 1, JT, false, true, true>), g, b, 0, c->stream,
                           c->d_col_ptr, d.d_cols, d.n, c->d_cpar, c->ncol, c->d_czf, xv, xi, xe,
                           iw, ie, ftw, fth, c->d_cfbw, c->d_cfbh, c->bc_bottom, dt, c->max_iters,
                           c->tol, iters, x.wso, x.u0out, c->bc_top, x.sf, h.d_cols, nb);
    else
        hipLaunchKernelGGL((k_eh_coop<64, 1, JT, false, false, true>), g, b, 0, c->stream,
                           c->d_col_ptr, d.d_cols, d.n, c->d_cpar, c->ncol, c->d_czf, xv, xi, xe,
                           iw, ie, ftw, fth, c->d_cfbw, c->d_cfbh, c->bc_bottom, dt, c->max_iters,
                           c->tol, iters, x.wso, x.u0out, c->bc_top, x.sf, h.d_cols, nb);
    if (h.n > nb) {  // the rest of the shallow bin on its own shape
        clm_ctx::Bin rest = h;
        rest.d_cols = h.d_cols + nb;
        rest.n = h.n - nb;
        launch_bin_any<JT>(c, rest, ftw, fth, dt, iters, x);
    }
}

template <typename JT>
void launch_group_t(clm_ctx* c, const double* ftw, const double* fth, double dt,
                    int32_t* iters, const StepExtra& x) {
    const bool pair = pairing_applies(c);
    const hipStream_t main = c->stream;
    if (c->rg_pending) {  // the last step's regrouping, before any bin reads its list
        (void)hipStreamWaitEvent(main, c->rg_done, 0);
        c->rg_pending = false;
    }
    (void)hipEventRecord(c->bin_fork, main);
    int k = 0;  // bins launched
    for (int b = 0; b < (int)c->bins.size(); ++b) {
        if (pair && b == kPairShallowBin) continue;
        if (c->bins[b].n == 0) continue;
        const int si = k % clm_ctx::kBinStreams;
        // a stream reused within the step waits for its previous bin anyway
        if (k < clm_ctx::kBinStreams) (void)hipStreamWaitEvent(c->bin_stream[si], c->bin_fork, 0);
        c->stream = c->bin_stream[si];
        if (pair && b == kPairDeepBin)
            launch_pair<JT>(c, ftw, fth, dt, iters, x);
        else
            launch_bin_any<JT>(c, c->bins[b], ftw, fth, dt, iters, x);
        c->stream = main;
        ++k;
    }
    for (int i = 0; i < std::min(k, clm_ctx::kBinStreams); ++i) {
        (void)hipEventRecord(c->bin_join[i], c->bin_stream[i]);
        (void)hipStreamWaitEvent(main, c->bin_join[i], 0);
    }
}

// ---- regrouping of a bin's columns by Newton iteration count ---------------------
// The lanes of a wave serve 64 / G columns and the wave runs until its slowest
// column has converged.  After each step every bin's column list is reordered
// by that step's iteration counts, largest first (a counting sort over 64
// buckets, ties in any order), so that the next step packs columns of similar
// counts into a wave; counts change slowly from step to step.  Each column's
// arithmetic does not depend on its neighbours, so results are unchanged.
// Three launches over all bins' lists at once (back to back in one array):
// count (block histograms, then one global atomic per bin and bucket), scan
// (fill pointers; clears the histogram for the next step) and scatter.
struct BinOffsets {
    int32_t off[kNBucket + 1];
};
__device__ __forceinline__ int bin_of_index(const BinOffsets& bo, int i) {
    int b = 0;
#pragma unroll
    for (int k = 1; k < kNBucket; ++k) b += i >= bo.off[k] ? 1 : 0;
    return b;
}
__device__ __forceinline__ int rg_bucket(const int32_t* __restrict__ iters, int32_t col) {
    return kRgBuckets - 1 - min(iters[col] & 0xFFFF, kRgBuckets - 1);
}

__global__ __launch_bounds__(256) void k_rg_count(const int32_t* __restrict__ cols,
                                                  int32_t* __restrict__ tmp, int n, BinOffsets bo,
                                                  const int32_t* __restrict__ iters,
                                                  int32_t* __restrict__ hist) {
    __shared__ int h[kNBucket * kRgBuckets];
    for (int k = threadIdx.x; k < kNBucket * kRgBuckets; k += 256) h[k] = 0;
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        int32_t col = cols[i];
        tmp[i] = col;
        atomicAdd(&h[bin_of_index(bo, i) * kRgBuckets + rg_bucket(iters, col)], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kNBucket * kRgBuckets; k += 256)
        if (h[k]) atomicAdd(&hist[k], h[k]);
}

// one wave per bin: lane q owns bucket q
__global__ __launch_bounds__(64 * kNBucket) void k_rg_scan(BinOffsets bo, int32_t* __restrict__ hist,
                                                           int32_t* __restrict__ fill) {
    const int b = threadIdx.x >> 6, q = threadIdx.x & 63;
    const int k = b * kRgBuckets + q;
    const int cnt = hist[k];
    int incl = cnt;
    for (int o = 1; o < kRgBuckets; o <<= 1) {
        int y = __shfl_up(incl, (unsigned)o, 64);
        if (q >= o) incl += y;
    }
    fill[k] = bo.off[b] + incl - cnt;
    hist[k] = 0;
}

__global__ __launch_bounds__(256) void k_rg_scatter(int32_t* __restrict__ cols,
                                                    const int32_t* __restrict__ tmp, int n,
                                                    BinOffsets bo,
                                                    const int32_t* __restrict__ iters,
                                                    int32_t* __restrict__ fill) {
    __shared__ int h[kNBucket * kRgBuckets];
    for (int k = threadIdx.x; k < kNBucket * kRgBuckets; k += 256) h[k] = 0;
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    int32_t col = 0;
    int key = 0, rank = 0;
    if (i < n) {
        col = tmp[i];
        key = bin_of_index(bo, i) * kRgBuckets + rg_bucket(iters, col);
        rank = atomicAdd(&h[key], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kNBucket * kRgBuckets; k += 256)
        if (h[k]) h[k] = atomicAdd(&fill[k], h[k]);  // this block's base in bucket k
    __syncthreads();
    if (i < n) cols[h[key] + rank] = col;
}

// tmp: ncol entries; hist, fill: kNBucket * kRgBuckets entries, hist zero on
// entry (k_rg_scan leaves it zero again)
void launch_regroup(clm_ctx* c, const int32_t* iters, int32_t* tmp, int32_t* hist,
                    int32_t* fill) {
    BinOffsets bo;
    for (int b = 0; b <= kNBucket; ++b) bo.off[b] = c->bin_off[b];
    const int n = (int)c->ncol;
    const dim3 g((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_rg_count, g, dim3(256), 0, c->stream, c->d_bincols, tmp, n, bo, iters,
                       hist);
    hipLaunchKernelGGL(k_rg_scan, dim3(1), dim3(64 * kNBucket), 0, c->stream, bo, hist, fill);
    hipLaunchKernelGGL(k_rg_scatter, g, dim3(256), 0, c->stream, c->d_bincols, tmp, n, bo, iters,
                       fill);
}
void launch_regroup(clm_ctx* c, const int32_t* iters) {
    launch_regroup(c, iters, c->d_bintmp, c->d_rghist, c->d_rgfill);
}

void ws_launch_guess(clm_ctx* c, const double* ftw, const double* fth, double dt,
                     StepExtra* x);

// defer_regroup (a run's steps but its last): the regrouping goes on
// bin_stream[0] and the next step's fork waits for it, so it runs beside
// that step's warm-start inference on the ctx stream
void launch_step(clm_ctx* c, const double* ftw, const double* fth, double dt, int32_t* iters,
                 double* u0out = nullptr, bool allow_warm = true, const Surface* sf = nullptr,
                 bool defer_regroup = false) {
    if (is_group(c)) {
        StepExtra x = current_extra(c);
        if (sf) x.sf = *sf;
        x.u0out = u0out;
        if (allow_warm) ws_launch_guess(c, ftw, fth, dt, &x);
        if (c->jac_prec == 32)
            launch_group_t<float>(c, ftw, fth, dt, iters, x);
        else
            launch_group_t<double>(c, ftw, fth, dt, iters, x);
        if (defer_regroup) {
            const hipStream_t main = c->stream;
            (void)hipEventRecord(c->rg_start, main);
            (void)hipStreamWaitEvent(c->bin_stream[0], c->rg_start, 0);
            c->stream = c->bin_stream[0];
            launch_regroup(c, iters);
            c->stream = main;
            (void)hipEventRecord(c->rg_done, c->bin_stream[0]);
            c->rg_pending = true;
        } else {
            launch_regroup(c, iters);
        }
        return;
    }
    if (c->jac_prec == 32)
        launch_step_t<float>(c, ftw, fth, dt, iters);
    else
        launch_step_t<double>(c, ftw, fth, dt, iters);
}

// ---- learned warm start (clm_ws_*) -----------------------------------------------
using ws::bf16;

}  // namespace

struct WsState {
    ws::Dims d;
    int ln[4], lk[4];          // layer output / input features
    int64_t off[4], offw[4];   // offsets: flat FP32 parameters (Wt then b) / BF16 weights
    int64_t P = 0, PW = 0;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    bf16 *wt = nullptr, *wn = nullptr;  // BF16 Wt [N x K] and W = Wt^T [K x N]
    bf16* wpk = nullptr;                // BF16 Wt chunk-major [K/32][N][32] (fused inference)
    bf16* tails = nullptr;              // per-column static tail features [ncol][8]
    int64_t adam_t = 0;
    bool enabled = false;
    // inference (one row per column)
    bf16 *X = nullptr, *Ha = nullptr, *Hb = nullptr;
    float* Y = nullptr;
    double *gv = nullptr, *ge = nullptr;
    // training set
    int64_t rows = 0, cap = 0;
    bf16* Xd = nullptr;
    float *Td = nullptr, *Md = nullptr;
    double *unv = nullptr, *uni = nullptr, *une = nullptr, *u0v = nullptr;
    // minibatch work
    int64_t Bm = 0, idx_cap = 0;
    bf16 *Xb = nullptr, *Hbuf[3] = {nullptr, nullptr, nullptr}, *dY = nullptr, *dZa = nullptr,
         *dZb = nullptr, *tA = nullptr, *tB = nullptr;
    float *Tb = nullptr, *Mk = nullptr, *Z[3] = {nullptr, nullptr, nullptr}, *Yb = nullptr,
          *part = nullptr, *loss = nullptr, *f32tmp = nullptr;
    int32_t* idx = nullptr;
};

namespace {

void ws_free_batch(WsState* w) {
    dfree(w->Xb);
    for (int l = 0; l < 3; ++l) {
        dfree(w->Hbuf[l]);
        dfree(w->Z[l]);
    }
    dfree(w->dY);
    dfree(w->dZa);
    dfree(w->dZb);
    dfree(w->tA);
    dfree(w->tB);
    dfree(w->Tb);
    dfree(w->Mk);
    dfree(w->Yb);
    dfree(w->part);
    dfree(w->f32tmp);
    w->Bm = 0;
}

void ws_free(WsState* w) {
    if (!w) return;
    ws_free_batch(w);
    dfree(w->p);
    dfree(w->g);
    dfree(w->m1);
    dfree(w->m2);
    dfree(w->wt);
    dfree(w->wn);
    dfree(w->wpk);
    dfree(w->tails);
    dfree(w->X);
    dfree(w->Ha);
    dfree(w->Hb);
    dfree(w->Y);
    dfree(w->gv);
    dfree(w->ge);
    dfree(w->Xd);
    dfree(w->Td);
    dfree(w->Md);
    dfree(w->unv);
    dfree(w->uni);
    dfree(w->une);
    dfree(w->u0v);
    dfree(w->loss);
    dfree(w->idx);
    delete w;
}

dim3 grid1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

template <int EPI>
void ws_gemm(clm_ctx* c, int M, int N, int K, const bf16* A, int lda, const bf16* B, int ldb,
             const float* bias, float* Cf, int ldcf, bf16* Cb, int ldcb, const float* Z, int ldz,
             int S = 1) {
    if (S == 1 && (EPI == ws::EPI_BIAS_GELU || EPI == ws::EPI_DGELU) && (N == 384 || N == 256)) {
        // hidden-layer width: one block owns whole output rows
        dim3 g((unsigned)((M + ws::kWBM - 1) / ws::kWBM));
        if (N == 384)
            hipLaunchKernelGGL((ws::k_gemm_wide<EPI, 3>), g, dim3(512), 0, c->stream, M, K, A,
                               lda, B, ldb, bias, Cf, ldcf, Cb, ldcb, Z, ldz);
        else
            hipLaunchKernelGGL((ws::k_gemm_wide<EPI, 2>), g, dim3(512), 0, c->stream, M, K, A,
                               lda, B, ldb, bias, Cf, ldcf, Cb, ldcb, Z, ldz);
        return;
    }
    dim3 g((unsigned)((M + ws::kBM - 1) / ws::kBM), (unsigned)((N + ws::kBN - 1) / ws::kBN),
           (unsigned)S);
    hipLaunchKernelGGL(ws::k_gemm_nt<EPI>, g, dim3(256), 0, c->stream, M, N, K, K / S, A, lda,
                       B, ldb, bias, Cf, ldcf, Cb, ldcb, Z, ldz);
}

// forward pass over M rows of X; train keeps pre-activations and activations
void ws_forward(clm_ctx* c, WsState* w, const bf16* X, int M, bool train, float* Yout) {
    const bf16* in = X;
    int kin = w->d.Fp;
    const int H = w->d.H;
    for (int l = 0; l < 3; ++l) {
        bf16* out = train ? w->Hbuf[l] : (l % 2 == 0 ? w->Ha : w->Hb);
        ws_gemm<ws::EPI_BIAS_GELU>(c, M, H, kin, in, kin, w->wt + w->offw[l], kin,
                                   w->p + w->off[l] + (int64_t)H * kin, train ? w->Z[l] : nullptr,
                                   H, out, H, nullptr, 0);
        in = out;
        kin = H;
    }
    ws_gemm<ws::EPI_BIAS_F32>(c, M, w->d.Nout, H, in, H, w->wt + w->offw[3], H,
                              w->p + w->off[3] + (int64_t)w->d.Nout * H, Yout, w->d.Nout, nullptr,
                              0, nullptr, 0);
}

// gradient of layer l's weights and bias from its output gradient D [M x N]
// and its input Hin [M x K]: dWt = D^T Hin, db = column sums of D
void ws_wgrad(clm_ctx* c, WsState* w, int l, const bf16* D, const bf16* Hin, int M) {
    const int N = w->ln[l], K = w->lk[l];
    hipLaunchKernelGGL(ws::k_transpose, dim3((unsigned)((N + 31) / 32), (unsigned)((M + 31) / 32)),
                       dim3(256), 0, c->stream, M, N, D, N, w->tA, M);
    hipLaunchKernelGGL(ws::k_transpose, dim3((unsigned)((K + 31) / 32), (unsigned)((M + 31) / 32)),
                       dim3(256), 0, c->stream, M, K, Hin, K, w->tB, M);
    int S = 1;
    while (S < 16 && M % (64 * S) == 0 && M / (2 * S) >= 512) S *= 2;
    ws_gemm<ws::EPI_F32>(c, N, K, M, w->tA, M, w->tB, M, nullptr, w->part, K, nullptr, 0,
                         nullptr, 0, S);
    int64_t nk = (int64_t)N * K;
    hipLaunchKernelGGL(ws::k_sum_slices, grid1(nk), dim3(256), 0, c->stream, nk, S, w->part,
                       w->g + w->off[l]);
    const int ry = (M + 255) / 256;
    hipLaunchKernelGGL(ws::k_colsum, dim3((unsigned)((N + 63) / 64), (unsigned)ry), dim3(256), 0,
                       c->stream, M, N, D, w->part);
    hipLaunchKernelGGL(ws::k_sum_slices, grid1(N), dim3(256), 0, c->stream, (int64_t)N, ry,
                       w->part, w->g + w->off[l] + nk);
}

// backward pass from dY (bf16 [M x Nout]) to every parameter gradient
void ws_backward(clm_ctx* c, WsState* w, int M) {
    const int H = w->d.H, No = w->d.Nout;
    ws_wgrad(c, w, 3, w->dY, w->Hbuf[2], M);
    ws_gemm<ws::EPI_DGELU>(c, M, H, No, w->dY, No, w->wn + w->offw[3], No, nullptr, nullptr, 0,
                           w->dZa, H, w->Z[2], H);
    ws_wgrad(c, w, 2, w->dZa, w->Hbuf[1], M);
    ws_gemm<ws::EPI_DGELU>(c, M, H, H, w->dZa, H, w->wn + w->offw[2], H, nullptr, nullptr, 0,
                           w->dZb, H, w->Z[1], H);
    ws_wgrad(c, w, 1, w->dZb, w->Hbuf[0], M);
    ws_gemm<ws::EPI_DGELU>(c, M, H, H, w->dZb, H, w->wn + w->offw[1], H, nullptr, nullptr, 0,
                           w->dZa, H, w->Z[0], H);
    ws_wgrad(c, w, 0, w->dZa, w->Xb, M);
}

void ws_adam(clm_ctx* c, WsState* w, float lr) {
    w->adam_t += 1;
    float b1t = (float)std::pow(0.9, (double)w->adam_t);
    float b2t = (float)std::pow(0.999, (double)w->adam_t);
    for (int l = 0; l < 4; ++l) {
        int N = w->ln[l], K = w->lk[l];
        int64_t n = (int64_t)N * K + N;
        float* P = w->p + w->off[l];
        hipLaunchKernelGGL(ws::k_adam, grid1(n), dim3(256), 0, c->stream, N, K, P,
                           P + (int64_t)N * K, w->g + w->off[l], w->g + w->off[l] + (int64_t)N * K,
                           w->m1 + w->off[l], w->m2 + w->off[l], lr, b1t, b2t,
                           w->wt + w->offw[l], w->wn + w->offw[l], w->wpk + w->offw[l]);
    }
}

void ws_cast(clm_ctx* c, WsState* w) {
    for (int l = 0; l < 4; ++l) {
        int64_t n = (int64_t)w->ln[l] * w->lk[l];
        hipLaunchKernelGGL(ws::k_cast_weights, grid1(n), dim3(256), 0, c->stream, w->ln[l],
                           w->lk[l], w->p + w->off[l], w->wt + w->offw[l], w->wn + w->offw[l],
                           w->wpk + w->offw[l]);
    }
}

int ws_batch_buffers(clm_ctx* c, WsState* w, int64_t M) {
    if (w->Bm == M) return CLM_OK;
    ws_free_batch(w);
    const int64_t H = w->d.H, Fp = w->d.Fp, No = w->d.Nout;
    int r = 0;
    const int64_t kmax = std::max(H, Fp), nmax = std::max(H, No);
    if ((r = dalloc(c, &w->Xb, M * Fp)) || (r = dalloc(c, &w->Tb, M * No)) ||
        (r = dalloc(c, &w->Mk, M)) || (r = dalloc(c, &w->Yb, M * No)) ||
        (r = dalloc(c, &w->dY, M * No)) || (r = dalloc(c, &w->dZa, M * H)) ||
        (r = dalloc(c, &w->dZb, M * H)) || (r = dalloc(c, &w->tA, nmax * M)) ||
        (r = dalloc(c, &w->tB, kmax * M)) || (r = dalloc(c, &w->part, 16 * nmax * kmax)) ||
        (r = dalloc(c, &w->f32tmp, M * std::max(Fp, No))))
        return r;
    for (int l = 0; l < 3; ++l)
        if ((r = dalloc(c, &w->Hbuf[l], M * H)) || (r = dalloc(c, &w->Z[l], M * H))) return r;
    w->Bm = M;
    return CLM_OK;
}

int ws_features_launch(clm_ctx* c, WsState* w, const double* ftw, const double* fth, double dt,
                       bf16* X) {
    hipLaunchKernelGGL(ws::k_features, grid1(c->ncol * ws::kFeatGroup), dim3(256), 0, c->stream,
                       is_eh(c), c->d_col_ptr, c->ncol, c->d_cpar, w->tails, c->d_cv, c->d_ci,
                       c->d_ce, ftw, fth ? fth : ftw, dt, w->d.Fp, X);
    return CLM_OK;
}

// network output rows of the current state into w->Y: the features
// (k_features), then the four layers in one fused kernel or as the
// layer-by-layer GEMMs (the same arithmetic)
void ws_infer(clm_ctx* c, WsState* w, const double* ftw, const double* fth, double dt,
              bool fused) {
    ws_features_launch(c, w, ftw, fth, dt, w->X);
    if (fused) {
        const bf16* W = w->wpk;
        const float* P = w->p;
        auto bias = [&](int l) { return P + w->off[l] + (int64_t)w->ln[l] * w->lk[l]; };
        dim3 g((unsigned)((c->ncol + ws::kIBM - 1) / ws::kIBM)), b(512);
        if (is_eh(c))
            hipLaunchKernelGGL(ws::k_ws_infer<true>, g, b, 0, c->stream, c->ncol, w->X,
                               W + w->offw[0], W + w->offw[1], W + w->offw[2], W + w->offw[3],
                               bias(0), bias(1), bias(2), bias(3), w->Y);
        else
            hipLaunchKernelGGL(ws::k_ws_infer<false>, g, b, 0, c->stream, c->ncol, w->X,
                               W + w->offw[0], W + w->offw[1], W + w->offw[2], W + w->offw[3],
                               bias(0), bias(1), bias(2), bias(3), w->Y);
        return;
    }
    ws_forward(c, w, w->X, (int)c->ncol, false, w->Y);
}

// per-cell guesses of the current state into w->gv / w->ge (the step kernels
// form the same guesses from w->Y themselves)
void ws_guess(clm_ctx* c, WsState* w, const double* ftw, const double* fth, double dt,
              bool fused) {
    ws_infer(c, w, ftw, fth, dt, fused);
    hipLaunchKernelGGL(ws::k_guess, grid1(c->ncol * ws::kGroup), dim3(256), 0, c->stream, is_eh(c),
                       c->d_col_ptr, c->ncol, c->d_cpar, c->d_cv, c->d_ci, w->Y, w->d.Nout,
                       w->gv, w->ge);
}

void ws_launch_guess(clm_ctx* c, const double* ftw, const double* fth, double dt,
                     StepExtra* x) {
    WsState* w = c->ws;
    if (!w || !w->enabled) return;
    ws_infer(c, w, ftw, fth, dt, true);
    x->wso = w->Y;
}

int upload_raw(clm_ctx* c, const double* h, double* d, int64_t n) {
    if (!h) {
        HIPCHK(c, hipMemsetAsync(d, 0, sizeof(double) * (size_t)n, c->stream));
    } else {
        HIPCHK(c, hipMemcpyAsync(d, h, sizeof(double) * (size_t)n, hipMemcpyHostToDevice,
                                 c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int download_raw(clm_ctx* c, const double* d, double* h, int64_t n) {
    if (!h) return CLM_OK;
    HIPCHK(c, hipMemcpyAsync(h, d, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

void destroy_graph(clm_ctx* c) {
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    if (c->graph) hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->g_nsteps = -1;
}

// The static scratch of clm_step_noalloc: the step's statistics record, then
// one Newton flag word per column (group solver) or column slot (lane
// solver).  Each GPU holds its own copy, at its own address, until the
// process ends; every context on a GPU shares it.  In host code `scratch`
// names a host-side stand-in, never the device memory: kernels and copies
// take the address hipGetSymbolAddress returns for the current device.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {nullptr};

// the current device's scratch address, looked up once per device
int scratch_address(clm_ctx* c, unsigned char** out) {
    int dev = -1;
    HIPCHK(c, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return set_err(c, CLM_E_HIP, "device %d is beyond the %d the scratch table holds", dev,
                       kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(c, CLM_E_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s",
                           dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[dev];
    return CLM_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

int clm_abi_version(void) { return CLM_ABI_VERSION; }

const char* clm_last_error(const clm_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int clm_create(clm_ctx** out, const clm_config* cfg, void* hip_stream) {
    if (!out || !cfg) return set_err(nullptr, CLM_E_ARG, "clm_create: null argument");
    *out = nullptr;
    if (cfg->model != CLM_MODEL_RICHARDS && cfg->model != CLM_MODEL_ENERGY_HYDROLOGY)
        return set_err(nullptr, CLM_E_ARG, "unknown model %d", cfg->model);
    if (cfg->bc_bottom != CLM_BC_FLUX && cfg->bc_bottom != CLM_BC_FREE_DRAINAGE &&
        cfg->bc_bottom != CLM_BC_MOISTURE_STATE)
        return set_err(nullptr, CLM_E_ARG, "unknown bottom boundary %d", cfg->bc_bottom);
    if (cfg->bc_top != CLM_BC_FLUX && cfg->bc_top != CLM_BC_MOISTURE_STATE &&
        cfg->bc_top != CLM_BC_TOPMODEL && cfg->bc_top != CLM_BC_SURFACE_RUNOFF &&
        cfg->bc_top != CLM_BC_ATMOS && cfg->bc_top != CLM_BC_ATMOS_TOPMODEL)
        return set_err(nullptr, CLM_E_ARG, "unknown top boundary %d", cfg->bc_top);
    const bool atmos = cfg->bc_top == CLM_BC_ATMOS || cfg->bc_top == CLM_BC_ATMOS_TOPMODEL;
    if (atmos && cfg->model != CLM_MODEL_ENERGY_HYDROLOGY)
        return set_err(nullptr, CLM_E_ARG,
                       "AtmosDrivenFluxBC is the EnergyHydrology top boundary; RichardsModel "
                       "takes RichardsAtmosDrivenFluxBC (CLM_BC_TOPMODEL / SURFACE_RUNOFF)");
    if (cfg->model == CLM_MODEL_ENERGY_HYDROLOGY &&
        (cfg->bc_top == CLM_BC_MOISTURE_STATE || cfg->bc_bottom == CLM_BC_MOISTURE_STATE))
        return set_err(nullptr, CLM_E_ARG, "MoistureStateBC is implemented for RichardsModel only");
    if (cfg->model == CLM_MODEL_ENERGY_HYDROLOGY &&
        (cfg->bc_top == CLM_BC_TOPMODEL || cfg->bc_top == CLM_BC_SURFACE_RUNOFF))
        return set_err(nullptr, CLM_E_ARG,
                       "RichardsAtmosDrivenFluxBC runoff is implemented for RichardsModel only");
    int need = cfg->model == CLM_MODEL_RICHARDS
                   ? (cfg->bc_top == CLM_BC_TOPMODEL ? CLM_NPARAM_RRE_TOPMODEL : CLM_NPARAM_RRE)
                   : (cfg->bc_top == CLM_BC_ATMOS_TOPMODEL ? CLM_NPARAM_EH_TOPMODEL
                                                           : CLM_NPARAM_EH);
    if (cfg->nparam < need)
        return set_err(nullptr, CLM_E_ARG, "nparam %d < %d required", cfg->nparam, need);
    if (cfg->ncol <= 0 || cfg->ncol > (int64_t)1 << 30 || !cfg->col_ptr || !cfg->z_face ||
        !cfg->params)
        return set_err(nullptr, CLM_E_ARG, "bad column batch (ncol=%lld)", (long long)cfg->ncol);
    if (cfg->jac_precision != 64 && cfg->jac_precision != 32)
        return set_err(nullptr, CLM_E_ARG, "jac_precision must be 64 or 32");
    if (cfg->newton_max_iters < 1 || cfg->newton_max_iters > 0xffff)
        return set_err(nullptr, CLM_E_ARG, "newton_max_iters out of range");
    if (cfg->solver < CLM_SOLVER_AUTO || cfg->solver > CLM_SOLVER_LANE)
        return set_err(nullptr, CLM_E_ARG, "unknown solver %d", cfg->solver);
    if (cfg->col_ptr[0] != 0) return set_err(nullptr, CLM_E_ARG, "col_ptr[0] must be 0");
    for (int64_t i = 0; i < cfg->ncol; ++i) {
        int32_t n = cfg->col_ptr[i + 1] - cfg->col_ptr[i];
        if (n < 2 || n > 4096)
            return set_err(nullptr, CLM_E_ARG, "column %lld has %d layers (need 2..4096)",
                           (long long)i, n);
    }
    for (int64_t i = 0; i < cfg->ncol; ++i) {
        int64_t a = cfg->col_ptr[i] + i, n = cfg->col_ptr[i + 1] - cfg->col_ptr[i];
        for (int64_t f = 0; f < n; ++f)
            if (!(cfg->z_face[a + f + 1] > cfg->z_face[a + f]))
                return set_err(nullptr, CLM_E_ARG, "column %lld: faces not increasing",
                               (long long)i);
    }

    int32_t nmax_all = 0;
    for (int64_t i = 0; i < cfg->ncol; ++i)
        nmax_all = std::max(nmax_all, cfg->col_ptr[i + 1] - cfg->col_ptr[i]);
    int solver = cfg->solver;
    if (solver == CLM_SOLVER_AUTO) solver = nmax_all <= 128 ? CLM_SOLVER_GROUP : CLM_SOLVER_LANE;
    if (solver == CLM_SOLVER_GROUP && nmax_all > 128)
        return set_err(nullptr, CLM_E_ARG, "group solver takes columns of up to 128 layers");
    if (atmos && solver != CLM_SOLVER_GROUP)
        return set_err(nullptr, CLM_E_ARG,
                       "AtmosDrivenFluxBC is implemented for the group solver (columns of up "
                       "to 128 layers)");
    if (cfg->jacobian != CLM_JACOBIAN_PICARD && cfg->jacobian != CLM_JACOBIAN_NEWTON)
        return set_err(nullptr, CLM_E_ARG, "unknown jacobian %d", cfg->jacobian);
    if (cfg->jacobian == CLM_JACOBIAN_NEWTON &&
        (cfg->model != CLM_MODEL_ENERGY_HYDROLOGY || solver != CLM_SOLVER_GROUP || atmos))
        return set_err(nullptr, CLM_E_ARG,
                       "the Newton matrix is implemented for EnergyHydrology on the group "
                       "solver with a prescribed-flux top boundary");

    clm_ctx* c = new clm_ctx();
    c->solver = solver;
    c->model = cfg->model;
    c->bc_bottom = cfg->bc_bottom;
    c->bc_top = cfg->bc_top;
    c->jac_prec = cfg->jac_precision;
    c->jacobian = cfg->jacobian;
    c->max_iters = cfg->newton_max_iters;
    c->tol = cfg->newton_tol;
    c->ncol = cfg->ncol;
    c->nparam = need;
    c->col_ptr.assign(cfg->col_ptr, cfg->col_ptr + cfg->ncol + 1);
    c->ncell = c->col_ptr.back();
    c->h_zf.assign(cfg->z_face, cfg->z_face + c->ncell + c->ncol);
    c->h_par.assign(cfg->params, cfg->params + (size_t)need * c->ncol);

    // ---- layout: sort by layer count (descending), packs of 64 ----
    std::vector<int32_t> order((size_t)c->ncol);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return (c->col_ptr[a + 1] - c->col_ptr[a]) > (c->col_ptr[b + 1] - c->col_ptr[b]);
    });
    c->npack = (int32_t)((c->ncol + 63) / 64);
    c->nslot = (int64_t)c->npack * 64;
    c->slot_col.assign((size_t)c->nslot, -1);
    c->slot_nl.assign((size_t)c->nslot, 0);
    for (int64_t k = 0; k < c->ncol; ++k) {
        c->slot_col[k] = order[k];
        c->slot_nl[k] = c->col_ptr[order[k] + 1] - c->col_ptr[order[k]];
    }
    c->cell_off.assign((size_t)c->npack + 1, 0);
    c->face_off.assign((size_t)c->npack + 1, 0);
    for (int32_t p = 0; p < c->npack; ++p) {
        int32_t nmax = 0;
        for (int l = 0; l < 64; ++l) nmax = std::max(nmax, c->slot_nl[(int64_t)p * 64 + l]);
        c->cell_off[p + 1] = c->cell_off[p] + 64LL * nmax;
        c->face_off[p + 1] = c->face_off[p] + 64LL * (nmax + 1);
    }
    c->npcell = c->cell_off.back();
    c->npface = c->face_off.back();

    if (hip_stream) {
        c->stream = (hipStream_t)hip_stream;
        c->own_stream = false;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return set_err(nullptr, CLM_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
        }
        c->own_stream = true;
    }
    int r = 0;
    auto fail = [&](int code) {
        g_create_error = c->err;
        clm_destroy(c);
        return code;
    };
#define ALLOC(p, n)                          \
    if ((r = dalloc(c, &(p), (n)))) return fail(r);
    ALLOC(c->d_cell_off, c->npack + 1);
    ALLOC(c->d_face_off, c->npack + 1);
    ALLOC(c->d_nl, c->nslot);
    ALLOC(c->d_slot_col, c->nslot);
    ALLOC(c->d_par, (int64_t)NPARAM_MAX * c->nslot);
    ALLOC(c->d_zf, c->npface);
    ALLOC(c->d_xv, c->npcell);
    ALLOC(c->d_iw, c->nslot);
    ALLOC(c->d_ftw, c->nslot);
    ALLOC(c->d_fth, c->nslot);
    ALLOC(c->d_fbw, c->nslot);
    ALLOC(c->d_fbh, c->nslot);
    ALLOC(c->d_K, c->npcell);
    ALLOC(c->d_psi, c->npcell);
    ALLOC(c->d_tw, c->nslot);
    if (c->bc_top == CLM_BC_TOPMODEL || c->bc_top == CLM_BC_ATMOS_TOPMODEL) {
        ALLOC(c->d_sink, c->npcell);
        ALLOC(c->d_rss, c->nslot);
    }
    if (c->bc_top == CLM_BC_ATMOS_TOPMODEL) {
        ALLOC(c->d_sinke, c->npcell);
        ALLOC(c->d_ress, c->nslot);
    }
    if (is_atmos(c)) {
        ALLOC(c->d_vice, c->nslot);
        for (int k = 0; k < 7; ++k) {
            ALLOC(c->d_sfc[k], c->nslot);
            ALLOC(c->d_csfc[k], c->ncol);
        }
    }
    ALLOC(c->d_th, c->nslot);
    ALLOC(c->d_bw, c->nslot);
    ALLOC(c->d_bh, c->nslot);
    ALLOC(c->d_bad, c->nslot);
    ALLOC(c->d_U0v, c->npcell);
    ALLOC(c->d_stats, 1);
    for (int i = 0; i < 5; ++i) {
        bool cell = i < 3;
        if (!is_eh(c) && (i == 1 || i == 2 || i == 4)) continue;
        ALLOC(c->d_wy[i], cell ? c->npcell : c->nslot);
        ALLOC(c->d_wd[i], cell ? c->npcell : c->nslot);
    }
    if (is_eh(c)) {
        ALLOC(c->d_xi, c->npcell);
        ALLOC(c->d_xe, c->npcell);
        ALLOC(c->d_ie, c->nslot);
        ALLOC(c->d_thl, c->npcell);
        ALLOC(c->d_T, c->npcell);
        ALLOC(c->d_kap, c->npcell);
        ALLOC(c->d_U0e, c->npcell);
        ALLOC(c->d_Kl, c->npcell);
        ALLOC(c->d_kl, c->npcell);
        ALLOC(c->d_tl, c->npcell);
        ALLOC(c->d_Tp, c->npcell);
        ALLOC(c->d_Pp, c->npcell);
        ALLOC(c->d_Re, c->npcell);
    }
    if (is_group(c)) {
        ALLOC(c->d_col_ptr, c->ncol + 1);
        ALLOC(c->d_cpar, (int64_t)NPARAM_MAX * c->ncol);
        ALLOC(c->d_czf, c->ncell + c->ncol);
        ALLOC(c->d_cv, c->ncell);
        ALLOC(c->d_ciw, c->ncol);
        ALLOC(c->d_cftw, c->ncol);
        ALLOC(c->d_cfth, c->ncol);
        ALLOC(c->d_cfbw, c->ncol);
        ALLOC(c->d_cfbh, c->ncol);
        if (is_eh(c)) {
            ALLOC(c->d_ci, c->ncell);
            ALLOC(c->d_ce, c->ncell);
            ALLOC(c->d_cie, c->ncol);
        }
        // depth buckets and their column lists
        std::vector<std::vector<int32_t>> cols(kNBucket);
        for (int b = 0; b < kNBucket; ++b) c->bucket_nmax[b] = 0;
        for (int64_t i = 0; i < c->ncol; ++i) {
            int Ni = c->col_ptr[i + 1] - c->col_ptr[i];
            int b = bucket_of(Ni);
            cols[b].push_back((int32_t)i);
            c->bucket_nmax[b] = std::max(c->bucket_nmax[b], Ni);
        }
        for (int b = 0; b < kNBucket; ++b) {
            c->bucket_G[b] = kDefaultShape[b].G;
            c->bucket_R[b] = kDefaultShape[b].R;
            clm_ctx::Bin bn{c->bucket_G[b], c->bucket_R[b], (int32_t)cols[b].size(), nullptr};
            c->bin_off[b + 1] = c->bin_off[b] + bn.n;
            c->bins.push_back(bn);
        }
        ALLOC(c->d_bincols, c->ncol);
        ALLOC(c->d_bintmp, c->ncol);
        ALLOC(c->d_rghist, kNBucket * kRgBuckets);
        ALLOC(c->d_rgfill, kNBucket * kRgBuckets);
        std::vector<int32_t> all;
        all.reserve((size_t)c->ncol);
        for (int b = 0; b < kNBucket; ++b) {
            c->bins[b].d_cols = c->d_bincols + c->bin_off[b];
            all.insert(all.end(), cols[b].begin(), cols[b].end());
        }
        if (hipMemcpyAsync(c->d_bincols, all.data(), 4 * all.size(), hipMemcpyHostToDevice,
                           c->stream) != hipSuccess ||
            hipMemsetAsync(c->d_rghist, 0, 4 * kNBucket * kRgBuckets, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            set_err(c, CLM_E_HIP, "bin upload failed");
            return fail(CLM_E_HIP);
        }
    }
#undef ALLOC
    {
        size_t wbytes = jt_size(c) * (size_t)c->npcell * (is_eh(c) ? 9 : 3);
        size_t sbytes = jt_size(c) * (size_t)c->npcell * 6;
        hipError_t e1 = hipMalloc(&c->d_W, wbytes);
        hipError_t e2 = hipMalloc(&c->d_scr, sbytes);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            set_err(c, CLM_E_HIP, "hipMalloc of Jacobian/scratch failed");
            return fail(CLM_E_HIP);
        }
        hipMemsetAsync(c->d_W, 0, wbytes, c->stream);
        hipMemsetAsync(c->d_scr, 0, sbytes, c->stream);
    }
    // ---- upload layout, parameters, faces ----
    {
        std::vector<double> par((size_t)NPARAM_MAX * c->nslot, 0.0);
        for (int p = 0; p < need; ++p)
            for (int64_t s = 0; s < c->nslot; ++s) {
                int32_t col = c->slot_col[s];
                par[(size_t)p * c->nslot + s] =
                    col >= 0 ? cfg->params[(int64_t)p * c->ncol + col] : 1.0;
            }
        std::vector<double> zf((size_t)c->npface, 0.0);
        for (int64_t s = 0; s < c->nslot; ++s) {
            int32_t col = c->slot_col[s];
            if (col < 0) continue;
            int64_t base = c->face_off[s >> 6] + (s & 63);
            const double* src = cfg->z_face + c->col_ptr[col] + col;
            for (int32_t f = 0; f <= c->slot_nl[s]; ++f) zf[base + 64 * f] = src[f];
        }
        hipError_t e = hipSuccess;
        auto up = [&](void* d, const void* h, size_t bytes) {
            if (e == hipSuccess) e = hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream);
        };
        up(c->d_cell_off, c->cell_off.data(), 8 * c->cell_off.size());
        up(c->d_face_off, c->face_off.data(), 8 * c->face_off.size());
        up(c->d_nl, c->slot_nl.data(), 4 * c->slot_nl.size());
        up(c->d_slot_col, c->slot_col.data(), 4 * c->slot_col.size());
        up(c->d_par, par.data(), 8 * par.size());
        up(c->d_zf, zf.data(), 8 * zf.size());
        if (is_group(c)) {
            std::vector<double> cpar((size_t)NPARAM_MAX * c->ncol, 1.0);
            for (int p = 0; p < need; ++p)
                std::memcpy(cpar.data() + (size_t)p * c->ncol, cfg->params + (size_t)p * c->ncol,
                            8 * (size_t)c->ncol);
            std::vector<double> czf(cfg->z_face, cfg->z_face + c->ncell + c->ncol);
            up(c->d_col_ptr, c->col_ptr.data(), 4 * c->col_ptr.size());
            up(c->d_cpar, cpar.data(), 8 * cpar.size());
            up(c->d_czf, czf.data(), 8 * czf.size());
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            set_err(c, CLM_E_HIP, "upload failed: %s", hipGetErrorString(e));
            return fail(CLM_E_HIP);
        }
    }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&c->bin_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->rg_start, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->rg_done, hipEventDisableTiming) != hipSuccess) {
        set_err(c, CLM_E_HIP, "hipEventCreate failed");
        return fail(CLM_E_HIP);
    }
    for (int i = 0; i < clm_ctx::kBinStreams; ++i) {
        if (hipStreamCreateWithFlags(&c->bin_stream[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->bin_join[i], hipEventDisableTiming) != hipSuccess) {
            set_err(c, CLM_E_HIP, "bin streams: create failed");
            return fail(CLM_E_HIP);
        }
    }
    {
        int r = scratch_address(c, &c->d_scratch);
        if (r) return fail(r);
    }
    *out = c;
    return CLM_OK;
}

int clm_destroy(clm_ctx* c) {
    if (!c) return CLM_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    destroy_graph(c);
    dfree(c->d_cell_off);
    dfree(c->d_face_off);
    dfree(c->d_nl);
    dfree(c->d_slot_col);
    dfree(c->d_par);
    dfree(c->d_zf);
    dfree(c->d_xv);
    dfree(c->d_xi);
    dfree(c->d_xe);
    dfree(c->d_iw);
    dfree(c->d_ie);
    dfree(c->d_ftw);
    dfree(c->d_fth);
    dfree(c->d_fbw);
    dfree(c->d_fbh);
    dfree(c->d_tab_w);
    dfree(c->d_tab_h);
    dfree(c->d_K);
    dfree(c->d_psi);
    dfree(c->d_thl);
    dfree(c->d_T);
    dfree(c->d_kap);
    dfree(c->d_tw);
    dfree(c->d_sink);
    dfree(c->d_rss);
    dfree(c->d_sinke);
    dfree(c->d_ress);
    dfree(c->d_vice);
    for (int k = 0; k < 7; ++k) {
        dfree(c->d_sfc[k]);
        dfree(c->d_csfc[k]);
        dfree(c->d_tab_sfc[k]);
    }
    dfree(c->d_th);
    dfree(c->d_bw);
    dfree(c->d_bh);
    for (int i = 0; i < 5; ++i) {
        dfree(c->d_wy[i]);
        dfree(c->d_wd[i]);
    }
    dfree(c->d_W);
    dfree(c->d_bad);
    dfree(c->d_U0v);
    dfree(c->d_U0e);
    dfree(c->d_Kl);
    dfree(c->d_kl);
    dfree(c->d_tl);
    dfree(c->d_Tp);
    dfree(c->d_Pp);
    dfree(c->d_Re);
    dfree(c->d_scr);
    dfree(c->d_iters);
    dfree(c->d_stats);
    dfree(c->d_col_ptr);
    dfree(c->d_cpar);
    dfree(c->d_czf);
    dfree(c->d_cv);
    dfree(c->d_ci);
    dfree(c->d_ce);
    dfree(c->d_ciw);
    dfree(c->d_cie);
    dfree(c->d_cftw);
    dfree(c->d_cfth);
    dfree(c->d_cfbw);
    dfree(c->d_cfbh);
    dfree(c->d_bincols);
    dfree(c->d_bintmp);
    dfree(c->d_rghist);
    dfree(c->d_rgfill);
    ws_free(c->ws);
    c->ws = nullptr;
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->bin_fork) hipEventDestroy(c->bin_fork);
    if (c->rg_start) hipEventDestroy(c->rg_start);
    if (c->rg_done) hipEventDestroy(c->rg_done);
    for (int i = 0; i < clm_ctx::kBinStreams; ++i) {
        if (c->bin_stream[i]) {
            hipStreamSynchronize(c->bin_stream[i]);
            hipStreamDestroy(c->bin_stream[i]);
        }
        if (c->bin_join[i]) hipEventDestroy(c->bin_join[i]);
    }
    if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
    delete c;
    return CLM_OK;
}

int clm_set_state(clm_ctx* c, const clm_state* Y) {
    if (!c) return CLM_E_ARG;
    if (is_group(c)) {
        if (!Y || !Y->vartheta_l) return set_err(c, CLM_E_ARG, "state: vartheta_l required");
        int r;
        if ((r = upload_raw(c, Y->vartheta_l, c->d_cv, c->ncell))) return r;
        if ((r = upload_raw(c, Y->intF_water, c->d_ciw, c->ncol))) return r;
        if (is_eh(c)) {
            if (!Y->theta_i || !Y->rho_e_int)
                return set_err(c, CLM_E_ARG, "EnergyHydrology state needs theta_i and rho_e_int");
            if ((r = upload_raw(c, Y->theta_i, c->d_ci, c->ncell))) return r;
            if ((r = upload_raw(c, Y->rho_e_int, c->d_ce, c->ncell))) return r;
            if ((r = upload_raw(c, Y->intF_energy, c->d_cie, c->ncol))) return r;
        }
        return CLM_OK;
    }
    double* dst[5] = {c->d_xv, c->d_xi, c->d_xe, c->d_iw, c->d_ie};
    return upload_state(c, Y, dst);
}

int clm_get_state(clm_ctx* c, clm_state* Y) {
    if (!c || !Y) return CLM_E_ARG;
    if (is_group(c)) {
        int r;
        if ((r = download_raw(c, c->d_cv, Y->vartheta_l, c->ncell))) return r;
        if ((r = download_raw(c, c->d_ciw, Y->intF_water, c->ncol))) return r;
        if (is_eh(c)) {
            if ((r = download_raw(c, c->d_ci, Y->theta_i, c->ncell))) return r;
            if ((r = download_raw(c, c->d_ce, Y->rho_e_int, c->ncell))) return r;
            if ((r = download_raw(c, c->d_cie, Y->intF_energy, c->ncol))) return r;
        }
        return CLM_OK;
    }
    double* src[5] = {c->d_xv, c->d_xi, c->d_xe, c->d_iw, c->d_ie};
    return download_state(c, src, Y);
}

int clm_set_forcing(clm_ctx* c, const double* tw, const double* th, const double* bw,
                    const double* bh) {
    if (!c) return CLM_E_ARG;
    int r;
    if ((r = upload_cols(c, tw, c->d_ftw))) return r;
    if ((r = upload_cols(c, th, c->d_fth))) return r;
    if ((r = upload_cols(c, bw, c->d_fbw))) return r;
    if ((r = upload_cols(c, bh, c->d_fbh))) return r;
    if (is_group(c)) {
        if ((r = upload_raw(c, tw, c->d_cftw, c->ncol))) return r;
        if ((r = upload_raw(c, th, c->d_cfth, c->ncol))) return r;
        if ((r = upload_raw(c, bw, c->d_cfbw, c->ncol))) return r;
        if ((r = upload_raw(c, bh, c->d_cfbh, c->ncol))) return r;
    }
    return CLM_OK;
}

int clm_set_forcing_table(clm_ctx* c, int32_t nslots, const double* tw, const double* th) {
    if (!c) return CLM_E_ARG;
    if (nslots < 1 || !tw) return set_err(c, CLM_E_ARG, "forcing table needs nslots >= 1");
    if (nslots != c->tab_slots) {
        // the captured run graph points at the old table: drop it, and leave
        // the context without a table until both allocations succeed (clm_run
        // then reports CLM_E_STATE instead of replaying freed memory)
        destroy_graph(c);
        c->tab_slots = 0;
        dfree(c->d_tab_w);
        dfree(c->d_tab_h);
        int r;
        if ((r = dalloc(c, &c->d_tab_w, (int64_t)nslots * ncolvec(c)))) return r;
        if ((r = dalloc(c, &c->d_tab_h, (int64_t)nslots * ncolvec(c)))) return r;
        c->tab_slots = nslots;
    }
    if (is_group(c)) {
        int r;
        if ((r = upload_raw(c, tw, c->d_tab_w, (int64_t)nslots * c->ncol))) return r;
        return upload_raw(c, th, c->d_tab_h, (int64_t)nslots * c->ncol);
    }
    for (int32_t k = 0; k < nslots; ++k) {
        int r;
        if ((r = upload_cols(c, tw + (int64_t)k * c->ncol, c->d_tab_w + (int64_t)k * c->nslot)))
            return r;
        if ((r = upload_cols(c, th ? th + (int64_t)k * c->ncol : nullptr,
                             c->d_tab_h + (int64_t)k * c->nslot)))
            return r;
    }
    return CLM_OK;
}

static const double* sfc_field(const clm_surface_fluxes* s, int k) {
    const double* f[7] = {s->P_liq, s->T_air, s->lhf, s->shf, s->vapor_flux_liq,
                          s->vapor_flux_ice, s->R_n};
    return f[k];
}

int clm_set_surface_fluxes(clm_ctx* c, const clm_surface_fluxes* s) {
    if (!c || !s) return CLM_E_ARG;
    if (!is_atmos(c)) return set_err(c, CLM_E_STATE, "the top boundary is not AtmosDrivenFluxBC");
    int r;
    for (int k = 0; k < 7; ++k) {
        if ((r = upload_cols(c, sfc_field(s, k), c->d_sfc[k]))) return r;
        if ((r = upload_raw(c, sfc_field(s, k), c->d_csfc[k], c->ncol))) return r;
    }
    return CLM_OK;
}

int clm_set_surface_flux_table(clm_ctx* c, int32_t nslots, const clm_surface_fluxes* s) {
    if (!c || !s) return CLM_E_ARG;
    if (!is_atmos(c)) return set_err(c, CLM_E_STATE, "the top boundary is not AtmosDrivenFluxBC");
    if (nslots < 1) return set_err(c, CLM_E_ARG, "surface flux table needs nslots >= 1");
    int r;
    if (nslots != c->sfc_slots) {
        destroy_graph(c);
        c->sfc_slots = 0;
        for (int k = 0; k < 7; ++k) dfree(c->d_tab_sfc[k]);
        for (int k = 0; k < 7; ++k)
            if ((r = dalloc(c, &c->d_tab_sfc[k], (int64_t)nslots * c->ncol))) return r;
        c->sfc_slots = nslots;
    }
    for (int k = 0; k < 7; ++k)
        if ((r = upload_raw(c, sfc_field(s, k), c->d_tab_sfc[k], (int64_t)nslots * c->ncol)))
            return r;
    return CLM_OK;
}

// ---- fine-grained hooks ------------------------------------------------------

int clm_update_cache(clm_ctx* c, const clm_state* Y, double t) {
    (void)t;
    if (!c) return CLM_E_ARG;
    int r = upload_state(c, Y, c->d_wy);
    if (r) return r;
    Layout L = layout(c);
    if (!is_eh(c)) {
        hipLaunchKernelGGL(k_rre_update_cache, grid_slots(c), dim3(256), 0, c->stream, L,
                           c->d_par, c->d_zf, c->d_wy[0], c->d_K, c->d_psi, c->d_ftw, c->d_fbw,
                           c->d_tw, c->d_bw, rre_bc(c), 1, c->d_sink, c->d_rss);
    } else {
        Surface sf{};
        if (is_atmos(c)) sf = surface_of(c->d_sfc, 0);
        hipLaunchKernelGGL(k_eh_update_cache, grid_slots(c), dim3(256), 0, c->stream, L,
                           c->d_par, c->d_zf, c->d_wy[0], c->d_wy[1], c->d_wy[2], c->d_thl,
                           c->d_kap, c->d_T, c->d_K, c->d_psi, c->d_ftw, c->d_fth, c->d_fbw,
                           c->d_fbh, c->d_tw, c->d_th, c->d_bw, c->d_bh, c->bc_bottom,
                           c->bc_top, sf, c->d_sink, c->d_sinke, c->d_rss, c->d_ress,
                           c->d_vice);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_cache = true;
    return CLM_OK;
}

int clm_update_implicit_cache(clm_ctx* c, const clm_state* Y, double t) {
    (void)t;
    if (!c) return CLM_E_ARG;
    if (!c->have_cache) return set_err(c, CLM_E_STATE, "clm_update_cache must run first");
    int r = upload_state(c, Y, c->d_wy);
    if (r) return r;
    Layout L = layout(c);
    if (!is_eh(c)) {
        // update_aux (models.jl:207-210); boundary fluxes stay lagged unless
        // the top is a MoistureStateBC (rre.jl:460-468)
        hipLaunchKernelGGL(k_rre_update_cache, grid_slots(c), dim3(256), 0, c->stream, L,
                           c->d_par, c->d_zf, c->d_wy[0], c->d_K, c->d_psi, c->d_ftw, c->d_fbw,
                           c->d_tw, c->d_bw, rre_bc(c),
                           c->bc_top == CLM_BC_MOISTURE_STATE ? 1 : 0, c->d_sink, c->d_rss);
    } else {
        hipLaunchKernelGGL(k_eh_imp_cache, grid_slots(c), dim3(256), 0, c->stream, L, c->d_par,
                           c->d_wy[0], c->d_wy[1], c->d_wy[2], c->d_T, c->d_psi);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int clm_exp_tendency(clm_ctx* c, const clm_state* Y, double t, clm_state* dY) {
    if (!c || !dY) return CLM_E_ARG;
    // make_exp_tendency = update_cache! + compute_exp_tendency! (models.jl:279-287)
    int r = clm_update_cache(c, Y, t);
    if (r) return r;
    Layout L = layout(c);
    if (!is_eh(c)) {
        HIPCHK(c, hipMemsetAsync(c->d_wd[0], 0, 8 * c->npcell, c->stream));
        HIPCHK(c, hipMemsetAsync(c->d_wd[3], 0, 8 * c->nslot, c->stream));
    } else {
        HIPCHK(c, hipMemsetAsync(c->d_bad, 0, 4 * c->nslot, c->stream));
        hipLaunchKernelGGL(k_eh_exp_tendency, grid_slots(c), dim3(256), 0, c->stream, L,
                           c->d_par, c->d_zf, c->d_wy[1], c->d_thl, c->d_T, c->d_kap,
                           c->d_wd[0], c->d_wd[1], c->d_wd[2], c->d_wd[3], c->d_wd[4],
                           c->d_bad, is_atmos(c) ? c->d_vice : nullptr);
        HIPCHK(c, hipGetLastError());
    }
    r = download_state(c, c->d_wd, dY);
    if (r) return r;
    if (is_eh(c)) {
        std::vector<int32_t> bad((size_t)c->nslot);
        HIPCHK(c, hipMemcpyAsync(bad.data(), c->d_bad, 4 * bad.size(), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (auto b : bad)
            if (b)
                return set_err(c, CLM_E_DOMAIN,
                               "inverse_matric_potential: matric potential is positive");
    }
    return CLM_OK;
}

int clm_imp_tendency(clm_ctx* c, const clm_state* Y, double t, clm_state* dY) {
    (void)t;
    (void)Y;
    if (!c || !dY) return CLM_E_ARG;
    if (!c->have_cache) return set_err(c, CLM_E_STATE, "clm_update_cache must run first");
    Layout L = layout(c);
    if (!is_eh(c)) {
        hipLaunchKernelGGL(k_rre_imp_tendency, grid_slots(c), dim3(256), 0, c->stream, L,
                           c->d_zf, c->d_K, c->d_psi, c->d_tw, c->d_bw, c->d_wd[0],
                           c->d_wd[3], c->d_sink, c->d_rss);
    } else {
        hipLaunchKernelGGL(k_eh_imp_tendency, grid_slots(c), dim3(256), 0, c->stream, L,
                           c->d_zf, c->d_K, c->d_psi, c->d_T, c->d_kap, c->d_tw, c->d_th,
                           c->d_bw, c->d_bh, c->d_wd[0], c->d_wd[1], c->d_wd[2], c->d_wd[3],
                           c->d_wd[4], c->d_sinke ? c->d_sink : nullptr, c->d_sinke, c->d_rss,
                           c->d_ress);
    }
    HIPCHK(c, hipGetLastError());
    return download_state(c, c->d_wd, dY);
}

int clm_jacobian(clm_ctx* c, const clm_state* Y, double dtgamma, double t) {
    (void)t;
    if (!c) return CLM_E_ARG;
    if (!c->have_cache) return set_err(c, CLM_E_STATE, "clm_update_cache must run first");
    int r = upload_state(c, Y, c->d_wy);
    if (r) return r;
    Layout L = layout(c);
    int64_t n = c->npcell;
    if (!is_eh(c)) {
        if (c->jac_prec == 32) {
            float* W = (float*)c->d_W;
            hipLaunchKernelGGL(k_rre_jacobian<float>, grid_slots(c), dim3(256), 0, c->stream, L,
                               c->d_par, c->d_zf, c->d_wy[0], c->d_K, dtgamma, W, W + n,
                               W + 2 * n, rre_bc(c));
        } else {
            double* W = (double*)c->d_W;
            hipLaunchKernelGGL(k_rre_jacobian<double>, grid_slots(c), dim3(256), 0, c->stream,
                               L, c->d_par, c->d_zf, c->d_wy[0], c->d_K, dtgamma, W, W + n,
                               W + 2 * n, rre_bc(c));
        }
    } else {
        if (c->jac_prec == 32) {
            float* W = (float*)c->d_W;
            hipLaunchKernelGGL(k_eh_jacobian<float>, grid_slots(c), dim3(256), 0, c->stream, L,
                               c->d_par, c->d_zf, c->d_wy[0], c->d_wy[1], c->d_K, c->d_T,
                               c->d_kap, c->d_thl, dtgamma, W, W + 3 * n, W + 6 * n, n);
        } else {
            double* W = (double*)c->d_W;
            hipLaunchKernelGGL(k_eh_jacobian<double>, grid_slots(c), dim3(256), 0, c->stream,
                               L, c->d_par, c->d_zf, c->d_wy[0], c->d_wy[1], c->d_K, c->d_T,
                               c->d_kap, c->d_thl, dtgamma, W, W + 3 * n, W + 6 * n, n);
        }
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_jac = true;
    return CLM_OK;
}

int clm_solve(clm_ctx* c, clm_state* dx, const clm_state* R) {
    if (!c || !dx || !R) return CLM_E_ARG;
    if (!c->have_jac) return set_err(c, CLM_E_STATE, "clm_jacobian must run first");
    int r = upload_state(c, R, c->d_wy);
    if (r) return r;
    Layout L = layout(c);
    int64_t n = c->npcell;
    if (!is_eh(c)) {
        if (c->jac_prec == 32) {
            float* W = (float*)c->d_W;
            float* s = (float*)c->d_scr;
            hipLaunchKernelGGL(k_rre_solve<float>, grid_slots(c), dim3(256), 0, c->stream, L, W,
                               W + n, W + 2 * n, c->d_wy[0], c->d_wy[3], c->d_wd[0],
                               c->d_wd[3], s, s + n);
        } else {
            double* W = (double*)c->d_W;
            double* s = (double*)c->d_scr;
            hipLaunchKernelGGL(k_rre_solve<double>, grid_slots(c), dim3(256), 0, c->stream, L,
                               W, W + n, W + 2 * n, c->d_wy[0], c->d_wy[3], c->d_wd[0],
                               c->d_wd[3], s, s + n);
        }
    } else {
        if (c->jac_prec == 32) {
            float* W = (float*)c->d_W;
            float* s = (float*)c->d_scr;
            hipLaunchKernelGGL(k_eh_solve<float>, grid_slots(c), dim3(256), 0, c->stream, L, W,
                               W + 3 * n, W + 6 * n, n, c->d_wy[0], c->d_wy[1], c->d_wy[2],
                               c->d_wy[3], c->d_wy[4], c->d_wd[0], c->d_wd[1], c->d_wd[2],
                               c->d_wd[3], c->d_wd[4], s, s + n, c->d_Re);
        } else {
            double* W = (double*)c->d_W;
            double* s = (double*)c->d_scr;
            hipLaunchKernelGGL(k_eh_solve<double>, grid_slots(c), dim3(256), 0, c->stream, L,
                               W, W + 3 * n, W + 6 * n, n, c->d_wy[0], c->d_wy[1], c->d_wy[2],
                               c->d_wy[3], c->d_wy[4], c->d_wd[0], c->d_wd[1], c->d_wd[2],
                               c->d_wd[3], c->d_wd[4], s, s + n, c->d_Re);
        }
    }
    HIPCHK(c, hipGetLastError());
    return download_state(c, c->d_wd, dx);
}

int clm_get_aux(clm_ctx* c, clm_aux* p) {
    if (!c || !p) return CLM_E_ARG;
    if (!c->have_cache) return set_err(c, CLM_E_STATE, "clm_update_cache must run first");
    int r;
    if ((r = download_cells(c, c->d_K, p->K))) return r;
    if ((r = download_cells(c, c->d_psi, p->psi))) return r;
    if ((r = download_cols(c, c->d_tw, p->top_water))) return r;
    if ((r = download_cols(c, c->d_bw, p->bottom_water))) return r;
    if (c->d_sink) {
        if ((r = download_cells(c, c->d_sink, p->sink))) return r;
        if ((r = download_cols(c, c->d_rss, p->r_ss))) return r;
    }
    if (c->d_sinke) {
        if ((r = download_cells(c, c->d_sinke, p->sink_energy))) return r;
        if ((r = download_cols(c, c->d_ress, p->r_ess))) return r;
    }
    if (is_eh(c)) {
        if ((r = download_cells(c, c->d_thl, p->theta_l))) return r;
        if ((r = download_cells(c, c->d_T, p->T))) return r;
        if ((r = download_cells(c, c->d_kap, p->kappa))) return r;
        if ((r = download_cols(c, c->d_th, p->top_heat))) return r;
        if ((r = download_cols(c, c->d_bh, p->bottom_heat))) return r;
    }
    return CLM_OK;
}

int clm_get_jacobian(clm_ctx* c, double* W_ww, double* W_ee, double* W_ew) {
    if (!c) return CLM_E_ARG;
    if (!c->have_jac) return set_err(c, CLM_E_STATE, "clm_jacobian must run first");
    int nplanes = is_eh(c) ? 9 : 3;
    size_t n = (size_t)c->npcell;
    std::vector<double> w(n * nplanes);
    if (c->jac_prec == 32) {
        std::vector<float> f(n * nplanes);
        HIPCHK(c, hipMemcpyAsync(f.data(), c->d_W, 4 * f.size(), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < f.size(); ++i) w[i] = f[i];
    } else {
        HIPCHK(c, hipMemcpyAsync(w.data(), c->d_W, 8 * w.size(), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    double* outs[3] = {W_ww, W_ee, W_ew};
    for (int b = 0; b < nplanes / 3; ++b) {
        double* o = outs[b];
        if (!o) continue;
        for (int64_t s = 0; s < c->nslot; ++s) {
            int32_t col = c->slot_col[s];
            if (col < 0) continue;
            int64_t base = c->cell_off[s >> 6] + (s & 63);
            for (int32_t j = 0; j < c->slot_nl[s]; ++j)
                for (int k = 0; k < 3; ++k)
                    o[3 * (int64_t)(c->col_ptr[col] + j) + k] =
                        w[(size_t)(3 * b + k) * n + base + 64 * j];
        }
    }
    return CLM_OK;
}

// ---- coarse path ---------------------------------------------------------------

int clm_step(clm_ctx* c, double t, double dt, clm_step_stats* stats) {
    (void)t;
    if (!c) return CLM_E_ARG;
    if (!(dt > 0.0)) return set_err(c, CLM_E_ARG, "dt must be > 0");
    int r = ensure_iters(c, 1);
    if (r) return r;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (is_group(c))
        launch_step(c, c->d_cftw, c->d_cfth, dt, c->d_iters);
    else
        launch_step(c, c->d_ftw, c->d_fth, dt, c->d_iters);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->have_cache = false;  // the fused step does not maintain the fine-hook cache
    c->have_jac = false;
    if (stats) {
        HIPCHK(c, hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_ms = ms;
        c->last_launches = 1;
        if ((r = collect_stats(c, 1, stats))) return r;
        stats->ms = ms;
        if (stats->n_domain_error)
            return set_err(c, CLM_E_DOMAIN,
                           "inverse_matric_potential: matric potential is positive");
    }
    return CLM_OK;
}

// clm_step with its Newton flags and statistics in the GPU's static scratch:
// the same kernels, and no allocation on any call (clm_step allocates its
// flag table on first use).  Calls that overlap on one GPU race on the
// scratch.  The step kernels write every column's flag word; the statistics
// record is cleared by a kernel before the reduction.
int clm_step_noalloc(clm_ctx* c, double t, double dt, clm_step_stats* stats) {
    (void)t;
    if (!c) return CLM_E_ARG;
    if (!(dt > 0.0)) return set_err(c, CLM_E_ARG, "dt must be > 0");
    if (!c->d_scratch) return set_err(c, CLM_E_STATE, "clm_step_noalloc: no scratch address");
    if (ncolvec(c) > kScratchFlags)
        return set_err(c, CLM_E_ARG,
                       "clm_step_noalloc: %lld column slots exceed the scratch's %lld; use "
                       "clm_step",
                       (long long)ncolvec(c), (long long)kScratchFlags);
    DevStats* dst = (DevStats*)c->d_scratch;
    int32_t* iters = (int32_t*)(c->d_scratch + kScratchHead);
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (is_group(c))
        launch_step(c, c->d_cftw, c->d_cfth, dt, iters);
    else
        launch_step(c, c->d_ftw, c->d_fth, dt, iters);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->have_cache = false;
    c->have_jac = false;
    if (stats) {
        HIPCHK(c, hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_ms = ms;
        c->last_launches = 1;
        int r = collect_stats(c, 1, stats, iters, dst);
        if (r) return r;
        stats->ms = ms;
        if (stats->n_domain_error)
            return set_err(c, CLM_E_DOMAIN,
                           "inverse_matric_potential: matric potential is positive");
    }
    return CLM_OK;
}

int clm_run(clm_ctx* c, double t0, double dt, int32_t nsteps, int32_t steps_per_slot,
            clm_step_stats* stats, int32_t* iters_out) {
    (void)t0;
    if (!c) return CLM_E_ARG;
    if (nsteps < 1 || steps_per_slot < 1 || !(dt > 0.0))
        return set_err(c, CLM_E_ARG, "clm_run: bad nsteps/steps_per_slot/dt");
    const bool atm = is_atmos(c);
    const int32_t nslots = atm ? c->sfc_slots : c->tab_slots;
    if (nslots < 1)
        return set_err(c, CLM_E_STATE, atm ? "clm_set_surface_flux_table must run first"
                                           : "clm_set_forcing_table must run first");
    int r = ensure_iters(c, nsteps);
    if (r) return r;
    if (!c->gexec || c->g_nsteps != nsteps || c->g_sps != steps_per_slot ||
        c->g_tab != nslots || c->g_dt != dt) {
        destroy_graph(c);
        c->rg_pending = false;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int32_t s = 0; s < nsteps; ++s) {
            int64_t slot = (s / steps_per_slot) % nslots;
            int64_t nv = ncolvec(c);
            if (atm) {
                Surface sf = surface_of(c->d_tab_sfc, slot * c->ncol);
                launch_step(c, nullptr, nullptr, dt, c->d_iters + (int64_t)s * nv, nullptr,
                            false, &sf, s + 1 < nsteps);
            } else {
                launch_step(c, c->d_tab_w + slot * nv, c->d_tab_h + slot * nv, dt,
                            c->d_iters + (int64_t)s * nv, nullptr, true, nullptr,
                            s + 1 < nsteps);
            }
        }
        c->rg_pending = false;  // the last step regroups on the ctx stream
        hipError_t e = hipStreamEndCapture(c->stream, &c->graph);
        if (e != hipSuccess) {
            c->graph = nullptr;
            return set_err(c, CLM_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
        }
        HIPCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        c->g_nsteps = nsteps;
        c->g_sps = steps_per_slot;
        c->g_tab = nslots;
        c->g_dt = dt;
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->have_cache = false;
    c->have_jac = false;
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = nsteps;
    if (stats) {
        if ((r = collect_stats(c, nsteps, stats))) return r;
        stats->ms = ms;
    }
    if (iters_out && is_group(c)) {
        HIPCHK(c, hipMemcpyAsync(iters_out, c->d_iters, 4 * (size_t)nsteps * c->ncol,
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else if (iters_out) {
        std::vector<int32_t> h((size_t)nsteps * c->nslot);
        HIPCHK(c, hipMemcpyAsync(h.data(), c->d_iters, 4 * h.size(), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int32_t s = 0; s < nsteps; ++s)
            for (int64_t k = 0; k < c->nslot; ++k)
                if (c->slot_col[k] >= 0)
                    iters_out[(int64_t)s * c->ncol + c->slot_col[k]] = h[(size_t)s * c->nslot + k];
    }
    if (stats && stats->n_domain_error)
        return set_err(c, CLM_E_DOMAIN, "inverse_matric_potential: matric potential is positive");
    return CLM_OK;
}

int clm_totals(clm_ctx* c, double* water, double* energy) {
    if (!c || !water) return CLM_E_ARG;
    const bool eh = is_eh(c);
    double* dw = c->d_wd[3];  // per-column work vectors of the fine hooks (ncol or nslot)
    double* de = eh ? c->d_wd[4] : nullptr;
    if (is_group(c)) {
        hipLaunchKernelGGL(k_totals_csr, dim3((unsigned)((c->ncol + 255) / 256)), dim3(256), 0,
                           c->stream, c->d_col_ptr, c->d_czf, c->d_cv, eh ? c->d_ci : nullptr,
                           eh ? c->d_ce : nullptr, c->ncol, dw, de);
        HIPCHK(c, hipGetLastError());
        int r;
        if ((r = download_raw(c, dw, water, c->ncol))) return r;
        return energy && eh ? download_raw(c, de, energy, c->ncol) : CLM_OK;
    }
    hipLaunchKernelGGL(k_totals_packed, grid_slots(c), dim3(256), 0, c->stream, layout(c),
                       c->d_zf, c->d_xv, eh ? c->d_xi : nullptr, eh ? c->d_xe : nullptr, dw, de);
    HIPCHK(c, hipGetLastError());
    int r;
    if ((r = download_cols(c, dw, water))) return r;
    return energy && eh ? download_cols(c, de, energy) : CLM_OK;
}

int clm_soil_conductance(clm_ctx* c, double d_ds, double evap_p, double evap_alpha,
                         double D_vapor, double* g_soil, double* dry_layer) {
    if (!c || !g_soil) return CLM_E_ARG;
    if (!is_eh(c) || !is_group(c))
        return set_err(c, CLM_E_STATE,
                       "clm_soil_conductance: EnergyHydrology on the group solver only");
    double* dg = c->d_wd[3];
    double* dd = dry_layer ? c->d_wd[4] : nullptr;
    hipLaunchKernelGGL(k_soil_conductance, dim3((unsigned)((c->ncol + 255) / 256)), dim3(256), 0,
                       c->stream, c->d_col_ptr, c->d_czf, c->d_cpar, c->ncol, c->d_cv, c->d_ci,
                       d_ds, evap_p, evap_alpha, D_vapor, dg, dd);
    HIPCHK(c, hipGetLastError());
    int r;
    if ((r = download_raw(c, dg, g_soil, c->ncol))) return r;
    return dd ? download_raw(c, dd, dry_layer, c->ncol) : CLM_OK;
}

int clm_get_shapes(clm_ctx* c, int32_t* rec, int32_t max_records, int32_t* nrec) {
    if (!c || !nrec) return CLM_E_ARG;
    int32_t n = 0;
    if (is_group(c)) {
        for (int b = 0; b < kNBucket; ++b) {
            if (rec && n < max_records) {
                rec[4 * n + 0] = kBucketMax[b];
                rec[4 * n + 1] = c->bins[b].G;
                rec[4 * n + 2] = c->bins[b].R;
                rec[4 * n + 3] = c->bins[b].n;
            }
            ++n;
        }
    }
    *nrec = n;
    return CLM_OK;
}

int clm_set_shape(clm_ctx* c, int32_t bucket_max_layers, int32_t G, int32_t R) {
    if (!c) return CLM_E_ARG;
    if (!is_group(c)) return set_err(c, CLM_E_STATE, "shapes apply to the group solver only");
    int b = -1;
    for (int k = 0; k < kNBucket; ++k)
        if (kBucketMax[k] == bucket_max_layers) b = k;
    if (b < 0) return set_err(c, CLM_E_ARG, "no depth bucket with max %d layers", bucket_max_layers);
    if (menu_index(G, R) < 0) return set_err(c, CLM_E_ARG, "shape (%d, %d) is not compiled", G, R);
    if (G * R < c->bucket_nmax[b])
        return set_err(c, CLM_E_ARG, "shape (%d, %d) holds %d rows < %d layers", G, R, G * R,
                       c->bucket_nmax[b]);
    c->bins[b].G = G;
    c->bins[b].R = R;
    destroy_graph(c);
    return CLM_OK;
}

constexpr int kTuneSteps = 4;  // consecutive steps per autotune trial

int clm_set_pairing(clm_ctx* c, int32_t on) {
    if (!c) return CLM_E_ARG;
    c->pairing = on != 0;
    destroy_graph(c);
    return CLM_OK;
}

int clm_get_pairing(const clm_ctx* c, int32_t* active) {
    if (!c || !active) return CLM_E_ARG;
    *active = pairing_applies(c) ? 1 : 0;
    return CLM_OK;
}

int clm_autotune(clm_ctx* c, double dt, int32_t reps) {
    if (!c || !(dt > 0.0) || reps < 1) return CLM_E_ARG;
    if (!is_group(c)) return CLM_OK;
    int r = ensure_iters(c, 1);
    if (r) return r;
    // save the state (every trial starts from it)
    const int64_t nc = c->ncell, nk = c->ncol;
    std::vector<std::pair<double*, int64_t>> st = {{c->d_cv, nc}, {c->d_ciw, nk}};
    if (is_eh(c)) {
        st.push_back({c->d_ci, nc});
        st.push_back({c->d_ce, nc});
        st.push_back({c->d_cie, nk});
    }
    int64_t tot = 0;
    for (auto& p : st) tot += p.second;
    double* save = nullptr;
    HIPCHK(c, hipMalloc(&save, 8 * (size_t)tot));
    auto copy_all = [&](bool to_save) -> hipError_t {
        int64_t off = 0;
        for (auto& p : st) {
            hipError_t e = to_save ? hipMemcpyAsync(save + off, p.first, 8 * p.second,
                                                    hipMemcpyDeviceToDevice, c->stream)
                                   : hipMemcpyAsync(p.first, save + off, 8 * p.second,
                                                    hipMemcpyDeviceToDevice, c->stream);
            if (e != hipSuccess) return e;
            off += p.second;
        }
        return hipSuccess;
    };
    hipError_t e = copy_all(true);
    for (int b = 0; b < kNBucket && e == hipSuccess; ++b) {
        clm_ctx::Bin& bn = c->bins[b];
        if (bn.n == 0) continue;
        int bestG = bn.G, bestR = bn.R;
        float best = 1e30f;
        for (int i = 0; i < kNMenu && e == hipSuccess; ++i) {
            int cap = kMenu[i].G * kMenu[i].R;
            if (cap < c->bucket_nmax[b] || cap > 2 * kBucketMax[b]) continue;
            bn.G = kMenu[i].G;
            bn.R = kMenu[i].R;
            float tmin = 1e30f;
            for (int k = 0; k < reps && e == hipSuccess; ++k) {
                e = copy_all(false);
                if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
                // a few consecutive steps, regrouped by Newton count in between as
                // the run is: the first step from the initial state iterates
                // more than the run does, and shapes with several columns per
                // wave are only as fast as the run makes them once a wave's
                // columns have similar counts
                for (int st = 0; st < kTuneSteps && e == hipSuccess; ++st) {
                    if (c->jac_prec == 32)
                        launch_bin_any<float>(c, bn, c->d_cftw, c->d_cfth, dt, c->d_iters,
                                              current_extra(c));
                    else
                        launch_bin_any<double>(c, bn, c->d_cftw, c->d_cfth, dt, c->d_iters,
                                               current_extra(c));
                    if (st + 1 < kTuneSteps) launch_regroup(c, c->d_iters);
                    e = hipGetLastError();
                }
                if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
                if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
                float ms = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
                tmin = std::min(tmin, ms);
            }
            if (tmin < best) {
                best = tmin;
                bestG = bn.G;
                bestR = bn.R;
            }
        }
        bn.G = bestG;
        bn.R = bestR;
    }
    // paired waves (launch_pair) against the per-bucket choice, over whole
    // steps: the buckets of <= 40 and <= 24 layers on (64, 1) and (32, 1)
    // sharing waves can beat the shapes each bucket wins on its own
    if (e == hipSuccess && c->pairing && is_eh(c) && !is_atmos(c) &&
        c->bins[kPairDeepBin].n > 0 && c->bins[kPairShallowBin].n > 0) {
        clm_ctx::Bin &d = c->bins[kPairDeepBin], &h = c->bins[kPairShallowBin];
        const int dG = d.G, dR = d.R, hG = h.G, hR = h.R;
        auto time_steps = [&](float* out) -> hipError_t {
            float tmin = 1e30f;
            hipError_t r2 = hipSuccess;
            for (int k = 0; k < reps && r2 == hipSuccess; ++k) {
                r2 = copy_all(false);
                if (r2 == hipSuccess) r2 = hipEventRecord(c->ev0, c->stream);
                for (int st = 0; st < kTuneSteps && r2 == hipSuccess; ++st) {
                    launch_step(c, c->d_cftw, c->d_cfth, dt, c->d_iters);
                    r2 = hipGetLastError();
                }
                if (r2 == hipSuccess) r2 = hipEventRecord(c->ev1, c->stream);
                if (r2 == hipSuccess) r2 = hipEventSynchronize(c->ev1);
                float ms = 0.f;
                if (r2 == hipSuccess) r2 = hipEventElapsedTime(&ms, c->ev0, c->ev1);
                tmin = std::min(tmin, ms);
            }
            *out = tmin;
            return r2;
        };
        float t_sep = 1e30f, t_pair = 1e30f;
        if (!(dG == 64 && dR == 1 && hG == 32 && hR == 1)) {
            c->pairing = false;
            e = time_steps(&t_sep);
            c->pairing = true;
            d.G = 64, d.R = 1, h.G = 32, h.R = 1;
            if (e == hipSuccess) e = time_steps(&t_pair);
            if (!(t_pair < t_sep)) d.G = dG, d.R = dR, h.G = hG, h.R = hR;
        }
    }
    if (e == hipSuccess) e = copy_all(false);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(save);
    destroy_graph(c);
    if (e != hipSuccess)
        return set_err(c, CLM_E_HIP, "autotune failed: %s", hipGetErrorString(e));
    return CLM_OK;
}

int clm_last_kernel_ms(const clm_ctx* c, double* ms_per_launch, int32_t* launches) {
    if (!c || !ms_per_launch) return CLM_E_ARG;
    *ms_per_launch = c->last_launches ? c->last_ms / c->last_launches : 0.0;
    if (launches) *launches = c->last_launches;
    return CLM_OK;
}


// ---- learned warm start ---------------------------------------------------------

int clm_ws_init(clm_ctx* c, uint64_t seed) {
    if (!c) return CLM_E_ARG;
    if (!is_group(c)) return set_err(c, CLM_E_STATE, "the warm start needs the group solver");
    if (is_atmos(c))
        return set_err(c, CLM_E_STATE, "the warm start takes prescribed top fluxes, not "
                                       "AtmosDrivenFluxBC");
    ws_free(c->ws);
    c->ws = nullptr;
    destroy_graph(c);
    WsState* w = new WsState();
    w->d = ws::dims_for(is_eh(c));
    const int H = w->d.H;
    const int ln[4] = {H, H, H, w->d.Nout}, lk[4] = {w->d.Fp, H, H, H};
    for (int l = 0; l < 4; ++l) {
        w->ln[l] = ln[l];
        w->lk[l] = lk[l];
        w->off[l] = w->P;
        w->offw[l] = w->PW;
        w->P += (int64_t)ln[l] * lk[l] + ln[l];
        w->PW += (int64_t)ln[l] * lk[l];
    }
    int r = 0;
    if ((r = dalloc(c, &w->p, w->P)) || (r = dalloc(c, &w->g, w->P)) ||
        (r = dalloc(c, &w->m1, w->P)) || (r = dalloc(c, &w->m2, w->P)) ||
        (r = dalloc(c, &w->wt, w->PW)) || (r = dalloc(c, &w->wn, w->PW)) ||
        (r = dalloc(c, &w->wpk, w->PW)) || (r = dalloc(c, &w->tails, c->ncol * ws::kTails)) ||
        (r = dalloc(c, &w->X, c->ncol * w->d.Fp)) || (r = dalloc(c, &w->Ha, c->ncol * H)) ||
        (r = dalloc(c, &w->Hb, c->ncol * H)) || (r = dalloc(c, &w->Y, c->ncol * w->d.Nout)) ||
        (r = dalloc(c, &w->gv, c->ncell)) || (r = dalloc(c, &w->ge, c->ncell)) ||
        (r = dalloc(c, &w->loss, 1))) {
        ws_free(w);
        return r;
    }
    // uniform(-a, a) weights, a = sqrt(6 / fan_in) (output layer x 0.1); zero biases
    std::vector<float> h((size_t)w->P, 0.0f);
    uint64_t st = seed ^ 0x5DEECE66DULL;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ULL;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    };
    for (int l = 0; l < 4; ++l) {
        double a = std::sqrt(6.0 / lk[l]) * (l == 3 ? 0.1 : 1.0);
        for (int64_t i = 0; i < (int64_t)ln[l] * lk[l]; ++i) {
            double u = (double)(next() >> 11) * (1.0 / 9007199254740992.0);
            h[(size_t)(w->off[l] + i)] = (float)((2.0 * u - 1.0) * a);
        }
    }
    hipError_t e = hipMemcpyAsync(w->p, h.data(), 4 * h.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(ws::k_static_tails, grid1(c->ncol), dim3(256), 0, c->stream, is_eh(c),
                           c->d_col_ptr, c->ncol, c->d_cpar, w->tails);
        ws_cast(c, w);
        e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) {
        ws_free(w);
        return set_err(c, CLM_E_HIP, "warm-start init failed: %s", hipGetErrorString(e));
    }
    c->ws = w;
    return CLM_OK;
}

int clm_ws_dims(clm_ctx* c, int32_t* dims, int64_t* nparams) {
    if (!c || !dims) return CLM_E_ARG;
    if (!c->ws) return set_err(c, CLM_E_STATE, "clm_ws_init must run first");
    dims[0] = c->ws->d.F;
    dims[1] = c->ws->d.Fp;
    dims[2] = c->ws->d.H;
    dims[3] = c->ws->d.Nout;
    if (nparams) *nparams = c->ws->P;
    return CLM_OK;
}

int clm_ws_get_params(clm_ctx* c, float* out) {
    if (!c || !out) return CLM_E_ARG;
    if (!c->ws) return set_err(c, CLM_E_STATE, "clm_ws_init must run first");
    HIPCHK(c, hipMemcpyAsync(out, c->ws->p, 4 * (size_t)c->ws->P, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int clm_ws_set_params(clm_ctx* c, const float* in) {
    if (!c || !in) return CLM_E_ARG;
    if (!c->ws) return set_err(c, CLM_E_STATE, "clm_ws_init must run first");
    HIPCHK(c, hipMemcpyAsync(c->ws->p, in, 4 * (size_t)c->ws->P, hipMemcpyHostToDevice,
                             c->stream));
    ws_cast(c, c->ws);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int clm_ws_enable(clm_ctx* c, int32_t on) {
    if (!c) return CLM_E_ARG;
    if (on && !c->ws) return set_err(c, CLM_E_STATE, "clm_ws_init must run first");
    if (c->ws && c->ws->enabled != (on != 0)) destroy_graph(c);
    if (c->ws) c->ws->enabled = on != 0;
    return CLM_OK;
}

int clm_ws_features(clm_ctx* c, double dt, float* out) {
    if (!c || !out) return CLM_E_ARG;
    if (!c->ws) return set_err(c, CLM_E_STATE, "clm_ws_init must run first");
    WsState* w = c->ws;
    ws_features_launch(c, w, c->d_cftw, is_eh(c) ? c->d_cfth : nullptr, dt, w->X);
    int64_t n = c->ncol * w->d.Fp;
    float* tmp = nullptr;
    HIPCHK(c, hipMalloc(&tmp, 4 * (size_t)n));
    hipLaunchKernelGGL(ws::k_bf16_to_f32, grid1(n), dim3(256), 0, c->stream, n, w->X, tmp);
    hipError_t e = hipMemcpyAsync(out, tmp, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return set_err(c, CLM_E_HIP, "features: %s", hipGetErrorString(e));
    return CLM_OK;
}

int clm_ws_guess(clm_ctx* c, double dt, int32_t fused, double* gv, double* ge) {
    if (!c || !gv) return CLM_E_ARG;
    if (!c->ws) return set_err(c, CLM_E_STATE, "clm_ws_init must run first");
    WsState* w = c->ws;
    ws_guess(c, w, c->d_cftw, is_eh(c) ? c->d_cfth : nullptr, dt, fused != 0);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(gv, w->gv, 8 * (size_t)c->ncell, hipMemcpyDeviceToHost, c->stream));
    if (ge && is_eh(c))
        HIPCHK(c, hipMemcpyAsync(ge, w->ge, 8 * (size_t)c->ncell, hipMemcpyDeviceToHost,
                                 c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int clm_ws_collect(clm_ctx* c, double dt, int32_t nsteps, int32_t steps_per_slot,
                   int32_t stride, int64_t* rows_out) {
    if (!c || nsteps < 1 || steps_per_slot < 1 || stride < 1 || !(dt > 0.0)) return CLM_E_ARG;
    if (!c->ws) return set_err(c, CLM_E_STATE, "clm_ws_init must run first");
    if (c->tab_slots < 1) return set_err(c, CLM_E_STATE, "clm_set_forcing_table must run first");
    WsState* w = c->ws;
    const int Fp = w->d.Fp, No = w->d.Nout;
    const int32_t ncollect = (nsteps + stride - 1) / stride;
    int64_t need = w->rows + (int64_t)ncollect * c->ncol;
    int r = 0;
    if (need > w->cap) {  // grow the training set, keeping its rows
        bf16* X2 = nullptr;
        float *T2 = nullptr, *M2 = nullptr;
        if ((r = dalloc(c, &X2, need * Fp)) || (r = dalloc(c, &T2, need * No)) ||
            (r = dalloc(c, &M2, need)))
            return r;
        if (w->rows) {
            HIPCHK(c, hipMemcpyAsync(X2, w->Xd, 2 * (size_t)(w->rows * Fp), hipMemcpyDeviceToDevice,
                                     c->stream));
            HIPCHK(c, hipMemcpyAsync(T2, w->Td, 4 * (size_t)(w->rows * No), hipMemcpyDeviceToDevice,
                                     c->stream));
            HIPCHK(c, hipMemcpyAsync(M2, w->Md, 4 * (size_t)w->rows, hipMemcpyDeviceToDevice,
                                     c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        dfree(w->Xd);
        dfree(w->Td);
        dfree(w->Md);
        w->Xd = X2;
        w->Td = T2;
        w->Md = M2;
        w->cap = need;
    }
    if (!w->u0v) {
        if ((r = dalloc(c, &w->u0v, c->ncell)) || (r = dalloc(c, &w->unv, c->ncell)) ||
            (r = dalloc(c, &w->uni, c->ncell)) || (r = dalloc(c, &w->une, c->ncell)))
            return r;
    }
    if ((r = ensure_iters(c, 1))) return r;
    const bool eh = is_eh(c);
    for (int32_t s = 0; s < nsteps; ++s) {
        int64_t slot = (s / steps_per_slot) % c->tab_slots;
        const double* ftw = c->d_tab_w + slot * c->ncol;
        const double* fth = eh ? c->d_tab_h + slot * c->ncol : nullptr;
        if (s % stride != 0) {  // a plain step between two collected ones
            launch_step(c, ftw, fth, dt, c->d_iters, nullptr, false);
            continue;
        }
        ws_features_launch(c, w, ftw, fth, dt, w->Xd + w->rows * Fp);
        HIPCHK(c, hipMemcpyAsync(w->unv, c->d_cv, 8 * (size_t)c->ncell, hipMemcpyDeviceToDevice,
                                 c->stream));
        if (eh) {
            HIPCHK(c, hipMemcpyAsync(w->uni, c->d_ci, 8 * (size_t)c->ncell,
                                     hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(w->une, c->d_ce, 8 * (size_t)c->ncell,
                                     hipMemcpyDeviceToDevice, c->stream));
        }
        launch_step(c, ftw, fth, dt, c->d_iters, w->u0v, false);
        hipLaunchKernelGGL(ws::k_targets, grid1(c->ncol * ws::kGroup), dim3(256), 0, c->stream, eh,
                           c->d_col_ptr, c->ncol, c->d_cpar, w->u0v, w->unv, w->uni, w->une,
                           c->d_cv, c->d_ce, c->d_iters, No, w->Td + w->rows * No,
                           w->Md + w->rows);
        HIPCHK(c, hipGetLastError());
        w->rows += c->ncol;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_cache = false;
    c->have_jac = false;
    if (rows_out) *rows_out = w->rows;
    return CLM_OK;
}

int clm_ws_train(clm_ctx* c, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                 float* loss_per_epoch) {
    if (!c || epochs < 1 || batch < 512 || batch % 512 != 0 || !(lr > 0.0f)) return CLM_E_ARG;
    if (!c->ws) return set_err(c, CLM_E_STATE, "clm_ws_init must run first");
    WsState* w = c->ws;
    int64_t nb = w->rows / batch;
    if (nb < 1) return set_err(c, CLM_E_STATE, "training set holds fewer rows than one batch");
    int r = ws_batch_buffers(c, w, batch);
    if (r) return r;
    if (w->idx_cap < w->rows) {
        dfree(w->idx);
        if ((r = dalloc(c, &w->idx, w->rows))) return r;
        w->idx_cap = w->rows;
    }
    std::vector<int32_t> perm((size_t)w->rows);
    std::iota(perm.begin(), perm.end(), 0);
    uint64_t st = seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ULL;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    };
    const int Fp = w->d.Fp, No = w->d.Nout;
    for (int32_t ep = 0; ep < epochs; ++ep) {
        // cosine decay of the step size over the epochs
        const float lr_ep =
            lr * 0.5f * (1.0f + (float)std::cos(3.141592653589793 * ep / (double)epochs));
        for (int64_t i = w->rows - 1; i > 0; --i) std::swap(perm[i], perm[next() % (i + 1)]);
        HIPCHK(c, hipMemcpyAsync(w->idx, perm.data(), 4 * perm.size(), hipMemcpyHostToDevice,
                                 c->stream));
        hipLaunchKernelGGL(ws::k_zero_f32, dim3(1), dim3(64), 0, c->stream, (int64_t)1, w->loss);
        for (int64_t b = 0; b < nb; ++b) {
            hipLaunchKernelGGL(ws::k_gather, dim3((unsigned)batch), dim3(128), 0, c->stream, batch,
                               w->idx + b * batch, Fp, w->Xd, No, w->Td, w->Md, w->Xb, w->Tb,
                               w->Mk);
            ws_forward(c, w, w->Xb, batch, true, w->Yb);
            hipLaunchKernelGGL(ws::k_mse, grid1((int64_t)batch * No), dim3(256), 0, c->stream,
                               batch, No, w->Yb, w->Tb, w->Mk, w->dY, w->loss);
            ws_backward(c, w, batch);
            ws_adam(c, w, lr_ep);
        }
        HIPCHK(c, hipGetLastError());
        float l = 0.0f;
        HIPCHK(c, hipMemcpyAsync(&l, w->loss, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (loss_per_epoch) loss_per_epoch[ep] = l / (float)nb;
    }
    return CLM_OK;
}

int clm_ws_forward(clm_ctx* c, const float* X, int64_t M, float* Y) {
    if (!c || !X || !Y || M < 1 || M % 32 != 0) return CLM_E_ARG;
    if (!c->ws) return set_err(c, CLM_E_STATE, "clm_ws_init must run first");
    WsState* w = c->ws;
    int r = ws_batch_buffers(c, w, M);
    if (r) return r;
    const int64_t nx = M * w->d.Fp, ny = M * w->d.Nout;
    HIPCHK(c, hipMemcpyAsync(w->f32tmp, X, 4 * (size_t)nx, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(ws::k_f32_to_bf16, grid1(nx), dim3(256), 0, c->stream, nx, w->f32tmp,
                       w->Xb);
    ws_forward(c, w, w->Xb, (int)M, true, w->Yb);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(Y, w->Yb, 4 * (size_t)ny, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int clm_ws_gradient(clm_ctx* c, const float* X, const float* T, const float* mask, int64_t M,
                    float* loss, float* grad) {
    if (!c || !X || !T || !mask || !grad || M < 32 || M % 32 != 0) return CLM_E_ARG;
    if (!c->ws) return set_err(c, CLM_E_STATE, "clm_ws_init must run first");
    WsState* w = c->ws;
    int r = ws_batch_buffers(c, w, M);
    if (r) return r;
    const int64_t nx = M * w->d.Fp, ny = M * w->d.Nout;
    HIPCHK(c, hipMemcpyAsync(w->f32tmp, X, 4 * (size_t)nx, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(ws::k_f32_to_bf16, grid1(nx), dim3(256), 0, c->stream, nx, w->f32tmp,
                       w->Xb);
    HIPCHK(c, hipMemcpyAsync(w->Tb, T, 4 * (size_t)ny, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(w->Mk, mask, 4 * (size_t)M, hipMemcpyHostToDevice, c->stream));
    ws_forward(c, w, w->Xb, (int)M, true, w->Yb);
    hipLaunchKernelGGL(ws::k_zero_f32, dim3(1), dim3(64), 0, c->stream, (int64_t)1, w->loss);
    hipLaunchKernelGGL(ws::k_mse, grid1(ny), dim3(256), 0, c->stream, (int)M, w->d.Nout, w->Yb,
                       w->Tb, w->Mk, w->dY, w->loss);
    ws_backward(c, w, (int)M);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(grad, w->g, 4 * (size_t)w->P, hipMemcpyDeviceToHost, c->stream));
    if (loss) HIPCHK(c, hipMemcpyAsync(loss, w->loss, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}


int clm_ws_time(clm_ctx* c, double dt, int32_t reps, double* ms_per_call) {
    if (!c || reps < 1 || !ms_per_call || !(dt > 0.0)) return CLM_E_ARG;
    if (!c->ws) return set_err(c, CLM_E_STATE, "clm_ws_init must run first");
    StepExtra x;
    bool was = c->ws->enabled;
    c->ws->enabled = true;
    ws_launch_guess(c, c->d_cftw, c->d_cfth, dt, &x);  // warm-up
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int32_t i = 0; i < reps; ++i) ws_launch_guess(c, c->d_cftw, c->d_cfth, dt, &x);
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->ws->enabled = was;
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *ms_per_call = ms / reps;
    return CLM_OK;
}


// ---- checkpoints (checkpoints.jl:125-195) ----------------------------------------
// File <dir>/day<D>.<S>.clmchk, D = floor(t / 86400 s), S = floor(t mod 86400 s).
// Layout (little endian): char magic[8] = "CLMCHK01"; int32 model, bc_bottom,
// bc_top, pad; int64 ncol, ncell; double t; uint64 model hash; then the state
// as float64 arrays: vartheta_l[ncell], intF_water[ncol] and, for
// EnergyHydrology, theta_i[ncell], rho_e_int[ncell], intF_energy[ncol].
// The model hash (FNV-1a over the model kind, boundary kinds, col_ptr,
// z_face and params) plays the part of the reference's land_model_hash.
namespace {
uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 0x100000001B3ULL;
    }
    return h;
}
struct ChkHeader {
    char magic[8];
    int32_t model, bc_bottom, bc_top, pad;
    int64_t ncol, ncell;
    double t;
    uint64_t hash;
};
}  // namespace

static uint64_t model_hash(const clm_ctx* c) {
    uint64_t h = 0xCBF29CE484222325ULL;
    int32_t kinds[3] = {c->model, c->bc_bottom, c->bc_top};
    h = fnv1a(h, kinds, sizeof kinds);
    h = fnv1a(h, c->col_ptr.data(), 4 * c->col_ptr.size());
    h = fnv1a(h, c->h_zf.data(), 8 * c->h_zf.size());
    h = fnv1a(h, c->h_par.data(), 8 * c->h_par.size());
    return h;
}

int clm_save_checkpoint(clm_ctx* c, const char* dir, double t, char* path, int32_t path_len) {
    if (!c || !dir || !(t >= 0.0)) return CLM_E_ARG;
    const bool eh = is_eh(c);
    std::vector<double> v((size_t)c->ncell), iw((size_t)c->ncol), ti, re, ie;
    clm_state Y{v.data(), nullptr, nullptr, iw.data(), nullptr};
    if (eh) {
        ti.resize((size_t)c->ncell);
        re.resize((size_t)c->ncell);
        ie.resize((size_t)c->ncol);
        Y.theta_i = ti.data();
        Y.rho_e_int = re.data();
        Y.intF_energy = ie.data();
    }
    int r = clm_get_state(c, &Y);
    if (r) return r;
    long long day = (long long)std::floor(t / 86400.0);
    long long sec = (long long)std::floor(std::fmod(t, 86400.0));
    std::string file = std::string(dir) + "/day" + std::to_string(day) + "." +
                       std::to_string(sec) + ".clmchk";
    FILE* f = std::fopen(file.c_str(), "wb");
    if (!f) return set_err(c, CLM_E_ARG, "cannot open %s for writing", file.c_str());
    ChkHeader h;
    std::memcpy(h.magic, "CLMCHK01", 8);
    h.model = c->model;
    h.bc_bottom = c->bc_bottom;
    h.bc_top = c->bc_top;
    h.pad = 0;
    h.ncol = c->ncol;
    h.ncell = c->ncell;
    h.t = t;
    h.hash = model_hash(c);
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
    ok = ok && std::fwrite(v.data(), 8, v.size(), f) == v.size();
    ok = ok && std::fwrite(iw.data(), 8, iw.size(), f) == iw.size();
    if (eh) {
        ok = ok && std::fwrite(ti.data(), 8, ti.size(), f) == ti.size();
        ok = ok && std::fwrite(re.data(), 8, re.size(), f) == re.size();
        ok = ok && std::fwrite(ie.data(), 8, ie.size(), f) == ie.size();
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return set_err(c, CLM_E_ARG, "writing %s failed", file.c_str());
    if (path && path_len > 0) std::snprintf(path, (size_t)path_len, "%s", file.c_str());
    return CLM_OK;
}

int clm_read_checkpoint(clm_ctx* c, const char* path, double* t, int32_t* same_model) {
    if (!c || !path) return CLM_E_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_err(c, CLM_E_ARG, "cannot open %s", path);
    ChkHeader h;
    bool ok = std::fread(&h, sizeof h, 1, f) == 1 && std::memcmp(h.magic, "CLMCHK01", 8) == 0;
    if (!ok) {
        std::fclose(f);
        return set_err(c, CLM_E_ARG, "%s is not a checkpoint", path);
    }
    if (h.model != c->model || h.ncol != c->ncol || h.ncell != c->ncell) {
        std::fclose(f);
        return set_err(c, CLM_E_ARG, "%s holds a %s state of %lld columns / %lld cells", path,
                       h.model ? "EnergyHydrology" : "RichardsModel", (long long)h.ncol,
                       (long long)h.ncell);
    }
    const bool eh = is_eh(c);
    std::vector<double> v((size_t)h.ncell), iw((size_t)h.ncol), ti, re, ie;
    ok = std::fread(v.data(), 8, v.size(), f) == v.size();
    ok = ok && std::fread(iw.data(), 8, iw.size(), f) == iw.size();
    clm_state Y{v.data(), nullptr, nullptr, iw.data(), nullptr};
    if (eh) {
        ti.resize(v.size());
        re.resize(v.size());
        ie.resize(iw.size());
        ok = ok && std::fread(ti.data(), 8, ti.size(), f) == ti.size();
        ok = ok && std::fread(re.data(), 8, re.size(), f) == re.size();
        ok = ok && std::fread(ie.data(), 8, ie.size(), f) == ie.size();
        Y.theta_i = ti.data();
        Y.rho_e_int = re.data();
        Y.intF_energy = ie.data();
    }
    std::fclose(f);
    if (!ok) return set_err(c, CLM_E_ARG, "%s is truncated", path);
    int r = clm_set_state(c, &Y);
    if (r) return r;
    if (t) *t = h.t;
    // the reference warns (and goes on) when the model differs
    if (same_model) *same_model = h.hash == model_hash(c) ? 1 : 0;
    return CLM_OK;
}

}  // extern "C"
