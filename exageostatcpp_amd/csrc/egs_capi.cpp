// C ABI of the gfx950 maximum-likelihood path (include/exageostat_amd.h).
//
// Host orchestration only: shape checks, HBM layout of a problem, the order of
// kernel launches on the context's stream.  All arithmetic on Sigma runs in the
// HIP kernels of egs_covgen.hip, egs_chol.hip and egs_solve.hip; the host adds
// up nt per-tile partials in tile order and evaluates the closing formula of
// ChameleonImplementation.cpp:257-265.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/exageostat_amd.h"
#include "egs_bobyqa.hpp"
#include "egs_common.hpp"
#include "egs_devmem.hpp"
#include "egs_kernels.hpp"
#include "egs_matern.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define EGS_HIP(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(EGS_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

constexpr double kPi = 3.141592653589793;   // Definitions.hpp:45
constexpr int kMaxTileSize = 2048;

// Device scratch of egs_covariance_tile_noalloc: the output tile and the
// location slices it stages, at most kScratchDoubles doubles (32 MiB), held by
// every GPU the library initializes until the process ends.  The array's name
// is a host-side stand-in; host code uses only the address that
// hipGetSymbolAddress returns for the current device (g_scratch).
constexpr int64_t kScratchDoubles = int64_t(1) << 22;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch[kMaxDevices] = {};

// Resolves this device's copy of the scratch array once (the device is current).
int resolve_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (g_scratch[device]) return EGS_SUCCESS;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(EGS_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch[device] = static_cast<double *>(ptr);
    return EGS_SUCCESS;
}

}  // namespace

namespace egs {
void set_last_error(const char *msg) { g_last_error = msg; }
}  // namespace egs

struct egs_problem_s;

struct egs_context_s {
    int device = 0;
    hipStream_t stream = nullptr;     // critical path: next panel, solves, copies
    hipStream_t bulk = nullptr;       // trailing updates that the next panel does not wait for
    std::vector<egs_problem_s *> problems;   // live problems, released by egs_finalize
    std::mutex mu;                           // guards problems (problems move between contexts)
    egs::BatchJob *d_jobs = nullptr;         // job records of egs_loglik_batch, jobs_cap entries
    int64_t jobs_cap = 0, dev_bytes = 0;
};

// Per-kernel timing with HIP events on the library stream (enabled by
// egs_problem_profile); collected after the stream synchronizes.
struct KernelStat {
    int64_t launches = 0;
    double ms = 0.0;
    double flops = 0.0;
};

struct Profiler {
    bool on = false;
    std::vector<hipEvent_t> ev;            // pairs: start, stop
    std::vector<std::pair<int, double>> pending;   // (kernel id, flops) per used pair
    KernelStat stats[EGS_PROFILE_KERNELS];
};

// Update classes of a mixed plan (see tile_class).
constexpr int kClasses = 6;
// Panel copies in flight: panel k is read by the bulk update of step k while
// the critical stream writes panel k+1 and then panel k+2.
constexpr int kPanelSlots = 3;

struct egs_problem_s {
    egs_context_s *ctx = nullptr;
    int kernel = 0, p = 1, metric = 0, dim = 2;
    int dts = 0, ts = 0, nt = 0, nsb = 0;
    int64_t n_loc = 0, n = 0, npad = 0;
    int mode = EGS_PRECISION_FP64, band64 = 2, band32 = 10;
    // One-product FP16 for blocks beyond band_single (-1: never) whose panel
    // rows have bmax(A) bmax(B) <= tau_single pmax^2.  With 2^-16 such a
    // block's product error, about 2^-12 K bmax(A) bmax(B), stays 16x under
    // 2^-24 K pmax^2, the split product's for the panel's largest block.
    int band_single = 0;
    // FP64 tiles with tile distance > band_prod64 (within band64) take split
    // FP16 products into their FP64 storage instead of FP64 products.  Set at
    // creation: 0 for univariate kernels (the diagonal tiles only), 2 for the
    // bivariate kernel, whose smoother second field is more sensitive (band 1
    // measured 1.2e-6 relative there under plan (2, 2), band 0 2.5e-6)
    int band_prod64 = 0;
    // The per-block choice also applies to the FP64 tiles that take split
    // products (class 5).  Set at creation: on for univariate kernels (at
    // n = 100k the error stays 2.3e-7 and the evaluation is 4.5% faster), off
    // for the bivariate kernel (plan (8, 8): 1.2e-6 with it, 2.2e-7 without).
    bool single_fp64 = true;
    // The panel solve of FP32 rows runs on FP32 MFMA (FP32 accumulation)
    // instead of FP64 MFMA.  Set at creation: on for univariate kernels
    // (n = 100k: 1021 -> 991 ms, relative error 2.3e-7 -> 2.7e-7), off for
    // the bivariate kernel (its sweep's errors grew to 1.4-3.0e-6).
    bool f32_panel_solve = true;
    double tau_single = 0x1p-16;
    // One FP8 (e4m3) product for FP32-tile blocks beyond band_single whose
    // panel rows have bmax(A) bmax(B) <= tau_fp8 pmax^2 (0: never).  2^-26 for
    // univariate kernels: an e4m3 product's error bound (2^-4 relative) there
    // is 8x under the FP16 product's (2^-11) at tau_single = 2^-16.
    double tau_fp8 = 0.0;
    // mixed plans: per-class update lists (descending column) and, for class c,
    // cls_count[c][j] = number of its tiles in columns >= j
    int2 *d_pairs_cls[kClasses] = {};
    std::vector<int2> h_pairs_cls[kClasses];
    std::vector<int> cls_count[kClasses];
    // panel k as FP64 column-major, FP32 column-major and 16-bit row-major, in
    // slot k % kPanelSlots: the critical stream writes panel k+2 while the bulk
    // stream may still read panel k (see run_potrf)
    double *d_p64[kPanelSlots] = {};
    float *d_p32[kPanelSlots] = {};
    uint16_t *d_p16h[kPanelSlots] = {}, *d_p16l[kPanelSlots] = {};   // FP16 or BF16 bits
    uint8_t *d_p8[kPanelSlots] = {};                                  // e4m3 of the row-scaled values
    float *d_pinv[kPanelSlots] = {};   // row scales of the FP16 parts (inverse, powers of two)
    float *d_bmax[kPanelSlots] = {};   // largest row scale per 192-row block of a panel
    unsigned *d_pmax = nullptr;              // largest row scale per panel (float bits), nt entries
    unsigned long long *d_counts = nullptr;  // split / single / FP8-product block updates of the last factorization
    double *d_cg = nullptr;                  // refinement vectors x, r, p, q, w (npad each)
    // FP64 copy of the lower tiles of Sigma written with every covariance
    // generation of a mixed plan (the refinement's Sigma v reads it), and the
    // per-tile partials of that product; null when the copy does not fit
    double *d_sig64 = nullptr, *d_prow = nullptr, *d_pcol = nullptr;
    // inverses X_k = L_kk^{-T} of the diagonal tiles (preconditioner solves)
    // and the solve's output vector
    double *d_tinv = nullptr, *d_ybuf = nullptr;
    double *d_mv = nullptr;                  // matrix-vector chunk partials
    double *d_dot = nullptr;                 // dot-product block partials
    double *d_pred = nullptr;                // prediction targets and outputs, pred_cap doubles
    int64_t pred_cap = 0;
    int64_t dev_bytes = 0;                   // device bytes held (size classes of egs_devmem)
    int refine_iters = 0;
    double refine_resid = 0.0;
    bool refine_converged = true;            // refine_resid <= kRefineAccept
    // mixed plans: an evaluation whose refinement needs more than
    // fallback_iters iterations (0: never), or whose mixed factor is not
    // positive definite, is repeated under the FP64 plan
    int fallback_iters = 4;
    // ... or whose estimated relative log-likelihood error exceeds fallback_tol:
    // |q0 - q| / (2 max(|llh|, 1)), q0 = z'(L L^T)^{-1} z with the mixed factor
    // and q = z'Sigma^{-1} z refined (see egs_loglik)
    double fallback_tol = 1e-5;
    double quad_unrefined = 0.0, error_estimate = 0.0;
    bool fell_back = false;
    int64_t fallbacks = 0;

    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr;
    char *d_arena = nullptr;
    size_t arena_bytes = 0;
    int64_t *d_off = nullptr;
    uint8_t *d_store = nullptr;
    int2 *d_pairs = nullptr;
    double *d_winv = nullptr;
    double *d_dscr = nullptr;   // ts * ts scratch of the diagonal-tile factorization
    double *d_obs = nullptr, *d_work = nullptr, *d_aux = nullptr, *d_partials = nullptr;
    int *d_info = nullptr;

    std::vector<int64_t> h_off;
    std::vector<uint8_t> h_store;
    std::vector<int2> h_pairs;
    std::vector<double> h_partials;

    bool has_obs = false;
    bool has_cov = false;    // the tiles hold Sigma(cov_theta), not yet factored
    bool factored = false;
    double factored_theta[6] = {0, 0, 0, 0, 0, 0};
    double cov_theta[6] = {0, 0, 0, 0, 0, 0};
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<hipEvent_t> ev_panel, ev_bulk;   // per step: panel solved / bulk update done
    double t_cov = 0, t_fact = 0, t_solve = 0;
    Profiler prof;
    // FP64 evaluations replayed from a captured graph (egs_problem_set_graph):
    // the covariance parameters live in d_prm so that the graph serves every theta
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    egs::MaternParams *d_prm = nullptr;
    egs::MaternParams h_prm;

    egs::TileMap map() const {
        egs::TileMap m;
        m.base = d_arena;
        m.off = d_off;
        m.store = d_store;
        m.nt = nt;
        m.ts = ts;
        return m;
    }
    int ntheta() const { return kernel == 1 ? 6 : 3; }
    double *winv(int k) const { return d_winv + (int64_t) k * nsb * egs::kBlk * egs::kBlk; }
};

namespace {

// Device buffers of a problem come from the block cache of egs_devmem and are
// counted in p->dev_bytes.
template <typename T>
hipError_t pmalloc(egs_problem_s *p, T **ptr, size_t bytes) {
    void *v = nullptr;
    const hipError_t e = egs::dev_alloc(&v, bytes, &p->dev_bytes);
    *ptr = static_cast<T *>(v);
    return e;
}

template <typename T>
void pfree(egs_problem_s *p, T *ptr) {
    egs::dev_free((void *) ptr, &p->dev_bytes);
}

int check_problem(egs_problem_t p) {
    if (!p) return fail(EGS_ERR_INVALID_ARG, "null problem");
    if (!p->ctx) return fail(EGS_ERR_STATE, "the problem's context was finalized");
    if (hipSetDevice(p->ctx->device) != hipSuccess) return fail(EGS_ERR_HIP, "hipSetDevice failed");
    return EGS_SUCCESS;
}

constexpr double kMaxNu = 100.0;

bool valid_nu(double nu) { return nu > 0.0 && nu <= kMaxNu; }

int make_params(int kernel, int metric, const double *th, egs::MaternParams &prm) {
    if (!th) return fail(EGS_ERR_INVALID_ARG, "theta is null");
    const int nth = kernel == EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS ? 6 : 3;
    for (int q = 0; q < nth; ++q)
        if (!std::isfinite(th[q])) return fail(EGS_ERR_INVALID_ARG, "theta[%d] is not finite", q);
    const double range = kernel == EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS ? th[2] : th[1];
    if (!(range > 0.0)) return fail(EGS_ERR_INVALID_ARG, "the range parameter must be positive");
    if (kernel == EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS ? !(valid_nu(th[3]) && valid_nu(th[4])) : !valid_nu(th[2]))
        return fail(EGS_ERR_INVALID_ARG, "smoothness must lie in (0, %g]", kMaxNu);
    prm.kernel = kernel;
    prm.metric = metric;
    if (kernel == EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY) {
        // UnivariateMaternStationary.cpp:47-49
        const double sigma2 = th[0], nu = th[2];
        prm.range = th[1];
        for (int q = 0; q < 4; ++q) {
            prm.sig[q] = sigma2;
            prm.nu[q] = nu;
            prm.con[q] = sigma2 * (1.0 / (std::pow(2.0, nu - 1) * std::tgamma(nu)));
        }
        return EGS_SUCCESS;
    }
    if (kernel == EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS) {
        // BivariateMaternParsimonious.cpp:50-66
        const double con1 = th[0] * (1.0 / (std::pow(2.0, th[3] - 1) * std::tgamma(th[3])));
        const double con2 = th[1] * (1.0 / (std::pow(2.0, th[4] - 1) * std::tgamma(th[4])));
        const double nu12 = 0.5 * (th[3] + th[4]);
        const double rho = th[5] *
                           std::sqrt((std::tgamma(th[3] + 1) * std::tgamma(th[4] + 1)) /
                                     (std::tgamma(th[3]) * std::tgamma(th[4]))) *
                           std::tgamma(nu12) / std::tgamma(nu12 + 1);
        const double con12 = rho * std::sqrt(th[0] * th[1]) * (1.0 / (std::pow(2.0, nu12 - 1) * std::tgamma(nu12)));
        prm.range = th[2];
        prm.sig[0] = th[0];
        prm.sig[1] = prm.sig[2] = rho * std::sqrt(th[0] * th[1]);
        prm.sig[3] = th[1];
        prm.nu[0] = th[3];
        prm.nu[1] = prm.nu[2] = nu12;
        prm.nu[3] = th[4];
        prm.con[0] = con1;
        prm.con[1] = prm.con[2] = con12;
        prm.con[3] = con2;
        return EGS_SUCCESS;
    }
    return fail(EGS_ERR_INVALID_ARG, "unknown kernel %d", kernel);
}

void free_sigma_copy(egs_problem_s *p) {
    pfree(p, p->d_sig64);
    pfree(p, p->d_prow);
    pfree(p, p->d_pcol);
    p->d_sig64 = p->d_prow = p->d_pcol = nullptr;
}

void free_tile_inverses(egs_problem_s *p) {
    pfree(p, p->d_tinv);
    pfree(p, p->d_ybuf);
    p->d_tinv = p->d_ybuf = nullptr;
}

// Drop the captured evaluation graph (the launch sequence it recorded no
// longer matches the problem's plan or buffers).
void invalidate_graph(egs_problem_s *p) {
    if (p->graph_exec) hipGraphExecDestroy(p->graph_exec);
    if (p->graph) hipGraphDestroy(p->graph);
    p->graph_exec = nullptr;
    p->graph = nullptr;
}

void free_problem_buffers(egs_problem_s *p) {
    if (p->ctx) {
        std::lock_guard<std::mutex> lk(p->ctx->mu);
        auto &v = p->ctx->problems;
        v.erase(std::remove(v.begin(), v.end(), p), v.end());
    }
    invalidate_graph(p);
    pfree(p, p->d_prm);
    p->d_prm = nullptr;
    pfree(p, p->d_x);
    pfree(p, p->d_y);
    pfree(p, p->d_z);
    pfree(p, p->d_arena);
    pfree(p, p->d_off);
    pfree(p, p->d_store);
    pfree(p, p->d_pairs);
    pfree(p, p->d_winv);
    pfree(p, p->d_dscr);
    pfree(p, p->d_obs);
    pfree(p, p->d_work);
    pfree(p, p->d_aux);
    pfree(p, p->d_partials);
    pfree(p, p->d_info);
    for (auto &q : p->d_pairs_cls) {
        pfree(p, q);
        q = nullptr;
    }
    for (int b = 0; b < kPanelSlots; ++b) {
        pfree(p, p->d_p64[b]);
        pfree(p, p->d_p32[b]);
        pfree(p, p->d_p16h[b]);
        pfree(p, p->d_p16l[b]);
        pfree(p, p->d_p8[b]);
        p->d_p8[b] = nullptr;
        pfree(p, p->d_pinv[b]);
        pfree(p, p->d_bmax[b]);
        p->d_pinv[b] = nullptr;
        p->d_bmax[b] = nullptr;
        p->d_p64[b] = nullptr;
        p->d_p32[b] = nullptr;
        p->d_p16h[b] = p->d_p16l[b] = nullptr;
    }
    pfree(p, p->d_pmax);
    pfree(p, p->d_counts);
    p->d_pmax = nullptr;
    p->d_counts = nullptr;
    pfree(p, p->d_cg);
    pfree(p, p->d_mv);
    pfree(p, p->d_dot);
    p->d_cg = p->d_mv = p->d_dot = nullptr;
    pfree(p, p->d_pred);
    p->d_pred = nullptr;
    p->pred_cap = 0;
    free_sigma_copy(p);
    free_tile_inverses(p);
    for (auto &e : p->ev)
        if (e) hipEventDestroy(e);
    for (auto &e : p->prof.ev)
        if (e) hipEventDestroy(e);
    p->prof.ev.clear();
    for (auto &e : p->ev_panel)
        if (e) hipEventDestroy(e);
    for (auto &e : p->ev_bulk)
        if (e) hipEventDestroy(e);
    p->ev_panel.clear();
    p->ev_bulk.clear();
    p->prof.pending.clear();
    p->d_x = p->d_y = p->d_z = nullptr;
    p->d_arena = nullptr;
    p->d_off = nullptr;
    p->d_store = nullptr;
    p->d_pairs = nullptr;
    p->d_winv = p->d_dscr = p->d_obs = p->d_work = p->d_aux = p->d_partials = nullptr;
    p->d_info = nullptr;
    for (auto &e : p->ev) e = nullptr;
}

// Opens a timed region for kernel `id` on the problem's stream; returns the
// event pair index or -1 when profiling is off.
int prof_begin(egs_problem_s *p, int id, double flops, hipStream_t stream = nullptr) {
    if (!p->prof.on) return -1;
    if (!stream) stream = p->ctx->stream;
    Profiler &pr = p->prof;
    const size_t idx = pr.pending.size();
    if (2 * idx + 2 > pr.ev.size()) {
        hipEvent_t a = nullptr, b = nullptr;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1;
        pr.ev.push_back(a);
        pr.ev.push_back(b);
    }
    if (hipEventRecord(pr.ev[2 * idx], stream) != hipSuccess) return -1;
    pr.pending.emplace_back(id, flops);
    return (int) idx;
}

void prof_end(egs_problem_s *p, int idx, hipStream_t stream = nullptr) {
    if (idx < 0) return;
    hipEventRecord(p->prof.ev[2 * (size_t) idx + 1], stream ? stream : p->ctx->stream);
}

// Call after the stream has synchronized.
void prof_collect(egs_problem_s *p) {
    Profiler &pr = p->prof;
    for (size_t i = 0; i < pr.pending.size(); ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pr.ev[2 * i], pr.ev[2 * i + 1]) != hipSuccess) continue;
        KernelStat &ks = pr.stats[pr.pending[i].first];
        ks.launches += 1;
        ks.ms += ms;
        ks.flops += pr.pending[i].second;
    }
    pr.pending.clear();
}

// Rows of tile i that belong to the matrix (the last tile row is short).
double real_rows(const egs_problem_s *p, int i) {
    const int64_t r = p->n - (int64_t) i * p->ts;
    return (double) std::min<int64_t>(p->ts, std::max<int64_t>(0, r));
}

// Storage precision of tile (i, j) under the problem's plan.
uint8_t tile_store(const egs_problem_s *p, int i, int j) {
    if (p->mode == EGS_PRECISION_FP64) return egs::kStoreF64;
    return (i - j) <= p->band64 ? egs::kStoreF64 : egs::kStoreF32;
}

// Update class of tile (i, j) under a mixed plan: 0 FP64, 1 FP32, 2 16-bit
// products (split FP16, or one BF16 product under MIXED_BF16), 3 one FP16
// product (beyond band_single under MIXED, per block when tau_single is
// finite), 4 split FP16 products into an FP64 tile (the FP64 band beyond the
// FP64-product band), 5 the same with one product per block where the panel
// rows are small (class 4 tiles beyond band_single, univariate kernels).
int tile_class(const egs_problem_s *p, int i, int j) {
    const int d = i - j;
    if (d <= p->band64) {
        if (p->mode != EGS_PRECISION_MIXED || d <= p->band_prod64) return 0;
        return (p->single_fp64 && p->band_single >= 0 && d > p->band_single) ? 5 : 4;
    }
    if (d <= p->band32) return 1;
    if (p->mode == EGS_PRECISION_MIXED && p->band_single >= 0 && d > p->band_single) return 3;
    return 2;
}

bool is_mixed(const egs_problem_s *p) { return p->mode != EGS_PRECISION_FP64; }

// Buffers of a mixed plan: class lists, panel copies, refinement vectors.
int setup_mixed(egs_problem_s *p) {
    const int nt = p->nt;
    for (int c = 0; c < kClasses; ++c) {
        p->h_pairs_cls[c].clear();
        p->cls_count[c].assign((size_t) nt + 1, 0);
    }
    for (int j = nt - 1; j >= 0; --j) {
        for (int i = j; i < nt; ++i) p->h_pairs_cls[tile_class(p, i, j)].push_back(int2{i, j});
        for (int c = 0; c < kClasses; ++c) p->cls_count[c][j] = (int) p->h_pairs_cls[c].size();
    }
    for (int c = 0; c < kClasses; ++c) {
        pfree(p, p->d_pairs_cls[c]);
        p->d_pairs_cls[c] = nullptr;
        const size_t bytes = sizeof(int2) * std::max<size_t>(1, p->h_pairs_cls[c].size());
        if (pmalloc(p, &p->d_pairs_cls[c], bytes) != hipSuccess)
            return fail(EGS_ERR_OUT_OF_MEMORY, "cannot allocate the update lists");
        if (!p->h_pairs_cls[c].empty())
            EGS_HIP(hipMemcpy(p->d_pairs_cls[c], p->h_pairs_cls[c].data(), sizeof(int2) * p->h_pairs_cls[c].size(),
                              hipMemcpyHostToDevice));
    }
    const size_t panel = (size_t) p->npad * p->ts;
    bool slots_ok = true;
    const bool need_slots = !p->d_p64[0];
    for (int b = 0; b < kPanelSlots && need_slots; ++b)
        slots_ok = slots_ok && pmalloc(p, &p->d_p64[b], panel * sizeof(double)) == hipSuccess &&
                   pmalloc(p, &p->d_p32[b], panel * sizeof(float)) == hipSuccess &&
                   pmalloc(p, &p->d_p16h[b], panel * sizeof(uint16_t)) == hipSuccess &&
                   pmalloc(p, &p->d_p16l[b], panel * sizeof(uint16_t)) == hipSuccess &&
                   pmalloc(p, &p->d_p8[b], panel) == hipSuccess &&
                   pmalloc(p, &p->d_pinv[b], sizeof(float) * (size_t) p->npad) == hipSuccess &&
                   pmalloc(p, &p->d_bmax[b], sizeof(float) * (size_t) (p->npad / 192)) == hipSuccess;
    if (!slots_ok) return fail(EGS_ERR_OUT_OF_MEMORY, "cannot allocate the mixed-precision panel copies");
    if (!p->d_cg) {
        if (pmalloc(p, &p->d_pmax, sizeof(unsigned) * (size_t) p->nt) != hipSuccess ||
            pmalloc(p, &p->d_counts, 3 * sizeof(unsigned long long)) != hipSuccess ||
            pmalloc(p, &p->d_cg, 5 * sizeof(double) * (size_t) p->npad) != hipSuccess ||
            pmalloc(p, &p->d_mv, sizeof(double) * (size_t) egs::matvec_chunks(p->n) * p->npad) != hipSuccess ||
            pmalloc(p, &p->d_dot, sizeof(double) * egs::dot_blocks()) != hipSuccess)
            return fail(EGS_ERR_OUT_OF_MEMORY, "cannot allocate the mixed-precision work buffers");
    }
    if (!p->d_tinv) {
        if (pmalloc(p, &p->d_tinv, sizeof(double) * (size_t) p->nt * p->ts * p->ts) != hipSuccess ||
            pmalloc(p, &p->d_ybuf, sizeof(double) * (size_t) p->npad) != hipSuccess) {
            free_tile_inverses(p);
            return fail(EGS_ERR_OUT_OF_MEMORY, "cannot allocate the diagonal-tile inverses");
        }
    }
    if (!p->d_sig64) {
        // Optional: without it Sigma v regenerates the covariance on the fly.
        const size_t ntiles = (size_t) p->nt * (p->nt + 1) / 2;
        const size_t tile = (size_t) p->ts * p->ts;
        if (pmalloc(p, &p->d_sig64, ntiles * tile * sizeof(double)) != hipSuccess ||
            pmalloc(p, &p->d_prow, ntiles * p->ts * sizeof(double)) != hipSuccess ||
            pmalloc(p, &p->d_pcol, ntiles * p->ts * sizeof(double)) != hipSuccess) {
            (void) hipGetLastError();
            free_sigma_copy(p);
        }
    }
    return EGS_SUCCESS;
}

// (Re)build the tile layout in HBM for the current plan.
int layout_tiles(egs_problem_s *p) {
    const int nt = p->nt;
    p->h_off.assign((size_t) nt * nt, -1);
    p->h_store.assign((size_t) nt * nt, 0);
    size_t off = 0;
    const size_t tile_elems = (size_t) p->ts * p->ts;
    for (int j = 0; j < nt; ++j) {
        for (int i = j; i < nt; ++i) {
            const uint8_t st = tile_store(p, i, j);
            p->h_store[(size_t) i * nt + j] = st;
            p->h_off[(size_t) i * nt + j] = (int64_t) off;
            off += tile_elems * (st == egs::kStoreF64 ? 8 : 4);
            off = (off + 255) & ~(size_t) 255;
        }
    }
    if (p->d_arena && p->arena_bytes != off) {
        pfree(p, p->d_arena);
        p->d_arena = nullptr;
    }
    if (!p->d_arena) {
        if (pmalloc(p, &p->d_arena, off) != hipSuccess) {
            p->d_arena = nullptr;
            return fail(EGS_ERR_OUT_OF_MEMORY, "cannot allocate %zu bytes for the covariance tiles", off);
        }
        p->arena_bytes = off;
    }
    EGS_HIP(hipMemcpy(p->d_off, p->h_off.data(), sizeof(int64_t) * p->h_off.size(), hipMemcpyHostToDevice));
    EGS_HIP(hipMemcpy(p->d_store, p->h_store.data(), p->h_store.size(), hipMemcpyHostToDevice));
    p->factored = false;
    p->has_cov = false;
    return EGS_SUCCESS;
}

int run_covariance(egs_problem_s *p, const double *theta) {
    egs::MaternParams prm;
    int st = make_params(p->kernel, p->metric, theta, prm);
    if (st) return st;
    const double *dz = (p->kernel == EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY) ? p->d_z : nullptr;
    const int pi = prof_begin(p, EGS_PROFILE_COVARIANCE, 0.0);
    egs::launch_covgen_tiles(p->map(), p->d_pairs, (int) p->h_pairs.size(), prm, p->d_x, p->d_y, dz, p->n,
                             is_mixed(p) ? p->d_sig64 : nullptr, p->ctx->stream);
    prof_end(p, pi);
    EGS_HIP(hipGetLastError());
    p->factored = false;
    p->has_cov = true;
    std::memcpy(p->cov_theta, theta, sizeof(double) * p->ntheta());
    return EGS_SUCCESS;
}

// Algorithmic flops of updating tile columns [j0, j1) with one panel (lower part).
double update_flops(const egs_problem_s *p, int j0, int j1) {
    const double ts = (double) p->ts;
    double f = 0.0;
    for (int j = j0; j < j1; ++j) {
        const double rj = real_rows(p, j);
        f += rj * (rj + 1.0) * ts;                                                  // SYRK on tile (j, j)
        for (int i = j + 1; i < p->nt; ++i) f += 2.0 * real_rows(p, i) * rj * ts;   // GEMM on tile (i, j)
    }
    return f;
}

// Diagonal tile k: 64 x 64 blocks left-looking, then the panel below it.
void factor_panel(egs_problem_s *p, int k) {
    hipStream_t s = p->ctx->stream;
    const egs::TileMap m = p->map();
    const int nt = p->nt, nsb = p->nsb;
    const double ts = (double) p->ts;
    double *wk = p->winv(k);
    const double kb = real_rows(p, k);
    int pi = prof_begin(p, EGS_PROFILE_DIAGONAL, kb * kb * kb / 3.0);
    egs::launch_factor_diag(m, k, wk, p->d_dscr, p->d_info, p->n, s);
    prof_end(p, pi);
    if (k + 1 < nt) {
        double rows_below = 0.0;
        for (int i = k + 1; i < nt; ++i) rows_below += real_rows(p, i);
        pi = prof_begin(p, EGS_PROFILE_PANEL_TRSM, rows_below * ts * ts);
        if (!is_mixed(p)) {
            egs::launch_trsm_rows(m, k, (int64_t) (k + 1) * nsb, (nt - k - 1) * nsb, 0, nsb, wk, s);
        } else {
            // rows within band64 of the diagonal are FP64 tiles, the rest FP32;
            // both in one launch when both are solved on FP64 MFMA
            const int last64 = std::min(nt - 1, k + p->band64);
            const int rows32 = std::max(0, nt - 1 - last64);
            if (!p->f32_panel_solve) {
                egs::launch_trsm_rows_mixed(m, k, (int64_t) (k + 1) * nsb, (last64 - k) * nsb, rows32 * nsb, wk, s);
            } else {
                if (last64 >= k + 1)
                    egs::launch_trsm_rows(m, k, (int64_t) (k + 1) * nsb, (last64 - k) * nsb, 0, nsb, wk, s, false);
                egs::launch_trsm_rows(m, k, (int64_t) (last64 + 1) * nsb, rows32 * nsb, 0, nsb, wk, s, true,
                                      p->f32_panel_solve);
            }
        }
        prof_end(p, pi);
        if (is_mixed(p)) {
            const bool split = p->mode == EGS_PRECISION_MIXED;
            const int sl = k % kPanelSlots;
            if (split) egs::launch_panel_rowscale(m, k, p->d_pinv[sl], p->d_bmax[sl], p->d_pmax, s);
            egs::launch_panel_convert(m, k, p->d_p64[sl], p->d_p32[sl], p->d_p16h[sl], p->d_p16l[sl],
                                      split && p->tau_fp8 > 0.0 ? p->d_p8[sl] : nullptr,
                                      split ? p->d_pinv[sl] : nullptr, s);
        }
    }
}

// Algorithmic flops of updating, with one panel, the class-c tiles (c < 0: all)
// of tile columns [j0, j1).
double update_flops_cls(const egs_problem_s *p, int c, int j0, int j1) {
    const double ts = (double) p->ts;
    double f = 0.0;
    for (int j = j0; j < j1; ++j) {
        const double rj = real_rows(p, j);
        for (int i = j; i < p->nt; ++i) {
            if (c >= 0 && tile_class(p, i, j) != c) continue;
            f += (i == j) ? rj * (rj + 1.0) * ts : 2.0 * real_rows(p, i) * rj * ts;
        }
    }
    return f;
}

// FP64 updates of at most this many tiles on the critical stream use 64 x 64
// workgroups: 225 per tile, each done in a ninth of the time of a 192 x 192 one.
constexpr int kShortUpdate = 8;

// Updates of step k under a mixed plan for tile columns [j0, nt) restricted to
// the column range encoded by the class prefix counts: entries [lo_c, hi_c) of
// each class list.  critical: the launches are on the critical path.
void launch_mixed_update(egs_problem_s *p, int k, int jcol_lo, int jcol_hi, hipStream_t st, bool critical = false) {
    const egs::TileMap m = p->map();
    static const int prof_id[kClasses] = {EGS_PROFILE_UPDATE, EGS_PROFILE_UPDATE_F32, EGS_PROFILE_UPDATE_16,
                                          EGS_PROFILE_UPDATE_16_SELECT, EGS_PROFILE_UPDATE_16, EGS_PROFILE_UPDATE_16};
    for (int c = 0; c < kClasses; ++c) {
        // class-c tiles with column in [jcol_lo, jcol_hi): list entries [count[jcol_hi], count[jcol_lo])
        const int beg = p->cls_count[c][jcol_hi];
        const int end = p->cls_count[c][jcol_lo];
        if (end <= beg) continue;
        const int pi = prof_begin(p, prof_id[c], update_flops_cls(p, c, jcol_lo, jcol_hi), st);
        const int2 *pr = p->d_pairs_cls[c] + beg;
        const int slot = k % kPanelSlots;
        if (c == 0 && critical && end - beg <= kShortUpdate)
            egs::launch_update64(m, p->d_p64[slot], pr, end - beg, k, st);
        else if (c == 0) egs::launch_update192_panel(m, p->d_p64[slot], pr, end - beg, k, st);
        else if (c == 1) egs::launch_update192_f32(m, p->d_p32[slot], pr, end - beg, k, st);
        else {
            const bool all_single = std::isinf(p->tau_single);
            const egs::Operands op =
                c == 5 ? (all_single ? egs::kSingleF16C64 : egs::kAdaptiveF16C64)
                : c == 4 ? egs::kSplitF16C64
                : c == 3 ? (all_single ? egs::kSingleF16 : egs::kAdaptiveF16)
                         : (p->mode == EGS_PRECISION_MIXED ? egs::kSplitF16 : egs::kSingleBF16);
            egs::launch_update192_16(m, p->d_p16h[slot], p->d_p16l[slot], p->d_pinv[slot], p->d_bmax[slot],
                                     p->d_pmax, (float) p->tau_single, p->d_p8[slot], (float) p->tau_fp8, pr,
                                     end - beg, k, op, p->d_counts, st);
        }
        prof_end(p, pi, st);
    }
}

// Enqueue the tiled Cholesky with a lookahead of one panel: the critical
// stream updates tile column k+1, factors it and solves its panel while the
// bulk stream applies panel k to the columns beyond k+1.  The update order
// lists tiles by descending column, so both tile sets are contiguous ranges of
// it.  The caller checks d_info after synchronizing the critical stream, which
// ends waiting for the bulk stream.
int run_potrf(egs_problem_s *p) {
    // EGS_NO_LOOKAHEAD=1 runs every launch on one stream (for measuring kernels
    // without the overlap of the two streams).
    static const bool no_lookahead = [] {
        const char *v = std::getenv("EGS_NO_LOOKAHEAD");
        return v && v[0] == '1';
    }();
    hipStream_t s = p->ctx->stream, b = no_lookahead ? p->ctx->stream : p->ctx->bulk;
    const egs::TileMap m = p->map();
    const int nt = p->nt;
    p->has_cov = false;   // the factorization overwrites the tiles
    for (int k = (int) p->ev_panel.size(); k < nt; ++k) {
        hipEvent_t e1 = nullptr, e2 = nullptr;
        EGS_HIP(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
        EGS_HIP(hipEventCreateWithFlags(&e2, hipEventDisableTiming));
        p->ev_panel.push_back(e1);
        p->ev_bulk.push_back(e2);
    }
    // a fill kernel, not a memset: this sequence is captured into a graph
    egs::launch_zero_u32(reinterpret_cast<unsigned *>(p->d_info), 1, s);
    if (p->mode == EGS_PRECISION_MIXED) {
        egs::launch_zero_u32(p->d_pmax, nt, s);
        egs::launch_zero_u32(reinterpret_cast<unsigned *>(p->d_counts), 6, s);
    }
    factor_panel(p, 0);
    EGS_HIP(hipEventRecord(p->ev_panel[0], s));
    for (int k = 0; k + 1 < nt; ++k) {
        const int beyond = nt - k - 2;                 // tile columns k+2 .. nt-1
        const int n_beyond = beyond * (beyond + 1) / 2;
        const int n_next = nt - k - 1;                 // tiles (i, k+1), i >= k+1
        // Bulk step k: column k+2 first, then ev_bulk[k], then columns k+3 ..:
        // the critical stream's step k+1 needs only column k+2, so it overlaps
        // the rest of this step.  (The list holds columns in descending order:
        // column k+2 is the last beyond - 1 ... entries of the prefix.)
        EGS_HIP(hipStreamWaitEvent(b, p->ev_panel[k], 0));
        if (is_mixed(p)) {
            if (k + 2 < nt) launch_mixed_update(p, k, k + 2, k + 3, b);
        } else if (n_beyond > 0) {
            const int pi = prof_begin(p, EGS_PROFILE_UPDATE, update_flops(p, k + 2, k + 3), b);
            egs::launch_update(m, p->d_pairs + (n_beyond - beyond), beyond, k, b);
            prof_end(p, pi, b);
        }
        EGS_HIP(hipEventRecord(p->ev_bulk[k], b));
        if (is_mixed(p)) {
            if (k + 3 < nt) launch_mixed_update(p, k, k + 3, nt, b);
        } else if (n_beyond > beyond) {
            const int pi = prof_begin(p, EGS_PROFILE_UPDATE, update_flops(p, k + 3, nt), b);
            egs::launch_update(m, p->d_pairs, n_beyond - beyond, k, b);
            prof_end(p, pi, b);
        }
        if (k > 0) EGS_HIP(hipStreamWaitEvent(s, p->ev_bulk[k - 1], 0));
        if (is_mixed(p)) {
            launch_mixed_update(p, k, k + 1, k + 2, s, true);
        } else {
            const int pi = prof_begin(p, EGS_PROFILE_UPDATE, update_flops(p, k + 1, k + 2), s);
            if (n_next <= kShortUpdate) egs::launch_update64(m, nullptr, p->d_pairs + n_beyond, n_next, k, s);
            else egs::launch_update(m, p->d_pairs + n_beyond, n_next, k, s);
            prof_end(p, pi, s);
        }
        factor_panel(p, k + 1);
        EGS_HIP(hipEventRecord(p->ev_panel[k + 1], s));
    }
    // ev_bulk[nt - 1] (unused by the steps) marks the end of the bulk stream;
    // a one-tile matrix never uses it (and a capture must not wait on an event
    // recorded outside it)
    if (nt >= 2) {
        EGS_HIP(hipEventRecord(p->ev_bulk[nt - 1], b));
        EGS_HIP(hipStreamWaitEvent(s, p->ev_bulk[nt - 1], 0));
    }
    if (is_mixed(p)) egs::launch_tinv(m, p->d_winv, p->d_tinv, s);
    EGS_HIP(hipGetLastError());
    return EGS_SUCCESS;
}

int read_info(egs_problem_s *p) {
    prof_collect(p);
    int info = 0;
    EGS_HIP(hipMemcpy(&info, p->d_info, sizeof(int), hipMemcpyDeviceToHost));
    if (info != 0) {
        p->factored = false;
        return fail(EGS_ERR_NOT_SPD, "dpotrf failed, Matrix is not positive definite (leading minor %d)", info);
    }
    return EGS_SUCCESS;
}

// Forward solve of d_work in place; per-tile log-det / dot partials go to
// `partials` (d_partials unless given).
void run_forward(egs_problem_s *p, double *partials = nullptr) {
    hipStream_t s = p->ctx->stream;
    const egs::TileMap m = p->map();
    if (!partials) partials = p->d_partials;
    const int pi = prof_begin(p, EGS_PROFILE_SOLVE, (double) p->n * (double) p->n);
    for (int k = 0; k < p->nt; ++k) {
        egs::launch_trsv_diag(m, k, p->winv(k), p->d_work, partials, s);
        egs::launch_fwd_panel(m, k, p->d_work + (int64_t) k * p->ts, p->d_work, s);
    }
    prof_end(p, pi);
}

void run_backward(egs_problem_s *p) {
    hipStream_t s = p->ctx->stream;
    const egs::TileMap m = p->map();
    for (int k = p->nt - 1; k >= 0; --k) {
        egs::launch_trsv_t_partial(m, k, p->d_work, p->d_aux, s);
        egs::launch_trsv_t_diag(m, k, p->winv(k), p->d_work, p->d_aux, s);
    }
}

// Preconditioner solves of mixed plans with the diagonal-tile inverses:
// d_ybuf = L^{-1} d_work (d_work is consumed), then d_ybuf = L^{-T} d_work.
void run_forward_inv(egs_problem_s *p) {
    hipStream_t s = p->ctx->stream;
    const egs::TileMap m = p->map();
    const int pi = prof_begin(p, EGS_PROFILE_SOLVE, (double) p->n * (double) p->n);
    for (int k = 0; k < p->nt; ++k) {
        const int64_t o = (int64_t) k * p->ts;
        egs::launch_fwd_diag_inv(p->d_tinv + o * p->ts, p->d_work + o, p->d_ybuf + o, p->ts, s);
        egs::launch_fwd_panel(m, k, p->d_ybuf + o, p->d_work, s);
    }
    prof_end(p, pi);
}

void run_backward_inv(egs_problem_s *p) {
    hipStream_t s = p->ctx->stream;
    const egs::TileMap m = p->map();
    const int pi = prof_begin(p, EGS_PROFILE_SOLVE, (double) p->n * (double) p->n);
    for (int k = p->nt - 1; k >= 0; --k) {
        const int64_t o = (int64_t) k * p->ts;
        egs::launch_bwd_diag_inv(p->d_tinv + o * p->ts, p->d_work + o, p->d_ybuf + o, p->ts, s);
        egs::launch_bwd_row(m, k, p->d_ybuf + o, p->d_work, s);
    }
    prof_end(p, pi);
}

// out = (L L^T)^{-1} in with the current factor (both vectors npad long).
void apply_factor_inverse(egs_problem_s *p, const double *in, double *out) {
    hipStream_t s = p->ctx->stream;
    const size_t bytes = sizeof(double) * (size_t) p->npad;
    hipMemcpyAsync(p->d_work, in, bytes, hipMemcpyDeviceToDevice, s);
    if (p->d_tinv) {
        run_forward_inv(p);
        hipMemcpyAsync(p->d_work, p->d_ybuf, bytes, hipMemcpyDeviceToDevice, s);
        run_backward_inv(p);
        hipMemcpyAsync(out, p->d_ybuf, bytes, hipMemcpyDeviceToDevice, s);
        return;
    }
    run_forward(p, p->d_mv);   // partials of a preconditioner solve are not needed
    run_backward(p);
    hipMemcpyAsync(out, p->d_work, bytes, hipMemcpyDeviceToDevice, s);
}

int host_dot(egs_problem_s *p, const double *x, const double *y, double &out) {
    egs::launch_dot(x, y, p->npad, p->d_dot, p->ctx->stream);
    double part[1024];
    const int nb = egs::dot_blocks();
    EGS_HIP(hipMemcpyAsync(part, p->d_dot, sizeof(double) * nb, hipMemcpyDeviceToHost, p->ctx->stream));
    EGS_HIP(hipStreamSynchronize(p->ctx->stream));
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) acc += part[b];
    out = acc;
    return EGS_SUCCESS;
}

// q = Sigma v in FP64 (covariance regenerated from the locations at the
// problem's current theta); q = base - Sigma v when base is given.
int sigma_times(egs_problem_s *p, const double *v, double *q, const double *base) {
    if (p->d_sig64) {
        // FP64 tiles written by the last covariance generation (at cov_theta)
        egs::launch_sigma_tile_matvec(p->d_sig64, p->d_pairs, (int) p->h_pairs.size(), p->ts, v, p->d_prow, p->d_pcol,
                                      p->ctx->stream);
        egs::launch_sigma_reduce(q, base, base ? -1.0 : 1.0, p->d_prow, p->d_pcol, p->nt, p->ts, p->n, p->npad,
                                 p->ctx->stream);
        EGS_HIP(hipGetLastError());
        return EGS_SUCCESS;
    }
    egs::MaternParams prm;
    int st = make_params(p->kernel, p->metric, p->cov_theta, prm);
    if (st) return st;
    const double *dz = (p->kernel == EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY) ? p->d_z : nullptr;
    egs::launch_cov_matvec(p->d_mv, v, p->n, p->npad, prm, p->d_x, p->d_y, dz, p->ctx->stream);
    egs::launch_matvec_reduce(q, base, base ? -1.0 : 1.0, p->d_mv, egs::matvec_chunks(p->n), p->n, p->npad,
                              p->ctx->stream);
    EGS_HIP(hipGetLastError());
    return EGS_SUCCESS;
}

// Iterative refinement of Sigma x = b for mixed plans: conjugate gradients
// preconditioned by the mixed factor, residuals b - Sigma x in FP64 with
// Sigma regenerated from the locations.  Stops at ||b - Sigma x|| <=
// tol ||b||.  x is left in d_cg[0 .. npad).
constexpr double kRefineTol = 1e-12;
constexpr int kRefineMaxIter = 60;
// A refinement whose true residual ends above this has not recovered the
// solve: the iteration broke down (p'Sigma p <= 0 in floating point, or a
// preconditioner far from Sigma) or ran out of iterations.
constexpr double kRefineAccept = 1e-10;

int refine_solve(egs_problem_s *p, const double *b) {
    hipStream_t s = p->ctx->stream;
    const int64_t np = p->npad;
    double *X = p->d_cg, *R = X + np, *P = R + np, *Q = P + np, *W = Q + np;
    const int pi = prof_begin(p, EGS_PROFILE_REFINE, 0.0);
    double bb = 0.0, rr = 0.0, rz = 0.0, pq = 0.0;
    int st = host_dot(p, b, b, bb);
    if (st) return st;
    apply_factor_inverse(p, b, X);
    // b' (L L^T)^{-1} b with the factor alone: the quadratic form before refinement
    if ((st = host_dot(p, b, X, p->quad_unrefined))) return st;
    st = sigma_times(p, X, R, b);
    if (st) return st;
    st = host_dot(p, R, R, rr);
    if (st) return st;
    int it = 0;
    if (bb > 0.0 && std::sqrt(rr / bb) > kRefineTol) {
        apply_factor_inverse(p, R, W);
        egs::launch_axpby(P, W, 1.0, 0.0, np, s);
        if ((st = host_dot(p, R, W, rz))) return st;
        for (it = 1; it <= kRefineMaxIter; ++it) {
            if ((st = sigma_times(p, P, Q, nullptr))) return st;
            if ((st = host_dot(p, P, Q, pq))) return st;
            if (!(pq > 0.0)) break;
            const double alpha = rz / pq;
            egs::launch_axpby(X, P, alpha, 1.0, np, s);
            egs::launch_axpby(R, Q, -alpha, 1.0, np, s);
            if ((st = host_dot(p, R, R, rr))) return st;
            if (std::sqrt(rr / bb) <= kRefineTol) break;
            apply_factor_inverse(p, R, W);
            double rz_new = 0.0;
            if ((st = host_dot(p, R, W, rz_new))) return st;
            const double beta = rz_new / rz;
            rz = rz_new;
            egs::launch_axpby(P, W, 1.0, beta, np, s);
        }
    }
    // true residual of the returned solution
    if ((st = sigma_times(p, X, R, b))) return st;
    if ((st = host_dot(p, R, R, rr))) return st;
    prof_end(p, pi);
    p->refine_iters = it;
    p->refine_resid = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
    p->refine_converged = p->refine_resid <= kRefineAccept;
    return EGS_SUCCESS;
}

int read_partials(egs_problem_s *p, double &logdet, double &quad) {
    p->h_partials.resize((size_t) 2 * p->nt);
    EGS_HIP(hipMemcpy(p->h_partials.data(), p->d_partials, sizeof(double) * 2 * p->nt, hipMemcpyDeviceToHost));
    double ld = 0.0, q = 0.0;
    for (int k = 0; k < p->nt; ++k) {
        ld += p->h_partials[2 * k];
        q += p->h_partials[2 * k + 1];
    }
    logdet = 2.0 * ld;
    quad = q;
    return EGS_SUCCESS;
}

// A covariance tile request of egs_covariance_tile: only the locations the
// tile touches are staged on the device, rows / P of the first set from
// row_off / P on and cols / P of the second from col_off / P on.
struct TileRequest {
    int rows = 0, cols = 0;
    int64_t s1 = 0, c1 = 0, s2 = 0, c2 = 0;   // first location and count of each set
    bool three_d = false;
    egs::MaternParams prm;
    // output tile, then x, y, z of each location slice
    int64_t scratch_doubles() const { return (int64_t) rows * cols + 3 * (c1 + c2); }
};

int check_tile_request(egs_context_t ctx, int kernel, const double *A, int rows, int cols, int row_off, int col_off,
                       const double *x1, const double *y1, const double *z1, int64_t n1, const double *x2,
                       const double *y2, const double *z2, int64_t n2, const double *theta, int metric,
                       TileRequest &t) {
    if (!ctx || !A || !x1 || !y1 || !x2 || !y2) return fail(EGS_ERR_INVALID_ARG, "null argument");
    if (rows < 0 || cols < 0 || row_off < 0 || col_off < 0) return fail(EGS_ERR_INVALID_ARG, "negative size");
    if (metric != 0 && metric != 1) return fail(EGS_ERR_INVALID_ARG, "unknown distance metric %d", metric);
    const int pvars = kernel == EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS ? 2 : 1;
    if (kernel == EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS && ((row_off | col_off) & 1))
        return fail(EGS_ERR_INVALID_ARG, "bivariate tiles need even row and column offsets");
    if ((int64_t) row_off + rows > n1 * pvars || (int64_t) col_off + cols > n2 * pvars)
        return fail(EGS_ERR_INVALID_ARG, "tile exceeds the location sets");
    t.three_d = (kernel == EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY) && z1 && z2;
    if (t.three_d && metric == 1) return fail(EGS_ERR_INVALID_ARG, "Great Circle (GC) distance is only valid for 2D!");
    int st = make_params(kernel, metric, theta, t.prm);
    if (st) return st;
    t.rows = rows;
    t.cols = cols;
    if (rows == 0 || cols == 0) return EGS_SUCCESS;
    // row_off and col_off are multiples of P, so the slices start at a
    // location and the tile's offsets inside them are 0
    t.s1 = row_off / pvars;
    t.c1 = ((int64_t) row_off + rows - 1) / pvars - t.s1 + 1;
    t.s2 = col_off / pvars;
    t.c2 = ((int64_t) col_off + cols - 1) / pvars - t.s2 + 1;
    return EGS_SUCCESS;
}

// Stages the location slices in `buf` (scratch_doubles() long), evaluates the
// tile into it and copies the tile to A; waits for the copy.
int covariance_tile_on(egs_context_t ctx, const TileRequest &t, double *A, const double *x1, const double *y1,
                       const double *z1, const double *x2, const double *y2, const double *z2, double *buf) {
    hipStream_t s = ctx->stream;
    double *d_out = buf;
    double *dx1 = buf + (int64_t) t.rows * t.cols, *dy1 = dx1 + t.c1, *dz1 = dy1 + t.c1;
    double *dx2 = dz1 + t.c1, *dy2 = dx2 + t.c2, *dz2 = dy2 + t.c2;
    const size_t nb1 = sizeof(double) * (size_t) t.c1, nb2 = sizeof(double) * (size_t) t.c2;
    hipError_t e = hipMemcpyAsync(dx1, x1 + t.s1, nb1, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dy1, y1 + t.s1, nb1, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dx2, x2 + t.s2, nb2, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dy2, y2 + t.s2, nb2, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && t.three_d) e = hipMemcpyAsync(dz1, z1 + t.s1, nb1, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && t.three_d) e = hipMemcpyAsync(dz2, z2 + t.s2, nb2, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        egs::launch_covgen_rect(d_out, t.rows, t.cols, 0, 0, t.prm, dx1, dy1, t.three_d ? dz1 : nullptr, dx2, dy2,
                                t.three_d ? dz2 : nullptr, s);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(A, d_out, sizeof(double) * (size_t) t.rows * t.cols, hipMemcpyDeviceToHost, s);
    // wait even after a failed enqueue: the caller may free buf next
    const hipError_t e2 = hipStreamSynchronize(s);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return fail(EGS_ERR_HIP, "covariance tile: %s", hipGetErrorString(e));
    return EGS_SUCCESS;
}

// Shape of a problem: tile size (dts rounded up to a multiple of 64), tile
// count and padded dimension.
void init_shape(egs_problem_s *p, int kernel, int64_t n_loc, int dim, int dts) {
    p->kernel = kernel;
    p->band_prod64 = kernel == EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS ? 2 : 0;
    p->single_fp64 = kernel != EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS;
    p->f32_panel_solve = kernel != EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS;
    p->tau_fp8 = kernel == EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS ? 0.0 : 0x1p-26;
    p->p = kernel == EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS ? 2 : 1;
    p->dim = dim;
    p->dts = dts;
    p->ts = std::min(((dts + egs::kBlk - 1) / egs::kBlk) * egs::kBlk, kMaxTileSize);
    p->nsb = p->ts / egs::kBlk;
    p->n_loc = n_loc;
    p->n = n_loc * p->p;
    p->nt = (int) ((p->n + p->ts - 1) / p->ts);
    p->npad = (int64_t) p->nt * p->ts;
}

// Device bytes a problem of q's shape and plan holds after an evaluation: the
// buffers of egs_problem_create, the tile arena, and under a mixed plan the
// buffers of setup_mixed (prediction and graph buffers, made on demand, are
// not included).  Each allocation counts with its size class.
int64_t footprint(const egs_problem_s &q) {
    using egs::dev_class_bytes;
    const int64_t nt = q.nt, ts = q.ts, npad = q.npad, ntiles = nt * (nt + 1) / 2;
    const int64_t nb = (int64_t) sizeof(double) * q.n_loc, vb = (int64_t) sizeof(double) * npad;
    int64_t b = dev_class_bytes(nb) * (q.dim == 3 ? 3 : 2) + dev_class_bytes(8 * nt * nt) + dev_class_bytes(nt * nt) +
                dev_class_bytes(8 * ntiles) + dev_class_bytes(8 * nt * q.nsb * egs::kBlk * egs::kBlk) +
                dev_class_bytes(8 * ts * ts) + 3 * dev_class_bytes(vb) + dev_class_bytes(16 * nt) + dev_class_bytes(4);
    int64_t arena = 0;
    for (int j = 0; j < q.nt; ++j)
        for (int i = j; i < q.nt; ++i) {
            arena += ts * ts * (tile_store(&q, i, j) == egs::kStoreF64 ? 8 : 4);
            arena = (arena + 255) & ~int64_t(255);
        }
    b += dev_class_bytes(arena);
    if (q.mode == EGS_PRECISION_FP64) return b;
    int64_t count[kClasses] = {};
    for (int j = 0; j < q.nt; ++j)
        for (int i = j; i < q.nt; ++i) ++count[tile_class(&q, i, j)];
    for (int c = 0; c < kClasses; ++c) b += dev_class_bytes(8 * std::max<int64_t>(1, count[c]));
    const int64_t panel = npad * ts;
    b += kPanelSlots * (dev_class_bytes(8 * panel) + dev_class_bytes(4 * panel) + 2 * dev_class_bytes(2 * panel) +
                        dev_class_bytes(panel) + dev_class_bytes(4 * npad) + dev_class_bytes(4 * (npad / 192)));
    b += dev_class_bytes(4 * nt) + dev_class_bytes(24) + dev_class_bytes(5 * vb) +
         dev_class_bytes(8 * (int64_t) egs::matvec_chunks(q.n) * npad) + dev_class_bytes(8 * egs::dot_blocks());
    b += dev_class_bytes(8 * nt * ts * ts) + dev_class_bytes(vb);
    b += dev_class_bytes(8 * ntiles * ts * ts) + 2 * dev_class_bytes(8 * ntiles * ts);
    return b;
}

uint64_t morton_spread(uint64_t v) {
    uint64_t out = 0;
    for (int b = 0; b < 16; ++b) out |= ((v >> b) & 1ull) << (4 * b);
    return out;
}

uint64_t morton_compact(uint64_t v) {
    uint64_t out = 0;
    for (int b = 0; b < 16; ++b) out |= ((v >> (4 * b)) & 1ull) << b;
    return out;
}

}  // namespace

extern "C" {

const char *egs_last_error(void) { return g_last_error.c_str(); }

int egs_version(void) { return 100; }

int egs_init(int device, egs_context_t *ctx) {
    if (!ctx) return fail(EGS_ERR_INVALID_ARG, "ctx is null");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(EGS_ERR_NO_DEVICE, "no HIP device is visible");
    if (device < 0 || device >= count) return fail(EGS_ERR_INVALID_ARG, "device %d out of range [0, %d)", device, count);
    if (device >= kMaxDevices) return fail(EGS_ERR_UNSUPPORTED, "device %d beyond the library's %d", device, kMaxDevices);
    EGS_HIP(hipSetDevice(device));
    const int sst = resolve_scratch(device);
    if (sst) return sst;
    auto *c = new egs_context_s();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->bulk, hipStreamNonBlocking) != hipSuccess) {
        if (c->stream) hipStreamDestroy(c->stream);
        if (c->bulk) hipStreamDestroy(c->bulk);
        delete c;
        return fail(EGS_ERR_HIP, "hipStreamCreate failed");
    }
    *ctx = c;
    return EGS_SUCCESS;
}

int egs_finalize(egs_context_t ctx) {
    if (!ctx) return EGS_SUCCESS;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    hipStreamSynchronize(ctx->bulk);
    // problems still alive lose their device memory; destroying them later is safe
    for (;;) {
        egs_problem_s *p = nullptr;
        {
            std::lock_guard<std::mutex> lk(ctx->mu);
            if (ctx->problems.empty()) break;
            p = ctx->problems.back();
        }
        free_problem_buffers(p);
        p->ctx = nullptr;
    }
    egs::dev_free(ctx->d_jobs, &ctx->dev_bytes);
    hipStreamDestroy(ctx->stream);
    hipStreamDestroy(ctx->bulk);
    delete ctx;
    return EGS_SUCCESS;
}

int egs_synchronize(egs_context_t ctx) {
    if (!ctx) return fail(EGS_ERR_INVALID_ARG, "ctx is null");
    EGS_HIP(hipSetDevice(ctx->device));
    EGS_HIP(hipStreamSynchronize(ctx->stream));
    EGS_HIP(hipStreamSynchronize(ctx->bulk));
    return EGS_SUCCESS;
}

int egs_covariance_tile(egs_context_t ctx, int kernel, double *A, int rows, int cols, int row_off, int col_off,
                        const double *x1, const double *y1, const double *z1, int64_t n1, const double *x2,
                        const double *y2, const double *z2, int64_t n2, const double *theta, int metric) {
    TileRequest t;
    int st = check_tile_request(ctx, kernel, A, rows, cols, row_off, col_off, x1, y1, z1, n1, x2, y2, z2, n2, theta,
                                metric, t);
    if (st || rows == 0 || cols == 0) return st;
    EGS_HIP(hipSetDevice(ctx->device));
    double *buf = nullptr;
    EGS_HIP(hipMalloc(&buf, sizeof(double) * (size_t) t.scratch_doubles()));
    st = covariance_tile_on(ctx, t, A, x1, y1, z1, x2, y2, z2, buf);
    const hipError_t e = hipFree(buf);
    if (st) return st;
    if (e != hipSuccess) return fail(EGS_ERR_HIP, "covariance tile: %s", hipGetErrorString(e));
    return EGS_SUCCESS;
}

int egs_covariance_tile_noalloc(egs_context_t ctx, int kernel, double *A, int rows, int cols, int row_off, int col_off,
                                const double *x1, const double *y1, const double *z1, int64_t n1, const double *x2,
                                const double *y2, const double *z2, int64_t n2, const double *theta, int metric) {
    TileRequest t;
    int st = check_tile_request(ctx, kernel, A, rows, cols, row_off, col_off, x1, y1, z1, n1, x2, y2, z2, n2, theta,
                                metric, t);
    if (st || rows == 0 || cols == 0) return st;
    if (t.scratch_doubles() > kScratchDoubles)
        return fail(EGS_ERR_INVALID_ARG,
                    "egs_covariance_tile_noalloc: the tile and its location slices need %lld doubles; the static "
                    "scratch holds %lld (rows * cols + 3 * (row and column locations))",
                    (long long) t.scratch_doubles(), (long long) kScratchDoubles);
    EGS_HIP(hipSetDevice(ctx->device));
    double *buf = g_scratch[ctx->device];   // resolved by egs_init on this device
    if (!buf) return fail(EGS_ERR_STATE, "egs_covariance_tile_noalloc: the device scratch is not resolved");
    // every element of the tile is written, so nothing in the scratch needs clearing
    return covariance_tile_on(ctx, t, A, x1, y1, z1, x2, y2, z2, buf);
}

int egs_problem_create(egs_context_t ctx, int kernel, int64_t n_loc, const double *x, const double *y,
                       const double *z, int metric, int dts, egs_problem_t *problem) {
    if (!ctx || !problem || !x || !y) return fail(EGS_ERR_INVALID_ARG, "null argument");
    if (kernel != EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY && kernel != EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS)
        return fail(EGS_ERR_INVALID_ARG, "unknown kernel %d", kernel);
    if (n_loc <= 0) return fail(EGS_ERR_INVALID_ARG, "n_loc must be positive");
    if (dts <= 0) return fail(EGS_ERR_INVALID_ARG, "dense tile size must be positive");
    if (metric != 0 && metric != 1) return fail(EGS_ERR_INVALID_ARG, "unknown distance metric %d", metric);
    if (kernel == EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY && z && metric == 1)
        return fail(EGS_ERR_INVALID_ARG, "Great Circle (GC) distance is only valid for 2D!");
    EGS_HIP(hipSetDevice(ctx->device));
    auto *p = new egs_problem_s();
    p->ctx = ctx;
    p->metric = metric;
    init_shape(p, kernel, n_loc, (z && kernel == EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY) ? 3 : 2, dts);
    const int nt = p->nt;
    for (int j = nt - 1; j >= 0; --j)
        for (int i = j; i < nt; ++i) p->h_pairs.push_back(int2{i, j});

    auto cleanup = [&](int code) {
        free_problem_buffers(p);
        delete p;
        return code;
    };
    const size_t nb = sizeof(double) * (size_t) n_loc;
    const size_t vb = sizeof(double) * (size_t) p->npad;
    bool ok = pmalloc(p, &p->d_x, nb) == hipSuccess && pmalloc(p, &p->d_y, nb) == hipSuccess &&
              (p->dim != 3 || pmalloc(p, &p->d_z, nb) == hipSuccess) &&
              pmalloc(p, &p->d_off, sizeof(int64_t) * (size_t) nt * nt) == hipSuccess &&
              pmalloc(p, &p->d_store, (size_t) nt * nt) == hipSuccess &&
              pmalloc(p, &p->d_pairs, sizeof(int2) * p->h_pairs.size()) == hipSuccess &&
              pmalloc(p, &p->d_winv, sizeof(double) * (size_t) nt * p->nsb * egs::kBlk * egs::kBlk) == hipSuccess &&
              pmalloc(p, &p->d_dscr, sizeof(double) * (size_t) p->ts * p->ts) == hipSuccess &&
              pmalloc(p, &p->d_obs, vb) == hipSuccess && pmalloc(p, &p->d_work, vb) == hipSuccess &&
              pmalloc(p, &p->d_aux, vb) == hipSuccess &&
              pmalloc(p, &p->d_partials, sizeof(double) * 2 * (size_t) nt) == hipSuccess &&
              pmalloc(p, &p->d_info, sizeof(int)) == hipSuccess;
    if (!ok) return cleanup(fail(EGS_ERR_OUT_OF_MEMORY, "cannot allocate the problem's device buffers"));
    for (auto &e : p->ev)
        if (hipEventCreate(&e) != hipSuccess) return cleanup(fail(EGS_ERR_HIP, "hipEventCreate failed"));
    if (hipMemcpy(p->d_x, x, nb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_y, y, nb, hipMemcpyHostToDevice) != hipSuccess ||
        (p->dim == 3 && hipMemcpy(p->d_z, z, nb, hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(p->d_pairs, p->h_pairs.data(), sizeof(int2) * p->h_pairs.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemset(p->d_obs, 0, vb) != hipSuccess)
        return cleanup(fail(EGS_ERR_HIP, "copying the locations to the device failed"));
    int st = layout_tiles(p);
    if (st) return cleanup(st);
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->problems.push_back(p);
    }
    *problem = p;
    return EGS_SUCCESS;
}

int egs_problem_set_context(egs_problem_t p, egs_context_t ctx) {
    int st = check_problem(p);
    if (st) return st;
    if (!ctx) return fail(EGS_ERR_INVALID_ARG, "ctx is null");
    if (ctx->device != p->ctx->device) return fail(EGS_ERR_INVALID_ARG, "the contexts are on different devices");
    if (ctx == p->ctx) return EGS_SUCCESS;
    // every entry point returns with the problem's work complete, so nothing of
    // it is pending on the old context's streams
    {
        std::lock_guard<std::mutex> lk(p->ctx->mu);
        auto &v = p->ctx->problems;
        v.erase(std::remove(v.begin(), v.end(), p), v.end());
    }
    invalidate_graph(p);   // recorded on the old context's streams
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->problems.push_back(p);
    }
    p->ctx = ctx;
    return EGS_SUCCESS;
}

int egs_problem_destroy(egs_problem_t p) {
    if (!p) return EGS_SUCCESS;
    if (p->ctx) {
        hipSetDevice(p->ctx->device);
        hipStreamSynchronize(p->ctx->stream);
        free_problem_buffers(p);
    }
    delete p;
    return EGS_SUCCESS;
}

int egs_problem_device_bytes(egs_problem_t p, int64_t *bytes) {
    int st = check_problem(p);
    if (st) return st;
    if (!bytes) return fail(EGS_ERR_INVALID_ARG, "bytes is null");
    *bytes = p->dev_bytes;
    return EGS_SUCCESS;
}

int egs_memory_cache(int64_t limit_bytes, int64_t *cached_bytes, int64_t *live_bytes) {
    int64_t c = 0, l = 0;
    egs::dev_cache(limit_bytes, &c, &l);
    if (cached_bytes) *cached_bytes = c;
    if (live_bytes) *live_bytes = l;
    return EGS_SUCCESS;
}

int egs_problem_estimate_bytes(int kernel, int64_t n_loc, int dim, int dts, int mode, int band_fp64, int band_fp32,
                               int64_t *bytes) {
    if (!bytes) return fail(EGS_ERR_INVALID_ARG, "bytes is null");
    if (kernel != EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY && kernel != EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS)
        return fail(EGS_ERR_INVALID_ARG, "unknown kernel %d", kernel);
    if (n_loc <= 0 || dts <= 0 || (dim != 2 && dim != 3)) return fail(EGS_ERR_INVALID_ARG, "invalid problem shape");
    if (mode != EGS_PRECISION_FP64 && mode != EGS_PRECISION_MIXED && mode != EGS_PRECISION_MIXED_BF16)
        return fail(EGS_ERR_INVALID_ARG, "unknown precision mode %d", mode);
    if (band_fp64 < 0 || band_fp32 < band_fp64) return fail(EGS_ERR_INVALID_ARG, "bands must satisfy 0 <= fp64 <= fp32");
    egs_problem_s q;
    init_shape(&q, kernel, n_loc, kernel == EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY ? dim : 2, dts);
    q.mode = mode;
    q.band64 = band_fp64;
    q.band32 = band_fp32;
    if (mode != EGS_PRECISION_FP64 && !egs::update192_supported(q.ts))
        return fail(EGS_ERR_UNSUPPORTED, "mixed precision needs a dense tile size that is a multiple of 192 (got %d)",
                    q.ts);
    *bytes = footprint(q);
    return EGS_SUCCESS;
}

int egs_problem_set_precision(egs_problem_t p, int mode, int band_fp64, int band_fp32) {
    int st = check_problem(p);
    if (st) return st;
    if (mode != EGS_PRECISION_FP64 && mode != EGS_PRECISION_MIXED && mode != EGS_PRECISION_MIXED_BF16)
        return fail(EGS_ERR_INVALID_ARG, "unknown precision mode %d", mode);
    if (band_fp64 < 0 || band_fp32 < band_fp64) return fail(EGS_ERR_INVALID_ARG, "bands must satisfy 0 <= fp64 <= fp32");
    if (mode != EGS_PRECISION_FP64 && !egs::update192_supported(p->ts))
        return fail(EGS_ERR_UNSUPPORTED, "mixed precision needs a dense tile size that is a multiple of 192 (got %d)",
                    p->ts);
    EGS_HIP(hipStreamSynchronize(p->ctx->stream));
    invalidate_graph(p);
    p->mode = mode;
    p->band64 = band_fp64;
    p->band32 = band_fp32;
    int st2 = layout_tiles(p);
    if (st2) return st2;
    if (!is_mixed(p)) {
        free_sigma_copy(p);
        free_tile_inverses(p);
    }
    return is_mixed(p) ? setup_mixed(p) : EGS_SUCCESS;
}

int egs_problem_set_graph(egs_problem_t p, int enable) {
    int st = check_problem(p);
    if (st) return st;
    EGS_HIP(hipStreamSynchronize(p->ctx->stream));
    invalidate_graph(p);
    p->use_graph = enable != 0;
    return EGS_SUCCESS;
}

int egs_problem_set_fp64_product_band(egs_problem_t p, int band) {
    int st = check_problem(p);
    if (st) return st;
    EGS_HIP(hipStreamSynchronize(p->ctx->stream));
    invalidate_graph(p);
    p->band_prod64 = band < 0 ? (1 << 30) : band;
    return is_mixed(p) ? setup_mixed(p) : EGS_SUCCESS;
}

int egs_problem_set_fp8_product(egs_problem_t p, double tau) {
    int st = check_problem(p);
    if (st) return st;
    if (std::isnan(tau) || tau < 0.0) return fail(EGS_ERR_INVALID_ARG, "tau must be >= 0 (got %g)", tau);
    EGS_HIP(hipStreamSynchronize(p->ctx->stream));
    invalidate_graph(p);
    p->tau_fp8 = tau;
    return EGS_SUCCESS;
}

int egs_problem_set_single_product(egs_problem_t p, int band, double tau) {
    int st = check_problem(p);
    if (st) return st;
    if (std::isnan(tau) || tau < 0.0) return fail(EGS_ERR_INVALID_ARG, "tau must be >= 0 (got %g)", tau);
    EGS_HIP(hipStreamSynchronize(p->ctx->stream));
    invalidate_graph(p);
    p->band_single = band < 0 ? -1 : band;
    p->tau_single = tau;
    return is_mixed(p) ? setup_mixed(p) : EGS_SUCCESS;
}

int egs_problem_set_observations(egs_problem_t p, const double *z, int64_t n) {
    int st = check_problem(p);
    if (st) return st;
    if (!z || n != p->n) return fail(EGS_ERR_INVALID_ARG, "observations must have n_loc * P = %lld entries", (long long) p->n);
    EGS_HIP(hipMemcpy(p->d_obs, z, sizeof(double) * (size_t) n, hipMemcpyHostToDevice));
    p->has_obs = true;
    return EGS_SUCCESS;
}

int egs_problem_get_observations(egs_problem_t p, double *z, int64_t n) {
    int st = check_problem(p);
    if (st) return st;
    if (!z || n != p->n) return fail(EGS_ERR_INVALID_ARG, "observations have n_loc * P = %lld entries", (long long) p->n);
    if (!p->has_obs) return fail(EGS_ERR_STATE, "the problem has no observations");
    EGS_HIP(hipMemcpy(z, p->d_obs, sizeof(double) * (size_t) n, hipMemcpyDeviceToHost));
    return EGS_SUCCESS;
}

int egs_covariance(egs_problem_t p, const double *theta) {
    int st = check_problem(p);
    if (st) return st;
    st = run_covariance(p, theta);
    if (st) return st;
    EGS_HIP(hipStreamSynchronize(p->ctx->stream));
    return EGS_SUCCESS;
}

int egs_potrf(egs_problem_t p) {
    int st = check_problem(p);
    if (st) return st;
    if (!p->has_cov) return fail(EGS_ERR_STATE, "no covariance matrix to factor: call egs_covariance first");
    st = run_potrf(p);
    if (st) return st;
    EGS_HIP(hipStreamSynchronize(p->ctx->stream));
    st = read_info(p);
    if (st) return st;
    p->factored = true;
    std::memcpy(p->factored_theta, p->cov_theta, sizeof(p->cov_theta));
    return EGS_SUCCESS;
}

int egs_solve_logdet_quad(egs_problem_t p, double *logdet, double *quad) {
    int st = check_problem(p);
    if (st) return st;
    if (!p->has_obs) return fail(EGS_ERR_STATE, "the problem has no observations");
    if (!p->factored) return fail(EGS_ERR_STATE, "no factored matrix: call egs_potrf first");
    hipStream_t s = p->ctx->stream;
    egs::launch_copy_pad(p->d_work, p->d_obs, p->n, p->npad, s);
    if (is_mixed(p)) egs::launch_diag_logdet(p->map(), p->n, p->d_partials, s);   // quad from the refinement
    else run_forward(p);
    EGS_HIP(hipGetLastError());
    EGS_HIP(hipStreamSynchronize(s));
    double ld, q;
    st = read_partials(p, ld, q);
    if (st) return st;
    if (is_mixed(p)) {
        if ((st = refine_solve(p, p->d_obs))) return st;
        if ((st = host_dot(p, p->d_obs, p->d_cg, q))) return st;
    }
    if (logdet) *logdet = ld;
    if (quad) *quad = q;
    return EGS_SUCCESS;
}

int egs_trsm_vector(egs_problem_t p, int trans, double alpha, double *z, int64_t n) {
    int st = check_problem(p);
    if (st) return st;
    if (!z || n != p->n) return fail(EGS_ERR_INVALID_ARG, "the vector must have n_loc * P = %lld entries", (long long) p->n);
    if (trans != 0 && trans != 1) return fail(EGS_ERR_INVALID_ARG, "trans must be 0 (L) or 1 (L^T)");
    if (!p->factored) return fail(EGS_ERR_STATE, "no factored matrix: call egs_potrf first");
    if (is_mixed(p))
        return fail(EGS_ERR_UNSUPPORTED, "the triangular solve with the stored factor runs under the FP64 plan");
    hipStream_t s = p->ctx->stream;
    EGS_HIP(hipMemcpyAsync(p->d_work, z, sizeof(double) * (size_t) n, hipMemcpyHostToDevice, s));
    egs::launch_copy_pad(p->d_work, p->d_work, p->n, p->npad, s);
    if (trans == 0) run_forward(p);
    else run_backward(p);
    EGS_HIP(hipGetLastError());
    EGS_HIP(hipMemcpyAsync(z, p->d_work, sizeof(double) * (size_t) n, hipMemcpyDeviceToHost, s));
    EGS_HIP(hipStreamSynchronize(s));
    if (alpha != 1.0)
        for (int64_t i = 0; i < n; ++i) z[i] *= alpha;
    return EGS_SUCCESS;
}

// One FP64 evaluation on the problem's streams, with the covariance parameters
// read from d_prm: observations copied, covariance, factorization, forward solve
// with the per-tile log-det and dot partials.  Everything it enqueues is
// capturable (no host synchronization, no memset).
int enqueue_fp64_eval(egs_problem_s *p) {
    hipStream_t s = p->ctx->stream;
    egs::launch_copy_pad(p->d_work, p->d_obs, p->n, p->npad, s);
    const double *dz = (p->kernel == EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY) ? p->d_z : nullptr;
    egs::launch_covgen_tiles(p->map(), p->d_pairs, (int) p->h_pairs.size(), p->h_prm, p->d_x, p->d_y, dz, p->n,
                             nullptr, s, p->d_prm);
    const int st = run_potrf(p);
    if (st) return st;
    run_forward(p);
    EGS_HIP(hipGetLastError());
    return EGS_SUCCESS;
}

// Captures the problem's FP64 evaluation into a graph (once per plan); theta
// reaches the graph through d_prm.
int graph_capture(egs_problem_s *p) {
    hipStream_t s = p->ctx->stream;
    if (!p->d_prm && pmalloc(p, &p->d_prm, sizeof(egs::MaternParams)) != hipSuccess)
        return fail(EGS_ERR_OUT_OF_MEMORY, "cannot allocate the covariance parameters");
    EGS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    const int st = enqueue_fp64_eval(p);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    if (st || ec != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return st ? st : fail(EGS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
    }
    p->graph = g;
    if (hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
        invalidate_graph(p);
        return fail(EGS_ERR_HIP, "graph instantiation failed");
    }
    return EGS_SUCCESS;
}

// Replays the captured evaluation for theta.
int graph_launch(egs_problem_s *p, const double *theta) {
    hipStream_t s = p->ctx->stream;
    const int st = make_params(p->kernel, p->metric, theta, p->h_prm);
    if (st) return st;
    // h_prm is not touched again before the stream synchronizes
    EGS_HIP(hipMemcpyAsync(p->d_prm, &p->h_prm, sizeof(egs::MaternParams), hipMemcpyHostToDevice, s));
    EGS_HIP(hipGraphLaunch(p->graph_exec, s));
    return EGS_SUCCESS;
}

// One evaluation under the problem's current plan (egs_loglik without the
// FP64 fallback of mixed plans).
static int loglik_once(egs_problem_s *p, const double *theta, double *loglik, double *logdet, double *quad) {
    int st = EGS_SUCCESS;
    if (!loglik) return fail(EGS_ERR_INVALID_ARG, "loglik is null");
    if (!p->has_obs) return fail(EGS_ERR_STATE, "the problem has no observations");
    hipStream_t s = p->ctx->stream;
    if (p->use_graph && !is_mixed(p) && !p->prof.on && !p->graph_exec) {
        // first use: capture (after checking theta, which the replay needs
        // valid: the capture itself leaves the tiles unchanged); if the
        // runtime refuses, evaluate directly from now on
        egs::MaternParams check;
        if ((st = make_params(p->kernel, p->metric, theta, check))) return st;
        st = graph_capture(p);
        if (st == EGS_ERR_HIP) {
            (void) hipGetLastError();
            invalidate_graph(p);
            p->use_graph = false;
        } else if (st) {
            return st;
        }
    }
    if (p->use_graph && !is_mixed(p) && !p->prof.on) {
        EGS_HIP(hipEventRecord(p->ev[0], s));
        st = graph_launch(p, theta);
        if (st) return st;
        p->has_cov = false;   // the replay factored the tiles
        EGS_HIP(hipEventRecord(p->ev[3], s));
        EGS_HIP(hipStreamSynchronize(s));
        std::memcpy(p->cov_theta, theta, sizeof(double) * p->ntheta());
        st = read_info(p);
        if (st) return st;
        p->factored = true;
        std::memcpy(p->factored_theta, theta, sizeof(double) * p->ntheta());
        float ms = 0;
        hipEventElapsedTime(&ms, p->ev[0], p->ev[3]);
        p->t_cov = p->t_solve = 0.0;   // one graph: only the total is timed
        p->t_fact = ms;
        double ld, q;
        st = read_partials(p, ld, q);
        if (st) return st;
        *loglik = -0.5 * q - 0.5 * ld - (double) p->n / 2.0 * std::log(2.0 * kPi);
        if (logdet) *logdet = ld;
        if (quad) *quad = q;
        return EGS_SUCCESS;
    }
    EGS_HIP(hipEventRecord(p->ev[0], s));
    egs::launch_copy_pad(p->d_work, p->d_obs, p->n, p->npad, s);
    st = run_covariance(p, theta);
    if (st) return st;
    EGS_HIP(hipEventRecord(p->ev[1], s));
    st = run_potrf(p);
    if (st) return st;
    EGS_HIP(hipEventRecord(p->ev[2], s));
    if (is_mixed(p)) egs::launch_diag_logdet(p->map(), p->n, p->d_partials, s);   // quad from the refinement
    else run_forward(p);
    EGS_HIP(hipGetLastError());
    EGS_HIP(hipEventRecord(p->ev[3], s));
    EGS_HIP(hipStreamSynchronize(s));
    st = read_info(p);
    if (st) return st;
    p->factored = true;
    std::memcpy(p->factored_theta, theta, sizeof(double) * p->ntheta());
    float ms = 0;
    hipEventElapsedTime(&ms, p->ev[0], p->ev[1]);
    p->t_cov = ms;
    hipEventElapsedTime(&ms, p->ev[1], p->ev[2]);
    p->t_fact = ms;
    hipEventElapsedTime(&ms, p->ev[2], p->ev[3]);
    p->t_solve = ms;
    double ld, q;
    st = read_partials(p, ld, q);
    if (st) return st;
    if (is_mixed(p)) {
        // z' Sigma^{-1} z from the refined FP64 solution instead of ||L^{-1} z||^2
        st = refine_solve(p, p->d_obs);
        if (st) return st;
        st = host_dot(p, p->d_obs, p->d_cg, q);
        if (st) return st;
        EGS_HIP(hipEventRecord(p->ev[3], s));
        EGS_HIP(hipStreamSynchronize(s));
        float ms2 = 0;
        hipEventElapsedTime(&ms2, p->ev[2], p->ev[3]);
        p->t_solve = ms2;
    }
    const double n = (double) p->n;
    *loglik = -0.5 * q - 0.5 * ld - (double) (n / 2.0) * std::log(2.0 * kPi);
    if (logdet) *logdet = ld;
    if (quad) *quad = q;
    return EGS_SUCCESS;
}

int egs_loglik_batch(egs_problem_t *problems, const double *thetas, int count, double *loglik, double *logdet,
                     double *quad, int *status) {
    if (!problems || !thetas || !loglik || !status || count <= 0)
        return fail(EGS_ERR_INVALID_ARG, "egs_loglik_batch: invalid arguments");
    egs_context_s *ctx = nullptr;
    int ts = 0;
    for (int b = 0; b < count; ++b) {
        egs_problem_s *p = problems[b];
        int st = check_problem(p);
        if (st) return st;
        if (!ctx) {
            ctx = p->ctx;
            ts = p->ts;
        }
        if (p->ctx != ctx) return fail(EGS_ERR_INVALID_ARG, "egs_loglik_batch: the problems must share one context");
        if (p->ts != ts) return fail(EGS_ERR_INVALID_ARG, "egs_loglik_batch: the problems must share one tile size");
        if (is_mixed(p)) return fail(EGS_ERR_UNSUPPORTED, "egs_loglik_batch: FP64 plans only");
        if (!p->has_obs) return fail(EGS_ERR_STATE, "egs_loglik_batch: problem %d has no observations", b);
        for (int c = 0; c < b; ++c)
            if (problems[c] == p) return fail(EGS_ERR_INVALID_ARG, "egs_loglik_batch: problem %d appears twice", b);
    }
    hipStream_t s = ctx->stream;
    // theta of each problem; a rejected theta fails that problem only
    std::vector<egs::MaternParams> prm(count);
    std::vector<int> run;
    for (int b = 0; b < count; ++b) {
        status[b] = make_params(problems[b]->kernel, problems[b]->metric, thetas + 6 * b, prm[b]);
        if (status[b] == EGS_SUCCESS) run.push_back(b);
    }
    // problems by descending tile count: at step k the active ones are a prefix
    std::stable_sort(run.begin(), run.end(), [&](int a, int c) { return problems[a]->nt > problems[c]->nt; });
    const int na = (int) run.size();
    const int max_nt = na ? problems[run[0]]->nt : 0;
    std::vector<int> active(max_nt + 1, 0);   // active[k]: problems with nt > k
    for (int k = 0; k <= max_nt; ++k)
        for (int b : run) active[k] += problems[b]->nt > k ? 1 : 0;
    std::vector<int64_t> first(max_nt + 1, 0);
    for (int k = 0; k < max_nt; ++k) first[k + 1] = first[k] + active[k];
    const int64_t njobs = first[max_nt];
    std::vector<egs::BatchJob> jobs((size_t) std::max<int64_t>(njobs, 1));
    for (int k = 0; k < max_nt; ++k)
        for (int a = 0; a < active[k]; ++a) {
            egs_problem_s *p = problems[run[a]];
            egs::BatchJob &J = jobs[(size_t) (first[k] + a)];
            J.map = p->map();
            J.k = k;
            J.pairs = p->d_pairs;
            J.winv_k = p->winv(k);
            J.scratch = p->d_dscr;
            J.info = p->d_info;
            J.n = p->n;
            J.z = p->d_work;
            J.partials = p->d_partials;
        }
    if (na > 0) {
        EGS_HIP(hipSetDevice(ctx->device));
        if (njobs > ctx->jobs_cap) {
            EGS_HIP(hipStreamSynchronize(s));
            egs::dev_free(ctx->d_jobs, &ctx->dev_bytes);
            ctx->d_jobs = nullptr;
            ctx->jobs_cap = 0;
            void *v = nullptr;
            if (egs::dev_alloc(&v, sizeof(egs::BatchJob) * (size_t) njobs, &ctx->dev_bytes) != hipSuccess)
                return fail(EGS_ERR_OUT_OF_MEMORY, "egs_loglik_batch: cannot allocate the job records");
            ctx->d_jobs = static_cast<egs::BatchJob *>(v);
            ctx->jobs_cap = njobs;
        }
        // the previous batch on this context has completed (every call returns synchronized)
        EGS_HIP(hipMemcpyAsync(ctx->d_jobs, jobs.data(), sizeof(egs::BatchJob) * (size_t) njobs,
                               hipMemcpyHostToDevice, s));
        for (int b : run) {
            egs_problem_s *p = problems[b];
            egs::launch_copy_pad(p->d_work, p->d_obs, p->n, p->npad, s);
            const double *dz = (p->kernel == EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY) ? p->d_z : nullptr;
            egs::launch_covgen_tiles(p->map(), p->d_pairs, (int) p->h_pairs.size(), prm[b], p->d_x, p->d_y, dz, p->n,
                                     nullptr, s);
            egs::launch_zero_u32(reinterpret_cast<unsigned *>(p->d_info), 1, s);
            p->has_cov = false;
            p->factored = false;
            invalidate_graph(p);
        }
        // lockstep right-looking factorization: one launch per kernel and step for all problems
        for (int k = 0; k < max_nt; ++k) {
            const egs::BatchJob *J = ctx->d_jobs + first[k];
            egs::launch_factor_diag_batch(J, active[k], ts, s);
            egs::launch_panel_update_batch(J, active[k + 1], ts, max_nt - k - 1, s);
        }
        for (int k = 0; k < max_nt; ++k)
            egs::launch_forward_batch(ctx->d_jobs + first[k], active[k], ts, max_nt - k - 1, s);
        EGS_HIP(hipGetLastError());
        EGS_HIP(hipStreamSynchronize(s));
    }
    for (int b : run) {
        egs_problem_s *p = problems[b];
        prof_collect(p);
        int info = 0;
        double ld = 0.0, q = 0.0;
        if (hipMemcpy(&info, p->d_info, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            status[b] = fail(EGS_ERR_HIP, "egs_loglik_batch: reading the factorization status failed");
            continue;
        }
        if (info != 0) {
            status[b] = fail(EGS_ERR_NOT_SPD, "dpotrf failed, Matrix is not positive definite (leading minor %d)", info);
            continue;
        }
        if ((status[b] = read_partials(p, ld, q)) != EGS_SUCCESS) continue;
        p->factored = true;
        std::memcpy(p->factored_theta, thetas + 6 * b, sizeof(double) * p->ntheta());
        std::memcpy(p->cov_theta, thetas + 6 * b, sizeof(double) * p->ntheta());
        loglik[b] = -0.5 * q - 0.5 * ld - (double) p->n / 2.0 * std::log(2.0 * kPi);
        if (logdet) logdet[b] = ld;
        if (quad) quad[b] = q;
    }
    return EGS_SUCCESS;
}

int egs_loglik(egs_problem_t p, const double *theta, double *loglik, double *logdet, double *quad) {
    int st = check_problem(p);
    if (st) return st;
    p->fell_back = false;
    double q_local = 0.0;
    if (!quad) quad = &q_local;
    st = loglik_once(p, theta, loglik, logdet, quad);
    if (!is_mixed(p)) return st;
    // The mixed factor gives log|Sigma| directly, and the refinement corrects
    // only z'Sigma^{-1}z; a refinement that needs many iterations means the
    // factor is far from Sigma's (smooth fields), so its log-determinant is
    // not trusted.  Such an evaluation, one whose refinement did not converge
    // (its quadratic form is not recovered, and q0 - q then says nothing), and
    // one whose mixed factor is not positive definite, is repeated under the
    // FP64 plan.
    // With w = L^{-1} z and S = L^{-1} Sigma L^{-T} - I, the mixed factor's
    // log-determinant is off by log det(I + S) ~ tr S, and q0 - q = w'w -
    // w'(I + S)^{-1} w ~ w'S w, whose mean is tr S when w ~ N(0, I) (z drawn
    // from the model): one probe of the log-determinant error that the
    // refinement already paid for.
    if (st == EGS_SUCCESS && quad)
        p->error_estimate = std::fabs(p->quad_unrefined - *quad) / (2.0 * std::max(std::fabs(*loglik), 1.0));
    else if (st == EGS_SUCCESS)
        p->error_estimate = 0.0;
    if (p->fallback_iters <= 0 && p->fallback_tol <= 0.0) return st;   // the guard is off
    const bool poor = st == EGS_SUCCESS && (!p->refine_converged ||
                                            (p->fallback_iters > 0 && p->refine_iters > p->fallback_iters) ||
                                            (p->fallback_tol > 0.0 && p->error_estimate > p->fallback_tol));
    if (!poor && st != EGS_ERR_NOT_SPD) return st;
    const int mode = p->mode, b64 = p->band64, b32 = p->band32;
    if ((st = egs_problem_set_precision(p, EGS_PRECISION_FP64, b64, b32))) return st;
    const int st_eval = loglik_once(p, theta, loglik, logdet, quad);
    const std::string err = g_last_error;
    if ((st = egs_problem_set_precision(p, mode, b64, b32))) return st;   // the plan again, not factored
    p->fell_back = true;
    ++p->fallbacks;
    if (st_eval) g_last_error = err;
    return st_eval;
}

int egs_problem_set_fallback(egs_problem_t p, int max_iterations, double tolerance) {
    int st = check_problem(p);
    if (st) return st;
    if (std::isnan(tolerance)) return fail(EGS_ERR_INVALID_ARG, "the fallback tolerance is NaN");
    p->fallback_iters = max_iterations < 0 ? 0 : max_iterations;
    p->fallback_tol = tolerance < 0.0 ? 0.0 : tolerance;
    return EGS_SUCCESS;
}

int egs_last_fallback(egs_problem_t p, int *fell_back, int64_t *total, double *error_estimate) {
    int st = check_problem(p);
    if (st) return st;
    if (fell_back) *fell_back = p->fell_back ? 1 : 0;
    if (total) *total = p->fallbacks;
    if (error_estimate) *error_estimate = p->error_estimate;
    return EGS_SUCCESS;
}

int egs_problem_profile(egs_problem_t p, int enable) {
    int st = check_problem(p);
    if (st) return st;
    EGS_HIP(hipStreamSynchronize(p->ctx->stream));
    prof_collect(p);
    p->prof.on = enable != 0;
    for (auto &k : p->prof.stats) k = KernelStat();
    return EGS_SUCCESS;
}

int egs_problem_kernel_stats(egs_problem_t p, int kernel, int64_t *launches, double *total_ms, double *flops) {
    int st = check_problem(p);
    if (st) return st;
    if (kernel < 0 || kernel >= EGS_PROFILE_KERNELS) return fail(EGS_ERR_INVALID_ARG, "unknown kernel id %d", kernel);
    const KernelStat &k = p->prof.stats[kernel];
    if (launches) *launches = k.launches;
    if (total_ms) *total_ms = k.ms;
    if (flops) *flops = k.flops;
    return EGS_SUCCESS;
}

int egs_last_refinement(egs_problem_t p, int *iterations, double *residual) {
    int st = check_problem(p);
    if (st) return st;
    if (iterations) *iterations = p->refine_iters;
    if (residual) *residual = p->refine_resid;
    return EGS_SUCCESS;
}

int egs_last_product_counts(egs_problem_t p, int64_t *split_blocks, int64_t *single_blocks) {
    return egs_last_product_counts_fp8(p, split_blocks, single_blocks, nullptr);
}

int egs_last_product_counts_fp8(egs_problem_t p, int64_t *split_blocks, int64_t *single_blocks, int64_t *fp8_blocks) {
    int st = check_problem(p);
    if (st) return st;
    unsigned long long c[3] = {0, 0, 0};
    if (p->mode == EGS_PRECISION_MIXED && p->d_counts) {
        EGS_HIP(hipStreamSynchronize(p->ctx->stream));
        EGS_HIP(hipMemcpy(c, p->d_counts, sizeof(c), hipMemcpyDeviceToHost));
    }
    if (split_blocks) *split_blocks = (int64_t) c[0];
    if (single_blocks) *single_blocks = (int64_t) c[1];
    if (fp8_blocks) *fp8_blocks = (int64_t) c[2];
    return EGS_SUCCESS;
}

int egs_last_timings(egs_problem_t p, double *t_cov, double *t_fact, double *t_solve) {
    int st = check_problem(p);
    if (st) return st;
    if (t_cov) *t_cov = p->t_cov;
    if (t_fact) *t_fact = p->t_fact;
    if (t_solve) *t_solve = p->t_solve;
    return EGS_SUCCESS;
}

int egs_get_matrix(egs_problem_t p, double *out, int64_t ld) {
    int st = check_problem(p);
    if (st) return st;
    if (!out || ld < p->n) return fail(EGS_ERR_INVALID_ARG, "output must be n x n with ld >= n");
    EGS_HIP(hipStreamSynchronize(p->ctx->stream));
    const int ts = p->ts;
    std::vector<char> buf((size_t) ts * ts * 8);
    for (int j = 0; j < p->nt; ++j) {
        for (int i = j; i < p->nt; ++i) {
            const size_t idx = (size_t) i * p->nt + j;
            const bool f64 = p->h_store[idx] == egs::kStoreF64;
            EGS_HIP(hipMemcpy(buf.data(), p->d_arena + p->h_off[idx], (size_t) ts * ts * (f64 ? 8 : 4),
                              hipMemcpyDeviceToHost));
            for (int c = 0; c < ts; ++c) {
                const int64_t gc = (int64_t) j * ts + c;
                if (gc >= p->n) break;
                for (int r = 0; r < ts; ++r) {
                    const int64_t gr = (int64_t) i * ts + r;
                    if (gr >= p->n) break;
                    if (gr < gc) continue;
                    const size_t e = (size_t) r + (size_t) c * ts;
                    out[gr + gc * ld] = f64 ? reinterpret_cast<double *>(buf.data())[e]
                                            : (double) reinterpret_cast<float *>(buf.data())[e];
                }
            }
        }
    }
    return EGS_SUCCESS;
}

int egs_generate_locations(int64_t n, int dim, double *x, double *y, double *z) {
    // LocationGenerator.cpp:25-123
    if (n <= 0 || !x || !y || (dim != 2 && dim != 3) || (dim == 3 && !z))
        return fail(EGS_ERR_INVALID_ARG, "invalid location request");
    const int root = dim == 3 ? (int) std::ceil(std::cbrt((double) n)) : (int) std::ceil(std::sqrt((double) n));
    auto uni = []() { return ((double) std::rand() / (1.0 + RAND_MAX)) * 0.8 - 0.4; };
    int64_t idx = 0;
    for (int i = 0; i < root && idx < n; ++i) {
        for (int j = 0; j < root && idx < n; ++j) {
            if (dim == 3) {
                for (int k = 0; k < root && idx < n; ++k) {
                    x[idx] = ((i + 1) - 0.5 + uni()) / root;
                    y[idx] = ((j + 1) - 0.5 + uni()) / root;
                    z[idx] = ((k + 1) - 0.5 + uni()) / root;
                    ++idx;
                }
            } else {
                x[idx] = ((i + 1) - 0.5 + uni()) / root;
                y[idx] = ((j + 1) - 0.5 + uni()) / root;
                ++idx;
            }
        }
    }
    std::vector<uint64_t> code((size_t) n);
    for (int64_t i = 0; i < n; ++i) {
        const uint16_t qx = (uint16_t) (x[i] * 65535.0 + .5);
        const uint16_t qy = (uint16_t) (y[i] * 65535.0 + .5);
        const uint16_t qz = dim == 3 ? (uint16_t) (z[i] * 65535.0 + .5) : 0;
        code[i] = (morton_spread(qz) << 2) + (morton_spread(qy) << 1) + morton_spread(qx);
    }
    std::sort(code.begin(), code.end());
    for (int64_t i = 0; i < n; ++i) {
        x[i] = (double) morton_compact(code[i]) / 65535.0;
        y[i] = (double) morton_compact(code[i] >> 1) / 65535.0;
        if (dim == 3) z[i] = (double) morton_compact(code[i] >> 2) / 65535.0;
    }
    return EGS_SUCCESS;
}

static int generate_observations_fp64(egs_problem_s *p, const double *theta, int seed, double *z_out);

int egs_generate_observations(egs_problem_t p, const double *theta, int seed, double *z_out) {
    int st = check_problem(p);
    if (st) return st;
    if (!is_mixed(p)) return generate_observations_fp64(p, theta, seed, z_out);
    // the data comes from the FP64 factor whatever the plan: FP64 for the
    // generation, then the plan again (unfactored)
    const int mode = p->mode, b64 = p->band64, b32 = p->band32;
    if ((st = egs_problem_set_precision(p, EGS_PRECISION_FP64, b64, b32))) return st;
    const int st_gen = generate_observations_fp64(p, theta, seed, z_out);
    const std::string err = g_last_error;
    if ((st = egs_problem_set_precision(p, mode, b64, b32))) return st;
    if (st_gen) g_last_error = err;
    return st_gen;
}

static int generate_observations_fp64(egs_problem_s *p, const double *theta, int seed, double *z_out) {
    int st = EGS_SUCCESS;
    hipStream_t s = p->ctx->stream;
    // iseed = {seed, seed, seed, 1} (LinearAlgebraMethods.cpp:348) as one 48-bit word
    const uint64_t sd = (uint64_t) (uint32_t) seed;
    const uint64_t seed48 = (((sd * 4096u + sd) * 4096u + sd) * 4096u + 1u) & ((1ull << 48) - 1);
    egs::launch_dlarnv_normal(p->d_aux, p->n, seed48, s);
    egs::launch_copy_pad(p->d_aux, p->d_aux, p->n, p->npad, s);
    st = run_covariance(p, theta);
    if (st) return st;
    st = run_potrf(p);
    if (st) return st;
    egs::launch_trmv(p->map(), p->d_aux, p->d_work, s);
    EGS_HIP(hipGetLastError());
    EGS_HIP(hipStreamSynchronize(s));
    st = read_info(p);
    if (st) return st;
    p->factored = true;
    std::memcpy(p->factored_theta, theta, sizeof(double) * p->ntheta());
    EGS_HIP(hipMemcpy(p->d_obs, p->d_work, sizeof(double) * (size_t) p->n, hipMemcpyDeviceToDevice));
    p->has_obs = true;
    if (z_out) EGS_HIP(hipMemcpy(z_out, p->d_obs, sizeof(double) * (size_t) p->n, hipMemcpyDeviceToHost));
    return EGS_SUCCESS;
}

// predict_once's status when a mixed plan's refinement did not converge and
// the guard is on (egs_predict_mspe then repeats the prediction in FP64)
constexpr int kRefineNotConverged = -1;

static bool guard_on(const egs_problem_s *p) { return p->fallback_iters > 0 || p->fallback_tol > 0.0; }

static int predict_once(egs_problem_s *p, const double *theta, int64_t n_miss, const double *xm, const double *ym,
                        const double *zm, double *zmiss, const double *z_actual, double *mspe3) {
    int st = EGS_SUCCESS;
    if (!theta || !xm || !ym || !zmiss || n_miss <= 0) return fail(EGS_ERR_INVALID_ARG, "invalid prediction request");
    if (!p->has_obs) return fail(EGS_ERR_STATE, "the problem has no observations");
    const bool three_d = p->dim == 3;
    if (three_d && !zm) return fail(EGS_ERR_INVALID_ARG, "3D problem needs z coordinates of the targets");
    egs::MaternParams prm;
    st = make_params(p->kernel, p->metric, theta, prm);
    if (st) return st;
    hipStream_t s = p->ctx->stream;
    // targets and predictions share one buffer that the problem keeps (grown
    // on demand), so repeated predictions allocate nothing
    const int64_t rows = n_miss * p->p;
    const int64_t need = rows + 3 * n_miss;
    if (need > p->pred_cap) {
        EGS_HIP(hipStreamSynchronize(s));
        pfree(p, p->d_pred);
        p->d_pred = nullptr;
        p->pred_cap = 0;
        if (pmalloc(p, &p->d_pred, sizeof(double) * (size_t) need) != hipSuccess) {
            p->d_pred = nullptr;
            return fail(EGS_ERR_OUT_OF_MEMORY, "cannot allocate the prediction buffer");
        }
        p->pred_cap = need;
    }
    const bool reuse = p->factored && std::memcmp(p->factored_theta, theta, sizeof(double) * p->ntheta()) == 0;
    if (!reuse) {
        st = run_covariance(p, theta);
        if (st) return st;
        st = run_potrf(p);
        if (st) return st;
    }
    const double *weights = p->d_work;
    if (is_mixed(p)) {
        st = refine_solve(p, p->d_obs);
        if (st) return st;
        if (!p->refine_converged && guard_on(p)) {
            if (!reuse && (st = read_info(p))) return st;   // NOT_SPD takes the same fallback
            return kRefineNotConverged;
        }
        weights = p->d_cg;
    } else {
        egs::launch_copy_pad(p->d_work, p->d_obs, p->n, p->npad, s);
        run_forward(p);
        run_backward(p);
    }
    double *d_out = p->d_pred, *dxm = d_out + rows, *dym = dxm + n_miss, *dzm = dym + n_miss;
    hipError_t e = hipMemcpyAsync(dxm, xm, sizeof(double) * n_miss, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dym, ym, sizeof(double) * n_miss, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && three_d) e = hipMemcpyAsync(dzm, zm, sizeof(double) * n_miss, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        egs::launch_cov_gemv(d_out, rows, p->n, prm, dxm, dym, three_d ? dzm : nullptr, p->d_x, p->d_y,
                             three_d ? p->d_z : nullptr, weights, s);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(zmiss, d_out, sizeof(double) * (size_t) rows, hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return fail(EGS_ERR_HIP, "prediction: %s", hipGetErrorString(e));
    if (!reuse) {
        st = read_info(p);
        if (st) return st;
        p->factored = true;
        std::memcpy(p->factored_theta, theta, sizeof(double) * p->ntheta());
    }
    if (mspe3) {
        if (z_actual) {
            // LinearAlgebraMethods.cpp:558-600 with the dmse and dmse-bivariate
            // codelets: all values, then the first (even entries) and second
            // (odd entries) variable of the bivariate kernel, 0 otherwise
            double acc = 0.0, acc1 = 0.0, acc2 = 0.0;
            for (int64_t i = 0; i < rows; ++i) {
                const double d = (z_actual[i] - zmiss[i]) * (z_actual[i] - zmiss[i]);
                acc += d;
                if (i & 1) acc2 += d;
                else acc1 += d;
            }
            mspe3[0] = acc / (double) rows;
            mspe3[1] = p->p == 2 ? acc1 / (double) n_miss : 0.0;
            mspe3[2] = p->p == 2 ? acc2 / (double) n_miss : 0.0;
        } else {
            mspe3[0] = -1.0;
            mspe3[1] = mspe3[2] = 0.0;
        }
    }
    return EGS_SUCCESS;
}

int egs_predict_mspe(egs_problem_t p, const double *theta, int64_t n_miss, const double *xm, const double *ym,
                     const double *zm, double *zmiss, const double *z_actual, double *mspe3) {
    int st = check_problem(p);
    if (st) return st;
    p->fell_back = false;
    st = predict_once(p, theta, n_miss, xm, ym, zm, zmiss, z_actual, mspe3);
    if ((st != EGS_ERR_NOT_SPD && st != kRefineNotConverged) || !is_mixed(p) || !guard_on(p)) return st;
    // a mixed factor that is not positive definite, or whose refinement did
    // not converge: the prediction under the FP64 plan, then the plan again
    // (unfactored), as egs_loglik does
    const int mode = p->mode, b64 = p->band64, b32 = p->band32;
    if ((st = egs_problem_set_precision(p, EGS_PRECISION_FP64, b64, b32))) return st;
    const int st_pred = predict_once(p, theta, n_miss, xm, ym, zm, zmiss, z_actual, mspe3);
    const std::string err = g_last_error;
    if ((st = egs_problem_set_precision(p, mode, b64, b32))) return st;
    p->fell_back = true;
    ++p->fallbacks;
    if (st_pred) g_last_error = err;
    return st_pred;
}

int egs_predict(egs_problem_t p, const double *theta, int64_t n_miss, const double *xm, const double *ym,
                const double *zm, double *zmiss, const double *z_actual, double *mspe) {
    double m3[3] = {-1.0, 0.0, 0.0};
    const int st = egs_predict_mspe(p, theta, n_miss, xm, ym, zm, zmiss, z_actual, m3);
    if (st == EGS_SUCCESS && mspe) *mspe = m3[0];
    return st;
}

int egs_bobyqa(int n, double *x, const double *lb, const double *ub, double ftol_abs, int maxeval, double rho_end,
               egs_objective_fn fn, void *user, double *fmin, int *nevals, int *reason) {
    if (n < 1 || n > 64 || !x || !lb || !ub || !fn || maxeval < 1)
        return fail(EGS_ERR_INVALID_ARG, "egs_bobyqa: invalid arguments (n=%d, maxeval=%d)", n, maxeval);
    for (int i = 0; i < n; ++i)
        if (!(std::isfinite(lb[i]) && std::isfinite(ub[i]) && ub[i] > lb[i]))
            return fail(EGS_ERR_INVALID_ARG, "egs_bobyqa: bound %d is not a finite interval [%g, %g]", i, lb[i], ub[i]);
    egs::BobyqaOptions opt;
    opt.ftol_abs = ftol_abs > 0.0 ? ftol_abs : 0.0;
    opt.maxeval = maxeval;
    if (rho_end > 0.0) opt.rho_end = rho_end;
    const egs::BobyqaResult r = egs::bobyqa_minimize(
        n, x, lb, ub, [&](const double *v) { return fn(v, user); }, opt);
    if (fmin) *fmin = r.fmin;
    if (nevals) *nevals = r.nevals;
    if (reason) *reason = static_cast<int>(r.status);
    if (r.status == egs::BobyqaStatus::kFailure) return fail(EGS_ERR_INVALID_ARG, "egs_bobyqa: the objective returned NaN");
    return EGS_SUCCESS;
}

int egs_data_modeling(egs_problem_t p, const double *lb, const double *ub, const double *start, double tolerance,
                      int max_iterations, double *theta, double *loglik, int *nevals) {
    int st = check_problem(p);
    if (st) return st;
    if (!lb || !ub || !theta || !loglik) return fail(EGS_ERR_INVALID_ARG, "egs_data_modeling: null argument");
    if (max_iterations < 1) return fail(EGS_ERR_INVALID_ARG, "egs_data_modeling: max_iterations must be >= 1");
    if (!p->has_obs) return fail(EGS_ERR_STATE, "the problem has no observations");
    const int nt = p->ntheta();
    std::vector<double> lo(lb, lb + nt), hi(ub, ub + nt), x(start ? start : lb, (start ? start : lb) + nt);
    // A parameter with equal bounds is held fixed (Configurations::InitializeAllTheta
    // pins estimated parameters this way, Configurations.cpp:204-211).
    std::vector<int> free_idx;
    for (int i = 0; i < nt; ++i) {
        if (!(std::isfinite(lo[i]) && std::isfinite(hi[i]) && hi[i] >= lo[i]))
            return fail(EGS_ERR_INVALID_ARG, "egs_data_modeling: bound %d is not a finite interval [%g, %g]", i, lo[i], hi[i]);
        if (hi[i] > lo[i]) free_idx.push_back(i);
        else x[i] = lo[i];
    }
    int eval_status = EGS_SUCCESS;
    std::string eval_error;
    double best = -INFINITY;
    int count = 0;
    auto llh_at = [&](const std::vector<double> &th) -> double {
        double v = 0.0;
        const int s2 = egs_loglik(p, th.data(), &v, nullptr, nullptr);
        ++count;
        if (s2 != EGS_SUCCESS) {
            if (eval_status == EGS_SUCCESS) {
                eval_status = s2;
                eval_error = g_last_error;
            }
            return NAN;
        }
        if (v > best) {
            best = v;
            std::copy(th.begin(), th.end(), theta);
        }
        return v;
    };
    const int nf = static_cast<int>(free_idx.size());
    // one BOBYQA search from `start` with the given initial step, at most budget evaluations
    auto search = [&](const std::vector<double> &start, double step_scale, int budget) {
        std::vector<double> xf(nf), lf(nf), hf(nf);
        for (int j = 0; j < nf; ++j) {
            xf[j] = start[free_idx[j]];
            lf[j] = lo[free_idx[j]];
            hf[j] = hi[free_idx[j]];
        }
        egs::BobyqaOptions opt;
        opt.ftol_abs = tolerance > 0.0 ? tolerance : 0.0;
        opt.maxeval = budget;
        opt.step_scale = step_scale;
        std::vector<double> th(start);
        egs::bobyqa_minimize(
            nf, xf.data(), lf.data(), hf.data(),
            [&](const double *v) {
                for (int j = 0; j < nf; ++j) th[free_idx[j]] = v[j];
                return -llh_at(th);
            },
            opt);
    };
    if (free_idx.empty()) llh_at(x);
    else search(x, 0.25, max_iterations);
    // Under a mixed plan the search ends with an FP64 polish: the mixed
    // log-likelihood carries an error of order 1e-7 relative that varies with
    // theta, which can stop BOBYQA early along the flat ridges of smooth fields;
    // a second search with the FP64 plan from the mixed optimum (smaller first
    // steps) finds the FP64 optimum.  The reported theta and value are FP64's.
    if (is_mixed(p) && nf > 0 && eval_status == EGS_SUCCESS && count < max_iterations) {
        const int mode = p->mode, b64 = p->band64, b32 = p->band32;
        int st_plan = egs_problem_set_precision(p, EGS_PRECISION_FP64, b64, b32);
        if (st_plan) return st_plan;
        const std::vector<double> start(theta, theta + nt);
        best = -INFINITY;
        search(start, 0.05, max_iterations - count);
        if ((st_plan = egs_problem_set_precision(p, mode, b64, b32))) return st_plan;
    }
    if (nevals) *nevals = count;
    if (eval_status != EGS_SUCCESS) {
        g_last_error = eval_error;
        return eval_status;
    }
    *loglik = best;
    return EGS_SUCCESS;
}

}  // extern "C"
