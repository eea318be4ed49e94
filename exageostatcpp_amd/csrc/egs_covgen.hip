// Covariance generation on gfx950: the device side of
// Kernel<T>::GenerateCovarianceMatrix (Kernel.hpp:93-97) for every lower tile of
// Sigma at once (the reference submits one dcmg task per tile,
// dcmg-codelet.cpp:47-74, and evaluates GSL's K_nu per element on CPU threads).
//
// One workgroup writes one 64 x 64 block of one tile.  The 64 row locations and
// 64 column locations are staged in LDS; each wave then writes whole 64-element
// column segments of the column-major tile (512 contiguous bytes per store
// instruction in FP64).  Arithmetic is FP64 throughout; the store converts to
// the tile's storage precision.
#include "egs_common.hpp"
#include "egs_kernels.hpp"
#include "egs_matern.hpp"

namespace egs {

struct CovTilesArgs {
    TileMap map;
    const int2 *pairs;      // lower tiles (i, j) to generate
    MaternParams prm;
    Locs loc;
    int64_t n;              // true matrix dimension; rows/cols >= n are identity padding
    int three_d;
    double *sig64;          // optional FP64 copy of every lower tile (tile i*(i+1)/2 + j, ts*ts each)
    const MaternParams *prm_dev;   // optional: read the parameters here instead of prm
};

// One workgroup per 64-row strip of a tile (all ts columns): the strip's row
// locations and the tile's column locations (3 ts doubles of dynamic LDS) are
// staged once, then each wave writes 64 consecutive rows (512 bytes in FP64)
// of one column at a time.

__global__ void __launch_bounds__(kThreads) covgen_tiles_kernel(CovTilesArgs a) {
    __shared__ double s_rx[kBlk], s_ry[kBlk], s_rz[kBlk];
    extern __shared__ double s_cols[];
    // parameters from device memory when given (a captured graph replays with new theta)
    const MaternParams prm = a.prm_dev ? *a.prm_dev : a.prm;
    const int ts = a.map.ts;
    double *s_cx = s_cols, *s_cy = s_cols + ts, *s_cz = s_cols + 2 * ts;
    const int2 ij = a.pairs[blockIdx.y];
    const int sbr = blockIdx.x;
    const int64_t r0 = (int64_t) ij.x * ts + sbr * kBlk;
    const int64_t c0 = (int64_t) ij.y * ts;
    const int t = threadIdx.x;
    const bool three_d = a.three_d != 0;
    const int shift = prm.kernel == 1 ? 1 : 0;

    // Stage the locations of this strip's rows and of the tile's columns.
    if (t < kBlk) {
        const int64_t li = (r0 + t) >> shift;
        const bool ok = (r0 + t) < a.n;
        s_rx[t] = ok ? a.loc.x[li] : 0.0;
        s_ry[t] = ok ? a.loc.y[li] : 0.0;
        s_rz[t] = (ok && three_d) ? a.loc.z[li] : 0.0;
    }
    for (int u = t; u < ts; u += kThreads) {
        const int64_t lj = (c0 + u) >> shift;
        const bool ok = (c0 + u) < a.n;
        s_cx[u] = ok ? a.loc.x[lj] : 0.0;
        s_cy[u] = ok ? a.loc.y[lj] : 0.0;
        s_cz[u] = (ok && three_d) ? a.loc.z[lj] : 0.0;
    }
    __syncthreads();

    const int r = t % kBlk;
    const int64_t gr = r0 + r;
    const bool row_pad = gr >= a.n;
    const double rx = s_rx[r], ry = s_ry[r], rz = s_rz[r];
    const int rvar = (int) (gr & 1);
    const uint8_t store = a.map.store[ij.x * a.map.nt + ij.y];
    char *tile = a.map.base + a.map.off[ij.x * a.map.nt + ij.y];
    const int64_t sig_base = ((int64_t) ij.x * (ij.x + 1) / 2 + ij.y) * ts * ts;
    for (int cc = t / kBlk; cc < ts; cc += kThreads / kBlk) {
        const int64_t gc = c0 + cc;
        double v;
        if (row_pad || gc >= a.n) {
            v = (gr == gc) ? 1.0 : 0.0;
        } else {
            const int pair = shift ? (rvar * 2 + (int) (gc & 1)) : 0;
            const double dist = distance(rx, ry, rz, s_cx[cc], s_cy[cc], s_cz[cc], prm.metric, three_d);
            const double d = dist / prm.range;
            v = (d == 0.0) ? prm.sig[pair] : matern_scaled(d, prm.con[pair], prm.sig[pair], prm.nu[pair]);
        }
        const int64_t idx = (int64_t) (sbr * kBlk + r) + (int64_t) cc * ts;
        if (store == kStoreF64) reinterpret_cast<double *>(tile)[idx] = v;
        else reinterpret_cast<float *>(tile)[idx] = (float) v;
        if (a.sig64) a.sig64[sig_base + idx] = v;
    }
}

// A rows x cols column-major block (ld = rows) of Sigma(l1, l2) starting at
// matrix row row_off and column col_off: the plugin interface
// Kernel<T>::GenerateCovarianceMatrix evaluated on the device.
struct CovRectArgs {
    double *out;
    int rows, cols;
    int64_t row_off, col_off;
    MaternParams prm;
    Locs l1, l2;
    int three_d;
};

__global__ void __launch_bounds__(kThreads) covgen_rect_kernel(CovRectArgs a) {
    const int64_t idx = (int64_t) blockIdx.x * kThreads + threadIdx.x;
    const int64_t total = (int64_t) a.rows * a.cols;
    if (idx >= total) return;
    const int i = (int) (idx % a.rows);
    const int j = (int) (idx / a.rows);
    a.out[idx] = covariance_entry(a.prm, a.l1, a.l2, a.row_off + i, a.col_off + j, a.three_d != 0);
}

void launch_covgen_tiles(const TileMap &map, const int2 *pairs, int npairs, const MaternParams &prm,
                         const double *x, const double *y, const double *z, int64_t n, double *sig64, hipStream_t s,
                         const MaternParams *prm_dev) {
    if (npairs <= 0) return;
    CovTilesArgs a;
    a.map = map;
    a.pairs = pairs;
    a.prm = prm;
    a.loc = Locs{x, y, z};
    a.n = n;
    a.three_d = z != nullptr;
    a.sig64 = sig64;
    a.prm_dev = prm_dev;
    const int nsb = map.ts / kBlk;
    dim3 grid(nsb, npairs);
    hipLaunchKernelGGL(covgen_tiles_kernel, grid, dim3(kThreads), 3 * sizeof(double) * (size_t) map.ts, s, a);
}

void launch_covgen_rect(double *out, int rows, int cols, int64_t row_off, int64_t col_off, const MaternParams &prm,
                        const double *x1, const double *y1, const double *z1, const double *x2, const double *y2,
                        const double *z2, hipStream_t s) {
    CovRectArgs a;
    a.out = out;
    a.rows = rows;
    a.cols = cols;
    a.row_off = row_off;
    a.col_off = col_off;
    a.prm = prm;
    a.l1 = Locs{x1, y1, z1};
    a.l2 = Locs{x2, y2, z2};
    a.three_d = (z1 != nullptr && z2 != nullptr);
    const int64_t total = (int64_t) rows * cols;
    if (total <= 0) return;
    const unsigned blocks = (unsigned) ((total + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(covgen_rect_kernel, dim3(blocks), dim3(kThreads), 0, s, a);
}

}  // namespace egs
