// Triangular solves, log-determinant / quadratic-form partials, kriging product
// and the synthetic-data helpers on gfx950.
//
// The forward solve replaces CHAMELEON_dtrsm_Tile(Left, Lower, NoTrans) on Z
// (ChameleonImplementation.cpp:216-217, 337-347) and fuses the per-tile
// reductions the reference runs as dmdet (dmdet-codelet.cpp:70-80) and ddotp
// (ddotp-codelet.cpp:62-72) tasks: the workgroup that solves diagonal tile k
// also writes sum(log L_ii) over L_ii > 0 and y_k . y_k for that tile.  The host
// sums the per-tile partials in tile order, so results do not depend on
// scheduling.  All reductions inside a workgroup use a fixed order.
#include "egs_common.hpp"
#include "egs_kernels.hpp"
#include "egs_matern.hpp"

namespace egs {

constexpr int kMaxTs = 2048;

// Fixed-order sum of one value per thread over the workgroup.
__device__ double wg_sum(double v, double *red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------------
// Forward solve of diagonal tile k, with the log-det and quad partials.  The
// 64-row blocks of the tile are solved in order; each block's GEMV over the
// already-solved part is spread over 16 waves (column groups) with four
// independent partial sums per thread so that many loads are in flight.
// ------------------------------------------------------------------------
constexpr int kSolveThreads = 1024;
constexpr int kGroups = kSolveThreads / kBlk;   // 16

__device__ double wg_sum_1024(double v, double *red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = kSolveThreads / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ void trsv_diag_body(const TileMap &map, int k, const double *winv_k, double *z,
                                               double *partials) {
    __shared__ double sy[kMaxTs];
    __shared__ double sp[kGroups][kBlk];
    __shared__ double st[kBlk];
    __shared__ double red[kSolveThreads];
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int t = threadIdx.x;
    const int r = t & 63, g = t >> 6;
    const double *L = tile_f64(map, k, k);
    double *zk = z + (int64_t) k * ts;
    for (int e = t; e < ts; e += kSolveThreads) sy[e] = zk[e];
    __syncthreads();
    for (int jb = 0; jb < nsb; ++jb) {
        // t = z_jb - L[jb, 0:jb] y[0:jb]
        const double *Lr = L + jb * kBlk + r;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        const int cend = jb * kBlk;
        int c = g;
        for (; c + 3 * kGroups < cend; c += 4 * kGroups) {
            s0 += Lr[(int64_t) c * ts] * sy[c];
            s1 += Lr[(int64_t) (c + kGroups) * ts] * sy[c + kGroups];
            s2 += Lr[(int64_t) (c + 2 * kGroups) * ts] * sy[c + 2 * kGroups];
            s3 += Lr[(int64_t) (c + 3 * kGroups) * ts] * sy[c + 3 * kGroups];
        }
        for (; c < cend; c += kGroups) s0 += Lr[(int64_t) c * ts] * sy[c];
        sp[g][r] = (s0 + s1) + (s2 + s3);
        __syncthreads();
        if (t < kBlk) {
            double a = 0.0;
            for (int q = 0; q < kGroups; ++q) a += sp[q][t];
            st[t] = sy[jb * kBlk + t] - a;
        }
        __syncthreads();
        // y_jb = W_jb t  (W lower triangular), 64 x 64 over 16 groups
        const double *W = winv_k + (int64_t) jb * kBlk * kBlk;
        double u = 0.0;
        for (int cc = g; cc <= r; cc += kGroups) u += W[r + cc * kBlk] * st[cc];
        sp[g][r] = u;
        __syncthreads();
        if (t < kBlk) {
            double a = 0.0;
            for (int q = 0; q < kGroups; ++q) a += sp[q][t];
            sy[jb * kBlk + t] = a;
        }
        __syncthreads();
    }
    double ld = 0.0, q = 0.0;
    for (int e = t; e < ts; e += kSolveThreads) {
        const double d = L[e + (int64_t) e * ts];
        if (d > 0.0) ld += log(d);
        q += sy[e] * sy[e];
        zk[e] = sy[e];
    }
    ld = wg_sum_1024(ld, red);
    q = wg_sum_1024(q, red);
    if (t == 0) {
        partials[2 * k] = ld;
        partials[2 * k + 1] = q;
    }
}

__global__ void __launch_bounds__(kSolveThreads) trsv_diag_kernel(TileMap map, int k, const double *winv_k, double *z,
                                                                  double *partials) {
    trsv_diag_body(map, k, winv_k, z, partials);
}

__global__ void __launch_bounds__(kSolveThreads) trsv_diag_batch_kernel(const BatchJob *jobs) {
    const BatchJob &J = jobs[blockIdx.x];
    trsv_diag_body(J.map, J.k, J.winv_k, J.z, J.partials);
}

// scratch[(i - k - 1) * ts + c] = sum_r L_ik[r][c] y_i[r] for every tile row i > k.
__global__ void __launch_bounds__(kThreads) trsv_t_partial_kernel(TileMap map, int k, const double *z,
                                                                  double *scratch) {
    __shared__ double sy[kMaxTs];
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int i = k + 1 + (int) (blockIdx.x / nsb);
    const int cb = (int) (blockIdx.x % nsb);
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const double *yi = z + (int64_t) i * ts;
    for (int e = t; e < ts; e += kThreads) sy[e] = yi[e];
    __syncthreads();
    const bool f32 = map.store[i * map.nt + k] == kStoreF32;
    const double *A = tile_f64(map, i, k);
    const float *Af = tile_f32(map, i, k);
    // wave w reduces columns cb*64 + w*16 .. +15, lanes stride the rows
    for (int cc = 0; cc < 16; ++cc) {
        const int c = cb * kBlk + w * 16 + cc;
        double s = 0.0;
        if (f32) {
            for (int r0 = lane; r0 < ts; r0 += 64) s += (double) Af[r0 + (int64_t) c * ts] * sy[r0];
        } else {
            for (int r0 = lane; r0 < ts; r0 += 64) s += A[r0 + (int64_t) c * ts] * sy[r0];
        }
        // fixed-order butterfly over the 64 lanes
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) scratch[(int64_t) (i - k - 1) * ts + c] = s;
    }
}

// Backward solve of diagonal tile k: y_k = L_kk^{-T} (z_k - sum_i partials_i),
// 64-column blocks from the last one back, 16 waves per block GEMV.
__global__ void __launch_bounds__(kSolveThreads) trsv_t_diag_kernel(TileMap map, int k, const double *winv_k,
                                                                    double *z, const double *scratch) {
    __shared__ double sy[kMaxTs];
    __shared__ double sp[kGroups][kBlk];
    __shared__ double st[kBlk];
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int nt = map.nt;
    const int t = threadIdx.x;
    const int c = t & 63, g = t >> 6;
    const double *L = tile_f64(map, k, k);
    double *zk = z + (int64_t) k * ts;
    for (int e = t; e < ts; e += kSolveThreads) {
        double v = zk[e];
        for (int i = k + 1; i < nt; ++i) v -= scratch[(int64_t) (i - k - 1) * ts + e];
        sy[e] = v;
    }
    __syncthreads();
    for (int jb = nsb - 1; jb >= 0; --jb) {
        // t[c] = sy[jb*64 + c] - sum_{r >= (jb+1)*64} L[r][jb*64 + c] y[r]; wave g takes rows r = base + g + 16 m
        const double *Lc = L + (int64_t) (jb * kBlk + c) * ts;
        double s0 = 0.0, s1 = 0.0;
        const int rbeg = (jb + 1) * kBlk;
        int rr = rbeg + g;
        for (; rr + kGroups < ts; rr += 2 * kGroups) {
            s0 += Lc[rr] * sy[rr];
            s1 += Lc[rr + kGroups] * sy[rr + kGroups];
        }
        for (; rr < ts; rr += kGroups) s0 += Lc[rr] * sy[rr];
        sp[g][c] = s0 + s1;
        __syncthreads();
        if (t < kBlk) {
            double a = 0.0;
            for (int q = 0; q < kGroups; ++q) a += sp[q][t];
            st[t] = sy[jb * kBlk + t] - a;
        }
        __syncthreads();
        // y[c] = sum_{r >= c} W[r][c] t[r]
        const double *W = winv_k + (int64_t) jb * kBlk * kBlk;
        double u = 0.0;
        for (int r2 = c + g; r2 < kBlk; r2 += kGroups) u += W[r2 + c * kBlk] * st[r2];
        sp[g][c] = u;
        __syncthreads();
        if (t < kBlk) {
            double a = 0.0;
            for (int q = 0; q < kGroups; ++q) a += sp[q][t];
            sy[jb * kBlk + t] = a;
        }
        __syncthreads();
    }
    for (int e = t; e < ts; e += kSolveThreads) zk[e] = sy[e];
}

// out = L e, one workgroup per 64 rows.
__global__ void __launch_bounds__(kThreads) trmv_kernel(TileMap map, const double *e, double *out) {
    __shared__ double sp[4][kBlk];
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int64_t gblk = blockIdx.x;
    const int i = (int) (gblk / nsb);
    const int rb = (int) (gblk % nsb);
    const int t = threadIdx.x;
    const int r = t & 63, g = t >> 6;
    double s = 0.0;
    for (int k = 0; k <= i; ++k) {
        const double *ek = e + (int64_t) k * ts;
        const int cend = (k < i) ? ts : (rb + 1) * kBlk;
        const int row_in_tile = rb * kBlk + r;
        if (map.store[i * map.nt + k] == kStoreF32) {
            const float *A = tile_f32(map, i, k) + rb * kBlk;
            for (int c = g; c < cend; c += 4) s += (double) A[r + (int64_t) c * ts] * ek[c];
        } else {
            const double *A = tile_f64(map, i, k) + rb * kBlk;
            for (int c = g; c < cend; c += 4) {
                if (k == i && c > row_in_tile) continue;
                s += A[r + (int64_t) c * ts] * ek[c];
            }
        }
    }
    sp[g][r] = s;
    __syncthreads();
    if (t < kBlk) out[gblk * kBlk + t] = ((sp[0][t] + sp[1][t]) + sp[2][t]) + sp[3][t];
}

// Kriging mean: out[r] = sum_c Sigma(l1_r, l2_c) x[c].
struct CovGemvArgs {
    double *out;
    int64_t rows, cols;
    MaternParams prm;
    Locs l1, l2;
    const double *x;
    int three_d;
};

__global__ void __launch_bounds__(kThreads) cov_gemv_kernel(CovGemvArgs a) {
    __shared__ double red[kThreads];
    const int64_t r = blockIdx.x;
    double s = 0.0;
    for (int64_t c = threadIdx.x; c < a.cols; c += kThreads)
        s += covariance_entry(a.prm, a.l1, a.l2, r, c, a.three_d != 0) * a.x[c];
    s = wg_sum(s, red);
    if (threadIdx.x == 0) a.out[r] = s;
}

// dlarnv(idist = 3): value m uses the uniforms number 2m and 2m+1 of the
// stream u_j = (s0 * A^(j+1) mod 2^48) / 2^48 with A = 33952834046453.
__device__ __forceinline__ uint64_t lcg_pow(uint64_t e) {
    const uint64_t mask = (1ull << 48) - 1;
    uint64_t base = 33952834046453ull, r = 1;
    while (e) {
        if (e & 1) r = (r * base) & mask;
        base = (base * base) & mask;
        e >>= 1;
    }
    return r;
}

__global__ void __launch_bounds__(kThreads) dlarnv_normal_kernel(double *e, int64_t n, uint64_t seed48) {
    const int64_t m = (int64_t) blockIdx.x * kThreads + threadIdx.x;
    if (m >= n) return;
    const uint64_t mask = (1ull << 48) - 1;
    const uint64_t s1 = (seed48 * lcg_pow((uint64_t) (2 * m + 1))) & mask;
    const uint64_t s2 = (s1 * 33952834046453ull) & mask;
    const double u1 = (double) s1 * 0x1p-48;
    const double u2 = (double) s2 * 0x1p-48;
    e[m] = sqrt(-2.0 * log(u1)) * cos(6.2831853071795864769252867663 * u2);
}

__global__ void __launch_bounds__(kThreads) copy_pad_kernel(double *dst, const double *src, int64_t n, int64_t npad) {
    const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
    if (i < npad) dst[i] = (i < n) ? src[i] : 0.0;
}

// ------------------------------------------------------------------------
// Matrix-free Sigma v in FP64 for the refinement residual: the covariance is
// regenerated from the locations, never stored.  Workgroup (row block rb,
// column chunk cb) writes partial[cb * npad + row] = sum over the chunk's
// columns; reduce kernels add the chunks in a fixed order.
// ------------------------------------------------------------------------
constexpr int kChunk = 2048;

struct MatvecArgs {
    double *partial;
    const double *v;
    int64_t n, npad;
    MaternParams prm;
    Locs loc;
    int three_d;
};

__global__ void __launch_bounds__(kThreads) cov_matvec_kernel(MatvecArgs a) {
    __shared__ double sx[kChunk], sy[kChunk], sz[kChunk], sv[kChunk];
    __shared__ double sp[4][kBlk];
    const int t = threadIdx.x;
    const int r = t & 63, g = t >> 6;
    const int64_t row = (int64_t) blockIdx.x * kBlk + r;
    const int64_t c0 = (int64_t) blockIdx.y * kChunk;
    const int shift = a.prm.kernel == 1 ? 1 : 0;
    const bool three_d = a.three_d != 0;
    for (int e = t; e < kChunk; e += kThreads) {
        const int64_t c = c0 + e;
        const bool ok = c < a.n;
        const int64_t l = c >> shift;
        sx[e] = ok ? a.loc.x[l] : 0.0;
        sy[e] = ok ? a.loc.y[l] : 0.0;
        sz[e] = (ok && three_d) ? a.loc.z[l] : 0.0;
        sv[e] = ok ? a.v[c] : 0.0;
    }
    __syncthreads();
    double s = 0.0;
    if (row < a.n) {
        const int64_t li = row >> shift;
        const double rx = a.loc.x[li], ry = a.loc.y[li], rz = three_d ? a.loc.z[li] : 0.0;
        const int rvar = (int) (row & 1);
        const int cend = (int) min<int64_t>(kChunk, a.n - c0);
        for (int e = g; e < cend; e += 4) {
            const int64_t c = c0 + e;
            const int pair = shift ? (rvar * 2 + (int) (c & 1)) : 0;
            const double d = distance(rx, ry, rz, sx[e], sy[e], sz[e], a.prm.metric, three_d) / a.prm.range;
            const double cov = (d == 0.0) ? a.prm.sig[pair]
                                          : matern_scaled(d, a.prm.con[pair], a.prm.sig[pair], a.prm.nu[pair]);
            s += cov * sv[e];
        }
    }
    sp[g][r] = s;
    __syncthreads();
    if (t < kBlk) {
        const int64_t rr = (int64_t) blockIdx.x * kBlk + t;
        a.partial[(int64_t) blockIdx.y * a.npad + rr] = ((sp[0][t] + sp[1][t]) + sp[2][t]) + sp[3][t];
    }
}

// out[i] = base[i] + scale * sum_cb partial[cb][i]  (fixed chunk order), for i < npad.
__global__ void __launch_bounds__(kThreads) matvec_reduce_kernel(double *out, const double *base, double scale,
                                                                 const double *partial, int nchunks, int64_t n,
                                                                 int64_t npad) {
    const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
    if (i >= npad) return;
    double s = 0.0;
    for (int cb = 0; cb < nchunks; ++cb) s += partial[(int64_t) cb * npad + i];
    const double b = base ? base[i] : 0.0;
    out[i] = (i < n) ? b + scale * s : 0.0;
}

// Fixed-order dot products: block b sums indices b, b + G, b + 2G, ... then the
// host adds the G block sums in order.
constexpr int kDotBlocks = 256;

__global__ void __launch_bounds__(kThreads) dot_kernel(const double *x, const double *y, int64_t n, double *partial) {
    __shared__ double red[kThreads];
    double s = 0.0;
    const int64_t stride = (int64_t) kDotBlocks * kThreads;
    for (int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) s += x[i] * y[i];
    s = wg_sum(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// y = a * x + b * y
__global__ void __launch_bounds__(kThreads) axpby_kernel(double *y, const double *x, double a, double b, int64_t n) {
    const int64_t i = (int64_t) blockIdx.x * kThreads + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

void launch_cov_matvec(double *partial, const double *v, int64_t n, int64_t npad, const MaternParams &prm,
                       const double *x, const double *y, const double *z, hipStream_t s) {
    MatvecArgs a;
    a.partial = partial;
    a.v = v;
    a.n = n;
    a.npad = npad;
    a.prm = prm;
    a.loc = Locs{x, y, z};
    a.three_d = z != nullptr;
    const unsigned rows = (unsigned) (npad / kBlk);
    const unsigned chunks = (unsigned) ((n + kChunk - 1) / kChunk);
    hipLaunchKernelGGL(cov_matvec_kernel, dim3(rows, chunks), dim3(kThreads), 0, s, a);
}

int matvec_chunks(int64_t n) { return (int) ((n + kChunk - 1) / kChunk); }

void launch_matvec_reduce(double *out, const double *base, double scale, const double *partial, int nchunks,
                          int64_t n, int64_t npad, hipStream_t s) {
    hipLaunchKernelGGL(matvec_reduce_kernel, dim3((unsigned) ((npad + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       s, out, base, scale, partial, nchunks, n, npad);
}

int dot_blocks() { return kDotBlocks; }

void launch_dot(const double *x, const double *y, int64_t n, double *partial, hipStream_t s) {
    hipLaunchKernelGGL(dot_kernel, dim3(kDotBlocks), dim3(kThreads), 0, s, x, y, n, partial);
}

void launch_axpby(double *y, const double *x, double a, double b, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(axpby_kernel, dim3((unsigned) ((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, y, x, a,
                       b, n);
}

void launch_trsv_diag(const TileMap &map, int k, const double *winv_k, double *z, double *partials, hipStream_t s) {
    hipLaunchKernelGGL(trsv_diag_kernel, dim3(1), dim3(kSolveThreads), 0, s, map, k, winv_k, z, partials);
}

void launch_trsv_t_partial(const TileMap &map, int k, const double *z, double *scratch, hipStream_t s) {
    const int nsb = map.ts / kBlk;
    const int rows_below = map.nt - k - 1;
    if (rows_below <= 0) return;
    hipLaunchKernelGGL(trsv_t_partial_kernel, dim3(rows_below * nsb), dim3(kThreads), 0, s, map, k, z, scratch);
}

void launch_trsv_t_diag(const TileMap &map, int k, const double *winv_k, double *z, const double *scratch,
                        hipStream_t s) {
    hipLaunchKernelGGL(trsv_t_diag_kernel, dim3(1), dim3(kSolveThreads), 0, s, map, k, winv_k, z, scratch);
}

void launch_trmv(const TileMap &map, const double *e, double *out, hipStream_t s) {
    const int nsb = map.ts / kBlk;
    hipLaunchKernelGGL(trmv_kernel, dim3(map.nt * nsb), dim3(kThreads), 0, s, map, e, out);
}

void launch_cov_gemv(double *out, int64_t rows, int64_t cols, const MaternParams &prm, const double *x1,
                     const double *y1, const double *z1, const double *x2, const double *y2, const double *z2,
                     const double *xv, hipStream_t s) {
    if (rows <= 0) return;
    CovGemvArgs a;
    a.out = out;
    a.rows = rows;
    a.cols = cols;
    a.prm = prm;
    a.l1 = Locs{x1, y1, z1};
    a.l2 = Locs{x2, y2, z2};
    a.x = xv;
    a.three_d = (z1 != nullptr && z2 != nullptr);
    hipLaunchKernelGGL(cov_gemv_kernel, dim3((unsigned) rows), dim3(kThreads), 0, s, a);
}

void launch_dlarnv_normal(double *e, int64_t n, uint64_t seed48, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(dlarnv_normal_kernel, dim3((unsigned) ((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                       e, n, seed48);
}

void launch_copy_pad(double *dst, const double *src, int64_t n, int64_t npad, hipStream_t s) {
    if (npad <= 0) return;
    hipLaunchKernelGGL(copy_pad_kernel, dim3((unsigned) ((npad + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                       dst, src, n, npad);
}

// ------------------------------------------------------------------------
// Sigma v from the FP64 lower tiles (refinement of mixed plans).  One
// workgroup per tile (i, j); the tile is read once: wave w walks the columns
// in groups of four, c = 4 w + 16 p + u, lane l holds rows l + 64 m of a
// 512-row chunk.  The row
// products T v_j go to prow, the column products T' v_i (off the diagonal) to
// pcol; both are summed in a fixed order, so results are reproducible.
// ------------------------------------------------------------------------
constexpr int kRowsPerLane = 8;
constexpr int kChunkRows = 64 * kRowsPerLane;

__global__ void __launch_bounds__(kThreads) sigma_tile_matvec_kernel(const double *__restrict__ sig64,
                                                                     const int2 *__restrict__ pairs, int ts,
                                                                     const double *__restrict__ v,
                                                                     double *__restrict__ prow,
                                                                     double *__restrict__ pcol) {
    __shared__ double s_acc[kChunkRows];
    __shared__ double s_col[kMaxTs];
    const int2 ij = pairs[blockIdx.x];
    const int i = ij.x, j = ij.y;
    const int64_t tid = (int64_t) i * (i + 1) / 2 + j;
    const double *T = sig64 + tid * ts * ts;
    const double *vi = v + (int64_t) i * ts;
    const double *vj = v + (int64_t) j * ts;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool offdiag = i != j;
    for (int c = threadIdx.x; c < ts; c += kThreads) s_col[c] = 0.0;
    __syncthreads();
    for (int r0 = 0; r0 < ts; r0 += kChunkRows) {
        double acc[kRowsPerLane], vr[kRowsPerLane];
#pragma unroll
        for (int m = 0; m < kRowsPerLane; ++m) {
            const int r = r0 + lane + 64 * m;
            vr[m] = r < ts ? vi[r] : 0.0;
            acc[m] = 0.0;
        }
        // four consecutive columns per wave and pass, so that 32 loads per
        // lane are in flight; their four column dot products are reduced
        // across the wave together (7 shuffles instead of 24)
        constexpr int kPass = 4;
        constexpr int kStride = kPass * (kThreads / 64);
        const int cmain = ts - ts % kStride;
        for (int c0 = w * kPass; c0 < cmain; c0 += kStride) {
            double vc[kPass], dot[kPass];
#pragma unroll
            for (int u = 0; u < kPass; ++u) {
                vc[u] = vj[c0 + u];
                dot[u] = 0.0;
            }
            const double *col = T + (int64_t) c0 * ts + r0 + lane;
#pragma unroll
            for (int m = 0; m < kRowsPerLane; ++m) {
                if (r0 + lane + 64 * m < ts) {
#pragma unroll
                    for (int u = 0; u < kPass; ++u) {
                        const double t = __builtin_nontemporal_load(col + (int64_t) u * ts + 64 * m);
                        acc[m] += t * vc[u];
                        dot[u] += t * vr[m];
                    }
                }
            }
            if (offdiag) {
                // lanes with bit 5 clear keep columns 0, 1, set 2, 3; then bit
                // 4 picks one of the two: lane 16 q ends with column q
                const bool h32 = lane & 32, h16 = lane & 16;
                const double k0 = (h32 ? dot[2] : dot[0]) + __shfl_xor(h32 ? dot[0] : dot[2], 32, 64);
                const double k1 = (h32 ? dot[3] : dot[1]) + __shfl_xor(h32 ? dot[1] : dot[3], 32, 64);
                double d = (h16 ? k1 : k0) + __shfl_xor(h16 ? k0 : k1, 16, 64);
#pragma unroll
                for (int off = 8; off > 0; off >>= 1) d += __shfl_xor(d, off, 64);
                if ((lane & 15) == 0) s_col[c0 + (lane >> 4)] += d;
            }
        }
        for (int c = cmain + w; c < ts; c += kThreads / 64) {
            const double vc = vj[c];
            const double *col = T + (int64_t) c * ts + r0 + lane;
            double dot = 0.0;
#pragma unroll
            for (int m = 0; m < kRowsPerLane; ++m) {
                if (r0 + lane + 64 * m < ts) {
                    const double t = __builtin_nontemporal_load(col + 64 * m);
                    acc[m] += t * vc;
                    dot += t * vr[m];
                }
            }
            if (offdiag) {
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
                if (lane == 0) s_col[c] += dot;
            }
        }
        // the four waves' row partials, added in wave order
        for (int q = 0; q < kThreads / 64; ++q) {
            if (w == q) {
#pragma unroll
                for (int m = 0; m < kRowsPerLane; ++m) {
                    const int e = lane + 64 * m;
                    s_acc[e] = (q == 0 ? 0.0 : s_acc[e]) + acc[m];
                }
            }
            __syncthreads();
        }
        for (int e = threadIdx.x; e < kChunkRows && r0 + e < ts; e += kThreads) prow[tid * ts + r0 + e] = s_acc[e];
        __syncthreads();
    }
    if (offdiag)
        for (int c = threadIdx.x; c < ts; c += kThreads) pcol[tid * ts + c] = s_col[c];
}

__global__ void __launch_bounds__(kThreads) sigma_reduce_kernel(double *out, const double *base, double scale,
                                                                const double *__restrict__ prow,
                                                                const double *__restrict__ pcol, int nt, int ts,
                                                                int64_t n, int64_t npad) {
    const int64_t g = (int64_t) blockIdx.x * kThreads + threadIdx.x;
    if (g >= npad) return;
    const int i = (int) (g / ts), r = (int) (g % ts);
    double acc = 0.0;
    for (int j = 0; j <= i; ++j) acc += prow[((int64_t) i * (i + 1) / 2 + j) * ts + r];
    for (int j = i + 1; j < nt; ++j) acc += pcol[((int64_t) j * (j + 1) / 2 + i) * ts + r];
    out[g] = g < n ? (base ? base[g] : 0.0) + scale * acc : 0.0;
}

void launch_sigma_tile_matvec(const double *sig64, const int2 *pairs, int npairs, int ts, const double *v,
                              double *prow, double *pcol, hipStream_t s) {
    if (npairs <= 0) return;
    hipLaunchKernelGGL(sigma_tile_matvec_kernel, dim3((unsigned) npairs), dim3(kThreads), 0, s, sig64, pairs, ts, v,
                       prow, pcol);
}

void launch_sigma_reduce(double *out, const double *base, double scale, const double *prow, const double *pcol,
                         int nt, int ts, int64_t n, int64_t npad, hipStream_t s) {
    if (npad <= 0) return;
    hipLaunchKernelGGL(sigma_reduce_kernel, dim3((unsigned) ((npad + kThreads - 1) / kThreads)), dim3(kThreads), 0, s,
                       out, base, scale, prow, pcol, nt, ts, n, npad);
}

// ------------------------------------------------------------------------
// Preconditioner solves of mixed plans with X_k = L_kk^{-T} (launch_tinv).
// Both directions are right-looking: a GEMV with the diagonal-tile inverse,
// then a GEMV of the tile column (forward) or tile row (backward) that
// updates the not-yet-solved blocks in place.  Every output element is
// written by one wave or one workgroup, sums run in a fixed order.
// ------------------------------------------------------------------------

// y[c] = sum_{r <= c} X[r][c] z[r]: one wave per column c.
__global__ void __launch_bounds__(kThreads) fwd_diag_inv_kernel(const double *__restrict__ X,
                                                                const double *__restrict__ z,
                                                                double *__restrict__ y, int ts) {
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    if (c >= ts) return;
    const double *col = X + (int64_t) c * ts;
    double s0 = 0.0, s1 = 0.0;
    int r = lane;
    for (; r + 64 <= c; r += 128) {
        s0 += col[r] * z[r];
        s1 += col[r + 64] * z[r + 64];
    }
    for (; r <= c; r += 64) s0 += col[r] * z[r];
    double s = s0 + s1;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[c] = s;
}

// z_i -= L_ik y_k for i > k; one workgroup per 64 rows, four independent
// partial sums per lane over the columns of its wave.
template <typename TA>
__device__ __forceinline__ double panel_row_dot(const TA *__restrict__ A, const double *sy, int ts, int w) {
    // 16 independent loads per lane and step (ts is a multiple of 64)
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = w; c < ts; c += 64) {
#pragma unroll
        for (int u = 0; u < 16; ++u) a[u & 3] += (double) A[(int64_t) (c + 4 * u) * ts] * sy[c + 4 * u];
    }
    return (a[0] + a[1]) + (a[2] + a[3]);
}

__device__ __forceinline__ void fwd_panel_body(const TileMap &map, int k, const double *__restrict__ yk,
                                               double *__restrict__ z, int bx) {
    __shared__ double sy[kMaxTs];
    __shared__ double sp[4][kBlk];
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int i = k + 1 + (int) (bx / nsb);
    const int rb = (int) (bx % nsb);
    for (int e = t; e < ts; e += kThreads) sy[e] = yk[e];
    __syncthreads();
    double s;
    if (map.store[i * map.nt + k] == kStoreF32) s = panel_row_dot(tile_f32(map, i, k) + rb * kBlk + lane, sy, ts, w);
    else s = panel_row_dot(tile_f64(map, i, k) + rb * kBlk + lane, sy, ts, w);
    sp[w][lane] = s;
    __syncthreads();
    if (t < kBlk) z[(int64_t) i * ts + rb * kBlk + t] -= ((sp[0][t] + sp[1][t]) + sp[2][t]) + sp[3][t];
}

__global__ void __launch_bounds__(kThreads) fwd_panel_kernel(TileMap map, int k, const double *__restrict__ yk,
                                                             double *__restrict__ z) {
    fwd_panel_body(map, k, yk, z, blockIdx.x);
}

__global__ void __launch_bounds__(kThreads) fwd_panel_batch_kernel(const BatchJob *jobs) {
    const BatchJob &J = jobs[blockIdx.y];
    const int nsb = J.map.ts / kBlk;
    if ((int) blockIdx.x >= (J.map.nt - J.k - 1) * nsb) return;
    fwd_panel_body(J.map, J.k, J.z + (int64_t) J.k * J.map.ts, J.z, blockIdx.x);
}

// y[r] = sum_{c >= r} X[r][c] x[c]: one workgroup of 16 waves per 64 rows,
// wave g takes the columns c = rb*64 + g + 16 m.
__global__ void __launch_bounds__(kSolveThreads) bwd_diag_inv_kernel(const double *__restrict__ X,
                                                                     const double *__restrict__ x,
                                                                     double *__restrict__ y, int ts) {
    __shared__ double sp[kGroups][kBlk];
    const int t = threadIdx.x;
    const int lane = t & 63, g = t >> 6;
    const int rb = blockIdx.x;
    const int r = rb * kBlk + lane;
    const double *Xr = X + r;
    double s0 = 0.0, s1 = 0.0;
    int c = rb * kBlk + g;
    for (; c + kGroups < ts; c += 2 * kGroups) {
        s0 += (c >= r ? Xr[(int64_t) c * ts] : 0.0) * x[c];
        s1 += Xr[(int64_t) (c + kGroups) * ts] * x[c + kGroups] * (c + kGroups >= r ? 1.0 : 0.0);
    }
    for (; c < ts; c += kGroups) s0 += (c >= r ? Xr[(int64_t) c * ts] : 0.0) * x[c];
    sp[g][lane] = s0 + s1;
    __syncthreads();
    if (t < kBlk) {
        double a = 0.0;
        for (int q = 0; q < kGroups; ++q) a += sp[q][t];
        y[rb * kBlk + t] = a;
    }
}

// x_j[c] -= sum_r L_kj[r][c] y_k[r] for j < k: workgroup (j, 64-column block),
// wave w reduces 16 columns, lanes stride the rows.
__global__ void __launch_bounds__(kThreads) bwd_row_kernel(TileMap map, int k, const double *__restrict__ yk,
                                                           double *__restrict__ x) {
    __shared__ double sy[kMaxTs];
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int j = (int) (blockIdx.x / nsb);
    const int cb = (int) (blockIdx.x % nsb);
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    for (int e = t; e < ts; e += kThreads) sy[e] = yk[e];
    __syncthreads();
    // the wave's 16 columns advance together over the rows: 16 independent
    // loads per lane and row step (one column at a time waits one memory
    // latency per 64 rows)
    const int c0 = cb * kBlk + w * 16;
    double acc[16];
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) acc[cc] = 0.0;
    auto sweep = [&](const auto *A) {
        const auto *col = A + (int64_t) c0 * ts;
        for (int r = lane; r < ts; r += 64) {
            const double yr = sy[r];
#pragma unroll
            for (int cc = 0; cc < 16; ++cc) acc[cc] += (double) col[r + (int64_t) cc * ts] * yr;
        }
    };
    if (map.store[k * map.nt + j] == kStoreF32) sweep(tile_f32(map, k, j));
    else sweep(tile_f64(map, k, j));
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) {
        double s = acc[cc];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) x[(int64_t) j * ts + c0 + cc] -= s;
    }
}

// partials[2k] = sum of log L_ii (L_ii > 0) over diagonal tile k, partials[2k+1] = 0.
__global__ void __launch_bounds__(kThreads) diag_logdet_kernel(TileMap map, double *partials) {
    __shared__ double red[kThreads];
    const int k = blockIdx.x;
    const int ts = map.ts;
    const double *L = tile_f64(map, k, k);
    double ld = 0.0;
    for (int e = threadIdx.x; e < ts; e += kThreads) {
        const double d = L[e + (int64_t) e * ts];
        if (d > 0.0) ld += log(d);
    }
    ld = wg_sum(ld, red);
    if (threadIdx.x == 0) {
        partials[2 * k] = ld;
        partials[2 * k + 1] = 0.0;
    }
}

void launch_fwd_diag_inv(const double *tinv_k, const double *z_k, double *y_k, int ts, hipStream_t s) {
    const int per = kThreads / 64;
    hipLaunchKernelGGL(fwd_diag_inv_kernel, dim3((unsigned) ((ts + per - 1) / per)), dim3(kThreads), 0, s, tinv_k, z_k,
                       y_k, ts);
}

void launch_fwd_panel(const TileMap &map, int k, const double *y_k, double *z, hipStream_t s) {
    const int rows = map.nt - k - 1;
    if (rows <= 0) return;
    const int nsb = map.ts / kBlk;
    hipLaunchKernelGGL(fwd_panel_kernel, dim3((unsigned) (rows * nsb)), dim3(kThreads), 0, s, map, k, y_k, z);
}

void launch_forward_batch(const BatchJob *jobs, int njobs, int ts, int rest, hipStream_t s) {
    if (njobs <= 0) return;
    hipLaunchKernelGGL(trsv_diag_batch_kernel, dim3((unsigned) njobs), dim3(kSolveThreads), 0, s, jobs);
    if (rest > 0)
        hipLaunchKernelGGL(fwd_panel_batch_kernel, dim3((unsigned) (rest * (ts / kBlk)), njobs), dim3(kThreads), 0, s,
                           jobs);
}

void launch_bwd_diag_inv(const double *tinv_k, const double *x_k, double *y_k, int ts, hipStream_t s) {
    hipLaunchKernelGGL(bwd_diag_inv_kernel, dim3((unsigned) (ts / kBlk)), dim3(kSolveThreads), 0, s, tinv_k, x_k, y_k,
                       ts);
}

void launch_bwd_row(const TileMap &map, int k, const double *y_k, double *x, hipStream_t s) {
    if (k <= 0) return;
    const int nsb = map.ts / kBlk;
    hipLaunchKernelGGL(bwd_row_kernel, dim3((unsigned) (k * nsb)), dim3(kThreads), 0, s, map, k, y_k, x);
}

void launch_diag_logdet(const TileMap &map, int64_t, double *partials, hipStream_t s) {
    hipLaunchKernelGGL(diag_logdet_kernel, dim3((unsigned) map.nt), dim3(kThreads), 0, s, map, partials);
}

}  // namespace egs
