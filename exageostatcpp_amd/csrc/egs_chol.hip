// Tiled Cholesky factorization on gfx950 (the reference calls
// CHAMELEON_dpotrf_Tile(Lower) at ChameleonDense.cpp:29: POTRF of the diagonal
// tile, TRSM of the panel below it, SYRK/GEMM of the trailing tiles).
//
// Step k of the factorization:
//   * the diagonal tile is factored left-looking in 64 x 64 blocks: one
//     workgroup per block updates it, factors it in LDS with a per-column
//     pivot broadcast, and inverts it (potrf_block_kernel); the rows below it
//     inside the tile are then solved against that block (trsm_rows_kernel);
//   * the panel below the diagonal tile is solved by one workgroup per 64-row
//     block, marching over the 64-column blocks of L_kk and applying the
//     64 x 64 inverses (trsm_rows_kernel);
//   * the trailing tiles are updated with FP64 MFMA (update_kernel), one
//     workgroup per 64 x 64 output block with K = ts.
// Every contraction uses v_mfma_f64_16x16x4f64 through wg_gemm_nt: four waves
// each own a 32 x 32 quadrant of the 64 x 64 output; K is staged through LDS in
// chunks of 32 columns, double buffered, with the next chunk's global loads
// issued before the current chunk's MFMAs (these kernels run on few
// workgroups, so each one has to hide its own memory latency).
#include "egs_common.hpp"
#include "egs_kernels.hpp"

namespace egs {

constexpr int kKc = 32;          // K chunk staged in LDS
constexpr int kLdsPad = 65;      // LDS row stride (doubles) for 64-wide panels
constexpr int kStage = kBlk * kKc / kThreads;   // elements of A (and of B) per thread per chunk

struct GemmSmem {
    double a[2][kKc][kLdsPad];
    double b[2][kKc][kLdsPad];
};

// acc[ni][mi][v] holds C[mq*32 + mi*16 + lane%16][nq*32 + ni*16 + lane/16 + 4*v]
// for wave w = threadIdx.x / 64, mq = w & 1, nq = w >> 1.
struct Acc64 {
    f64x4 v[2][2];
};

__device__ __forceinline__ void acc_zero(Acc64 &acc) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc.v[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
}

// acc += A (64 x K, column-major, lda) * B (64 x K, column-major, ldb)^T.
// K must be a multiple of kKc.
template <typename TA>
__device__ void wg_gemm_nt(Acc64 &acc, const TA *__restrict__ A, int64_t lda, const double *__restrict__ B,
                           int64_t ldb, int K, GemmSmem &sm) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int mq = w & 1, nq = w >> 1;
    const int lr = lane & 15, lk = lane >> 4;
    // element e = t + q * 256 of a 64 x kKc chunk: row e % 64, column e / 64
    const int srow = t & 63;
    double ra[kStage], rb[kStage];
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
        const int col = (t >> 6) + q * (kThreads / kBlk);
        ra[q] = (double) A[srow + (int64_t) col * lda];
        rb[q] = B[srow + (int64_t) col * ldb];
    }
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
        const int col = (t >> 6) + q * (kThreads / kBlk);
        sm.a[0][col][srow] = ra[q];
        sm.b[0][col][srow] = rb[q];
    }
    __syncthreads();
    const int nch = K / kKc;
    for (int c = 0; c < nch; ++c) {
        const int buf = c & 1;
        const bool more = c + 1 < nch;
        if (more) {
            const int64_t k0 = (int64_t) (c + 1) * kKc;
#pragma unroll
            for (int q = 0; q < kStage; ++q) {
                const int col = (t >> 6) + q * (kThreads / kBlk);
                ra[q] = (double) A[srow + (k0 + col) * lda];
                rb[q] = B[srow + (k0 + col) * ldb];
            }
        }
#pragma unroll
        for (int kk = 0; kk < kKc; kk += 4) {
            const double a0 = sm.a[buf][kk + lk][mq * 32 + lr];
            const double a1 = sm.a[buf][kk + lk][mq * 32 + 16 + lr];
            const double b0 = sm.b[buf][kk + lk][nq * 32 + lr];
            const double b1 = sm.b[buf][kk + lk][nq * 32 + 16 + lr];
            acc.v[0][0] = mfma_f64(b0, a0, acc.v[0][0]);
            acc.v[0][1] = mfma_f64(b0, a1, acc.v[0][1]);
            acc.v[1][0] = mfma_f64(b1, a0, acc.v[1][0]);
            acc.v[1][1] = mfma_f64(b1, a1, acc.v[1][1]);
        }
        if (more) {
#pragma unroll
            for (int q = 0; q < kStage; ++q) {
                const int col = (t >> 6) + q * (kThreads / kBlk);
                sm.a[buf ^ 1][col][srow] = ra[q];
                sm.b[buf ^ 1][col][srow] = rb[q];
            }
        }
        __syncthreads();
    }
}

// Visit every (row, col, value) of the accumulator.
template <typename F>
__device__ __forceinline__ void acc_foreach(const Acc64 &acc, F f) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int mq = w & 1, nq = w >> 1;
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = mq * 32 + mi * 16 + (lane & 15);
                const int c = nq * 32 + ni * 16 + (lane >> 4) + 4 * v;
                f(r, c, acc.v[ni][mi][v]);
            }
}

// ------------------------------------------------------------------------
// Diagonal 64 x 64 block: update, factor, invert.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) potrf_block_kernel(TileMap map, int k, int jb, double *winv_k, int *info,
                                                               int64_t n) {
    __shared__ GemmSmem sm;
    __shared__ double sT[kBlk][kLdsPad];
    __shared__ double sW[kBlk][kLdsPad];
    const int ts = map.ts;
    const int t = threadIdx.x;
    double *T = tile_f64(map, k, k);
    double *Ajj = T + (int64_t) jb * kBlk + (int64_t) jb * kBlk * ts;

    Acc64 acc;
    acc_zero(acc);
    if (jb > 0) wg_gemm_nt(acc, T + jb * kBlk, ts, T + jb * kBlk, ts, jb * kBlk, sm);
    acc_foreach(acc, [&](int r, int c, double v) { sT[r][c] = Ajj[r + (int64_t) c * ts] - v; });
    __syncthreads();

    // Right-looking Cholesky of sT (lower).
    for (int j = 0; j < kBlk; ++j) {
        const double d = sT[j][j];
        const double s = sqrt(d);
        const double inv = 1.0 / s;
        if (t == 0 && !(d > 0.0)) {
            const int64_t col = (int64_t) k * ts + (int64_t) jb * kBlk + j;
            if (col < n) atomicCAS(info, 0, (int) (col + 1));
        }
        for (int e = t; e < kBlk * kBlk; e += kThreads) {
            const int r = e & 63, c = e >> 6;
            if (c > j && r >= c) sT[r][c] -= (sT[r][j] * inv) * (sT[c][j] * inv);
        }
        __syncthreads();
        if (t >= j && t < kBlk) sT[t][j] = (t == j) ? s : sT[t][j] * inv;
        __syncthreads();
    }

    // Inverse of the lower factor, one column per thread.
    if (t < kBlk) {
        const int c = t;
        for (int r = 0; r < kBlk; ++r) {
            if (r < c) {
                sW[r][c] = 0.0;
                continue;
            }
            double acc_s = (r == c) ? 1.0 : 0.0;
            for (int l = c; l < r; ++l) acc_s -= sT[r][l] * sW[l][c];
            sW[r][c] = acc_s / sT[r][r];
        }
    }
    __syncthreads();

    double *W = winv_k + (int64_t) jb * kBlk * kBlk;
    for (int e = t; e < kBlk * kBlk; e += kThreads) {
        const int r = e & 63, c = e >> 6;
        Ajj[r + (int64_t) c * ts] = (r >= c) ? sT[r][c] : 0.0;
        W[r + c * kBlk] = sW[r][c];
    }
}

// ------------------------------------------------------------------------
// Row-block triangular solve against L_kk, column blocks [jb0, jb1).
// ------------------------------------------------------------------------
template <typename TX>
__global__ void __launch_bounds__(kThreads) trsm_rows_kernel(TileMap map, int k, int64_t g0, int jb0, int jb1,
                                                             const double *winv_k) {
    __shared__ GemmSmem sm;
    __shared__ double sT[kBlk][kLdsPad];
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int64_t g = g0 + blockIdx.x;
    const int i = (int) (g / nsb);
    const int rb = (int) (g % nsb);
    TX *X = reinterpret_cast<TX *>(map.base + map.off[i * map.nt + k]) + rb * kBlk;
    const double *L = tile_f64(map, k, k);
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int mq = w & 1, nq = w >> 1;
    const int lr = lane & 15, lk = lane >> 4;

    for (int jb = jb0; jb < jb1; ++jb) {
        Acc64 acc;
        acc_zero(acc);
        if (jb > 0) wg_gemm_nt(acc, X, ts, L + jb * kBlk, ts, jb * kBlk, sm);
        TX *Xj = X + (int64_t) jb * kBlk * ts;
        acc_foreach(acc, [&](int r, int c, double v) { sT[r][c] = (double) Xj[r + (int64_t) c * ts] - v; });
        __syncthreads();
        // X_j = T W_j^T :  acc2[r][c] = sum_l T[r][l] W[c][l]
        const double *W = winv_k + (int64_t) jb * kBlk * kBlk;
        Acc64 acc2;
        acc_zero(acc2);
#pragma unroll 4
        for (int kk = 0; kk < kBlk; kk += 4) {
            const double a0 = sT[mq * 32 + lr][kk + lk];
            const double a1 = sT[mq * 32 + 16 + lr][kk + lk];
            const double b0 = W[(nq * 32 + lr) + (kk + lk) * kBlk];
            const double b1 = W[(nq * 32 + 16 + lr) + (kk + lk) * kBlk];
            acc2.v[0][0] = mfma_f64(b0, a0, acc2.v[0][0]);
            acc2.v[0][1] = mfma_f64(b0, a1, acc2.v[0][1]);
            acc2.v[1][0] = mfma_f64(b1, a0, acc2.v[1][0]);
            acc2.v[1][1] = mfma_f64(b1, a1, acc2.v[1][1]);
        }
        acc_foreach(acc2, [&](int r, int c, double v) { Xj[r + (int64_t) c * ts] = (TX) v; });
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// Trailing update, FP64.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) update_kernel(TileMap map, const int2 *pairs, int k) {
    __shared__ GemmSmem sm;
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int2 ij = pairs[blockIdx.y];
    const int sbm = blockIdx.x % nsb;
    const int sbn = blockIdx.x / nsb;
    if (ij.x == ij.y && sbm < sbn) return;
    double *C = tile_f64(map, ij.x, ij.y) + sbm * kBlk + (int64_t) sbn * kBlk * ts;
    const double *A = tile_f64(map, ij.x, k) + sbm * kBlk;
    const double *B = tile_f64(map, ij.y, k) + sbn * kBlk;
    Acc64 acc;
    acc_zero(acc);
    wg_gemm_nt(acc, A, ts, B, ts, ts, sm);
    // read all 16 values of C first: one memory round trip for the read-modify-write
    double cv[16];
    int q = 0;
    acc_foreach(acc, [&](int r, int c, double) { cv[q++] = C[r + (int64_t) c * ts]; });
    q = 0;
    acc_foreach(acc, [&](int r, int c, double v) { C[r + (int64_t) c * ts] = cv[q++] - v; });
}

void launch_potrf_block(const TileMap &map, int k, int jb, double *winv_k, int *info, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(potrf_block_kernel, dim3(1), dim3(kThreads), 0, s, map, k, jb, winv_k, info, n);
}

void launch_trsm_rows(const TileMap &map, int k, int64_t g0, int count, int jb0, int jb1, const double *winv_k,
                      hipStream_t s, bool x_f32) {
    if (count <= 0 || jb1 <= jb0) return;
    if (x_f32)
        hipLaunchKernelGGL(trsm_rows_kernel<float>, dim3(count), dim3(kThreads), 0, s, map, k, g0, jb0, jb1, winv_k);
    else
        hipLaunchKernelGGL(trsm_rows_kernel<double>, dim3(count), dim3(kThreads), 0, s, map, k, g0, jb0, jb1, winv_k);
}

void launch_update(const TileMap &map, const int2 *pairs, int npairs, int k, hipStream_t s) {
    if (npairs <= 0) return;
    if (update192_supported(map.ts)) {
        launch_update192(map, pairs, npairs, k, s);
        return;
    }
    const int nsb = map.ts / kBlk;
    hipLaunchKernelGGL(update_kernel, dim3(nsb * nsb, npairs), dim3(kThreads), 0, s, map, pairs, k);
}

}  // namespace egs
