// Tiled Cholesky factorization on gfx950 (the reference calls
// CHAMELEON_dpotrf_Tile(Lower) at ChameleonDense.cpp:29: POTRF of the diagonal
// tile, TRSM of the panel below it, SYRK/GEMM of the trailing tiles).
//
// Step k of the factorization:
//   * the diagonal tile is factored right-looking in 64 x 64 blocks
//     (diag_step_kernel, diag_copy_kernel below), which also leaves the
//     64 x 64 inverses of its diagonal blocks;
//   * the panel below the diagonal tile is solved by one workgroup per 64-row
//     block, marching over the 64-column blocks of L_kk and applying the
//     64 x 64 inverses (trsm_rows_kernel);
//   * the trailing tiles are updated with FP64 MFMA: update192 (egs_update.hip)
//     when ts is a multiple of 192, else update_kernel here, one workgroup per
//     64 x 64 output block with K = ts.
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
// Row-block triangular solve against L_kk, column blocks [jb0, jb1).
// ------------------------------------------------------------------------
// 66.5 KB of LDS (sm): T is formed after each GEMM through sm (it reuses
// sm.a), so that a workgroup fits beside one workgroup of the trailing update.
template <typename TX>
__device__ __forceinline__ void trsm_rows_body(const TileMap &map, int k, int64_t g0, int jb0, int jb1,
                                               const double *winv_k, int bx, GemmSmem &sm) {
    static_assert(sizeof(GemmSmem::a) >= sizeof(double) * kBlk * kLdsPad, "T must fit in the staging buffer");
    double (*sT)[kLdsPad] = reinterpret_cast<double (*)[kLdsPad]>(&sm.a[0][0][0]);
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int64_t g = g0 + bx;
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

template <typename TX>
__global__ void __launch_bounds__(kThreads) trsm_rows_kernel(TileMap map, int k, int64_t g0, int jb0, int jb1,
                                                             const double *winv_k) {
    __shared__ GemmSmem sm;
    trsm_rows_body<TX>(map, k, g0, jb0, jb1, winv_k, blockIdx.x, sm);
}

// ------------------------------------------------------------------------
// The same solve in 32-row blocks.  A row block's solve is a chain of ts / 64
// dependent steps bound by its CU's FP64 MFMA rate; in 32-row blocks a short
// panel (the end of a factorization, small problems) spreads over twice the
// CUs, each with half the MFMA work (and the same loads of L_kk).
// ------------------------------------------------------------------------
namespace {
constexpr int kHalf = 32;   // rows per block

// acc.v[mi] holds C[mi*16 + lane%16][w*16 + lane/16 + 4*v], w = threadIdx.x / 64.
struct AccHalf {
    f64x4 v[2];
};

__device__ __forceinline__ void acc_zero(AccHalf &acc) {
    acc.v[0] = f64x4{0.0, 0.0, 0.0, 0.0};
    acc.v[1] = f64x4{0.0, 0.0, 0.0, 0.0};
}

template <typename F>
__device__ __forceinline__ void acc_foreach(const AccHalf &acc, F f) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int v = 0; v < 4; ++v) f(mi * 16 + (lane & 15), w * 16 + (lane >> 4) + 4 * v, acc.v[mi][v]);
}

// acc += A (32 x K, column-major, lda) * B (64 x K, column-major, ldb)^T,
// staged like wg_gemm_nt (A in the first 32 rows of sm.a).  K must be a
// multiple of kKc.
template <typename TA>
__device__ void wg_gemm_nt_half(AccHalf &acc, const TA *__restrict__ A, int64_t lda, const double *__restrict__ B,
                                int64_t ldb, int K, GemmSmem &sm) {
    constexpr int kStageA = kHalf * kKc / kThreads;   // 4
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    // A chunk element (row t % 32, column t / 32 + 8 q); B chunk element
    // (row t % 64, column t / 64 + 4 q)
    const int arow = t & (kHalf - 1), acol = t / kHalf;
    const int brow = t & 63, bcol = t >> 6;
    double ra[kStageA], rb[kStage];
#pragma unroll
    for (int q = 0; q < kStageA; ++q) ra[q] = (double) A[arow + (int64_t) (acol + q * (kThreads / kHalf)) * lda];
#pragma unroll
    for (int q = 0; q < kStage; ++q) rb[q] = B[brow + (int64_t) (bcol + q * (kThreads / kBlk)) * ldb];
#pragma unroll
    for (int q = 0; q < kStageA; ++q) sm.a[0][acol + q * (kThreads / kHalf)][arow] = ra[q];
#pragma unroll
    for (int q = 0; q < kStage; ++q) sm.b[0][bcol + q * (kThreads / kBlk)][brow] = rb[q];
    __syncthreads();
    const int nch = K / kKc;
    for (int c = 0; c < nch; ++c) {
        const int buf = c & 1;
        const bool more = c + 1 < nch;
        if (more) {
            const int64_t k0 = (int64_t) (c + 1) * kKc;
#pragma unroll
            for (int q = 0; q < kStageA; ++q) ra[q] = (double) A[arow + (k0 + acol + q * (kThreads / kHalf)) * lda];
#pragma unroll
            for (int q = 0; q < kStage; ++q) rb[q] = B[brow + (k0 + bcol + q * (kThreads / kBlk)) * ldb];
        }
#pragma unroll
        for (int kk = 0; kk < kKc; kk += 4) {
            const double a0 = sm.a[buf][kk + lk][lr];
            const double a1 = sm.a[buf][kk + lk][16 + lr];
            const double b0 = sm.b[buf][kk + lk][w * 16 + lr];
            acc.v[0] = mfma_f64(b0, a0, acc.v[0]);
            acc.v[1] = mfma_f64(b0, a1, acc.v[1]);
        }
        if (more) {
#pragma unroll
            for (int q = 0; q < kStageA; ++q) sm.a[buf ^ 1][acol + q * (kThreads / kHalf)][arow] = ra[q];
#pragma unroll
            for (int q = 0; q < kStage; ++q) sm.b[buf ^ 1][bcol + q * (kThreads / kBlk)][brow] = rb[q];
        }
        __syncthreads();
    }
}
}  // namespace

// Row block g0 + bx of 32 rows (ts / 32 per tile row), all column blocks.
template <typename TX>
__device__ __forceinline__ void trsm_rows_half_body(const TileMap &map, int k, int64_t g0, const double *winv_k,
                                                    int bx, GemmSmem &sm) {
    double (*sT)[kLdsPad] = reinterpret_cast<double (*)[kLdsPad]>(&sm.a[0][0][0]);
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int nrb = ts / kHalf;
    const int64_t g = g0 + bx;
    const int i = (int) (g / nrb);
    const int rb = (int) (g % nrb);
    TX *X = reinterpret_cast<TX *>(map.base + map.off[i * map.nt + k]) + rb * kHalf;
    const double *L = tile_f64(map, k, k);
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;

    for (int jb = 0; jb < nsb; ++jb) {
        AccHalf acc;
        acc_zero(acc);
        if (jb > 0) wg_gemm_nt_half(acc, X, ts, L + jb * kBlk, ts, jb * kBlk, sm);
        TX *Xj = X + (int64_t) jb * kBlk * ts;
        acc_foreach(acc, [&](int r, int c, double v) { sT[r][c] = (double) Xj[r + (int64_t) c * ts] - v; });
        __syncthreads();
        // X_j = T W_j^T
        const double *W = winv_k + (int64_t) jb * kBlk * kBlk;
        AccHalf acc2;
        acc_zero(acc2);
#pragma unroll 4
        for (int kk = 0; kk < kBlk; kk += 4) {
            const double a0 = sT[lr][kk + lk];
            const double a1 = sT[16 + lr][kk + lk];
            const double b0 = W[(w * 16 + lr) + (kk + lk) * kBlk];
            acc2.v[0] = mfma_f64(b0, a0, acc2.v[0]);
            acc2.v[1] = mfma_f64(b0, a1, acc2.v[1]);
        }
        acc_foreach(acc2, [&](int r, int c, double v) { Xj[r + (int64_t) c * ts] = (TX) v; });
        __syncthreads();
    }
}

template <typename TX>
__global__ void __launch_bounds__(kThreads) trsm_rows_half_kernel(TileMap map, int k, int64_t g0,
                                                                  const double *winv_k) {
    __shared__ GemmSmem sm;
    trsm_rows_half_body<TX>(map, k, g0, winv_k, blockIdx.x, sm);
}

// ------------------------------------------------------------------------
// The same solve for FP32 panel rows with v_mfma_f32_16x16x4f32 (FP32
// accumulation; L and W converted to FP32 as they are staged, 50 KB of LDS):
// 1.8x the rate of the FP64 form, for the univariate kernel's mixed plans
// (the bivariate field's error grew 3-14x with it).  FP32 result element v of
// lane l is D[4 (l / 16) + v][l % 16].
// ------------------------------------------------------------------------
struct GemmSmem32 {
    float a[2][kKc][kLdsPad];
    float b[2][kKc][kLdsPad];
};
struct Acc32 {
    f32x4 v[2][2];
};

__device__ void wg_gemm_nt_f32(Acc32 &acc, const float *__restrict__ A, int64_t lda, const double *__restrict__ B,
                               int64_t ldb, int K, GemmSmem32 &sm) {
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int mq = w & 1, nq = w >> 1;
    const int lr = lane & 15, lk = lane >> 4;
    const int srow = t & 63;
    float ra[kStage], rb[kStage];
#pragma unroll
    for (int q = 0; q < kStage; ++q) {
        const int col = (t >> 6) + q * (kThreads / kBlk);
        ra[q] = A[srow + (int64_t) col * lda];
        rb[q] = (float) B[srow + (int64_t) col * ldb];
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
                ra[q] = A[srow + (k0 + col) * lda];
                rb[q] = (float) B[srow + (k0 + col) * ldb];
            }
        }
#pragma unroll
        for (int kk = 0; kk < kKc; kk += 4) {
            const float a0 = sm.a[buf][kk + lk][mq * 32 + lr];
            const float a1 = sm.a[buf][kk + lk][mq * 32 + 16 + lr];
            const float b0 = sm.b[buf][kk + lk][nq * 32 + lr];
            const float b1 = sm.b[buf][kk + lk][nq * 32 + 16 + lr];
            acc.v[0][0] = mfma_f32(b0, a0, acc.v[0][0]);
            acc.v[0][1] = mfma_f32(b0, a1, acc.v[0][1]);
            acc.v[1][0] = mfma_f32(b1, a0, acc.v[1][0]);
            acc.v[1][1] = mfma_f32(b1, a1, acc.v[1][1]);
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

template <typename F>
__device__ __forceinline__ void acc32_foreach(const Acc32 &acc, F f) {
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
                const int c = nq * 32 + ni * 16 + 4 * (lane >> 4) + v;
                f(r, c, acc.v[ni][mi][v]);
            }
}

__global__ void __launch_bounds__(kThreads) trsm_rows_f32mfma_kernel(TileMap map, int k, int64_t g0, int jb0,
                                                                     int jb1, const double *winv_k) {
    __shared__ GemmSmem32 sm;
    __shared__ float sW[kBlk][kLdsPad];
    static_assert(sizeof(GemmSmem32::a) >= sizeof(float) * kBlk * kLdsPad, "T must fit in the staging buffer");
    float (*sT)[kLdsPad] = reinterpret_cast<float (*)[kLdsPad]>(&sm.a[0][0][0]);
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int64_t g = g0 + blockIdx.x;
    const int i = (int) (g / nsb);
    const int rb = (int) (g % nsb);
    float *X = reinterpret_cast<float *>(map.base + map.off[i * map.nt + k]) + rb * kBlk;
    const double *L = tile_f64(map, k, k);
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int mq = w & 1, nq = w >> 1;
    const int lr = lane & 15, lk = lane >> 4;
    for (int jb = jb0; jb < jb1; ++jb) {
        Acc32 acc;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc.v[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        if (jb > 0) wg_gemm_nt_f32(acc, X, ts, L + jb * kBlk, ts, jb * kBlk, sm);
        float *Xj = X + (int64_t) jb * kBlk * ts;
        acc32_foreach(acc, [&](int r, int c, float v) { sT[r][c] = Xj[r + (int64_t) c * ts] - v; });
        // W_jb (64 x 64 lower inverse, ld 64) as FP32 in LDS: sW[c][l] = W[c][l]
        const double *W = winv_k + (int64_t) jb * kBlk * kBlk;
        for (int e = threadIdx.x; e < kBlk * kBlk; e += kThreads) sW[e & 63][e >> 6] = (float) W[e];
        __syncthreads();
        Acc32 acc2;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc2.v[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
        for (int kk = 0; kk < kBlk; kk += 4) {
            const float a0 = sT[mq * 32 + lr][kk + lk];
            const float a1 = sT[mq * 32 + 16 + lr][kk + lk];
            const float b0 = sW[nq * 32 + lr][kk + lk];
            const float b1 = sW[nq * 32 + 16 + lr][kk + lk];
            acc2.v[0][0] = mfma_f32(b0, a0, acc2.v[0][0]);
            acc2.v[0][1] = mfma_f32(b0, a1, acc2.v[0][1]);
            acc2.v[1][0] = mfma_f32(b1, a0, acc2.v[1][0]);
            acc2.v[1][1] = mfma_f32(b1, a1, acc2.v[1][1]);
        }
        acc32_foreach(acc2, [&](int r, int c, float v) { Xj[r + (int64_t) c * ts] = v; });
        __syncthreads();
    }
}

// The FP32-MFMA solve in 32-row blocks (as trsm_rows_half_body): wave w
// holds rows 0-31 and columns 16 w .. 16 w + 15; acc.v[mi] element v of lane
// l is C[mi*16 + l%16][16 w + 4 (l/16) + v].
struct AccHalf32 {
    f32x4 v[2];
};

__device__ void wg_gemm_nt_f32_half(AccHalf32 &acc, const float *__restrict__ A, int64_t lda,
                                    const double *__restrict__ B, int64_t ldb, int K, GemmSmem32 &sm) {
    constexpr int kStageA = kHalf * kKc / kThreads;   // 4
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int arow = t & (kHalf - 1), acol = t / kHalf;
    const int brow = t & 63, bcol = t >> 6;
    float ra[kStageA], rb[kStage];
#pragma unroll
    for (int q = 0; q < kStageA; ++q) ra[q] = A[arow + (int64_t) (acol + q * (kThreads / kHalf)) * lda];
#pragma unroll
    for (int q = 0; q < kStage; ++q) rb[q] = (float) B[brow + (int64_t) (bcol + q * (kThreads / kBlk)) * ldb];
#pragma unroll
    for (int q = 0; q < kStageA; ++q) sm.a[0][acol + q * (kThreads / kHalf)][arow] = ra[q];
#pragma unroll
    for (int q = 0; q < kStage; ++q) sm.b[0][bcol + q * (kThreads / kBlk)][brow] = rb[q];
    __syncthreads();
    const int nch = K / kKc;
    for (int c = 0; c < nch; ++c) {
        const int buf = c & 1;
        const bool more = c + 1 < nch;
        if (more) {
            const int64_t k0 = (int64_t) (c + 1) * kKc;
#pragma unroll
            for (int q = 0; q < kStageA; ++q) ra[q] = A[arow + (k0 + acol + q * (kThreads / kHalf)) * lda];
#pragma unroll
            for (int q = 0; q < kStage; ++q) rb[q] = (float) B[brow + (k0 + bcol + q * (kThreads / kBlk)) * ldb];
        }
#pragma unroll
        for (int kk = 0; kk < kKc; kk += 4) {
            const float a0 = sm.a[buf][kk + lk][lr];
            const float a1 = sm.a[buf][kk + lk][16 + lr];
            const float b0 = sm.b[buf][kk + lk][w * 16 + lr];
            acc.v[0] = mfma_f32(b0, a0, acc.v[0]);
            acc.v[1] = mfma_f32(b0, a1, acc.v[1]);
        }
        if (more) {
#pragma unroll
            for (int q = 0; q < kStageA; ++q) sm.a[buf ^ 1][acol + q * (kThreads / kHalf)][arow] = ra[q];
#pragma unroll
            for (int q = 0; q < kStage; ++q) sm.b[buf ^ 1][bcol + q * (kThreads / kBlk)][brow] = rb[q];
        }
        __syncthreads();
    }
}

template <typename F>
__device__ __forceinline__ void acc32_foreach(const AccHalf32 &acc, F f) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int v = 0; v < 4; ++v) f(mi * 16 + (lane & 15), w * 16 + 4 * (lane >> 4) + v, acc.v[mi][v]);
}

// g0 in 32-row blocks; all column blocks.
__global__ void __launch_bounds__(kThreads) trsm_rows_f32mfma_half_kernel(TileMap map, int k, int64_t g0,
                                                                          const double *winv_k) {
    __shared__ GemmSmem32 sm;
    __shared__ float sW[kBlk][kLdsPad];
    float (*sT)[kLdsPad] = reinterpret_cast<float (*)[kLdsPad]>(&sm.a[0][0][0]);
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int nrb = ts / kHalf;
    const int64_t g = g0 + blockIdx.x;
    const int i = (int) (g / nrb);
    const int rb = (int) (g % nrb);
    float *X = reinterpret_cast<float *>(map.base + map.off[i * map.nt + k]) + rb * kHalf;
    const double *L = tile_f64(map, k, k);
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    for (int jb = 0; jb < nsb; ++jb) {
        AccHalf32 acc;
        acc.v[0] = f32x4{0.f, 0.f, 0.f, 0.f};
        acc.v[1] = f32x4{0.f, 0.f, 0.f, 0.f};
        if (jb > 0) wg_gemm_nt_f32_half(acc, X, ts, L + jb * kBlk, ts, jb * kBlk, sm);
        float *Xj = X + (int64_t) jb * kBlk * ts;
        acc32_foreach(acc, [&](int r, int c, float v) { sT[r][c] = Xj[r + (int64_t) c * ts] - v; });
        const double *W = winv_k + (int64_t) jb * kBlk * kBlk;
        for (int e = threadIdx.x; e < kBlk * kBlk; e += kThreads) sW[e & 63][e >> 6] = (float) W[e];
        __syncthreads();
        AccHalf32 acc2;
        acc2.v[0] = f32x4{0.f, 0.f, 0.f, 0.f};
        acc2.v[1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
        for (int kk = 0; kk < kBlk; kk += 4) {
            const float a0 = sT[lr][kk + lk];
            const float a1 = sT[16 + lr][kk + lk];
            const float b0 = sW[w * 16 + lr][kk + lk];
            acc2.v[0] = mfma_f32(b0, a0, acc2.v[0]);
            acc2.v[1] = mfma_f32(b0, a1, acc2.v[1]);
        }
        acc32_foreach(acc2, [&](int r, int c, float v) { Xj[r + (int64_t) c * ts] = v; });
        __syncthreads();
    }
}

// FP64 rows [g0, g0 + count64) and FP32 rows after them in one launch, both
// on FP64 MFMA, so that the two panel solves of a mixed plan run side by side
// rather than one after the other (each is a chain of ts / 64 dependent steps
// per workgroup).  Plans whose FP32 rows are solved on FP32 MFMA keep two
// launches: that kernel's lighter workgroups (50 KB of LDS) would otherwise
// carry the FP64 body's LDS and registers.
__global__ void __launch_bounds__(kThreads) trsm_rows_mixed_kernel(TileMap map, int k, int64_t g0, int count64,
                                                                   const double *winv_k) {
    __shared__ GemmSmem sm;
    const int nsb = map.ts / kBlk;
    if ((int) blockIdx.x < count64)
        trsm_rows_body<double>(map, k, g0, 0, nsb, winv_k, blockIdx.x, sm);
    else
        trsm_rows_body<float>(map, k, g0, 0, nsb, winv_k, blockIdx.x, sm);
}

// The same in 32-row blocks (g0 and count64 in 32-row blocks).
__global__ void __launch_bounds__(kThreads) trsm_rows_half_mixed_kernel(TileMap map, int k, int64_t g0, int count64,
                                                                        const double *winv_k) {
    __shared__ GemmSmem sm;
    if ((int) blockIdx.x < count64)
        trsm_rows_half_body<double>(map, k, g0, winv_k, blockIdx.x, sm);
    else
        trsm_rows_half_body<float>(map, k, g0, winv_k, blockIdx.x, sm);
}

// ------------------------------------------------------------------------
// Trailing update, FP64.
// ------------------------------------------------------------------------
// p64 (may be null): read panel k from its contiguous FP64 copy (tile i at
// p64 + (i - k - 1) ts^2) instead of from the tiles.
__device__ __forceinline__ void update_body(const TileMap &map, const double *p64, const int2 *pairs, int k, int bx,
                                            int by) {
    __shared__ GemmSmem sm;
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int2 ij = pairs[by];
    const int sbm = bx % nsb;
    const int sbn = bx / nsb;
    if (ij.x == ij.y && sbm < sbn) return;
    double *C = tile_f64(map, ij.x, ij.y) + sbm * kBlk + (int64_t) sbn * kBlk * ts;
    const int64_t tt = (int64_t) ts * ts;
    const double *A = (p64 ? p64 + (ij.x - k - 1) * tt : tile_f64(map, ij.x, k)) + sbm * kBlk;
    const double *B = (p64 ? p64 + (ij.y - k - 1) * tt : tile_f64(map, ij.y, k)) + sbn * kBlk;
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

__global__ void __launch_bounds__(kThreads) update_kernel(TileMap map, const double *p64, const int2 *pairs, int k) {
    update_body(map, p64, pairs, k, blockIdx.x, blockIdx.y);
}

// A panel of fewer 64-row blocks than this is solved in 32-row blocks (with
// more, every CU already holds a row block, and 64-row blocks load L_kk half
// as often per flop).
constexpr int kHalfRowsBelow = 256;

void launch_trsm_rows(const TileMap &map, int k, int64_t g0, int count, int jb0, int jb1, const double *winv_k,
                      hipStream_t s, bool x_f32, bool f32mfma) {
    if (count <= 0 || jb1 <= jb0) return;
    const bool halves = count < kHalfRowsBelow && jb0 == 0 && jb1 == map.ts / kBlk;
    if (halves && x_f32 && f32mfma)
        hipLaunchKernelGGL(trsm_rows_f32mfma_half_kernel, dim3(2 * count), dim3(kThreads), 0, s, map, k, 2 * g0,
                           winv_k);
    else if (x_f32 && f32mfma)
        hipLaunchKernelGGL(trsm_rows_f32mfma_kernel, dim3(count), dim3(kThreads), 0, s, map, k, g0, jb0, jb1, winv_k);
    else if (halves && x_f32)
        hipLaunchKernelGGL(trsm_rows_half_kernel<float>, dim3(2 * count), dim3(kThreads), 0, s, map, k, 2 * g0,
                           winv_k);
    else if (halves)
        hipLaunchKernelGGL(trsm_rows_half_kernel<double>, dim3(2 * count), dim3(kThreads), 0, s, map, k, 2 * g0,
                           winv_k);
    else if (x_f32)
        hipLaunchKernelGGL(trsm_rows_kernel<float>, dim3(count), dim3(kThreads), 0, s, map, k, g0, jb0, jb1, winv_k);
    else
        hipLaunchKernelGGL(trsm_rows_kernel<double>, dim3(count), dim3(kThreads), 0, s, map, k, g0, jb0, jb1, winv_k);
}

void launch_trsm_rows_mixed(const TileMap &map, int k, int64_t g0, int count64, int count32, const double *winv_k,
                            hipStream_t s) {
    if (count64 + count32 <= 0) return;
    if (count64 + count32 < kHalfRowsBelow)
        hipLaunchKernelGGL(trsm_rows_half_mixed_kernel, dim3((unsigned) (2 * (count64 + count32))), dim3(kThreads), 0,
                           s, map, k, 2 * g0, 2 * count64, winv_k);
    else
        hipLaunchKernelGGL(trsm_rows_mixed_kernel, dim3((unsigned) (count64 + count32)), dim3(kThreads), 0, s, map, k,
                           g0, count64, winv_k);
}

void launch_update(const TileMap &map, const int2 *pairs, int npairs, int k, hipStream_t s) {
    if (npairs <= 0) return;
    if (update192_supported(map.ts)) {
        launch_update192(map, pairs, npairs, k, s);
        return;
    }
    launch_update64(map, nullptr, pairs, npairs, k, s);
}

void launch_update64(const TileMap &map, const double *p64, const int2 *pairs, int npairs, int k, hipStream_t s) {
    if (npairs <= 0) return;
    const int nsb = map.ts / kBlk;
    hipLaunchKernelGGL(update_kernel, dim3(nsb * nsb, npairs), dim3(kThreads), 0, s, map, p64, pairs, k);
}

// ------------------------------------------------------------------------
// Inverses of the diagonal tiles, X_k = L_kk^{-T} (upper triangular,
// column-major, ts x ts), for the preconditioner solves of mixed plans.
// Workgroup (rb, k) solves row block rb of X_k L_kk^T = I left-looking over
// the 64-column blocks jb >= rb (blocks left of rb are zero), with the 64 x 64
// inverses W of the diagonal blocks of L_kk.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) tinv_kernel(TileMap map, const double *winv, double *tinv) {
    __shared__ GemmSmem sm;
    __shared__ double sT[kBlk][kLdsPad];
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int rb = blockIdx.x, k = blockIdx.y;
    double *X = tinv + (int64_t) k * ts * ts + rb * kBlk;
    const double *L = tile_f64(map, k, k);
    const double *winv_k = winv + (int64_t) k * nsb * kBlk * kBlk;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int mq = w & 1, nq = w >> 1;
    const int lr = lane & 15, lk = lane >> 4;
    for (int e = threadIdx.x; e < kBlk * ts; e += kThreads) {
        const int r = e & 63, c = e >> 6;
        X[r + (int64_t) c * ts] = (c == rb * kBlk + r) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int jb = rb; jb < nsb; ++jb) {
        Acc64 acc;
        acc_zero(acc);
        if (jb > rb)
            wg_gemm_nt(acc, X + (int64_t) rb * kBlk * ts, ts, L + jb * kBlk + (int64_t) rb * kBlk * ts, ts,
                       (jb - rb) * kBlk, sm);
        double *Xj = X + (int64_t) jb * kBlk * ts;
        acc_foreach(acc, [&](int r, int c, double v) { sT[r][c] = Xj[r + (int64_t) c * ts] - v; });
        __syncthreads();
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
        acc_foreach(acc2, [&](int r, int c, double v) { Xj[r + (int64_t) c * ts] = v; });
        __syncthreads();
    }
}

void launch_tinv(const TileMap &map, const double *winv, double *tinv, hipStream_t s) {
    const int nsb = map.ts / kBlk;
    hipLaunchKernelGGL(tinv_kernel, dim3(nsb, map.nt), dim3(kThreads), 0, s, map, winv, tinv);
}

// ------------------------------------------------------------------------
// Diagonal tile, right-looking in 64 x 64 blocks, one launch per block
// column (diag_step_kernel).  Launch jb:
//   workgroup 0         applies column jb - 1 to block (jb, jb), then factors
//                       that block and inverts it (potrf64_body);
//   workgroups 1 ..     apply column jb - 1 to the other blocks (r, c),
//                       jb <= c <= r (diag_update_body): form
//                       L_r = A_r,jb-1 W_jb-1^T and L_c = A_c,jb-1 W_jb-1^T and
//                       apply A_rc -= L_r L_c^T; the diagonal blocks (r, r)
//                       also write L_r to a scratch copy of the tile (the
//                       tile's own block (r, jb - 1) is still being read by
//                       other workgroups);
//   diag_copy_kernel    at the end moves the solved blocks below the diagonal
//                       from the scratch copy into the tile.
// Block (jb, jb) needs only column jb - 1's update before it is factored, so
// the rest of that update runs beside the factorization instead of before it.
// ------------------------------------------------------------------------
namespace {
// value of v in lane l (l uniform across the wave)
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
}  // namespace

// Block (jb, jb) of diagonal tile k: L = chol(A) in place and W = L^{-1} into
// winv_k, both lower (zeros above the diagonal).  The block lives in registers:
// thread (w, i) = (threadIdx.x / 64, lane) holds row i of columns c = w + kW q,
// q < 64 / kW, in y[q]: column c of A until step c, then column c of X, which
// becomes L^{-1}.  Column step j (64 steps, one barrier each) is the
// right-looking Cholesky step A_ic -= L_ij L_cj (c > j) fused with the
// column-oriented inversion X_jc /= L_jj, X_ic -= L_ij X_jc (i > j, c <= j).
// X is kept with its rows unscaled (row j of X times L_jj); the rows are
// scaled once at the end.  Both updates are then y_ic -= r_i g_c with
// r_i = A_ij / A_jj (i > j, else 0) and g_c the uniform operand: A_cj from LDS
// (column j, published by its owner wave at the end of step j - 1, double
// buffered) for c > j, the unscaled X_jc read from lane j for c <= j.  At step
// j the owner wave of column j stores L's column j and restarts its register
// as column j of X.
// kW = 4 waves, one per SIMD: the block size of the update workgroups it
// launches with (8 waves, two per SIMD, factor a block 2.5% faster alone).
constexpr int kPotrfWaves = kThreads / 64;

__device__ __forceinline__ void potrf64_body(const TileMap &map, int k, int jb, double *winv_k, int *info, int64_t n) {
    constexpr int kW = kPotrfWaves;
    constexpr int kCols = kBlk / kW;   // columns per thread
    __shared__ double s_col[2][kBlk];
    __shared__ double s_rs[kBlk];     // 1 / L_jj
    const int ts = map.ts;
    const int t = threadIdx.x;
    const int i = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);   // wave index, held in a scalar register
    double *Ajj = tile_f64(map, k, k) + (int64_t) jb * kBlk + (int64_t) jb * kBlk * ts;
    double y[kCols];
#pragma unroll
    for (int q = 0; q < kCols; ++q) y[q] = Ajj[i + (int64_t) (w + kW * q) * ts];
    if (w == 0) s_col[0][i] = y[0];
    __syncthreads();
    // j = kW p + jj: with p unrolled, column c = w + kW q is known to lie
    // before column j for q < p and after it for q > p; only q = p compares
    // the wave index with jj
#pragma unroll
    for (int p = 0; p < kCols; ++p) {
#pragma unroll 1
        for (int jj = 0; jj < kW; ++jj) {
            const int j = kW * p + jj;
            const int buf = jj & 1;
            // every LDS operand of the step is read here, ahead of the pivot's
            // arithmetic, so that the step waits on LDS once
            const double piv = s_col[buf][j];
            const double aij = s_col[buf][i];
            double gl[kCols];
#pragma unroll
            for (int q = p; q < kCols; ++q) gl[q] = s_col[buf][w + kW * q];
            if (t == 0 && !(piv > 0.0)) {
                const int64_t col = (int64_t) k * ts + (int64_t) jb * kBlk + j;
                if (col < n) atomicCAS(info, 0, (int) (col + 1));
            }
            // 1 / sqrt(piv): the hardware estimate and two Newton steps (each
            // doubles the correct bits; the chain is 7 dependent operations
            // instead of a correctly rounded square root followed by a division)
            double rs = __builtin_amdgcn_rsq(piv);
            const double hp = 0.5 * piv;
            rs = fma(rs, fma(-hp * rs, rs, 0.5), rs);
            rs = fma(rs, fma(-hp * rs, rs, 0.5), rs);
            const double sq = piv * rs;
            if (t == 0) s_rs[j] = rs;
            const double lo = i > j ? aij * rs : 0.0;   // L_ij below the diagonal
            const double r = lo * rs;
            // column j of the unscaled X after its own step
            const double xj0 = i == j ? 1.0 : -r;
#pragma unroll
            for (int q = 0; q < kCols; ++q) {
                if (q < p) {
                    // c < j: X's column, operand X_jc from lane j
                    y[q] = fma(-r, readlane_f64(y[q], j), y[q]);
                } else if (q > p) {
                    // c > j: A's column, operand A_cj from LDS
                    y[q] = fma(-r, gl[q], y[q]);
                    if (q == p + 1 && w == 0 && jj == kW - 1) s_col[buf ^ 1][i] = y[q];
                } else {
                    // c = w + kW p against j = kW p + jj: selects rather than
                    // branches, so that y does not pass through branch joins
                    const double g = w < jj ? readlane_f64(y[q], j) : gl[q];
                    const double upd = fma(-r, g, y[q]);
                    if (w == jj) Ajj[i + (int64_t) j * ts] = i > j ? lo : (i == j ? sq : 0.0);
                    if (w == jj + 1) s_col[buf ^ 1][i] = upd;
                    y[q] = w == jj ? xj0 : upd;   // c == j: L's column stored, X's column starts
                }
            }
            __syncthreads();
        }
    }
    double *W = winv_k + (int64_t) jb * kBlk * kBlk;
    const double rsi = s_rs[i];
#pragma unroll
    for (int q = 0; q < kCols; ++q) {
        const int c = w + kW * q;
        W[i + c * kBlk] = i >= c ? y[q] * rsi : 0.0;
    }
}

// acc = X (64 x 64, column-major, ldx) W^T, W the 64 x 64 lower inverse (ld 64).
__device__ __forceinline__ void times_wt(Acc64 &acc, const double *X, int64_t ldx, const double *W, GemmSmem &sm) {
    acc_zero(acc);
    wg_gemm_nt(acc, X, ldx, W, kBlk, kBlk, sm);
}

__device__ __forceinline__ void diag_update_body(const TileMap &map, int k, int jb, const double *winv_k,
                                                 double *scratch, int bx) {
    // 66.5 KB of LDS: L_r and L_c stay in registers until both GEMMs are done,
    // then go through sm (sm.a and sm.b), so that a workgroup fits beside one
    // workgroup of the trailing update
    __shared__ GemmSmem sm;
    static_assert(sizeof(GemmSmem::a) >= sizeof(double) * kBlk * kLdsPad, "L_r must fit in the staging buffer");
    static_assert(sizeof(GemmSmem::b) >= sizeof(double) * kBlk * kLdsPad, "L_c must fit in the staging buffer");
    double (*sLr)[kLdsPad] = reinterpret_cast<double (*)[kLdsPad]>(&sm.a[0][0][0]);
    double (*sLc)[kLdsPad] = reinterpret_cast<double (*)[kLdsPad]>(&sm.b[0][0][0]);
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    // bx enumerates the pairs jb < c <= r < nsb column by column
    int idx = bx, c = jb + 1;
    while (idx >= nsb - c) {
        idx -= nsb - c;
        ++c;
    }
    const int r = c + idx;
    double *T = tile_f64(map, k, k);
    const double *W = winv_k + (int64_t) jb * kBlk * kBlk;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int mq = wv & 1, nq = wv >> 1;
    const int lr = lane & 15, lk = lane >> 4;
    Acc64 acc_r, acc_c;
    times_wt(acc_r, T + r * kBlk + (int64_t) jb * kBlk * ts, ts, W, sm);
    if (c != r) times_wt(acc_c, T + c * kBlk + (int64_t) jb * kBlk * ts, ts, W, sm);
    // the GEMMs end with a barrier: sm is free
    acc_foreach(acc_r, [&](int rr, int cc, double v) { sLr[rr][cc] = v; });
    if (c != r) acc_foreach(acc_c, [&](int rr, int cc, double v) { sLc[rr][cc] = v; });
    __syncthreads();
    const double (*Lc)[kLdsPad] = (c != r) ? sLc : sLr;
    // A_rc -= L_r L_c^T
    Acc64 upd;
    acc_zero(upd);
#pragma unroll 4
    for (int kk = 0; kk < kBlk; kk += 4) {
        const double a0 = sLr[mq * 32 + lr][kk + lk];
        const double a1 = sLr[mq * 32 + 16 + lr][kk + lk];
        const double b0 = Lc[nq * 32 + lr][kk + lk];
        const double b1 = Lc[nq * 32 + 16 + lr][kk + lk];
        upd.v[0][0] = mfma_f64(b0, a0, upd.v[0][0]);
        upd.v[0][1] = mfma_f64(b0, a1, upd.v[0][1]);
        upd.v[1][0] = mfma_f64(b1, a0, upd.v[1][0]);
        upd.v[1][1] = mfma_f64(b1, a1, upd.v[1][1]);
    }
    double *C = T + r * kBlk + (int64_t) c * kBlk * ts;
    double cv[16];
    int q = 0;
    acc_foreach(upd, [&](int rr, int cc, double) { cv[q++] = C[rr + (int64_t) cc * ts]; });
    q = 0;
    acc_foreach(upd, [&](int rr, int cc, double v) { C[rr + (int64_t) cc * ts] = cv[q++] - v; });
    if (c == r) {
        double *S = scratch + r * kBlk + (int64_t) jb * kBlk * ts;
        for (int e = threadIdx.x; e < kBlk * kBlk; e += kThreads) {
            const int rr = e & 63, cc = e >> 6;
            S[rr + (int64_t) cc * ts] = sLr[rr][cc];
        }
    }
}

// Launch jb of the diagonal tile (see above); rem (rem + 1) / 2 workgroups
// for jb > 0, rem = nsb - jb, and one for jb = 0.
__device__ __forceinline__ void diag_step_body(const TileMap &map, int k, int jb, double *winv_k, double *scratch,
                                               int *info, int64_t n, int bx) {
    if (jb > 0) {
        // workgroup 0 takes the first pair, (jb, jb)
        diag_update_body(map, k, jb - 1, winv_k, scratch, bx);
        if (bx != 0) return;
        __syncthreads();   // block (jb, jb) updated by every thread
    }
    potrf64_body(map, k, jb, winv_k, info, n);
}

__global__ void __launch_bounds__(kThreads) diag_step_kernel(TileMap map, int k, int jb, double *winv_k,
                                                             double *scratch, int *info, int64_t n) {
    diag_step_body(map, k, jb, winv_k, scratch, info, n, blockIdx.x);
}

namespace {
unsigned diag_step_blocks(int nsb, int jb) {
    const int rem = nsb - jb;
    return jb == 0 ? 1u : (unsigned) (rem * (rem + 1) / 2);
}
}  // namespace

// Blocks (r, c), r > c, of the scratch copy back into diagonal tile k.
__device__ __forceinline__ void diag_copy_body(const TileMap &map, int k, const double *scratch, int bx) {
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    int idx = bx, c = 0;
    while (idx >= nsb - 1 - c) {
        idx -= nsb - 1 - c;
        ++c;
    }
    const int r = c + 1 + idx;
    double *T = tile_f64(map, k, k);
    const int64_t base = r * kBlk + (int64_t) c * kBlk * ts;
    for (int e = threadIdx.x; e < kBlk * kBlk; e += kThreads) {
        const int rr = e & 63, cc = e >> 6;
        T[base + rr + (int64_t) cc * ts] = scratch[base + rr + (int64_t) cc * ts];
    }
}

__global__ void __launch_bounds__(kThreads) diag_copy_kernel(TileMap map, int k, const double *scratch) {
    diag_copy_body(map, k, scratch, blockIdx.x);
}

// ------------------------------------------------------------------------
// Batched forms: one launch applies the same step to several problems of the
// same tile size, problem b described by jobs[b] (blockIdx.y, or blockIdx.z
// for the update); workgroups beyond a problem's extent exit at once.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) diag_step_batch_kernel(const BatchJob *jobs, int jb) {
    const BatchJob &J = jobs[blockIdx.y];
    diag_step_body(J.map, J.k, jb, J.winv_k, J.scratch, J.info, J.n, blockIdx.x);
}

__global__ void __launch_bounds__(kThreads) diag_copy_batch_kernel(const BatchJob *jobs) {
    const BatchJob &J = jobs[blockIdx.y];
    diag_copy_body(J.map, J.k, J.scratch, blockIdx.x);
}

__global__ void __launch_bounds__(kThreads) trsm_rows_batch_kernel(const BatchJob *jobs) {
    const BatchJob &J = jobs[blockIdx.y];
    const int nsb = J.map.ts / kBlk;
    if ((int) blockIdx.x >= (J.map.nt - J.k - 1) * nsb) return;
    __shared__ GemmSmem sm;
    trsm_rows_body<double>(J.map, J.k, (int64_t) (J.k + 1) * nsb, 0, nsb, J.winv_k, blockIdx.x, sm);
}

__global__ void __launch_bounds__(kThreads) trsm_rows_half_batch_kernel(const BatchJob *jobs) {
    const BatchJob &J = jobs[blockIdx.y];
    const int nrb = J.map.ts / kHalf;
    if ((int) blockIdx.x >= (J.map.nt - J.k - 1) * nrb) return;
    __shared__ GemmSmem sm;
    trsm_rows_half_body<double>(J.map, J.k, (int64_t) (J.k + 1) * nrb, J.winv_k, blockIdx.x, sm);
}

__global__ void __launch_bounds__(kThreads) update_batch_kernel(const BatchJob *jobs) {
    const BatchJob &J = jobs[blockIdx.z];
    const int rest = J.map.nt - J.k - 1;   // tile columns k+1 .. nt-1: the first rest (rest + 1) / 2 pairs
    if ((int) blockIdx.y >= rest * (rest + 1) / 2) return;
    update_body(J.map, nullptr, J.pairs, J.k, blockIdx.x, blockIdx.y);
}

void launch_factor_diag_batch(const BatchJob *jobs, int njobs, int ts, hipStream_t s) {
    if (njobs <= 0) return;
    const int nsb = ts / kBlk;
    for (int jb = 0; jb < nsb; ++jb)
        hipLaunchKernelGGL(diag_step_batch_kernel, dim3(diag_step_blocks(nsb, jb), njobs), dim3(kThreads), 0, s, jobs,
                           jb);
    if (nsb > 1)
        hipLaunchKernelGGL(diag_copy_batch_kernel, dim3((unsigned) (nsb * (nsb - 1) / 2), njobs), dim3(kThreads), 0, s,
                           jobs);
}

void launch_panel_update_batch(const BatchJob *jobs, int njobs, int ts, int rest, hipStream_t s) {
    if (njobs <= 0 || rest <= 0) return;
    const int nsb = ts / kBlk;
    if (rest * nsb * njobs < kHalfRowsBelow)
        hipLaunchKernelGGL(trsm_rows_half_batch_kernel, dim3((unsigned) (2 * rest * nsb), njobs), dim3(kThreads), 0, s,
                           jobs);
    else
        hipLaunchKernelGGL(trsm_rows_batch_kernel, dim3((unsigned) (rest * nsb), njobs), dim3(kThreads), 0, s, jobs);
    hipLaunchKernelGGL(update_batch_kernel, dim3((unsigned) (nsb * nsb), (unsigned) (rest * (rest + 1) / 2), njobs),
                       dim3(kThreads), 0, s, jobs);
}

void launch_factor_diag(const TileMap &map, int k, double *winv_k, double *scratch, int *info, int64_t n,
                        hipStream_t s) {
    const int nsb = map.ts / kBlk;
    for (int jb = 0; jb < nsb; ++jb)
        hipLaunchKernelGGL(diag_step_kernel, dim3(diag_step_blocks(nsb, jb)), dim3(kThreads), 0, s, map, k, jb, winv_k,
                           scratch, info, n);
    if (nsb > 1)
        hipLaunchKernelGGL(diag_copy_kernel, dim3((unsigned) (nsb * (nsb - 1) / 2)), dim3(kThreads), 0, s, map, k,
                           scratch);
}

}  // namespace egs
