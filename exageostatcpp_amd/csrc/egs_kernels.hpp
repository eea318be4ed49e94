// Host-side launchers of the gfx950 kernels.  All launches are asynchronous on
// the given stream; shapes are validated by the caller (egs_capi.cpp) before
// any launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "egs_common.hpp"
#include "egs_matern.hpp"

namespace egs {

// One problem's share of a batched step (FP64 plan): step k of the tiled
// Cholesky or of the forward solve on the tiles of `map`, with that problem's
// update order (pairs), diagonal-block inverses of tile k, diagonal scratch,
// not-SPD flag (n: its real dimension) and solve vector with its partials.
struct BatchJob {
    TileMap map;
    int k;
    const int2 *pairs;
    double *winv_k;
    double *scratch;
    int *info;
    int64_t n;
    double *z;
    double *partials;
};

// ---- covariance generation (egs_covgen.hip)
// sig64 (may be null): also writes every generated tile in FP64 to
// sig64 + (i*(i+1)/2 + j) * ts * ts, column-major.
// prm_dev, when given, holds the parameters in device memory (read instead of
// prm, so that a captured launch can be replayed for another theta).
void launch_covgen_tiles(const TileMap &map, const int2 *pairs, int npairs, const MaternParams &prm,
                         const double *x, const double *y, const double *z, int64_t n, double *sig64, hipStream_t s,
                         const MaternParams *prm_dev = nullptr);
void launch_covgen_rect(double *out, int rows, int cols, int64_t row_off, int64_t col_off, const MaternParams &prm,
                        const double *x1, const double *y1, const double *z1, const double *x2, const double *y2,
                        const double *z2, hipStream_t s);

// ---- tiled Cholesky (egs_chol.hip)
// Triangular solve X L_kk^T = B for row blocks [g0, g0 + count) of tile column k
// (global 64-row block g lives in tile g / nsb at row (g % nsb) * 64), column
// blocks [jb0, jb1).
void launch_trsm_rows(const TileMap &map, int k, int64_t g0, int count, int jb0, int jb1, const double *winv_k,
                      hipStream_t s, bool x_f32 = false, bool f32mfma = false);
// Both panel solves of a mixed plan in one launch on FP64 MFMA: count64 FP64
// row blocks from g0 on, then count32 FP32 ones.
void launch_trsm_rows_mixed(const TileMap &map, int k, int64_t g0, int count64, int count32, const double *winv_k,
                            hipStream_t s);
// Trailing update C_ij -= L_ik L_jk^T for the first npairs lower tiles of the
// update order (all (i, j) with i >= j > k).
void launch_update(const TileMap &map, const int2 *pairs, int npairs, int k, hipStream_t s);
// The 64 x 64-block update whatever ts (more, shorter workgroups than the
// 192 x 192 form: lower latency for a few tiles); p64 as launch_update192_panel
// when given, else panel k is read from the tiles.
void launch_update64(const TileMap &map, const double *p64, const int2 *pairs, int npairs, int k, hipStream_t s);
// 192 x 192-block variant of the update (egs_update.hip), used when ts % 192 == 0.
bool update192_supported(int ts);
void launch_update192(const TileMap &map, const int2 *pairs, int npairs, int k, hipStream_t s);

// ---- triangular solves and reductions (egs_solve.hip)
// Forward solve of tile row k: y_k = L_kk^{-1} z_k in place, with per-tile
// partials[2k] = sum log L_ii (L_ii > 0), partials[2k+1] = y_k . y_k.
void launch_trsv_diag(const TileMap &map, int k, const double *winv_k, double *z, double *partials, hipStream_t s);
// Backward solve with L^T of tile row k (after rows > k are solved):
// partial sums of L_ik^T y_i into scratch, then the diagonal solve.
void launch_trsv_t_partial(const TileMap &map, int k, const double *z, double *scratch, hipStream_t s);
void launch_trsv_t_diag(const TileMap &map, int k, const double *winv_k, double *z, const double *scratch,
                        hipStream_t s);
// out = L e (lower-triangular product over the tiles).
void launch_trmv(const TileMap &map, const double *e, double *out, hipStream_t s);
// out[r] = sum_c Sigma(l1_r, l2_c) x[c] for r < rows, c < cols (kriging mean).
void launch_cov_gemv(double *out, int64_t rows, int64_t cols, const MaternParams &prm, const double *x1,
                     const double *y1, const double *z1, const double *x2, const double *y2, const double *z2,
                     const double *xv, hipStream_t s);
// partial[cb * npad + r] = sum over column chunk cb of Sigma(r, c) v[c] (FP64,
// regenerated from the locations); matvec_chunks(n) chunks.
void launch_cov_matvec(double *partial, const double *v, int64_t n, int64_t npad, const MaternParams &prm,
                       const double *x, const double *y, const double *z, hipStream_t s);
int matvec_chunks(int64_t n);
// Cholesky of diagonal tile k in place, right-looking in 64 x 64 blocks, with
// the inverses of its diagonal blocks in winv_k; scratch holds ts * ts doubles.
void launch_factor_diag(const TileMap &map, int k, double *winv_k, double *scratch, int *info, int64_t n,
                        hipStream_t s);
// Batched steps over njobs problems of tile size ts (jobs in device memory):
// factor diagonal tile k of each; then solve panel k and apply it to tile
// columns k+1 .. of each (rest = the largest nt - k - 1 among the jobs; jobs
// with fewer tiles exit early); forward-solve tile row k with its partials
// and subtract it from the rows below.
void launch_factor_diag_batch(const BatchJob *jobs, int njobs, int ts, hipStream_t s);
void launch_panel_update_batch(const BatchJob *jobs, int njobs, int ts, int rest, hipStream_t s);
void launch_forward_batch(const BatchJob *jobs, int njobs, int ts, int rest, hipStream_t s);
// X_k = L_kk^{-T} for every diagonal tile k (tinv: nt * ts * ts doubles).
void launch_tinv(const TileMap &map, const double *winv, double *tinv, hipStream_t s);
// Preconditioner solves with the diagonal-tile inverses (right-looking):
// forward step k: y_k = X_k' z_k, then z_i -= L_ik y_k for i > k;
// backward step k: y_k = X_k x_k, then x_j -= L_kj' y_k for j < k.
void launch_fwd_diag_inv(const double *tinv_k, const double *z_k, double *y_k, int ts, hipStream_t s);
void launch_fwd_panel(const TileMap &map, int k, const double *y_k, double *z, hipStream_t s);
void launch_bwd_diag_inv(const double *tinv_k, const double *x_k, double *y_k, int ts, hipStream_t s);
void launch_bwd_row(const TileMap &map, int k, const double *y_k, double *x, hipStream_t s);
// logdet partials: partials[2k] = sum log L_ii over diagonal tile k (L_ii > 0).
void launch_diag_logdet(const TileMap &map, int64_t n, double *partials, hipStream_t s);
// Sigma v from the FP64 lower tiles written by launch_covgen_tiles: one
// workgroup per tile (i, j) writes T_ij v_j to prow[t * ts ..] and, off the
// diagonal, T_ij' v_i to pcol[t * ts ..] (t = i*(i+1)/2 + j).
void launch_sigma_tile_matvec(const double *sig64, const int2 *pairs, int npairs, int ts, const double *v,
                              double *prow, double *pcol, hipStream_t s);
// out[g] = base[g] + scale * (Sigma v)[g] for g < n (0 for padding rows), summing
// the tile partials of row block g / ts in a fixed order.
void launch_sigma_reduce(double *out, const double *base, double scale, const double *prow, const double *pcol,
                         int nt, int ts, int64_t n, int64_t npad, hipStream_t s);
// out = base + scale * (sum of the chunk partials), padding rows set to zero.
void launch_matvec_reduce(double *out, const double *base, double scale, const double *partial, int nchunks,
                          int64_t n, int64_t npad, hipStream_t s);
// partial[b] (b < dot_blocks()) = fixed-order partial sums of x . y.
int dot_blocks();
void launch_dot(const double *x, const double *y, int64_t n, double *partial, hipStream_t s);
// y = a x + b y
void launch_axpby(double *y, const double *x, double a, double b, int64_t n, hipStream_t s);
// Mixed-precision factorization pieces (egs_mixed.hip).
// inv_scale[slot * ts + r] = 2^(e+1) for the largest |x| = f 2^e of row r of panel
// slot; bmax[slot * (ts/192) + b] the largest of 192-row block b; pmax[k] raised
// to the largest of the panel (bits of a float, must start at zero).
void launch_panel_rowscale(const TileMap &map, int k, float *inv_scale, float *bmax, unsigned *pmax,
                           hipStream_t s);
void launch_zero_u32(unsigned *p, int n, hipStream_t s);
// Panel k to FP64 and FP32 column-major copies and 16-bit hi/lo row-major parts:
// FP16 of the row-scaled values when inv_scale is given, BF16 of the values otherwise.
// With inv_scale, p8 (optional) receives the OCP e4m3 byte of each row-scaled
// value, row-major like the 16-bit parts.
void launch_panel_convert(const TileMap &map, int k, double *p64, float *p32, uint16_t *p16h, uint16_t *p16l,
                          uint8_t *p8, const float *inv_scale, hipStream_t s);
void launch_update192_f32(const TileMap &map, const float *p32, const int2 *pairs, int npairs, int k, hipStream_t s);
// Operands of the 16-bit update: split FP16 (hi*hi + hi*lo + lo*hi of the
// row-scaled parts), one FP16 product of the row-scaled hi parts, the choice
// between the two per 192 x 192 block (one product when bmax(A rows) *
// bmax(B rows) <= tau * pmax[k]^2), or one BF16 product of unscaled values.
// FP16 forms undo the row scales with inv_scale.
// The C64 forms update an FP64 tile: the split product, one product, or the
// choice per block.
enum Operands {
    kSplitF16 = 0, kSingleF16 = 1, kAdaptiveF16 = 2, kSingleBF16 = 3, kSplitF16C64 = 4, kSingleF16C64 = 5,
    kAdaptiveF16C64 = 6
};
// C -= A B^T from the 16-bit parts of panel k.  The single and per-block forms
// add the number of 192 x 192 blocks they update with the split product to
// counts[0], with one product to counts[1] and with one FP8 product to
// counts[2].  kAdaptiveF16 with tau8 > 0 takes one e4m3 product from p8 (the
// panel's FP8 copy) for blocks with bmax(A) bmax(B) <= tau8 pmax[k]^2.
void launch_update192_16(const TileMap &map, const uint16_t *p16h, const uint16_t *p16l, const float *inv_scale,
                         const float *bmax, const unsigned *pmax, float tau, const uint8_t *p8, float tau8,
                         const int2 *pairs, int npairs, int k, Operands op, unsigned long long *counts,
                         hipStream_t s);
// FP64 update reading panel k from a contiguous FP64 copy (tile i at p64 + (i-k-1) ts^2).
void launch_update192_panel(const TileMap &map, const double *p64, const int2 *pairs, int npairs, int k,
                            hipStream_t s);
// e = LAPACK dlarnv(3, iseed) N(0,1) values (48-bit LCG stream, Box-Muller).
void launch_dlarnv_normal(double *e, int64_t n, uint64_t seed48, hipStream_t s);
// dst = src (n doubles), padding [n, npad) with zeros.
void launch_copy_pad(double *dst, const double *src, int64_t n, int64_t npad, hipStream_t s);

}  // namespace egs
