// Mixed-precision pieces of the tiled Cholesky on gfx950.
//
// Plan: tile (i, j) with d = i - j is
//   d <= b64        : stored FP64, updated with FP64 MFMA (egs_update.hip);
//   b64 < d <= b32  : stored FP32, updated with v_mfma_f32_16x16x4_f32;
//   d > b32         : stored FP32, updated with v_mfma_f32_16x16x32_bf16 on
//                     BF16 operands with FP32 accumulation.  By default each
//                     FP32 operand is split into hi = bf16(x) and lo = bf16(x - hi)
//                     and the product is hi*hi + hi*lo + lo*hi (three BF16
//                     MFMAs, ~16 significant bits); a single hi*hi product is the
//                     plain-BF16 option.
// After panel k is solved, panel_convert_kernel writes it once in the layouts
// the three update kernels read: FP64 column-major, FP32 column-major, and FP32
// row-major (K-contiguous, for the BF16 operands that pack 8 consecutive K
// values per lane).  Diagonal tiles, their factorization, the log-determinant
// and every residual stay FP64.
#include "egs_common.hpp"
#include "egs_kernels.hpp"

namespace egs {

namespace {

constexpr int kBM = 192;
constexpr int kNT = 512;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma_bf16(const bf16x8 &x, const bf16x8 &y, f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, y, acc, 0, 0, 0);
}

// ------------------------------------------------------------------------
// Panel conversion: tile (i, k) of panel k -> P64, P32 (column-major) and
// P32T (row-major) at slot i - k - 1.  One workgroup per 64 x 64 block.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) panel_convert_kernel(TileMap map, int k, double *p64, float *p32,
                                                                 float *p32t) {
    __shared__ float sT[kBlk][kBlk + 1];
    const int ts = map.ts;
    const int nsb = ts / kBlk;
    const int per_tile = nsb * nsb;
    const int slot = blockIdx.x / per_tile;
    const int sub = blockIdx.x % per_tile;
    const int br = sub % nsb, bc = sub / nsb;
    const int i = k + 1 + slot;
    const bool f32 = map.store[i * map.nt + k] == kStoreF32;
    const char *src = map.base + map.off[i * map.nt + k];
    const int64_t slot_off = (int64_t) slot * ts * ts;
    const int t = threadIdx.x;
    const int r = t & 63;
    for (int c = t >> 6; c < kBlk; c += kThreads / kBlk) {
        const int64_t idx = (int64_t) (br * kBlk + r) + (int64_t) (bc * kBlk + c) * ts;
        const double v = f32 ? (double) reinterpret_cast<const float *>(src)[idx]
                             : reinterpret_cast<const double *>(src)[idx];
        p64[slot_off + idx] = v;
        p32[slot_off + idx] = (float) v;
        sT[c][r] = (float) v;
    }
    __syncthreads();
    // row-major: element (row, col) at row * ts + col; consecutive threads walk columns
    for (int rr = t >> 6; rr < kBlk; rr += kThreads / kBlk) {
        const int cc = t & 63;
        p32t[slot_off + (int64_t) (br * kBlk + rr) * ts + bc * kBlk + cc] = sT[cc][rr];
    }
}

// ------------------------------------------------------------------------
// FP32-class update: C_ij (FP32) -= A B^T with A, B from the FP32 column-major
// panel, v_mfma_f32_16x16x4_f32.  Same block geometry as the FP64 kernel:
// 192 x 192 per workgroup, wave tile 96 x 48, K chunk 32 double buffered.
// FP32 MFMA output: lane l, element v is D[4*(l/16) + v][l%16].
// ------------------------------------------------------------------------
constexpr int kKC32 = 32;
constexpr int kLds32 = kBM + 4;
constexpr int kPer32 = kBM * kKC32 / kNT;   // 12

__global__ void __launch_bounds__(kNT, 1) update192_f32_kernel(TileMap map, const float *p32, const int2 *pairs,
                                                               int k) {
    __shared__ float sA[2][kKC32][kLds32];
    __shared__ float sB[2][kKC32][kLds32];
    const int ts = map.ts;
    const int nsb = ts / kBM;
    const int nsub = nsb * nsb;
    const int pair = blockIdx.x / nsub;
    const int sub = blockIdx.x % nsub;
    const int sbm = sub % nsb, sbn = sub / nsb;
    const int2 ij = pairs[pair];
    const int64_t tsz = (int64_t) ts * ts;
    const float *A = p32 + (int64_t) (ij.x - k - 1) * tsz + sbm * kBM;
    const float *B = p32 + (int64_t) (ij.y - k - 1) * tsz + sbn * kBM;
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;
    const int lr = lane & 15, lk = lane >> 4;

    uint32_t goff[kPer32];
    int loff[kPer32];
#pragma unroll
    for (int q = 0; q < kPer32; ++q) {
        const int e = t + q * kNT;
        goff[q] = (uint32_t) ((e % kBM) + (e / kBM) * ts);
        loff[q] = (e / kBM) * kLds32 + (e % kBM);
    }
    const uint32_t kstep = (uint32_t) (kKC32 * ts);

    f32x4 acc[6][3];
#pragma unroll
    for (int mi = 0; mi < 6; ++mi)
#pragma unroll
        for (int ni = 0; ni < 3; ++ni) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};

    float ra[kPer32], rb[kPer32];
#pragma unroll
    for (int q = 0; q < kPer32; ++q) {
        ra[q] = A[goff[q]];
        rb[q] = B[goff[q]];
    }
#pragma unroll
    for (int q = 0; q < kPer32; ++q) {
        (&sA[0][0][0])[loff[q]] = ra[q];
        (&sB[0][0][0])[loff[q]] = rb[q];
    }
    __syncthreads();
    const int nchunks = ts / kKC32;
    uint32_t kofs = 0;
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        const bool more = c + 1 < nchunks;
        kofs += kstep;
        if (more) {
#pragma unroll
            for (int q = 0; q < kPer32; ++q) {
                ra[q] = A[goff[q] + kofs];
                rb[q] = B[goff[q] + kofs];
            }
        }
#pragma unroll
        for (int kk = 0; kk < kKC32; kk += 4) {
            float a[6], b[3];
#pragma unroll
            for (int mi = 0; mi < 6; ++mi) a[mi] = sA[buf][kk + lk][wm * 96 + mi * 16 + lr];
#pragma unroll
            for (int ni = 0; ni < 3; ++ni) b[ni] = sB[buf][kk + lk][wn * 48 + ni * 16 + lr];
#pragma unroll
            for (int mi = 0; mi < 6; ++mi)
#pragma unroll
                for (int ni = 0; ni < 3; ++ni) acc[mi][ni] = mfma_f32(b[ni], a[mi], acc[mi][ni]);
        }
        if (more) {
            float *dA = &sA[buf ^ 1][0][0];
            float *dB = &sB[buf ^ 1][0][0];
#pragma unroll
            for (int q = 0; q < kPer32; ++q) {
                dA[loff[q]] = ra[q];
                dB[loff[q]] = rb[q];
            }
        }
        __syncthreads();
    }
    float *C = tile_f32(map, ij.x, ij.y);
    const uint32_t cbase = (uint32_t) ((sbm * kBM + wm * 96 + lr) + (sbn * kBM + wn * 48 + 4 * lk) * ts);
#pragma unroll
    for (int mi = 0; mi < 6; ++mi) {
        float cv[3][4];
#pragma unroll
        for (int ni = 0; ni < 3; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) cv[ni][v] = C[cbase + (uint32_t) (mi * 16 + (ni * 16 + v) * ts)];
#pragma unroll
        for (int ni = 0; ni < 3; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) C[cbase + (uint32_t) (mi * 16 + (ni * 16 + v) * ts)] = cv[ni][v] - acc[mi][ni][v];
    }
}

// ------------------------------------------------------------------------
// BF16-class update: C_ij (FP32) -= A B^T with BF16 products on the
// row-major FP32 panel.  K chunk 32 = one v_mfma_f32_16x16x32_bf16 step; LDS
// holds the chunk m-major (32 K-values per row) so that a lane reads its 8
// consecutive K values with two 16-byte loads.  Lane l supplies row l%16 and
// K values 8*(l/16) .. +7 of both operands.
// ------------------------------------------------------------------------
constexpr int kKC16 = 32;
constexpr int kRow16 = kKC16 + 4;          // floats per LDS row (16-byte aligned)
constexpr int kPer16 = kBM * kKC16 / 4 / kNT;   // float4 loads per operand per thread: 3

__device__ __forceinline__ void split_bf16(const f32x4v &x0, const f32x4v &x1, bf16x8 &hi, bf16x8 &lo) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const __bf16 h0 = (__bf16) x0[e];
        const __bf16 h1 = (__bf16) x1[e];
        hi[e] = h0;
        hi[e + 4] = h1;
        lo[e] = (__bf16) (x0[e] - (float) h0);
        lo[e + 4] = (__bf16) (x1[e] - (float) h1);
    }
}

template <bool kSplit>
__global__ void __launch_bounds__(kNT, 1) update192_bf16_kernel(TileMap map, const float *p32t, const int2 *pairs,
                                                                int k) {
    __shared__ f32x4v sA[2][kBM][kRow16 / 4];
    __shared__ f32x4v sB[2][kBM][kRow16 / 4];
    const int ts = map.ts;
    const int nsb = ts / kBM;
    const int nsub = nsb * nsb;
    const int pair = blockIdx.x / nsub;
    const int sub = blockIdx.x % nsub;
    const int sbm = sub % nsb, sbn = sub / nsb;
    const int2 ij = pairs[pair];
    const int64_t tsz = (int64_t) ts * ts;
    // row-major tiles: element (row, col) at row * ts + col
    const float *A = p32t + (int64_t) (ij.x - k - 1) * tsz + (int64_t) sbm * kBM * ts;
    const float *B = p32t + (int64_t) (ij.y - k - 1) * tsz + (int64_t) sbn * kBM * ts;
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;
    const int lr = lane & 15, lk = lane >> 4;

    uint32_t goff[kPer16];
    int lrow[kPer16], lc4[kPer16];
#pragma unroll
    for (int q = 0; q < kPer16; ++q) {
        const int e = t + q * kNT;          // 0 .. 1535: row e / 8, float4 column e % 8
        lrow[q] = e >> 3;
        lc4[q] = e & 7;
        goff[q] = (uint32_t) ((e >> 3) * ts + (e & 7) * 4);
    }

    f32x4 acc[6][3];
#pragma unroll
    for (int mi = 0; mi < 6; ++mi)
#pragma unroll
        for (int ni = 0; ni < 3; ++ni) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};

    f32x4v ra[kPer16], rb[kPer16];
#pragma unroll
    for (int q = 0; q < kPer16; ++q) {
        ra[q] = *reinterpret_cast<const f32x4v *>(A + goff[q]);
        rb[q] = *reinterpret_cast<const f32x4v *>(B + goff[q]);
    }
#pragma unroll
    for (int q = 0; q < kPer16; ++q) {
        sA[0][lrow[q]][lc4[q]] = ra[q];
        sB[0][lrow[q]][lc4[q]] = rb[q];
    }
    __syncthreads();
    const int nchunks = ts / kKC16;
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        const bool more = c + 1 < nchunks;
        if (more) {
            const uint32_t kofs = (uint32_t) ((c + 1) * kKC16);
#pragma unroll
            for (int q = 0; q < kPer16; ++q) {
                ra[q] = *reinterpret_cast<const f32x4v *>(A + goff[q] + kofs);
                rb[q] = *reinterpret_cast<const f32x4v *>(B + goff[q] + kofs);
            }
        }
        bf16x8 ah[6], al[6], bh[3], bl[3];
#pragma unroll
        for (int mi = 0; mi < 6; ++mi) {
            const int row = wm * 96 + mi * 16 + lr;
            split_bf16(sA[buf][row][2 * lk], sA[buf][row][2 * lk + 1], ah[mi], al[mi]);
        }
#pragma unroll
        for (int ni = 0; ni < 3; ++ni) {
            const int row = wn * 48 + ni * 16 + lr;
            split_bf16(sB[buf][row][2 * lk], sB[buf][row][2 * lk + 1], bh[ni], bl[ni]);
        }
#pragma unroll
        for (int mi = 0; mi < 6; ++mi)
#pragma unroll
            for (int ni = 0; ni < 3; ++ni) {
                acc[mi][ni] = mfma_bf16(bh[ni], ah[mi], acc[mi][ni]);
                if (kSplit) {
                    acc[mi][ni] = mfma_bf16(bh[ni], al[mi], acc[mi][ni]);
                    acc[mi][ni] = mfma_bf16(bl[ni], ah[mi], acc[mi][ni]);
                }
            }
        if (more) {
#pragma unroll
            for (int q = 0; q < kPer16; ++q) {
                sA[buf ^ 1][lrow[q]][lc4[q]] = ra[q];
                sB[buf ^ 1][lrow[q]][lc4[q]] = rb[q];
            }
        }
        __syncthreads();
    }
    float *C = tile_f32(map, ij.x, ij.y);
    const uint32_t cbase = (uint32_t) ((sbm * kBM + wm * 96 + lr) + (sbn * kBM + wn * 48 + 4 * lk) * ts);
#pragma unroll
    for (int mi = 0; mi < 6; ++mi) {
        float cv[3][4];
#pragma unroll
        for (int ni = 0; ni < 3; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) cv[ni][v] = C[cbase + (uint32_t) (mi * 16 + (ni * 16 + v) * ts)];
#pragma unroll
        for (int ni = 0; ni < 3; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) C[cbase + (uint32_t) (mi * 16 + (ni * 16 + v) * ts)] = cv[ni][v] - acc[mi][ni][v];
    }
}

}  // namespace

void launch_panel_convert(const TileMap &map, int k, double *p64, float *p32, float *p32t, hipStream_t s) {
    const int rows = map.nt - k - 1;
    if (rows <= 0) return;
    const int nsb = map.ts / kBlk;
    hipLaunchKernelGGL(panel_convert_kernel, dim3((unsigned) (rows * nsb * nsb)), dim3(kThreads), 0, s, map, k, p64,
                       p32, p32t);
}

void launch_update192_f32(const TileMap &map, const float *p32, const int2 *pairs, int npairs, int k, hipStream_t s) {
    if (npairs <= 0) return;
    const int nsb = map.ts / kBM;
    hipLaunchKernelGGL(update192_f32_kernel, dim3((unsigned) npairs * (unsigned) (nsb * nsb)), dim3(kNT), 0, s, map,
                       p32, pairs, k);
}

void launch_update192_bf16(const TileMap &map, const float *p32t, const int2 *pairs, int npairs, int k, bool split,
                           hipStream_t s) {
    if (npairs <= 0) return;
    const int nsb = map.ts / kBM;
    const dim3 grid((unsigned) npairs * (unsigned) (nsb * nsb));
    if (split) hipLaunchKernelGGL(update192_bf16_kernel<true>, grid, dim3(kNT), 0, s, map, p32t, pairs, k);
    else hipLaunchKernelGGL(update192_bf16_kernel<false>, grid, dim3(kNT), 0, s, map, p32t, pairs, k);
}

}  // namespace egs
