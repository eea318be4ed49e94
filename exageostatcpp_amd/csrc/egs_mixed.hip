// Mixed-precision pieces of the tiled Cholesky on gfx950.
//
// Plan: tile (i, j) with d = i - j is
//   d <= b64        : stored FP64, updated with FP64 MFMA (egs_update.hip);
//   b64 < d <= b32  : stored FP32, updated with v_mfma_f32_16x16x4_f32;
//   d > b32         : stored FP32, updated with 32x32x16 MFMAs on 16-bit
//                     operands with FP32 accumulation.  By default every row of
//                     the panel is scaled by a power of two so that its largest
//                     entry lies in [1/2, 1), and each scaled value is split into
//                     hi = fp16(x) and lo = fp16(x - hi); the product is
//                     hi*hi + hi*lo + lo*hi (three FP16 MFMAs, ~22 significant
//                     bits per operand) and the epilogue undoes the two row
//                     scales exactly.  A single BF16 product of the unscaled
//                     values is the plain-BF16 option.
// After panel k is solved, panel_convert_kernel writes it once in the layouts
// the three update kernels read: FP64 column-major, FP32 column-major, and the
// 16-bit hi and lo parts row-major (K-contiguous: an MFMA operand is 8
// consecutive K values of one row, one 16-byte load).  The split costs no
// arithmetic in the GEMM.  Diagonal tiles, their factorization, the
// log-determinant and every residual stay FP64.
#include <cstdint>
#include <type_traits>

#include "egs_common.hpp"
#include "egs_kernels.hpp"

namespace egs {

namespace {

constexpr int kBM = 192;
constexpr int kNT = 512;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x4v __attribute__((ext_vector_type(4)));


// ------------------------------------------------------------------------
// Row scales of panel k: inv_scale[slot * ts + r] = 2^(e+1) for the largest
// |x| = f 2^e (1 <= f < 2) of row r of tile (k+1+slot, k), 1 for a zero row.
// Also the largest scale of each 192-row block, bmax[slot * (ts/192) + b],
// and of the whole panel, pmax[k] (zero rows count 0; pmax is the bits of a
// non-negative float, raised with atomicMax, and must start at zero).  One
// workgroup per 192 rows; four column groups reduced through LDS.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) panel_rowscale_kernel(TileMap map, int k, float *inv_scale, float *bmax,
                                                                  unsigned *pmax) {
    __shared__ float sm[4][kBM];
    __shared__ float sw[kThreads / 64];
    const int ts = map.ts;
    const int nb = ts / kBM;
    const int slot = blockIdx.x / nb;
    const int b = blockIdx.x % nb;
    const int i = k + 1 + slot;
    const int t = threadIdx.x;
    const int r = t & 63, g = t >> 6;
    const bool f32 = map.store[i * map.nt + k] == kStoreF32;
    // 48 independent loads in flight per thread (3 rows x 16 columns): a
    // column-at-a-time loop waits one memory latency per element
    constexpr int kRows = kBM / 64, kCols = 16;
    float m[kRows];
#pragma unroll
    for (int rs = 0; rs < kRows; ++rs) m[rs] = 0.0f;
    auto scan = [&](const auto *A) {
        for (int c0 = g; c0 < ts; c0 += 4 * kCols) {   // ts is a multiple of 64
            float v[kRows][kCols];
#pragma unroll
            for (int rs = 0; rs < kRows; ++rs)
#pragma unroll
                for (int u = 0; u < kCols; ++u)
                    v[rs][u] = (float) A[(int64_t) b * kBM + rs * 64 + r + (int64_t) (c0 + 4 * u) * ts];
#pragma unroll
            for (int rs = 0; rs < kRows; ++rs)
#pragma unroll
                for (int u = 0; u < kCols; ++u) m[rs] = fmaxf(m[rs], fabsf(v[rs][u]));
        }
    };
    if (f32) scan(tile_f32(map, i, k));
    else scan(tile_f64(map, i, k));
#pragma unroll
    for (int rs = 0; rs < kRows; ++rs) sm[g][rs * 64 + r] = m[rs];
    __syncthreads();
    float mag = 0.0f;
    if (t < kBM) {
        const float mx = fmaxf(fmaxf(sm[0][t], sm[1][t]), fmaxf(sm[2][t], sm[3][t]));
        const bool finite = mx > 0.0f && mx < 3.0e38f;
        const float sc = finite ? ldexpf(1.0f, ilogbf(mx) + 1) : 1.0f;
        inv_scale[(int64_t) slot * ts + b * kBM + t] = sc;
        mag = mx > 0.0f ? sc : 0.0f;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mag = fmaxf(mag, __shfl_xor(mag, off));
    if (r == 0) sw[g] = mag;
    __syncthreads();
    if (t == 0) {
        float bm = sw[0];
#pragma unroll
        for (int q = 1; q < kThreads / 64; ++q) bm = fmaxf(bm, sw[q]);
        bmax[slot * nb + b] = bm;
        atomicMax(pmax + k, __float_as_uint(bm));
    }
}

__global__ void zero_u32_kernel(unsigned *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ------------------------------------------------------------------------
// Panel conversion: tile (i, k) of panel k -> P64, P32 (column-major) and the
// 16-bit hi/lo parts (row-major) at slot i - k - 1.  With inv_scale, the parts
// are FP16 of the row-scaled values; without, BF16 of the values.  One
// workgroup per 64 x 64 block.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) panel_convert_kernel(TileMap map, int k, double *p64, float *p32,
                                                                 uint16_t *p16h, uint16_t *p16l, uint8_t *p8,
                                                                 const float *inv_scale) {
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
        const float x = sT[cc][rr];
        const int64_t o = slot_off + (int64_t) (br * kBlk + rr) * ts + bc * kBlk + cc;
        if (inv_scale) {
            const float xs = x / inv_scale[(int64_t) slot * ts + br * kBlk + rr];   // exact: a power of two
            const _Float16 hi = (_Float16) xs;
            const _Float16 lo = (_Float16) (xs - (float) hi);
            p16h[o] = __builtin_bit_cast(uint16_t, hi);
            p16l[o] = __builtin_bit_cast(uint16_t, lo);
            // OCP e4m3 of the same row-scaled value (|xs| < 1, nearest even)
            if (p8) p8[o] = (uint8_t) (__builtin_amdgcn_cvt_pk_fp8_f32(xs, 0.0f, 0, false) & 0xff);
        } else {
            const __bf16 hi = (__bf16) x;
            const __bf16 lo = (__bf16) (x - (float) hi);
            p16h[o] = __builtin_bit_cast(uint16_t, hi);
            p16l[o] = __builtin_bit_cast(uint16_t, lo);
        }
    }
}

// ------------------------------------------------------------------------
// FP32-class update: C_ij (FP32) -= A B^T with A, B from the FP32 column-major
// panel, v_mfma_f32_16x16x4_f32.  Same block geometry as the FP64 kernel:
// 192 x 192 per workgroup, wave tile 96 x 48, K chunk 32 double buffered.
// FP32 MFMA output: lane l, element v is D[4*(l/16) + v][l%16].
// ------------------------------------------------------------------------
constexpr int kKC32 = 32;
constexpr int kLds32 = kBM + 16;   // 208 floats: the 4 K-rows a wave reads use disjoint banks
constexpr int kPer32 = kBM * kKC32 / kNT;   // 12

__global__ void __launch_bounds__(kNT, 1) update192_f32_kernel(TileMap map, const float *p32, const int2 *pairs,
                                                               int k) {
    __shared__ float sA[2][kKC32][kLds32];
    __shared__ float sB[2][kKC32][kLds32];
    const int ts = map.ts;
    const int nsb = ts / kBM;
    const int nsub = nsb * nsb;
    const int bid = xcd_remap((int) blockIdx.x, (int) gridDim.x);
    const int pair = bid / nsub;
    const int sub = bid % nsub;
    const int sbm = sub % nsb, sbn = sub / nsb;
    const int2 ij = pairs[pair];
    const int64_t tsz = (int64_t) ts * ts;
    const float *A = p32 + (int64_t) (ij.x - k - 1) * tsz + sbm * kBM;
    const float *B = p32 + (int64_t) (ij.y - k - 1) * tsz + sbn * kBM;
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;
    const int lr = lane & 15, lk = lane >> 4;

    auto goff = [&](int q) { const int e = t + q * kNT; return (uint32_t) ((e % kBM) + (e / kBM) * ts); };
    auto loff = [&](int q) { const int e = t + q * kNT; return (e / kBM) * kLds32 + (e % kBM); };
    const uint32_t kstep = (uint32_t) (kKC32 * ts);

    f32x4 acc[6][3];
#pragma unroll
    for (int mi = 0; mi < 6; ++mi)
#pragma unroll
        for (int ni = 0; ni < 3; ++ni) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};

    constexpr int kDepth = 2;   // chunks of loads in flight (register ring)
    float ra[kDepth][kPer32], rb[kDepth][kPer32];
    const int nchunks = ts / kKC32;
    auto issue = [&](int set, int chunk) {
        const uint32_t kofs = (uint32_t) chunk * kstep;
#pragma unroll
        for (int q = 0; q < kPer32; ++q) {
            ra[set][q] = A[goff(q) + kofs];
            rb[set][q] = B[goff(q) + kofs];
        }
    };
    auto stage = [&](int set, int buf) {
        float *dA = &sA[buf][0][0];
        float *dB = &sB[buf][0][0];
#pragma unroll
        for (int q = 0; q < kPer32; ++q) {
            dA[loff(q)] = ra[set][q];
            dB[loff(q)] = rb[set][q];
        }
    };
#pragma unroll
    for (int u = 0; u < kDepth; ++u)
        if (u < nchunks) issue(u, u);
    stage(0, 0);
    __syncthreads();
    for (int c0 = 0; c0 < nchunks; c0 += kDepth) {
#pragma unroll
        for (int u = 0; u < kDepth; ++u) {
            const int c = c0 + u;
            const int buf = c & 1;
            if (c + kDepth < nchunks) issue(u, c + kDepth);
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
            if (c + 1 < nchunks) stage((u + 1) % kDepth, buf ^ 1);
            __syncthreads();
        }
    }
    // C -= acc in two batches of 36 values (one memory round trip each)
    float *C = tile_f32(map, ij.x, ij.y);
    const uint32_t cbase = (uint32_t) ((sbm * kBM + wm * 96 + lr) + (sbn * kBM + wn * 48 + 4 * lk) * ts);
#pragma unroll
    for (int m0 = 0; m0 < 6; m0 += 3) {
        float cv[3][3][4];
#pragma unroll
        for (int mi = 0; mi < 3; ++mi)
#pragma unroll
            for (int ni = 0; ni < 3; ++ni)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    cv[mi][ni][v] = __builtin_nontemporal_load(
                        &C[cbase + (uint32_t) ((m0 + mi) * 16 + (ni * 16 + v) * ts)]);
#pragma unroll
        for (int mi = 0; mi < 3; ++mi)
#pragma unroll
            for (int ni = 0; ni < 3; ++ni)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    __builtin_nontemporal_store(cv[mi][ni][v] - acc[m0 + mi][ni][v],
                                                &C[cbase + (uint32_t) ((m0 + mi) * 16 + (ni * 16 + v) * ts)]);
    }
}

// ------------------------------------------------------------------------
// 16-bit-operand update with 32x32x16 MFMAs: four waves per 192 x 192 block,
// each owning a 96 x 96 quadrant as 3 x 3 tiles of 32 x 32 (144 FP32
// accumulators per lane).  Per 16-deep K step a wave reads 12 operand
// fragments for 27 MFMAs.  Output slot i of lane l holds row (l % 32) of the A
// operand and column 8 (i / 4) + 4 (l / 32) + i % 4 of the B operand (measured
// on gfx950).  T16 = _Float16 with row scales, or __bf16 without (the
// plain-BF16 option, single product).  The single-product and per-block
// kernels fit 256 registers and 50 KB of LDS, so two workgroups share a CU
// (measured 1.5x faster than one for the single product); the split-only
// kernel double-buffers hi and lo (100 KB) and runs one per CU.
// ------------------------------------------------------------------------
constexpr int kNTw = 256;
constexpr int kKCw = 32;                  // K per chunk: two MFMA steps
constexpr int kPlanesW = kKCw / 8;        // 16-byte vectors per row per chunk
// rows per LDS plane, padded: at 194 x 16 B the four planes of a staged row
// start 32 B apart in the banks, so eight lanes' 16-byte writes do not collide
constexpr int kPlaneW = kBM + 2;
constexpr int kVecsW = kBM * kPlanesW;    // 768 vectors per operand part per chunk
constexpr int kPerThreadW = kVecsW / kNTw;   // 3

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

template <typename T16>
__device__ __forceinline__ f32x16 mfma_32(const u16x8 &x, const u16x8 &y, f32x16 acc) {
    if constexpr (std::is_same<T16, _Float16>::value)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(half8, x), __builtin_bit_cast(half8, y), acc,
                                                      0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, x), __builtin_bit_cast(bf16x8, y),
                                                       acc, 0, 0, 0);
}

typedef int i32x8 __attribute__((ext_vector_type(8)));

// One e4m3 x e4m3 product over 64 K values, FP32 accumulation (scales 2^0).
// x and y each hold 32 one-byte values per lane: lanes 0-31 K 0-31 of the
// chunk, lanes 32-63 K 32-63, from the two 16-byte planes p, p + 1.
__device__ __forceinline__ f32x16 mfma_fp8(const u16x8 &x0, const u16x8 &x1, const u16x8 &y0, const u16x8 &y1,
                                           f32x16 acc) {
    struct Pair {
        u16x8 a, b;
    };
    const i32x8 x = __builtin_bit_cast(i32x8, Pair{x0, x1});
    const i32x8 y = __builtin_bit_cast(i32x8, Pair{y0, y1});
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(x, y, acc, 0, 0, 0, 127, 0, 127);
}

// kMode: 0 split product, 1 single product, 2 chosen per block: single when
// bmax(A rows) * bmax(B rows) <= tau * pmax(panel k)^2, split otherwise; with
// an FP8 copy of the panel (p8, FP32 tiles only), one e4m3 product where that
// ratio is also <= tau8.
enum { kModeSplit = 0, kModeSingle = 1, kModeAdaptive = 2 };

template <typename T16, int kMode, int kDepth, bool kC64 = false>
__global__ void __launch_bounds__(kNTw, kMode == kModeSplit ? 1 : 2) update192_w_kernel(TileMap map, const uint16_t *p16h, const uint16_t *p16l,
                                                              const float *inv_scale, const float *bmax,
                                                              const unsigned *pmax, float tau, const uint16_t *p8,
                                                              float tau8, const int2 *pairs, int k,
                                                              unsigned long long *counts) {
    // [buffer][part]: the split kernel double-buffers hi and lo; the others
    // double-buffer one part, and the per-block kernel's split path uses the
    // two buffers for hi and lo, single-buffered (50 KB: two workgroups per CU)
    __shared__ u16x8 sA[2][kMode == kModeSplit ? 2 : 1][kPlanesW][kPlaneW];
    __shared__ u16x8 sB[2][kMode == kModeSplit ? 2 : 1][kPlanesW][kPlaneW];
    const int ts = map.ts;
    const int nsb = ts / kBM;
    const int nsub = nsb * nsb;
    // per-block choice: split blocks (about 2.3x the cost of single ones)
    // gather where the list ends (the columns next to the panel), so whole
    // tiles are dealt to the XCDs in turn
    const int bid = xcd_deal((int) blockIdx.x, (int) gridDim.x, nsub);
    const int pair = bid / nsub;
    const int sub = bid % nsub;
    const int sbm = sub % nsb, sbn = sub / nsb;
    const int2 ij = pairs[pair];
    if (ij.x == ij.y && sbm < sbn) return;   // a diagonal tile's upper blocks (uniform over the workgroup)
    const int64_t tsz = (int64_t) ts * ts;
    const int64_t aoff = (int64_t) (ij.x - k - 1) * tsz + (int64_t) sbm * kBM * ts;
    const int64_t boff = (int64_t) (ij.y - k - 1) * tsz + (int64_t) sbn * kBM * ts;
    const uint16_t *src[4] = {p16h + aoff, p16l + aoff, p16h + boff, p16l + boff};
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;
    const int lr = lane & 31, lk = lane >> 5;

    // staging: vector e = t + 256 q of a chunk is row e / 4, plane e % 4
    uint32_t goff[kPerThreadW];
#pragma unroll
    for (int q = 0; q < kPerThreadW; ++q) {
        const int e = t + kNTw * q;
        goff[q] = (uint32_t) ((e >> 2) * ts + (e & 3) * 8);
    }

    f32x16 acc[3][3];
#pragma unroll
    for (int mi = 0; mi < 3; ++mi)
#pragma unroll
        for (int ni = 0; ni < 3; ++ni)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[mi][ni][v] = 0.f;

    const int nchunks = ts / kKCw;
    // The K loop with P = 2 (hi and lo parts, three products) or P = 1 (hi
    // parts, one product); part p of a chunk sits in LDS plane set p.
    auto run = [&](auto split_tag) {
        constexpr bool kSplit = decltype(split_tag)::value;
        constexpr int P = kSplit ? 2 : 1;
        u16x8 rg[kDepth][kPerThreadW][2 * P];
        auto issue = [&](int set, int chunk) {
            const uint32_t kofs = (uint32_t) (chunk * kKCw);
#pragma unroll
            for (int q = 0; q < kPerThreadW; ++q)
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    rg[set][q][p] = *reinterpret_cast<const u16x8 *>(src[p] + (goff[q] + kofs));
                    rg[set][q][P + p] = *reinterpret_cast<const u16x8 *>(src[2 + p] + (goff[q] + kofs));
                }
        };
        auto stage = [&](int set, int buf) {
#pragma unroll
            for (int q = 0; q < kPerThreadW; ++q) {
                const int e = t + kNTw * q;
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    sA[buf][p][e & 3][e >> 2] = rg[set][q][p];
                    sB[buf][p][e & 3][e >> 2] = rg[set][q][P + p];
                }
            }
        };
#pragma unroll
        for (int u = 0; u < kDepth; ++u)
            if (u < nchunks) issue(u, u);
        stage(0, 0);
        __syncthreads();
        for (int c0 = 0; c0 < nchunks; c0 += kDepth) {
#pragma unroll
            for (int u = 0; u < kDepth; ++u) {
                const int c = c0 + u;
                const int buf = c & 1;
                if (c + kDepth < nchunks) issue(u, c + kDepth);
#pragma unroll
                for (int ks = 0; ks < kKCw / 16; ++ks) {
                    const int pl = 2 * ks + lk;
                    u16x8 ah[3], bh[3];
#pragma unroll
                    for (int mi = 0; mi < 3; ++mi) ah[mi] = sA[buf][0][pl][wm * 96 + mi * 32 + lr];
#pragma unroll
                    for (int ni = 0; ni < 3; ++ni) bh[ni] = sB[buf][0][pl][wn * 96 + ni * 32 + lr];
                    if constexpr (kSplit) {
                        u16x8 al[3];
#pragma unroll
                        for (int mi = 0; mi < 3; ++mi) al[mi] = sA[buf][P - 1][pl][wm * 96 + mi * 32 + lr];
#pragma unroll
                        for (int ni = 0; ni < 3; ++ni) {
                            const u16x8 bl = sB[buf][P - 1][pl][wn * 96 + ni * 32 + lr];
#pragma unroll
                            for (int mi = 0; mi < 3; ++mi) {
                                acc[mi][ni] = mfma_32<T16>(bh[ni], ah[mi], acc[mi][ni]);
                                acc[mi][ni] = mfma_32<T16>(bh[ni], al[mi], acc[mi][ni]);
                                acc[mi][ni] = mfma_32<T16>(bl, ah[mi], acc[mi][ni]);
                            }
                        }
                    } else {
#pragma unroll
                        for (int mi = 0; mi < 3; ++mi)
#pragma unroll
                            for (int ni = 0; ni < 3; ++ni) acc[mi][ni] = mfma_32<T16>(bh[ni], ah[mi], acc[mi][ni]);
                    }
                }
                if (c + 1 < nchunks) stage((u + 1) % kDepth, buf ^ 1);
                __syncthreads();
            }
        }
    };
    // One FP8 product: the FP8 copy of the panel has one byte per value, so a
    // chunk of the same 64 bytes per row (four 16-byte planes, staged as in
    // run) holds 64 K values, and ts / 64 chunks cover K.  The panel is read
    // as 16-bit words: row stride ts / 2, chunk stride 32 words.
    auto run_fp8 = [&]() {
        const int64_t a8 = ((int64_t) (ij.x - k - 1) * tsz + (int64_t) sbm * kBM * ts) / 2;
        const int64_t b8 = ((int64_t) (ij.y - k - 1) * tsz + (int64_t) sbn * kBM * ts) / 2;
        const uint16_t *s8a = p8 + a8, *s8b = p8 + b8;
        const int nch8 = ts / (2 * kKCw);
        // one chunk of loads in flight (a chunk is twice the FP16 path's K):
        // two would not fit beside the accumulators without spilling
        constexpr int kD8 = 1;
        uint32_t g8[kPerThreadW];
#pragma unroll
        for (int q = 0; q < kPerThreadW; ++q) {
            const int e = t + kNTw * q;
            g8[q] = (uint32_t) ((e >> 2) * (ts / 2) + (e & 3) * 8);
        }
        u16x8 rg[kD8][kPerThreadW][2];
        auto issue = [&](int set, int chunk) {
            const uint32_t kofs = (uint32_t) (chunk * kKCw);
#pragma unroll
            for (int q = 0; q < kPerThreadW; ++q) {
                rg[set][q][0] = *reinterpret_cast<const u16x8 *>(s8a + (g8[q] + kofs));
                rg[set][q][1] = *reinterpret_cast<const u16x8 *>(s8b + (g8[q] + kofs));
            }
        };
        auto stage = [&](int set, int buf) {
#pragma unroll
            for (int q = 0; q < kPerThreadW; ++q) {
                const int e = t + kNTw * q;
                sA[buf][0][e & 3][e >> 2] = rg[set][q][0];
                sB[buf][0][e & 3][e >> 2] = rg[set][q][1];
            }
        };
#pragma unroll
        for (int u = 0; u < kD8; ++u)
            if (u < nch8) issue(u, u);
        stage(0, 0);
        __syncthreads();
        for (int c0 = 0; c0 < nch8; c0 += kD8) {
#pragma unroll
            for (int u = 0; u < kD8; ++u) {
                const int c = c0 + u;
                if (c >= nch8) break;
                const int buf = c & 1;
                if (c + kD8 < nch8) issue(u, c + kD8);
                u16x8 a0[3], a1[3];
#pragma unroll
                for (int mi = 0; mi < 3; ++mi) {
                    a0[mi] = sA[buf][0][2 * lk][wm * 96 + mi * 32 + lr];
                    a1[mi] = sA[buf][0][2 * lk + 1][wm * 96 + mi * 32 + lr];
                }
#pragma unroll
                for (int ni = 0; ni < 3; ++ni) {
                    const u16x8 b0 = sB[buf][0][2 * lk][wn * 96 + ni * 32 + lr];
                    const u16x8 b1 = sB[buf][0][2 * lk + 1][wn * 96 + ni * 32 + lr];
#pragma unroll
                    for (int mi = 0; mi < 3; ++mi) acc[mi][ni] = mfma_fp8(b0, b1, a0[mi], a1[mi], acc[mi][ni]);
                }
                if (c + 1 < nch8) stage((u + 1) % kD8, buf ^ 1);
                __syncthreads();
            }
        }
    };
    // Split product with hi and lo in the two LDS buffers, one chunk at a time
    // (the per-block kernel): loads of chunk c+1 are in flight while chunk c
    // is multiplied.
    auto run_split_1buf = [&]() {
        u16x8 rg[kPerThreadW][4];
        auto issue = [&](int chunk) {
            const uint32_t kofs = (uint32_t) (chunk * kKCw);
#pragma unroll
            for (int q = 0; q < kPerThreadW; ++q)
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    rg[q][p] = *reinterpret_cast<const u16x8 *>(src[p] + (goff[q] + kofs));
                    rg[q][2 + p] = *reinterpret_cast<const u16x8 *>(src[2 + p] + (goff[q] + kofs));
                }
        };
        issue(0);
        for (int c = 0; c < nchunks; ++c) {
            __syncthreads();   // every wave is done with chunk c-1
#pragma unroll
            for (int q = 0; q < kPerThreadW; ++q) {
                const int e = t + kNTw * q;
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    sA[p][0][e & 3][e >> 2] = rg[q][p];
                    sB[p][0][e & 3][e >> 2] = rg[q][2 + p];
                }
            }
            if (c + 1 < nchunks) issue(c + 1);
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < kKCw / 16; ++ks) {
                const int pl = 2 * ks + lk;
                u16x8 ah[3], al[3];
#pragma unroll
                for (int mi = 0; mi < 3; ++mi) {
                    ah[mi] = sA[0][0][pl][wm * 96 + mi * 32 + lr];
                    al[mi] = sA[1][0][pl][wm * 96 + mi * 32 + lr];
                }
#pragma unroll
                for (int ni = 0; ni < 3; ++ni) {
                    const u16x8 bh = sB[0][0][pl][wn * 96 + ni * 32 + lr];
                    const u16x8 bl = sB[1][0][pl][wn * 96 + ni * 32 + lr];
#pragma unroll
                    for (int mi = 0; mi < 3; ++mi) {
                        acc[mi][ni] = mfma_32<T16>(bh, ah[mi], acc[mi][ni]);
                        acc[mi][ni] = mfma_32<T16>(bh, al[mi], acc[mi][ni]);
                        acc[mi][ni] = mfma_32<T16>(bl, ah[mi], acc[mi][ni]);
                    }
                }
            }
        }
    };
    bool single = kMode == kModeSingle;
    bool fp8 = false;
    if constexpr (kMode == kModeAdaptive) {
        // uniform over the workgroup
        const float a = bmax[(ij.x - k - 1) * nsb + sbm];
        const float b = bmax[(ij.y - k - 1) * nsb + sbn];
        const float pm = __uint_as_float(pmax[k]);
        single = a * b <= tau * pm * pm;
        if constexpr (!kC64) fp8 = single && p8 != nullptr && a * b <= tau8 * pm * pm;
    }
    if (counts && t == 0) atomicAdd(counts + (fp8 ? 2 : single ? 1 : 0), 1ull);
    if constexpr (kMode == kModeSplit) {
        run(std::true_type{});
    } else if constexpr (kMode == kModeSingle) {
        run(std::false_type{});
    } else {
        if (fp8)
            run_fp8();
        else if (single)
            run(std::false_type{});
        else
            run_split_1buf();
    }
    // C -= acc (times the two row scales' inverses, exact powers of two): lane
    // row (A side) wm*96 + mi*32 + lr, column (B side) wn*96 + ni*32 + 8 (v / 4)
    // + 4 lk + v % 4; one 3 x 16 batch per mi
    // C is FP32, or FP64 for tiles of the FP64 band updated with 16-bit products
    using TC = typename std::conditional<kC64, double, float>::type;
    TC *C;
    if constexpr (kC64)
        C = tile_f64(map, ij.x, ij.y);
    else
        C = tile_f32(map, ij.x, ij.y);
    const uint32_t cbase = (uint32_t) ((sbm * kBM + wm * 96 + lr) + (sbn * kBM + wn * 96 + 4 * lk) * ts);
    constexpr bool kScaled = std::is_same<T16, _Float16>::value;
    const float *invA = kScaled ? inv_scale + (int64_t) (ij.x - k - 1) * ts + sbm * kBM + wm * 96 + lr : nullptr;
    const float *invB = kScaled ? inv_scale + (int64_t) (ij.y - k - 1) * ts + sbn * kBM + wn * 96 + 4 * lk : nullptr;
#pragma unroll
    for (int mi = 0; mi < 3; ++mi) {
        const float sa = kScaled ? invA[mi * 32] : 1.0f;
#pragma unroll
        for (int ni = 0; ni < 3; ++ni) {
            TC cv[16];
            f32x4v sb[4];
#pragma unroll
            for (int g = 0; g < 4; ++g)
                sb[g] = kScaled ? *reinterpret_cast<const f32x4v *>(invB + ni * 32 + 8 * g) : f32x4v{1.f, 1.f, 1.f, 1.f};
#pragma unroll
            for (int v = 0; v < 16; ++v)
                cv[v] = __builtin_nontemporal_load(
                    &C[cbase + (uint32_t) (mi * 32 + (ni * 32 + 8 * (v >> 2) + (v & 3)) * ts)]);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                float a = acc[mi][ni][v];
                if constexpr (kScaled) a *= sa * sb[v >> 2][v & 3];
                __builtin_nontemporal_store(cv[v] - (TC) a,
                                            &C[cbase + (uint32_t) (mi * 32 + (ni * 32 + 8 * (v >> 2) + (v & 3)) * ts)]);
            }
        }
    }
}

}  // namespace

void launch_panel_rowscale(const TileMap &map, int k, float *inv_scale, float *bmax, unsigned *pmax,
                           hipStream_t s) {
    const int rows = map.nt - k - 1;
    if (rows <= 0) return;
    const int nb = map.ts / kBM;
    hipLaunchKernelGGL(panel_rowscale_kernel, dim3((unsigned) (rows * nb)), dim3(kThreads), 0, s, map, k, inv_scale,
                       bmax, pmax);
}

void launch_zero_u32(unsigned *p, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, p, n);
}

void launch_panel_convert(const TileMap &map, int k, double *p64, float *p32, uint16_t *p16h, uint16_t *p16l,
                          uint8_t *p8, const float *inv_scale, hipStream_t s) {
    const int rows = map.nt - k - 1;
    if (rows <= 0) return;
    const int nsb = map.ts / kBlk;
    hipLaunchKernelGGL(panel_convert_kernel, dim3((unsigned) (rows * nsb * nsb)), dim3(kThreads), 0, s, map, k, p64,
                       p32, p16h, p16l, p8, inv_scale);
}

void launch_update192_f32(const TileMap &map, const float *p32, const int2 *pairs, int npairs, int k, hipStream_t s) {
    if (npairs <= 0) return;
    const int nsb = map.ts / kBM;
    hipLaunchKernelGGL(update192_f32_kernel, dim3((unsigned) npairs * (unsigned) (nsb * nsb)), dim3(kNT), 0, s, map,
                       p32, pairs, k);
}

void launch_update192_16(const TileMap &map, const uint16_t *p16h, const uint16_t *p16l, const float *inv_scale,
                         const float *bmax, const unsigned *pmax, float tau, const uint8_t *p8, float tau8,
                         const int2 *pairs, int npairs, int k, Operands op, unsigned long long *counts,
                         hipStream_t s) {
    if (npairs <= 0) return;
    const int nsb = map.ts / kBM;
    const dim3 grid((unsigned) npairs * (unsigned) (nsb * nsb));
    const uint16_t *w8 = reinterpret_cast<const uint16_t *>(p8);
    // two chunks of loads in flight: a third measured no faster
    switch (op) {
    case kSplitF16:
        hipLaunchKernelGGL((update192_w_kernel<_Float16, kModeSplit, 2>), grid, dim3(kNTw), 0, s, map, p16h, p16l,
                           inv_scale, bmax, pmax, tau, nullptr, 0.0f, pairs, k, nullptr);
        break;
    case kSingleF16:
        hipLaunchKernelGGL((update192_w_kernel<_Float16, kModeSingle, 2>), grid, dim3(kNTw), 0, s, map, p16h, p16l,
                           inv_scale, bmax, pmax, tau, nullptr, 0.0f, pairs, k, counts);
        break;
    case kAdaptiveF16:
        hipLaunchKernelGGL((update192_w_kernel<_Float16, kModeAdaptive, 2>), grid, dim3(kNTw), 0, s, map, p16h,
                           p16l, inv_scale, bmax, pmax, tau, tau8 > 0.0f ? w8 : nullptr, tau8, pairs, k, counts);
        break;
    case kSplitF16C64:
        hipLaunchKernelGGL((update192_w_kernel<_Float16, kModeSplit, 2, true>), grid, dim3(kNTw), 0, s, map, p16h,
                           p16l, inv_scale, bmax, pmax, tau, nullptr, 0.0f, pairs, k, nullptr);
        break;
    case kSingleF16C64:
        hipLaunchKernelGGL((update192_w_kernel<_Float16, kModeSingle, 2, true>), grid, dim3(kNTw), 0, s, map, p16h,
                           p16l, inv_scale, bmax, pmax, tau, nullptr, 0.0f, pairs, k, counts);
        break;
    case kAdaptiveF16C64:
        hipLaunchKernelGGL((update192_w_kernel<_Float16, kModeAdaptive, 2, true>), grid, dim3(kNTw), 0, s, map, p16h,
                           p16l, inv_scale, bmax, pmax, tau, nullptr, 0.0f, pairs, k, counts);
        break;
    default:
        hipLaunchKernelGGL((update192_w_kernel<__bf16, kModeSingle, 2>), grid, dim3(kNTw), 0, s, map, p16h, p16l,
                           nullptr, nullptr, nullptr, 0.0f, nullptr, 0.0f, pairs, k, nullptr);
    }
}

}  // namespace egs
