// Trailing SYRK/GEMM update of the tiled Cholesky on gfx950, the kernel that
// carries ~95% of the factorization flops (the reference's Chameleon runs the
// same tile update C_ij -= L_ik L_jk^T as dsyrk/dgemm tasks).
//
// One workgroup of 512 threads (8 waves) owns a 192 x 192 block of C_ij and
// contracts over K = ts with v_mfma_f64_16x16x4f64:
//   * wave w covers rows (w % 2) * 96 .. +96 and columns (w / 2) * 48 .. +48,
//     i.e. 6 x 3 independent 16 x 16 accumulators (72 FP64 values per lane);
//   * A and B are staged through LDS in K-chunks of 16 columns, double
//     buffered, fed by a ring of registers that keeps the global loads of the
//     next two chunks in flight while a chunk is multiplied, so each chunk
//     costs one barrier and the HBM/L2 latency hides behind 144 MFMAs per wave;
//   * global loads read 192 contiguous doubles per K-column (coalesced).
// The operand panel is either the tile arena itself or a contiguous FP64 copy
// of panel k (PanelRef), so the inner loop has no type dispatch.
#include "egs_common.hpp"
#include "egs_kernels.hpp"

namespace egs {

namespace {

constexpr int kBM = 192;           // block rows (and columns)
constexpr int kKC = 16;            // K chunk
constexpr int kNT = 512;           // threads per workgroup
constexpr int kLds = kBM + 16;     // LDS row stride in doubles: the 4 K-rows a wave reads land on disjoint bank halves
constexpr int kPer = kBM * kKC / kNT;   // elements of A (and of B) each thread stages per chunk

// Where the operand tiles of panel k live: in the tile arena itself (FP64 plan)
// or in a contiguous FP64 copy of the panel (mixed plans), tile i of the panel
// at pbase + (i - k - 1) * pstride.
struct PanelRef {
    const char *pbase;
    int64_t pstride;
    int contiguous;
};

__device__ __forceinline__ const double *panel_tile(const TileMap &map, const PanelRef &pr, int i, int k) {
    if (pr.contiguous) return reinterpret_cast<const double *>(pr.pbase + (int64_t) (i - k - 1) * pr.pstride);
    return reinterpret_cast<const double *>(map.base + map.off[i * map.nt + k]);
}

__global__ void __launch_bounds__(kNT, 1) update192_kernel(TileMap map, PanelRef pr, const int2 *pairs, int k) {
    __shared__ double sA[2][kKC][kLds];
    __shared__ double sB[2][kKC][kLds];
    const int ts = map.ts;
    const int nsb = ts / kBM;
    const int nsub = nsb * nsb;
    const int bid = xcd_remap((int) blockIdx.x, (int) gridDim.x);
    const int pair = bid / nsub;
    const int sub = bid % nsub;
    const int sbm = sub % nsb, sbn = sub / nsb;
    const int2 ij = pairs[pair];
    if (ij.x == ij.y && sbm < sbn) return;   // strictly upper block of a diagonal tile

    const double *A = panel_tile(map, pr, ij.x, k) + sbm * kBM;
    const double *B = panel_tile(map, pr, ij.y, k) + sbn * kBM;

    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;
    const int lr = lane & 15, lk = lane >> 4;

    // staging: element e = t + q * kNT of the 192 x 16 chunk sits at row e % 192,
    // column e / 192; its offset inside the tile (in elements) fits in 32 bits.
    // The offsets are recomputed where used (cheap integer ops) to save registers.
    auto goff = [&](int q) { const int e = t + q * kNT; return (uint32_t) ((e % kBM) + (e / kBM) * ts); };
    auto loff = [&](int q) { const int e = t + q * kNT; return (e / kBM) * kLds + (e % kBM); };
    const uint32_t kstep = (uint32_t) (kKC * ts);

    f64x4 acc[6][3];
#pragma unroll
    for (int mi = 0; mi < 6; ++mi)
#pragma unroll
        for (int ni = 0; ni < 3; ++ni) acc[mi][ni] = f64x4{0.0, 0.0, 0.0, 0.0};

    // Register ring: chunks c+1 .. c+kDepth are in flight while chunk c is
    // multiplied from LDS; set u of the ring holds chunks u, u + kDepth, ...
    constexpr int kDepth = 2;
    double ra[kDepth][kPer], rb[kDepth][kPer];
    const int nchunks = ts / kKC;
    auto issue = [&](int set, int chunk) {
        const uint32_t kofs = (uint32_t) chunk * kstep;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            ra[set][q] = A[goff(q) + kofs];
            rb[set][q] = B[goff(q) + kofs];
        }
    };
    auto stage = [&](int set, int buf) {
        double *dA = &sA[buf][0][0];
        double *dB = &sB[buf][0][0];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
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
            for (int kk = 0; kk < kKC; kk += 4) {
                double a[6], b[3];
#pragma unroll
                for (int mi = 0; mi < 6; ++mi) a[mi] = sA[buf][kk + lk][wm * 96 + mi * 16 + lr];
#pragma unroll
                for (int ni = 0; ni < 3; ++ni) b[ni] = sB[buf][kk + lk][wn * 48 + ni * 16 + lr];
#pragma unroll
                for (int mi = 0; mi < 6; ++mi)
#pragma unroll
                    for (int ni = 0; ni < 3; ++ni) acc[mi][ni] = mfma_f64(b[ni], a[mi], acc[mi][ni]);
            }
            if (c + 1 < nchunks) stage((u + 1) % kDepth, buf ^ 1);
            __syncthreads();
        }
    }

    // C -= acc.  Lane l, element v of acc[mi][ni] is
    // C[wm*96 + mi*16 + l%16][wn*48 + ni*16 + l/16 + 4v]  (FP64 MFMA output order).
    // Read-modify-write of C in three batches of 24 values (one memory round
    // trip per batch).
    double *C = tile_f64(map, ij.x, ij.y);
    const uint32_t cbase = (uint32_t) ((sbm * kBM + wm * 96 + lr) + (sbn * kBM + wn * 48 + lk) * ts);
#pragma unroll
    for (int m0 = 0; m0 < 6; m0 += 2) {
        double cv[2][3][4];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 3; ++ni)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    cv[mi][ni][v] = __builtin_nontemporal_load(
                        &C[cbase + (uint32_t) ((m0 + mi) * 16 + (ni * 16 + 4 * v) * ts)]);
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 3; ++ni)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    __builtin_nontemporal_store(cv[mi][ni][v] - acc[m0 + mi][ni][v],
                                                &C[cbase + (uint32_t) ((m0 + mi) * 16 + (ni * 16 + 4 * v) * ts)]);
    }
}

}  // namespace

bool update192_supported(int ts) { return ts % kBM == 0; }

void launch_update192(const TileMap &map, const int2 *pairs, int npairs, int k, hipStream_t s) {
    if (npairs <= 0) return;
    const int nsb = map.ts / kBM;
    const unsigned blocks = (unsigned) npairs * (unsigned) (nsb * nsb);
    PanelRef pr{nullptr, 0, 0};
    hipLaunchKernelGGL(update192_kernel, dim3(blocks), dim3(kNT), 0, s, map, pr, pairs, k);
}

void launch_update192_panel(const TileMap &map, const double *p64, const int2 *pairs, int npairs, int k,
                            hipStream_t s) {
    if (npairs <= 0) return;
    const int nsb = map.ts / kBM;
    const unsigned blocks = (unsigned) npairs * (unsigned) (nsb * nsb);
    PanelRef pr{reinterpret_cast<const char *>(p64), (int64_t) map.ts * map.ts * (int64_t) sizeof(double), 1};
    hipLaunchKernelGGL(update192_kernel, dim3(blocks), dim3(kNT), 0, s, map, pr, pairs, k);
}

}  // namespace egs
