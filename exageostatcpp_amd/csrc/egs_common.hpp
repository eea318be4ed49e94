// Shared device/host definitions for the gfx950 kernels of exageostatcpp_amd.
//
// Matrix storage: the covariance matrix of dimension n (= locations x variables)
// is held as lower tiles of ts x ts elements (ts = the dense tile size rounded up
// to a multiple of 64).  Each tile is column-major with leading dimension ts and
// owns its own block of one HBM arena, so every tile can carry its own storage
// precision.  The matrix is padded to npad = nt * ts with the identity, which
// keeps it SPD, leaves log|Sigma| unchanged and lets every kernel work on full
// 64-row blocks without bounds checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace egs {

constexpr int kBlk = 64;        // sub-block edge used by every factorization kernel
constexpr int kThreads = 256;   // 4 waves of 64 lanes per workgroup

enum StoragePrecision : uint8_t { kStoreF64 = 0, kStoreF32 = 1 };
enum ComputeClass : uint8_t { kComputeF64 = 0, kComputeF32 = 1, kComputeBF16 = 2 };

// Location of every lower tile in the arena.  off[i * nt + j] is the byte offset
// of tile (i, j), i >= j; store[i * nt + j] its storage precision.
struct TileMap {
    char *base;
    const int64_t *off;
    const uint8_t *store;
    int nt;
    int ts;
};

__device__ __forceinline__ double *tile_f64(const TileMap &m, int i, int j) {
    return reinterpret_cast<double *>(m.base + m.off[i * m.nt + j]);
}
__device__ __forceinline__ float *tile_f32(const TileMap &m, int i, int j) {
    return reinterpret_cast<float *>(m.base + m.off[i * m.nt + j]);
}

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin (workgroup w
// runs on XCD w % 8), and each XCD has its own L2.  xcd_remap gives XCD x a
// contiguous range of the logical block order, so blocks that share operand
// rows (consecutive logical blocks) meet in the same L2.  It is a bijection
// of [0, n).
__device__ __forceinline__ int xcd_remap(int w, int n) {
    constexpr int kXcd = 8;
    const int q = n / kXcd, r = n % kXcd;
    const int x = w % kXcd, i = w / kXcd;
    return x * q + min(x, r) + i;
}

// The same idea for blocks of uneven cost: groups of g consecutive logical
// blocks (one tile's blocks) are dealt to the XCDs in turn, XCD x taking
// groups x, x + 8, x + 16, ...  A group's blocks meet in one L2, and every
// XCD gets a share of each part of the order, so a run of expensive tiles
// does not all land on one XCD (with xcd_remap's contiguous ranges the last
// XCD took every tile at the end of the list).  Workgroups past the last full
// round of 8 g keep their index.  A bijection of [0, n).
__device__ __forceinline__ int xcd_deal(int w, int n, int g) {
    constexpr int kXcd = 8;
    const int full = n / (kXcd * g) * (kXcd * g);
    if (w >= full) return w;
    const int x = w % kXcd, i = w / kXcd;
    return (i / g) * (kXcd * g) + x * g + i % g;
}

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// D[i][j] += sum_k X[i][k] Y[k][j] on a 16x16x4 tile.  Lane l supplies X[l%16][l/16]
// and Y[l/16][l%16].  Result element v (0..3) of lane l is D[l/16 + 4*v][l%16] for
// the FP64 instruction and D[4*(l/16) + v][l%16] for the FP32 one (both measured
// on gfx950).
__device__ __forceinline__ f64x4 mfma_f64(double x, double y, f64x4 acc) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_f32(float x, float y, f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, acc, 0, 0, 0);
}

}  // namespace egs
