// Device building blocks of the dense FP64 kernels (single matrix and batched):
// the 32x32 MFMA tile, its element layout, and block sums.
#pragma once
#include <hip/hip_runtime.h>

namespace pf {

typedef double double4_t __attribute__((ext_vector_type(4)));

// Row-major: element (r, c) of A is A[r*lda + c]; op(A) is m x k.
__device__ __forceinline__ double ld_op(const double *__restrict__ M, int ld, int r, int c, int nr, int nc, int t) {
    // element (r, c) of op(M) (op(M) is nr x nc); t: M stored transposed
    if (r >= nr || c >= nc) return 0.0;
    return t ? M[(size_t)c * ld + r] : M[(size_t)r * ld + c];
}

// The 32x32 tile (r0, c0) of op(A) op(B) computed by W wavefronts (64 W
// threads) that split K: wavefront w takes runs of KS MFMA k-steps (4 k
// each) starting at k = 4 KS (w + W c), issues every operand load of a run
// before the run's MFMAs (so the loads are in flight together instead of one
// k-step's latency at a time), and accumulates the 2x2 grid of 16x16
// v_mfma_f64_16x16x4f64 blocks.  The W partial tiles land in red[W]; the
// caller sums red[0..W-1][reg][ln] for tile element idx = 64 reg + ln.
template <int W, int KS>
__device__ __forceinline__ void mfma_tile_w(int m, int n, int k, int ta, int tb, const double *__restrict__ A,
                                            int lda, const double *__restrict__ B, int ldb, int r0, int c0,
                                            double (*red)[16][64]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = lane & 15, kq = lane >> 4;
    double4_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = double4_t{0, 0, 0, 0};
    for (int kb = 4 * KS * w; kb < k; kb += 4 * KS * W) {
        // operand lanes: A lane l = (row l & 15, k l >> 4); B lane l = (k l >> 4, col l & 15)
        double a0[KS], a1[KS], b0[KS], b1[KS];
#pragma unroll
        for (int j = 0; j < KS; ++j) {
            const int kk = kb + 4 * j + kq;
            a0[j] = ld_op(A, lda, r0 + i, kk, m, k, ta);
            a1[j] = ld_op(A, lda, r0 + 16 + i, kk, m, k, ta);
            b0[j] = ld_op(B, ldb, kk, c0 + i, k, n, tb);
            b1[j] = ld_op(B, ldb, kk, c0 + 16 + i, k, n, tb);
        }
#pragma unroll
        for (int j = 0; j < KS; ++j) {
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[j], b0[j], acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[j], b1[j], acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[j], b0[j], acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[j], b1[j], acc[1][1], 0, 0, 0);
        }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) red[w][(a * 2 + b) * 4 + r][lane] = acc[a][b][r];
    __syncthreads();
}

// The four-wavefront form used by the batched kernels (256 threads).
__device__ __forceinline__ void mfma_tile(int m, int n, int k, int ta, int tb, const double *__restrict__ A, int lda,
                                          const double *__restrict__ B, int ldb, int r0, int c0,
                                          double (*red)[16][64]) {
    mfma_tile_w<4, 4>(m, n, k, ta, tb, A, lda, B, ldb, r0, c0, red);
}

// Sum over the W partial tiles of tile element (reg, ln).
template <int W>
__device__ __forceinline__ double red_sum(const double (*red)[16][64], int reg, int ln) {
    double v = red[0][reg][ln];
#pragma unroll
    for (int q = 1; q < W; ++q) v += red[q][reg][ln];
    return v;
}

// Row and column in C of tile element idx (0..1023) of the tile at (r0, c0).
// D layout of the MFMA: register r of lane l holds row (l >> 4) + 4 r, column l & 15.
__device__ __forceinline__ void tile_pos(int idx, int r0, int c0, int &gr, int &gc) {
    const int reg = idx >> 6, ln = idx & 63;
    const int a = reg >> 3, b = (reg >> 2) & 1, r = reg & 3;
    gr = r0 + 16 * a + (ln >> 4) + 4 * r;
    gc = c0 + 16 * b + (ln & 15);
}

// Block sum of one value per thread (64 W threads); every thread gets the sum.
template <int W>
__device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = sh[0];
#pragma unroll
    for (int q = 1; q < W; ++q) t += sh[q];
    __syncthreads();
    return t;
}

// Block sum of one value per thread (256 threads); every thread gets the sum.
__device__ __forceinline__ double block_sum256(double v, double *sh) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return t;
}

}  // namespace pf
