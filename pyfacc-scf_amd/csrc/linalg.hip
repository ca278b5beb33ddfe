// Dense FP64 linear algebra for the SCF step on one GPU.
//
// Replaces orthonormal_diagonalize / diagonalize_symmetric (la.f90:7-57,
// LAPACK dsyev on the host, S re-diagonalised every call) and the host
// matmuls of rscf.f90:113-151:
//   * gemm: FP64 MFMA (v_mfma_f64_16x16x4f64), 32x32 tile per wavefront,
//     K staged through LDS in chunks of 16;
//   * jacobi: two-sided parallel Jacobi (round-robin pairing, n/2 disjoint
//     rotations per round, one kernel per round), used for S once and for
//     orbital energies;
//   * SP2 spectral projector (Niklasson, PRB 66, 155115 (2002)): the density
//     of the lowest n_occ eigenvectors of F' = X^T F X built from GEMMs only,
//     X <- X^2 or 2X - X^2 chosen on the device from the two traces;
//   * reductions (energy, ||dD||_F, traces, dot products) and element-wise
//     Fock assembly.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "kernels.h"
#include "linalg.h"

namespace pf {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ GEMM
// Row-major: element (r, c) of A is A[r*lda + c]; op(A) is m x k.
// C = alpha op(A) op(B) + beta Cin (Cin may alias C).  alpha_p / beta_p, when
// non-null, are read from device memory (device-side branch selection).
// One 256-thread workgroup per 32x32 tile of C: the four wavefronts split K
// (interleaved steps of 4), each accumulating the 2x2 grid of 16x16 MFMA
// blocks from operands loaded straight from L2, then the four partial tiles
// are summed through LDS.  For n = 448 this is 196 tiles x 4 wavefronts.
__device__ __forceinline__ double ld_op(const double *__restrict__ M, int ld, int r, int c, int nr, int nc, int t) {
    // element (r, c) of op(M) (op(M) is nr x nc); t: M stored transposed
    if (r >= nr || c >= nc) return 0.0;
    return t ? M[(size_t)c * ld + r] : M[(size_t)r * ld + c];
}

__global__ void __launch_bounds__(256) gemm_f64_kernel(int m, int n, int k, int ta, int tb, double alpha,
                                                        const double *__restrict__ alpha_p,
                                                        const double *__restrict__ A, int lda,
                                                        const double *__restrict__ B, int ldb, double beta,
                                                        const double *__restrict__ beta_p, const double *Cin,
                                                        double *C, int ldc) {
    __shared__ double red[4][16][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    const int i = lane & 15, kq = lane >> 4;
    double4_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = double4_t{0, 0, 0, 0};
#pragma unroll 4
    for (int kb = 4 * w; kb < k; kb += 16) {
        const int kk = kb + kq;
        // v_mfma_f64_16x16x4f64 operands: A lane l = (row l & 15, k l >> 4); B lane l = (k l >> 4, col l & 15)
        const double a0 = ld_op(A, lda, r0 + i, kk, m, k, ta), a1 = ld_op(A, lda, r0 + 16 + i, kk, m, k, ta);
        const double b0 = ld_op(B, ldb, kk, c0 + i, k, n, tb), b1 = ld_op(B, ldb, kk, c0 + 16 + i, k, n, tb);
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) red[w][(a * 2 + b) * 4 + r][lane] = acc[a][b][r];
    __syncthreads();
    const double al = alpha_p ? *alpha_p : alpha, be = beta_p ? *beta_p : beta;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int idx = threadIdx.x + 256 * q;  // 0..1023 = (reg, lane)
        const int reg = idx >> 6, ln = idx & 63;
        const double v = red[0][reg][ln] + red[1][reg][ln] + red[2][reg][ln] + red[3][reg][ln];
        // D layout: register r of lane l holds row (l >> 4) + 4 r, column l & 15
        const int a = reg >> 3, b = (reg >> 2) & 1, r = reg & 3;
        const int gr = r0 + 16 * a + (ln >> 4) + 4 * r, gc = c0 + 16 * b + (ln & 15);
        if (gr < m && gc < n) {
            const size_t o = (size_t)gr * ldc + gc;
            C[o] = be == 0.0 ? al * v : al * v + be * Cin[o];
        }
    }
}

void gemm_f64_ex(int m, int n, int k, bool ta, bool tb, double alpha, const double *alpha_p, const double *A, int lda,
                 const double *B, int ldb, double beta, const double *beta_p, const double *Cin, double *C, int ldc,
                 hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    dim3 grid((n + 31) / 32, (m + 31) / 32);
    gemm_f64_kernel<<<grid, 256, 0, st>>>(m, n, k, ta, tb, alpha, alpha_p, A, lda, B, ldb, beta, beta_p, Cin, C, ldc);
    PF_HIP(hipGetLastError());
}

void gemm_f64_mnk(int m, int n, int k, bool ta, bool tb, double alpha, const double *A, int lda, const double *B,
                  int ldb, double beta, double *C, int ldc, hipStream_t st) {
    gemm_f64_ex(m, n, k, ta, tb, alpha, nullptr, A, lda, B, ldb, beta, nullptr, C, C, ldc, st);
}

void gemm_f64(int n, bool ta, bool tb, double alpha, const double *A, const double *B, double beta, double *C,
              hipStream_t st) {
    gemm_f64_mnk(n, n, n, ta, tb, alpha, A, n, B, n, beta, C, n, st);
}

// ------------------------------------------------------------------ reductions
// out[slot] = sum_i f(i); single block of 1024 threads (n <= a few million).
template <typename Fn>
__global__ void __launch_bounds__(1024) reduce_kernel(int64_t n, Fn f, double *out) {
    __shared__ double part[16];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += f(i);
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x < 64) {
        s = threadIdx.x < (int)(blockDim.x >> 6) ? part[threadIdx.x] : 0.0;
        for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
        if (threadIdx.x == 0) *out = s;
    }
}

struct DotF {
    const double *a, *b;
    __device__ double operator()(int64_t i) const { return a[i] * b[i]; }
};
struct DiffSqF {
    const double *a, *b;
    __device__ double operator()(int64_t i) const { const double d = a[i] - b[i]; return d * d; }
};
struct TraceF {
    const double *a;
    int n;
    __device__ double operator()(int64_t i) const { return a[i * (n + 1)]; }
};
struct EnergyRF {  // D (H + F)
    const double *D, *H, *F;
    __device__ double operator()(int64_t i) const { return D[i] * (H[i] + F[i]); }
};
struct EnergyUF {  // 1/2 (Da+Db) H + 1/2 (Da Fa + Db Fb)
    const double *Da, *Db, *H, *Fa, *Fb;
    __device__ double operator()(int64_t i) const {
        return 0.5 * (Da[i] + Db[i]) * H[i] + 0.5 * (Da[i] * Fa[i] + Db[i] * Fb[i]);
    }
};

void dev_dot(int64_t n, const double *a, const double *b, double *out, hipStream_t st) {
    reduce_kernel<<<1, 1024, 0, st>>>(n, DotF{a, b}, out);
    PF_HIP(hipGetLastError());
}
void dev_diff_sq(int64_t n, const double *a, const double *b, double *out, hipStream_t st) {
    reduce_kernel<<<1, 1024, 0, st>>>(n, DiffSqF{a, b}, out);
    PF_HIP(hipGetLastError());
}
void dev_trace(int n, const double *a, double *out, hipStream_t st) {
    reduce_kernel<<<1, 1024, 0, st>>>(n, TraceF{a, n}, out);
    PF_HIP(hipGetLastError());
}
void dev_energy_rhf(int n, const double *D, const double *H, const double *F, double *out, hipStream_t st) {
    reduce_kernel<<<1, 1024, 0, st>>>((int64_t)n * n, EnergyRF{D, H, F}, out);
    PF_HIP(hipGetLastError());
}
void dev_energy_uhf(int n, const double *Da, const double *Db, const double *H, const double *Fa, const double *Fb,
                    double *out, hipStream_t st) {
    reduce_kernel<<<1, 1024, 0, st>>>((int64_t)n * n, EnergyUF{Da, Db, H, Fa, Fb}, out);
    PF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ element-wise
__global__ void fill_kernel(int64_t n, double v, double *x) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}
__global__ void fillf_kernel(int64_t n, float v, float *x) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}
void dev_fill(int64_t n, double v, double *x, hipStream_t st) {
    if (n <= 0) return;
    fill_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, v, x);
    PF_HIP(hipGetLastError());
}
void dev_fillf(int64_t n, float v, float *x, hipStream_t st) {
    if (n <= 0) return;
    fillf_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, v, x);
    PF_HIP(hipGetLastError());
}

// G from non-symmetric accumulators: RHF G = (J'+J'^T) - 1/2 (K'+K'^T);
// UHF Ga = Jt - Ka, Gb = Jt - Kb.  Template on the accumulator type; the
// result is added (scale 1) or stored (scale 0) into G (FP64).
template <typename R>
__global__ void assemble_g_kernel(int n, const R *J, const R *K0, const R *K1, double *G0, double *G1, int uhf,
                                  double keep) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const size_t ij = (size_t)i * n + j, ji = (size_t)j * n + i;
    const double jt = double(J[ij]) + double(J[ji]);
    if (!uhf) {
        const double v = jt - 0.5 * (double(K0[ij]) + double(K0[ji]));
        G0[ij] = keep * G0[ij] + v;
    } else {
        G0[ij] = keep * G0[ij] + jt - (double(K0[ij]) + double(K0[ji]));
        G1[ij] = keep * G1[ij] + jt - (double(K1[ij]) + double(K1[ji]));
    }
}
template <typename R>
void dev_assemble_g(int n, const R *J, const R *K0, const R *K1, double *G0, double *G1, bool uhf, bool accumulate,
                    hipStream_t st) {
    dim3 grid((n + 127) / 128, n);
    assemble_g_kernel<R><<<grid, 128, 0, st>>>(n, J, K0, K1, G0, G1, uhf ? 1 : 0, accumulate ? 1.0 : 0.0);
    PF_HIP(hipGetLastError());
}
template void dev_assemble_g<double>(int, const double *, const double *, const double *, double *, double *, bool,
                                     bool, hipStream_t);
template void dev_assemble_g<float>(int, const float *, const float *, const float *, double *, double *, bool, bool,
                                    hipStream_t);

__global__ void axpby_kernel(int64_t n, double a, const double *x, double b, const double *y, double *z) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) z[i] = a * x[i] + b * y[i];
}
void dev_axpby(int64_t n, double a, const double *x, double b, const double *y, double *z, hipStream_t st) {
    if (n <= 0) return;
    axpby_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, a, x, b, y, z);
    PF_HIP(hipGetLastError());
}

__global__ void to_float_kernel(int64_t n, double scale, const double *x, const double *y, float *z) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) z[i] = float(scale * (x[i] + (y ? y[i] : 0.0)));
}
void dev_to_float(int64_t n, double scale, const double *x, const double *y, float *z, hipStream_t st) {
    if (n <= 0) return;
    to_float_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, scale, x, y, z);
    PF_HIP(hipGetLastError());
}

// antisymmetric part: E = M - M^T
__global__ void antisym_kernel(int n, const double *M, double *E) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    E[(size_t)i * n + j] = M[(size_t)i * n + j] - M[(size_t)j * n + i];
}
void dev_antisym(int n, const double *M, double *E, hipStream_t st) {
    dim3 grid((n + 127) / 128, n);
    antisym_kernel<<<grid, 128, 0, st>>>(n, M, E);
    PF_HIP(hipGetLastError());
}

// Shell-block density maxima for screening: Dsh[a][b] = max |D_mu nu| over
// mu in a, nu in b, over up to two densities; then dmax = 2 max(Dsh).
template <typename R>
__global__ void shell_dmax_kernel(int nsh, int nbf, const int *sh_bf, const int *sh_nf, const double *D0,
                                  const double *D1, const double *D2, R *Dsh) {
    const int a = blockIdx.y, b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nsh) return;
    double m = 0.0;
    for (int i = 0; i < sh_nf[a]; ++i)
        for (int j = 0; j < sh_nf[b]; ++j) {
            const size_t ij = (size_t)(sh_bf[a] + i) * nbf + sh_bf[b] + j;
            m = fmax(m, fabs(D0[ij]));
            if (D1) m = fmax(m, fabs(D1[ij]));
            if (D2) m = fmax(m, fabs(D2[ij]));
        }
    Dsh[(size_t)a * nsh + b] = R(m);
}
template <typename R>
__global__ void __launch_bounds__(1024) max_kernel(int64_t n, const R *x, R scale, R *out) {
    __shared__ R part[16];
    R s = R(0);
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s = max(s, x[i]);
    for (int o = 32; o >= 1; o >>= 1) s = max(s, __shfl_xor(s, o, 64));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x < 64) {
        s = threadIdx.x < (int)(blockDim.x >> 6) ? part[threadIdx.x] : R(0);
        for (int o = 32; o >= 1; o >>= 1) s = max(s, __shfl_xor(s, o, 64));
        if (threadIdx.x == 0) *out = s * scale;
    }
}
template <typename R>
void dev_shell_dmax(int nsh, int nbf, const int *sh_bf, const int *sh_nf, const double *D0, const double *D1, R *Dsh,
                    R *dmax, hipStream_t st, const double *D2) {
    dim3 grid((nsh + 63) / 64, nsh);
    shell_dmax_kernel<R><<<grid, 64, 0, st>>>(nsh, nbf, sh_bf, sh_nf, D0, D1, D2, Dsh);
    max_kernel<R><<<1, 1024, 0, st>>>((int64_t)nsh * nsh, Dsh, R(2), dmax);
    PF_HIP(hipGetLastError());
}
template void dev_shell_dmax<double>(int, int, const int *, const int *, const double *, const double *, double *,
                                     double *, hipStream_t, const double *);
template void dev_shell_dmax<float>(int, int, const int *, const int *, const double *, const double *, float *,
                                    float *, hipStream_t, const double *);

// ------------------------------------------------------------------ Jacobi
// Round-robin (circle method) pairing over N = even padded size: player N-1
// is fixed; in round r, pair 0 is (r, N-1) and pair k is
// ((r+k) mod (N-1), (r-k) mod (N-1)) for k = 1..N/2-1.
__device__ __forceinline__ void rr_pair(int N, int r, int k, int &p, int &q) {
    if (k == 0) { p = r; q = N - 1; }
    else { p = (r + k) % (N - 1); q = (r - k + (N - 1)) % (N - 1); }
    if (p > q) { const int t = p; p = q; q = t; }
}

__device__ __forceinline__ void jacobi_rot(const double *A, int n, int p, int q, double &c, double &s) {
    c = 1.0; s = 0.0;
    if (q >= n) return;  // padding index
    const double apq = A[(size_t)p * n + q];
    if (apq == 0.0) return;
    const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
    const double theta = (aqq - app) / (2.0 * apq);
    const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    c = 1.0 / sqrt(t * t + 1.0);
    s = t * c;
}

// One round: A_out = G^T A_in G, V_out = V_in G with the round's N/2 disjoint
// rotations G.  Thread (k1, k2) updates the 2x2 block rows {p1,q1} x cols {p2,q2}.
__global__ void jacobi_round_kernel(int n, int N, int r, const double *__restrict__ A, double *__restrict__ Ao,
                                    const double *__restrict__ V, double *__restrict__ Vo) {
    const int m = N / 2;
    const int k2 = blockIdx.x * blockDim.x + threadIdx.x, k1 = blockIdx.y;
    if (k2 < m && k1 < m) {
        int p1, q1, p2, q2;
        rr_pair(N, r, k1, p1, q1);
        rr_pair(N, r, k2, p2, q2);
        double c1, s1, c2, s2;
        jacobi_rot(A, n, p1, q1, c1, s1);
        jacobi_rot(A, n, p2, q2, c2, s2);
        const bool q1ok = q1 < n, q2ok = q2 < n;
        const double a_pp = A[(size_t)p1 * n + p2];
        const double a_pq = q2ok ? A[(size_t)p1 * n + q2] : 0.0;
        const double a_qp = q1ok ? A[(size_t)q1 * n + p2] : 0.0;
        const double a_qq = (q1ok && q2ok) ? A[(size_t)q1 * n + q2] : 0.0;
        // right multiply by G2 (columns p2, q2): col_p' = c col_p - s col_q ; col_q' = s col_p + c col_q
        const double b_pp = c2 * a_pp - s2 * a_pq, b_pq = s2 * a_pp + c2 * a_pq;
        const double b_qp = c2 * a_qp - s2 * a_qq, b_qq = s2 * a_qp + c2 * a_qq;
        // left multiply by G1^T (rows p1, q1)
        double o_pp = c1 * b_pp - s1 * b_qp, o_pq = c1 * b_pq - s1 * b_qq;
        double o_qp = s1 * b_pp + c1 * b_qp, o_qq = s1 * b_pq + c1 * b_qq;
        if (k1 == k2) { o_pq = 0.0; o_qp = 0.0; }
        Ao[(size_t)p1 * n + p2] = o_pp;
        if (q2ok) Ao[(size_t)p1 * n + q2] = o_pq;
        if (q1ok) Ao[(size_t)q1 * n + p2] = o_qp;
        if (q1ok && q2ok) Ao[(size_t)q1 * n + q2] = o_qq;
    }
    // eigenvectors: blockIdx.y also indexes rows of V (grid.y = max(m, n))
    const int row = blockIdx.y;
    if (row < n && k2 < m) {
        int p, q;
        rr_pair(N, r, k2, p, q);
        double c, s;
        jacobi_rot(A, n, p, q, c, s);
        const double vp = V[(size_t)row * n + p];
        const double vq = q < n ? V[(size_t)row * n + q] : 0.0;
        Vo[(size_t)row * n + p] = c * vp - s * vq;
        if (q < n) Vo[(size_t)row * n + q] = s * vp + c * vq;
    }
}

struct OffF {
    const double *a;
    int n;
    __device__ double operator()(int64_t i) const {
        const int r = int(i / n), c = int(i % n);
        return r == c ? 0.0 : a[i] * a[i];
    }
};
struct SqF {
    const double *a;
    __device__ double operator()(int64_t i) const { return a[i] * a[i]; }
};

__global__ void eye_kernel(int n, double *V) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) V[(size_t)i * n + j] = (i == j) ? 1.0 : 0.0;
}
__global__ void diag_kernel(int n, const double *A, double *w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = A[(size_t)i * n + i];
}
__global__ void symmetrize_kernel(int n, const double *A, double *B) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) B[(size_t)i * n + j] = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
}
__global__ void permute_cols_kernel(int n, const int *perm, const double *V, double *Vo) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) Vo[(size_t)i * n + j] = V[(size_t)i * n + perm[j]];
}

void jacobi_eigh(int n, const double *A_in, double *w, double *V, hipStream_t st, int max_sweeps, double tol) {
    const int N = n + (n & 1);
    const int m = N / 2;
    DevBuf<double> A(size_t(n) * n), A2(size_t(n) * n), V2(size_t(n) * n), red(2);
    dim3 g2((n + 127) / 128, n);
    symmetrize_kernel<<<g2, 128, 0, st>>>(n, A_in, A.p);
    eye_kernel<<<g2, 128, 0, st>>>(n, V);
    double *a = A.p, *ao = A2.p, *v = V, *vo = V2.p;
    dim3 grid((m + 63) / 64, std::max(m, n));
    reduce_kernel<<<1, 1024, 0, st>>>((int64_t)n * n, SqF{a}, red.p + 1);
    double tot = 0.0;
    PF_HIP(hipMemcpyAsync(&tot, red.p + 1, sizeof(double), hipMemcpyDeviceToHost, st));
    PF_HIP(hipStreamSynchronize(st));
    for (int sweep = 0; sweep < max_sweeps && n > 1; ++sweep) {
        reduce_kernel<<<1, 1024, 0, st>>>((int64_t)n * n, OffF{a, n}, red.p);
        double off = 0.0;
        PF_HIP(hipMemcpyAsync(&off, red.p, sizeof(double), hipMemcpyDeviceToHost, st));
        PF_HIP(hipStreamSynchronize(st));
        if (off <= tol * tol * tot) break;
        for (int r = 0; r < N - 1; ++r) {
            // Ao/Vo entries not touched by this round keep their old value: copy first is unnecessary
            // because every element of A and V is written by exactly one (k1,k2)/(row,k2) thread.
            jacobi_round_kernel<<<grid, 64, 0, st>>>(n, N, r, a, ao, v, vo);
            std::swap(a, ao);
            std::swap(v, vo);
        }
    }
    PF_HIP(hipGetLastError());
    if (v != V) PF_HIP(hipMemcpyAsync(V, v, sizeof(double) * n * n, hipMemcpyDeviceToDevice, st));
    // ascending order
    std::vector<double> hw(n);
    DevBuf<double> dw(n);
    diag_kernel<<<(n + 63) / 64, 64, 0, st>>>(n, a, dw.p);
    PF_HIP(hipMemcpyAsync(hw.data(), dw.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    PF_HIP(hipStreamSynchronize(st));
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return hw[x] < hw[y]; });
    std::vector<double> ws(n);
    for (int i = 0; i < n; ++i) ws[i] = hw[perm[i]];
    PF_HIP(hipMemcpyAsync(w, ws.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
    DevBuf<int> dp(n);
    PF_HIP(hipMemcpyAsync(dp.p, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    permute_cols_kernel<<<g2, 128, 0, st>>>(n, dp.p, V, V2.p);
    PF_HIP(hipMemcpyAsync(V, V2.p, sizeof(double) * n * n, hipMemcpyDeviceToDevice, st));
    PF_HIP(hipStreamSynchronize(st));
}

// ------------------------------------------------------------------ SP2
// Gershgorin bounds of a symmetric matrix -> bounds[0] = min, bounds[1] = max.
__global__ void __launch_bounds__(1024) gershgorin_kernel(int n, const double *A, double *bounds) {
    __shared__ double lo_s[16], hi_s[16];
    double lo = 1e300, hi = -1e300;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double r = 0.0;
        for (int j = 0; j < n; ++j)
            if (j != i) r += fabs(A[(size_t)i * n + j]);
        const double d = A[(size_t)i * n + i];
        lo = fmin(lo, d - r);
        hi = fmax(hi, d + r);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, o, 64));
        hi = fmax(hi, __shfl_xor(hi, o, 64));
    }
    if ((threadIdx.x & 63) == 0) { lo_s[threadIdx.x >> 6] = lo; hi_s[threadIdx.x >> 6] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { lo = fmin(lo, lo_s[w]); hi = fmax(hi, hi_s[w]); }
        lo = fmin(lo, lo_s[0]);
        hi = fmax(hi, hi_s[0]);
        bounds[0] = lo;
        bounds[1] = hi;
    }
}
// X = (hi I - F) / (hi - lo)
__global__ void sp2_init_kernel(int n, const double *F, const double *bounds, double *X) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double lo = bounds[0], hi = bounds[1];
    const double s = 1.0 / (hi - lo);
    X[(size_t)i * n + j] = ((i == j ? hi : 0.0) - F[(size_t)i * n + j]) * s;
}
// One block: t1 = tr X, t2 = sum X_ij^2 (= tr X^2 for symmetric X).  Chooses
// the next step X <- X^2 (alpha 1, beta 0) when |t2 - nocc| < |2 t1 - t2 - nocc|,
// else X <- 2X - X^2 (alpha -1, beta 2); once the idempotency error t1 - t2 is
// below tol (or stops shrinking below 1e-8) the step becomes X <- X
// (alpha 0, beta 1) and state[2] flags completion.
// state: 0 t1, 1 t2, 2 done, 3 previous error, 4 alpha, 5 beta, 6/7 bounds
__global__ void __launch_bounds__(1024) sp2_step_kernel(int n, double nocc, const double *X, double *state,
                                                         int iter, double tol) {
    __shared__ double p1[16], p2[16];
    double t1 = 0.0, t2 = 0.0;
    const size_t n2 = (size_t)n * n;
    for (size_t e = threadIdx.x; e < n2; e += blockDim.x) {
        const double v = X[e];
        t2 += v * v;
        if (e % (n + 1) == 0) t1 += v;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        t1 += __shfl_xor(t1, o, 64);
        t2 += __shfl_xor(t2, o, 64);
    }
    if ((threadIdx.x & 63) == 0) { p1[threadIdx.x >> 6] = t1; p2[threadIdx.x >> 6] = t2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        t1 = 0.0; t2 = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { t1 += p1[w]; t2 += p2[w]; }
        state[0] = t1;
        state[1] = t2;
        const double err = fabs(t1 - t2);
        if (state[2] == 0.0 && iter > 2 && (err < tol || (err < 1e-8 && err >= state[3]))) state[2] = 1.0;
        state[3] = err;
        if (state[2] != 0.0) { state[4] = 0.0; state[5] = 1.0; }
        else if (fabs(t2 - nocc) < fabs(2.0 * t1 - t2 - nocc)) { state[4] = 1.0; state[5] = 0.0; }
        else { state[4] = -1.0; state[5] = 2.0; }
    }
}

int sp2_density(int n, const double *Fp, double nocc, double *X, double *work, double *state, hipStream_t st,
                int max_iter) {
    dim3 g2((n + 127) / 128, n);
    gershgorin_kernel<<<1, 1024, 0, st>>>(n, Fp, state + 6);
    sp2_init_kernel<<<g2, 128, 0, st>>>(n, Fp, state + 6, X);
    dev_fill(3, 0.0, state, st);
    dev_fill(1, 1e300, state + 3, st);
    const double tol = 1e-13 * std::max(1.0, nocc);
    double *cur = X, *nxt = work;
    int it = 0;
    for (; it < max_iter; ++it) {
        sp2_step_kernel<<<1, 1024, 0, st>>>(n, nocc, cur, state, it, tol);
        gemm_f64_ex(n, n, n, false, false, 0.0, state + 4, cur, n, cur, n, 0.0, state + 5, cur, nxt, n, st);
        std::swap(cur, nxt);
        if ((it & 7) == 7) {
            double done = 0.0;
            PF_HIP(hipMemcpyAsync(&done, state + 2, sizeof(double), hipMemcpyDeviceToHost, st));
            PF_HIP(hipStreamSynchronize(st));
            if (done != 0.0) break;
        }
    }
    // symmetrise (X + X^T)/2 into X
    if (cur == X) {
        symmetrize_kernel<<<g2, 128, 0, st>>>(n, X, work);
        PF_HIP(hipMemcpyAsync(X, work, sizeof(double) * n * n, hipMemcpyDeviceToDevice, st));
    } else {
        symmetrize_kernel<<<g2, 128, 0, st>>>(n, work, X);
    }
    PF_HIP(hipGetLastError());
    return it;
}

}  // namespace pf
