// Dense FP64 linear algebra for the SCF step on one GPU.
//
// Replaces orthonormal_diagonalize / diagonalize_symmetric (la.f90:7-57,
// LAPACK dsyev on the host, S re-diagonalised every call) and the host
// matmuls of rscf.f90:113-151:
//   * gemm: FP64 MFMA (v_mfma_f64_16x16x4f64), one 32x32 tile per
//     workgroup whose wavefronts split K, partial tiles summed through LDS;
//   * jacobi: two-sided parallel Jacobi (round-robin pairing, n/2 disjoint
//     rotations per round, one kernel per round), used for S once and for
//     orbital energies;
//   * SP2 spectral projector (Niklasson, PRB 66, 155115 (2002)): the density
//     of the lowest n_occ eigenvectors of F' = X^T F X built from GEMMs only,
//     X <- X^2 or 2X - X^2 chosen on the device from the two traces, one
//     kernel per step over the upper-triangle tiles of X^2 (16x16 tiles up
//     to n = 1024, 32x32 above; the traces of the next iterate come from its
//     epilogue);
//   * reductions (energy, ||dD||_F, traces, dot products) and element-wise
//     Fock assembly.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "bdense.h"
#include "kernels.h"
#include "linalg.h"

namespace pf {

// ------------------------------------------------------------------ GEMM
// Row-major: element (r, c) of A is A[r*lda + c]; op(A) is m x k.
// C = alpha op(A) op(B) + beta Cin (Cin may alias C).  alpha_p / beta_p, when
// non-null, are read from device memory (device-side branch selection).
// One workgroup of GW wavefronts per 32x32 tile of C; the wavefronts split K
// (runs of 4 MFMA k-steps, operand loads of a run issued together) and their
// partial tiles are summed through LDS.  For n = 448 this is 196 tiles x 8
// wavefronts.
constexpr int GW = 8;
__global__ void __launch_bounds__(64 * GW) gemm_f64_kernel(int m, int n, int k, int ta, int tb, double alpha,
                                                            const double *__restrict__ alpha_p,
                                                            const double *__restrict__ A, int lda,
                                                            const double *__restrict__ B, int ldb, double beta,
                                                            const double *__restrict__ beta_p, const double *Cin,
                                                            double *C, int ldc) {
    __shared__ double red[GW][16][64];
    const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    mfma_tile_w<GW, 4>(m, n, k, ta, tb, A, lda, B, ldb, r0, c0, red);
    const double al = alpha_p ? *alpha_p : alpha, be = beta_p ? *beta_p : beta;
    for (int idx = threadIdx.x; idx < 1024; idx += 64 * GW) {
        const int reg = idx >> 6, ln = idx & 63;
        const double v = red_sum<GW>(red, reg, ln);
        int gr, gc;
        tile_pos(idx, r0, c0, gr, gc);
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
    gemm_f64_kernel<<<grid, 64 * GW, 0, st>>>(m, n, k, ta, tb, alpha, alpha_p, A, lda, B, ldb, beta, beta_p, Cin, C, ldc);
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
// Two-stage deterministic reductions: up to RED_BLOCKS workgroups write one
// partial sum each (part[y * RED_BLOCKS + block] for result y), then one
// workgroup per result adds the partials in a fixed order.
template <typename Fn>
__global__ void __launch_bounds__(256) reduce_part_kernel(int64_t n, Fn f, double *part) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) s += f(i, blockIdx.y);
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * RED_BLOCKS + blockIdx.x] = s;
}
__global__ void __launch_bounds__(256) reduce_final_kernel(int nb, const double *part, double *out) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) s += part[(size_t)blockIdx.x * RED_BLOCKS + b];
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}
// out[y] = sum_i f(i, y), y < ny
template <typename Fn>
static void reduce2(int64_t n, int ny, Fn f, double *out, double *part, hipStream_t st) {
    const int nb = (int)std::min<int64_t>(RED_BLOCKS, std::max<int64_t>(1, (n + 1023) / 1024));
    reduce_part_kernel<<<dim3(nb, ny), 256, 0, st>>>(n, f, part);
    reduce_final_kernel<<<ny, 256, 0, st>>>(nb, part, out);
    PF_HIP(hipGetLastError());
}

struct DotsF {  // a . b_y
    const double *a;
    DotPtrs b;
    __device__ double operator()(int64_t i, int y) const { return a[i] * b.p[y][i]; }
};
struct DiffSqF {
    const double *a, *b;
    __device__ double operator()(int64_t i, int) const { const double d = a[i] - b[i]; return d * d; }
};
struct EnergyRF {  // D (H + F)
    const double *D, *H, *F;
    __device__ double operator()(int64_t i, int) const { return D[i] * (H[i] + F[i]); }
};
struct EnergyUF {  // 1/2 (Da+Db) H + 1/2 (Da Fa + Db Fb)
    const double *Da, *Db, *H, *Fa, *Fb;
    __device__ double operator()(int64_t i, int) const {
        return 0.5 * (Da[i] + Db[i]) * H[i] + 0.5 * (Da[i] * Fa[i] + Db[i] * Fb[i]);
    }
};

void dev_dots(int64_t n, const double *a, const DotPtrs &b, int m, double *out, double *part, hipStream_t st) {
    if (m < 1 || m > DotPtrs::MAX) throw Error("dev_dots: 1..8 vectors");
    reduce2(n, m, DotsF{a, b}, out, part, st);
}
void dev_diff_sq(int64_t n, const double *a, const double *b, double *out, double *part, hipStream_t st) {
    reduce2(n, 1, DiffSqF{a, b}, out, part, st);
}
void dev_energy_rhf(int n, const double *D, const double *H, const double *F, double *out, double *part,
                    hipStream_t st) {
    reduce2((int64_t)n * n, 1, EnergyRF{D, H, F}, out, part, st);
}
void dev_energy_uhf(int n, const double *Da, const double *Db, const double *H, const double *Fa, const double *Fb,
                    double *out, double *part, hipStream_t st) {
    reduce2((int64_t)n * n, 1, EnergyUF{Da, Db, H, Fa, Fb}, out, part, st);
}

// ---- DIIS
__global__ void __launch_bounds__(64) diis_solve_kernel(int ndiis, int nstore, int slot, const double *dots,
                                                        double *B, double *coef, int *ok) {
    constexpr int NM = DiisPtrs::MAX + 1;
    __shared__ double M[NM][NM + 1];
    __shared__ int piv_s;
    const int t = threadIdx.x, m = nstore, N = m + 1;
    if (t < m) {
        B[slot * ndiis + t] = dots[t];
        B[t * ndiis + slot] = dots[t];
    }
    __syncthreads();
    if (m < 2) {
        if (t == 0) ok[0] = 0;
        return;
    }
    // augmented system, row r = lane r: [B -1 | 0] and [-1 .. -1 0 | -1]
    if (t < N) {
        for (int c = 0; c < N; ++c)
            M[t][c] = t < m ? (c < m ? B[t * ndiis + c] : -1.0) : (c < m ? -1.0 : 0.0);
        M[t][N] = t < m ? 0.0 : -1.0;
    }
    __syncthreads();
    bool good = true;
    for (int col = 0; col < N; ++col) {
        // pivot: the largest |M[r][col]|, r >= col (lowest r on ties)
        double v = (t >= col && t < N) ? fabs(M[t][col]) : -1.0;
        int r = t;
        for (int o = 32; o >= 1; o >>= 1) {
            const double v2 = __shfl_xor(v, o, 64);
            const int r2 = __shfl_xor(r, o, 64);
            if (v2 > v || (v2 == v && r2 < r)) { v = v2; r = r2; }
        }
        if (t == 0) piv_s = r;
        __syncthreads();
        const int pr = piv_s;
        // uniform (every lane reads the same values); pr is outside col..N-1 only
        // when the column holds NaNs, and is then not used as an index
        if (pr < col || pr >= N || !(fabs(M[pr][col]) >= 1e-300)) { good = false; break; }
        if (pr != col && t <= N) {
            const double a = M[col][t], b = M[pr][t];
            M[col][t] = b;
            M[pr][t] = a;
        }
        __syncthreads();
        const double f = (t < N && t != col) ? M[t][col] / M[col][col] : 0.0;
        __syncthreads();
        if (t < N && t != col)
            for (int k = col; k <= N; ++k) M[t][k] -= f * M[col][k];
        __syncthreads();
    }
    if (t < m) coef[t] = good ? M[t][N] / M[t][t] : 0.0;
    if (t == 0) ok[0] = good ? 1 : 0;
}

__global__ void diis_extrap_kernel(int64_t n, int nstore, const double *coef, const int *ok, DiisPtrs v,
                                   const double *F, double *Fx) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (!ok[0]) {
        Fx[i] = F[i];
        return;
    }
    double s = 0.0;
    for (int k = 0; k < nstore; ++k) s += coef[k] * v.p[k][i];
    Fx[i] = s;
}

void dev_diis_solve(int ndiis, int nstore, int slot, const double *dots, double *B, double *coef, int *ok,
                    hipStream_t st) {
    if (ndiis > DiisPtrs::MAX || nstore > ndiis || slot < 0 || slot >= ndiis) throw Error("dev_diis_solve: sizes");
    diis_solve_kernel<<<1, 64, 0, st>>>(ndiis, nstore, slot, dots, B, coef, ok);
    PF_HIP(hipGetLastError());
}

void dev_diis_extrap(int64_t n, int nstore, const double *coef, const int *ok, const DiisPtrs &v, const double *F,
                     double *Fx, hipStream_t st) {
    if (nstore < 1 || nstore > DiisPtrs::MAX) throw Error("dev_diis_extrap: 1..32 vectors");
    diis_extrap_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, nstore, coef, ok, v, F, Fx);
    PF_HIP(hipGetLastError());
}

// Single-workgroup reduction for the Jacobi solver's small helper sums.
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
__global__ void __launch_bounds__(1024) fx_scale_kernel(int64_t n, const double *D0, const double *D1, double qmax,
                                                         double *fx) {
    __shared__ double sh[16];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) s += fabs(D0[i]) + (D1 ? fabs(D1[i]) : 0.0);
    s = block_sum<16>(s, sh);
    if (threadIdx.x == 0) {
        int e = 0;
        (void)frexp(2.0 * qmax * qmax * s, &e);  // bound < 2^e
        fx[0] = ldexp(1.0, min(61 - e, 200));
    }
}

__global__ void fx_to_double_kernel(int64_t n, const double *fx, double *x) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const long long v = reinterpret_cast<const long long *>(x)[i];
    x[i] = double(v) * (1.0 / fx[0]);
}

void dev_fx_scale(int64_t n, const double *D0, const double *D1, double qmax, double *fx, hipStream_t st) {
    fx_scale_kernel<<<1, 1024, 0, st>>>(n, D0, D1, qmax, fx);
    PF_HIP(hipGetLastError());
}

void dev_fx_to_double(int64_t n, const double *fx, double *x, hipStream_t st) {
    if (n <= 0) return;
    fx_to_double_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, fx, x);
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
// mu in a, nu in b, over up to three densities; and dmax = 2 max(Dsh), each
// wavefront's maximum merged with an integer atomic max on the bits (the
// order of non-negative IEEE values is that of their bit patterns), dmax
// cleared beforehand.
__device__ __forceinline__ void atomic_max_nonneg(double *p, double v) {
    atomicMax(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(__double_as_longlong(v)));
}
__device__ __forceinline__ void atomic_max_nonneg(float *p, float v) {
    atomicMax(reinterpret_cast<unsigned int *>(p), __float_as_uint(v));
}
template <typename R>
__global__ void __launch_bounds__(256) shell_dmax_kernel(int nsh, int nbf, const int *sh_bf, const int *sh_nf,
                                                         const double *D0, const double *D1, const double *D2, R *Dsh,
                                                         R *dmax) {
    // one wavefront per shell row a, its lanes over the column shells b; one
    // device atomic per workgroup of four rows
    __shared__ R wmax[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int a = blockIdx.x * 4 + w;
    double m = 0.0;
    if (a < nsh) {
        const int fa = sh_bf[a], na = sh_nf[a];
        for (int b = lane; b < nsh; b += 64) {
            const int fb = sh_bf[b], nb = sh_nf[b];
            double mb = 0.0;
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < nb; ++j) {
                    const size_t ij = (size_t)(fa + i) * nbf + fb + j;
                    mb = fmax(mb, fabs(D0[ij]));
                    if (D1) mb = fmax(mb, fabs(D1[ij]));
                    if (D2) mb = fmax(mb, fabs(D2[ij]));
                }
            Dsh[(size_t)a * nsh + b] = R(mb);
            m = fmax(m, mb);
        }
    }
    R v = R(m);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    if (lane == 0) wmax[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomic_max_nonneg(dmax, R(2) * max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3])));
}
template <typename R>
void dev_shell_dmax(int nsh, int nbf, const int *sh_bf, const int *sh_nf, const double *D0, const double *D1, R *Dsh,
                    R *dmax, hipStream_t st, const double *D2) {
    if constexpr (std::is_same<R, double>::value) dev_fill(1, 0.0, dmax, st);
    else dev_fillf(1, 0.0f, dmax, st);
    if (nsh <= 0) return;
    shell_dmax_kernel<R><<<(nsh + 3) / 4, 256, 0, st>>>(nsh, nbf, sh_bf, sh_nf, D0, D1, D2, Dsh, dmax);
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

// Gershgorin discs, one wavefront per row: rl[i] = A_ii - r_i, rh[i] = A_ii + r_i.
__global__ void __launch_bounds__(256) gershgorin_rows_kernel(int n, const double *A, double *rl, double *rh) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    double r = 0.0;
    for (int j = lane; j < n; j += 64)
        if (j != i) r += fabs(A[(size_t)i * n + j]);
    for (int o = 32; o >= 1; o >>= 1) r += __shfl_xor(r, o, 64);
    if (lane == 0) {
        const double d = A[(size_t)i * n + i];
        rl[i] = d - r;
        rh[i] = d + r;
    }
}
__global__ void __launch_bounds__(1024) gershgorin_bounds_kernel(int n, const double *rl, const double *rh,
                                                                 double *bounds) {
    __shared__ double lo_s[16], hi_s[16];
    double lo = 1e300, hi = -1e300;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        lo = fmin(lo, rl[i]);
        hi = fmax(hi, rh[i]);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, o, 64));
        hi = fmax(hi, __shfl_xor(hi, o, 64));
    }
    if ((threadIdx.x & 63) == 0) { lo_s[threadIdx.x >> 6] = lo; hi_s[threadIdx.x >> 6] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { lo = fmin(lo, lo_s[w]); hi = fmax(hi, hi_s[w]); }
        bounds[0] = lo;
        bounds[1] = hi;
    }
}

// ------------------------------------------------------------------ S^{-1/2}
// Coupled Newton-Schulz iteration (Higham, Numer. Algorithms 15 (1997) 227):
// Y0 = S/c, Z0 = I, T = (3I - Z Y)/2, Y <- Y T, Z <- T Z, with c >= lambda_max(S)
// from Gershgorin discs, converges quadratically to Y = (S/c)^{1/2},
// Z = (S/c)^{-1/2} for positive definite S; X = Z / sqrt(c) = S^{-1/2}.
// X X^T = S^{-1} as for the reference's canonical U s^{-1/2} (la.f90:36-50),
// so densities and orbital energies are the same.
__global__ void ns_init_kernel(int n, const double *S, const double *bounds, double *Y, double *Z, double *I) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const size_t o = (size_t)i * n + j;
    Y[o] = S[o] / bounds[1];
    Z[o] = I[o] = (i == j) ? 1.0 : 0.0;
}
__global__ void ns_finish_kernel(int n, const double *Z, const double *bounds, double *X) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    X[(size_t)i * n + j] = Z[(size_t)i * n + j] / sqrt(bounds[1]);
}
// out = max_ij |T_ij - delta_ij|
__global__ void __launch_bounds__(1024) resid_kernel(int n, const double *T, double *out) {
    __shared__ double sh[16];
    double m = 0.0;
    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += 1024) {
        const int64_t i = e / n, j = e - i * n;
        m = fmax(m, fabs(T[e] - (i == j ? 1.0 : 0.0)));
    }
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < 16; ++w) m = fmax(m, sh[w]);
        *out = m;
    }
}
// Repeated squaring M <- M^2 / tr(M)^2 for lambda_max: t = tr M, alpha = 1/t^2, logt = log t.
__global__ void __launch_bounds__(1024) trace_log_kernel(int n, const double *M, double *alpha, double *logt) {
    __shared__ double sh[16];
    double t = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) t += M[(size_t)i * n + i];
    for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        t = 0.0;
        for (int w = 0; w < 16; ++w) t += sh[w];
        *alpha = 1.0 / (t * t);
        *logt = log(t);
    }
}

bool inv_sqrt_spd(int n, const double *S, double *X, double *lam_min, hipStream_t st) {
    const size_t n2 = (size_t)n * n;
    DevBuf<double> buf(6 * n2), small(2 * (size_t)n + 64);
    double *Y = buf.p, *Z = Y + n2, *T = Z + n2, *Y2 = T + n2, *Z2 = Y2 + n2, *I = Z2 + n2;
    double *bounds = small.p, *resid = small.p + 2, *alpha = small.p + 3, *logs = small.p + 8, *rows = small.p + 64;
    gershgorin_rows_kernel<<<(n + 3) / 4, 256, 0, st>>>(n, S, rows, rows + n);
    gershgorin_bounds_kernel<<<1, 1024, 0, st>>>(n, rows, rows + n, bounds);
    const dim3 g2((n + 127) / 128, n);
    ns_init_kernel<<<g2, 128, 0, st>>>(n, S, bounds, Y, Z, I);
    bool ok = false;
    for (int round = 0; round < 6 && !ok; ++round) {
        for (int k = 0; k < 8; ++k) {
            gemm_f64_ex(n, n, n, false, false, -0.5, nullptr, Z, n, Y, n, 1.5, nullptr, I, T, n, st);  // (3I - ZY)/2
            gemm_f64_ex(n, n, n, false, false, 1.0, nullptr, Y, n, T, n, 0.0, nullptr, Y2, Y2, n, st);
            gemm_f64_ex(n, n, n, false, false, 1.0, nullptr, T, n, Z, n, 0.0, nullptr, Z2, Z2, n, st);
            std::swap(Y, Y2);
            std::swap(Z, Z2);
        }
        gemm_f64_ex(n, n, n, false, false, 1.0, nullptr, Z, n, Y, n, 0.0, nullptr, T, T, n, st);
        resid_kernel<<<1, 1024, 0, st>>>(n, T, resid);
        double r = 0.0;
        PF_HIP(hipMemcpyAsync(&r, resid, sizeof(double), hipMemcpyDeviceToHost, st));
        PF_HIP(hipStreamSynchronize(st));
        ok = std::isfinite(r) && r < 1e-11;
    }
    if (!ok) return false;
    ns_finish_kernel<<<g2, 128, 0, st>>>(n, Z, bounds, X);
    // lambda_min(S) = c / lambda_max(Z)^2, lambda_max(Z) = lim tr(Z^(2^k))^(2^-k): with
    // M_{k+1} = M_k^2 / t_k^2, t_k = tr M_k, log tr Z^(2^K) = log P_K + log t_K where
    // log P_{k+1} = 2 (log P_k + log t_k); relative error below n^(2^-K) - 1.
    constexpr int KSQ = 24;
    static_assert(8 + KSQ + 1 <= 64, "log slots");
    PF_HIP(hipMemcpyAsync(T, Z, sizeof(double) * n2, hipMemcpyDeviceToDevice, st));
    double *M = T, *M2 = Y2;
    for (int k = 0; k < KSQ; ++k) {
        trace_log_kernel<<<1, 1024, 0, st>>>(n, M, alpha, logs + k);
        gemm_f64_ex(n, n, n, false, false, 0.0, alpha, M, n, M, n, 0.0, nullptr, M2, M2, n, st);
        std::swap(M, M2);
    }
    trace_log_kernel<<<1, 1024, 0, st>>>(n, M, alpha, logs + KSQ);
    double hl[KSQ + 1], c[2];
    PF_HIP(hipMemcpyAsync(hl, logs, sizeof(hl), hipMemcpyDeviceToHost, st));
    PF_HIP(hipMemcpyAsync(c, bounds, sizeof(c), hipMemcpyDeviceToHost, st));
    PF_HIP(hipStreamSynchronize(st));
    double logP = 0.0;
    for (int k = 0; k < KSQ; ++k) logP = 2.0 * (logP + hl[k]);
    const double log_lmax_z = (logP + hl[KSQ]) / std::ldexp(1.0, KSQ);
    *lam_min = c[1] * std::exp(-2.0 * log_lmax_z);
    PF_HIP(hipGetLastError());
    return true;
}

// ------------------------------------------------------------------ SP2
// Work layout (sp2_state_size doubles): [0] lo, [1] hi Gershgorin bounds;
// [2 + 2s] idempotency error and [3 + 2s] steps since convergence (0: not
// converged) recorded with slot s;
// from [8]: per-tile partial sums P[s][q][tile], q = 0 trace, q = 1 sum of
// squares, for three rotating slots s; then 2n doubles of Gershgorin row
// bounds.  Step k reads slot k % 3 (written by
// step k-1, or by the initial kernel for k = 0) and writes slot (k+1) % 3, so
// one kernel per step needs no separate reduction launch, and every tile sums
// the partials in the same order (the decisions are deterministic).

// Tile t of the upper triangle of the nt x nt tile grid (row-major) -> (ty, tx), ty <= tx.
__device__ __forceinline__ void tri_tile(int t, int nt, int &ty, int &tx) {
    ty = 0;
    while (t >= nt - ty) { t -= nt - ty; ++ty; }
    tx = ty + t;
}

// X0 = (hi I - F) / (hi - lo) on the upper-triangle T x T tiles (each written to
// both halves, so X0 is exactly symmetric), with slot-0 partials and records.
template <int T>
__global__ void __launch_bounds__(256) sp2_init_kernel(int n, const double *F, double *state, double *X) {
    __shared__ double sh[4];
    const int nt = (n + T - 1) / T, tile = blockIdx.x, ntile = nt * (nt + 1) / 2;
    int ty, tx;
    tri_tile(tile, nt, ty, tx);
    const int r0 = ty * T, c0 = tx * T;
    const double lo = state[0], hi = state[1], sc = 1.0 / (hi - lo);
    double t1 = 0.0, t2 = 0.0;
    for (int e = threadIdx.x; e < T * T; e += 256) {
        const int r = r0 + e / T, c = c0 + e % T;
        if (r < n && c < n && (ty != tx || r <= c)) {
            const double v = ((r == c ? hi : 0.0) - F[(size_t)r * n + c]) * sc;
            X[(size_t)r * n + c] = v;
            X[(size_t)c * n + r] = v;
            t2 += r == c ? v * v : 2.0 * v * v;
            if (r == c) t1 += v;
        }
    }
    t1 = block_sum256(t1, sh);
    t2 = block_sum256(t2, sh);
    if (threadIdx.x == 0) {
        state[8 + 0 * ntile + tile] = t1;
        state[8 + 1 * ntile + tile] = t2;
        if (tile == 0) { state[2] = 1e300; state[3] = 0.0; }
    }
}

// The 16x16 tile (r0, c0) of X X for symmetric X (n x n, row-major) by W
// wavefronts that split k into runs of 4 KS.  Both MFMA operands are read
// from rows of X (B[k][c] = X[c][k]), and inside a run the k index is
// relabelled: lane (i, q) supplies k = kb + 8 (j / 2) + 2 q + (j % 2) to MFMA
// step j, for A (row r0 + i) and B (row c0 + i) alike.  The sum over k is
// unchanged, each lane loads pairs of doubles (16 bytes), the four lanes of a
// row read 64 contiguous bytes per load, and every load of the run is in
// flight before the first MFMA.  At n <= 4 KS W (448 for KS = 14, W = 8)
// each wavefront makes one run.
template <int KS>
struct SymRun16 {
    static_assert(KS % 2 == 0, "pairs of doubles per load");
    double a[KS], b[KS];
    // operands of the run at kb (rows and k past n read as 0)
    __device__ __forceinline__ void load(int n, const double *__restrict__ X, int r0, int c0, int kb) {
        const int lane = threadIdx.x & 63, i = lane & 15, q = lane >> 4;
        const int ra = r0 + i, rb = c0 + i;
        const double *pa = X + (size_t)min(ra, n - 1) * n, *pb = X + (size_t)min(rb, n - 1) * n;
        const int k0 = kb + 2 * q;
        if ((n & 1) == 0 && kb + 4 * KS <= n) {
            // n and k0 even: 16-byte aligned pairs, the whole run in bounds
#pragma unroll
            for (int m = 0; m < KS / 2; ++m) {
                const double2 va = *reinterpret_cast<const double2 *>(pa + k0 + 8 * m);
                const double2 vb = *reinterpret_cast<const double2 *>(pb + k0 + 8 * m);
                a[2 * m] = va.x; a[2 * m + 1] = va.y;
                b[2 * m] = vb.x; b[2 * m + 1] = vb.y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < KS; ++j) {
                const int k = k0 + 8 * (j >> 1) + (j & 1);
                a[j] = k < n ? pa[k] : 0.0;
                b[j] = k < n ? pb[k] : 0.0;
            }
        }
        if (ra >= n)
#pragma unroll
            for (int j = 0; j < KS; ++j) a[j] = 0.0;
        if (rb >= n)
#pragma unroll
            for (int j = 0; j < KS; ++j) b[j] = 0.0;
    }
    __device__ __forceinline__ void mma(double4_t &acc) const {
#pragma unroll
        for (int j = 0; j < KS; ++j) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[j], b[j], acc, 0, 0, 0);
    }
};

// One purification step k: from slot k % 3 (t1 = tr X, t2 = sum X_ij^2 = tr X^2
// for symmetric X, the previous error and done flag), X' = X^2 when
// |t2 - nocc| < |2 t1 - t2 - nocc|, else 2X - X^2; once the idempotency error
// |t1 - t2| is below tol (or stops shrinking below 1e-8, after step 2) the
// step copies X and the done flag is raised.  X' partials and records go to
// slot (k+1) % 3.  X' is symmetric: only the upper-triangle T x T tiles are
// computed, each written to both halves.  With T = 16 the first k run's
// operands and each thread's element of X are loaded before the step's
// record and partials are read, so the step waits for memory once.
constexpr int SW = GW;  // wavefronts per tile of the purification step
constexpr int SP2_KS = 14;
template <int T>
__global__ void __launch_bounds__(64 * SW) sp2_step_kernel(int n, double nocc, double tol, int k, const double *X,
                                                        double *Xo, double *state) {
    constexpr int NREG = T * T / 64;  // tile doubles per lane of one partial
    __shared__ double red[SW][NREG][64];
    __shared__ double sh[SW];
    const int nt = (n + T - 1) / T, tile = blockIdx.x, ntile = nt * (nt + 1) / 2;
    const int s0 = k % 3, s1 = (k + 1) % 3;
    int ty, tx;
    tri_tile(tile, nt, ty, tx);
    const int r0 = ty * T, c0 = tx * T;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    SymRun16<SP2_KS> run;
    double xo = 0.0;  // X at this thread's element of a 16 x 16 tile
    if constexpr (T == 16) {
        if (4 * SP2_KS * w < n) run.load(n, X, r0, c0, 4 * SP2_KS * w);
        if (threadIdx.x < 256) {
            const int gr = r0 + (lane >> 4) + 4 * w, gc = c0 + (lane & 15);
            if (gr < n && gc < n) xo = X[(size_t)gr * n + gc];
        }
    }
    // record [3 + 2s]: steps since convergence (0 = not converged).  The first
    // two converged steps copy X into both buffers; later ones do nothing.
    const double age = state[3 + 2 * s0];
    if (age >= 2.0) {
        if (tile == 0 && threadIdx.x == 0) {
            state[2 + 2 * s1] = state[2 + 2 * s0];
            state[3 + 2 * s1] = age + 1.0;
        }
        return;
    }
    const double *P = state + 8;
    double t1 = 0.0, t2 = 0.0;
    for (int b = threadIdx.x; b < ntile; b += 64 * SW) {  // fixed order: deterministic sums
        t1 += P[(2 * s0 + 0) * ntile + b];
        t2 += P[(2 * s0 + 1) * ntile + b];
    }
    t1 = block_sum<SW>(t1, sh);
    t2 = block_sum<SW>(t2, sh);
    const double err = fabs(t1 - t2), prev = state[2 + 2 * s0];
    bool done = age != 0.0;
    if (!done && k > 2 && (err < tol || (err < 1e-8 && err >= prev))) done = true;
    const bool square = fabs(t2 - nocc) < fabs(2.0 * t1 - t2 - nocc);
    if (!done) {
        if constexpr (T == 16) {
            double4_t acc = double4_t{0, 0, 0, 0};
            for (int kb = 4 * SP2_KS * w; kb < n; kb += 4 * SP2_KS * SW) {
                if (kb != 4 * SP2_KS * w) run.load(n, X, r0, c0, kb);
                run.mma(acc);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) red[w][r][lane] = acc[r];
            __syncthreads();
        } else {
            mfma_tile_w<SW, 4>(n, n, n, 0, 0, X, n, X, n, r0, c0, red);
        }
    }
    double u1 = 0.0, u2 = 0.0;
    for (int idx = threadIdx.x; idx < T * T; idx += 64 * SW) {
        const int reg = idx >> 6, ln = idx & 63;
        int gr, gc;
        if constexpr (T == 16) {
            gr = r0 + (ln >> 4) + 4 * reg;
            gc = c0 + (ln & 15);
        } else {
            tile_pos(idx, r0, c0, gr, gc);
        }
        if (gr < n && gc < n && (ty != tx || gr <= gc)) {
            const size_t o = (size_t)gr * n + gc;
            const double x = T == 16 ? xo : X[o];
            double v;
            if (done) {
                v = x;
            } else {
                double x2 = red[0][reg][ln];
#pragma unroll
                for (int q = 1; q < SW; ++q) x2 += red[q][reg][ln];
                v = square ? x2 : 2.0 * x - x2;
            }
            Xo[o] = v;
            Xo[(size_t)gc * n + gr] = v;
            u2 += gr == gc ? v * v : 2.0 * v * v;
            if (gr == gc) u1 += v;
        }
    }
    u1 = block_sum<SW>(u1, sh);
    u2 = block_sum<SW>(u2, sh);
    if (threadIdx.x == 0) {
        state[8 + (2 * s1 + 0) * ntile + tile] = u1;
        state[8 + (2 * s1 + 1) * ntile + tile] = u2;
        if (tile == 0) {
            state[2 + 2 * s1] = err;
            state[3 + 2 * s1] = done ? age + 1.0 : 0.0;
        }
    }
}

namespace {
// Gershgorin bounds and X0, the first part of both purification drivers.
void sp2_begin(int n, const double *Fp, double *X, double *state, hipStream_t st) {
    double *rows = state + 8 + 6 * sp2_tiles(n);
    gershgorin_rows_kernel<<<(n + 3) / 4, 256, 0, st>>>(n, Fp, rows, rows + n);
    gershgorin_bounds_kernel<<<1, 1024, 0, st>>>(n, rows, rows + n, state);
    const dim3 tiles((unsigned)sp2_tiles(n));
    if (sp2_tile_edge(n) == 16) sp2_init_kernel<16><<<tiles, 256, 0, st>>>(n, Fp, state, X);
    else sp2_init_kernel<32><<<tiles, 256, 0, st>>>(n, Fp, state, X);
}

void sp2_step(int n, double nocc, double tol, int k, const double *X, double *Xo, double *state, hipStream_t st) {
    const dim3 tiles((unsigned)sp2_tiles(n));
    if (sp2_tile_edge(n) == 16) sp2_step_kernel<16><<<tiles, 64 * SW, 0, st>>>(n, nocc, tol, k, X, Xo, state);
    else sp2_step_kernel<32><<<tiles, 64 * SW, 0, st>>>(n, nocc, tol, k, X, Xo, state);
}
}  // namespace

int sp2_density(int n, const double *Fp, double nocc, double *X, double *work, double *state, hipStream_t st,
                int max_iter, int first_check) {
    sp2_begin(n, Fp, X, state, st);
    const double tol = 1e-13 * std::max(1.0, nocc);
    double *cur = X, *nxt = work;
    int used = max_iter, check = std::max(1, std::min(first_check, max_iter));
    for (int it = 0; it < max_iter; ++it) {
        sp2_step(n, nocc, tol, it, cur, nxt, state, st);
        std::swap(cur, nxt);
        if (it + 1 == check) {
            // record written by step it: steps since convergence (0: not yet)
            double age = 0.0;
            PF_HIP(hipMemcpyAsync(&age, state + 3 + 2 * ((it + 1) % 3), sizeof(double), hipMemcpyDeviceToHost, st));
            PF_HIP(hipStreamSynchronize(st));
            if (age != 0.0) {
                used = it + 1 - (int)age;
                break;
            }
            check += 4;
        }
    }
    // symmetrise (X + X^T)/2 into X
    const dim3 g2((n + 127) / 128, n);
    if (cur == X) {
        symmetrize_kernel<<<g2, 128, 0, st>>>(n, X, work);
        PF_HIP(hipMemcpyAsync(X, work, sizeof(double) * n * n, hipMemcpyDeviceToDevice, st));
    } else {
        symmetrize_kernel<<<g2, 128, 0, st>>>(n, work, X);
    }
    PF_HIP(hipGetLastError());
    return used;
}

// The same purification with no host round trip and no allocation: every
// buffer comes from the caller (X and work n x n, state sp2_state_size(n)
// doubles), exactly max_steps step kernels are launched (steps after
// convergence only copy or do nothing), and the symmetrised projector goes to
// P.  Nothing here synchronises, so the sequence can be captured in a graph.
void sp2_project_fixed(int n, const double *Fp, double nocc, double *P, double *X, double *work, double *state,
                       int max_steps, hipStream_t st) {
    sp2_begin(n, Fp, X, state, st);
    const double tol = 1e-13 * std::max(1.0, nocc);
    double *cur = X, *nxt = work;
    for (int it = 0; it < max_steps; ++it) {
        sp2_step(n, nocc, tol, it, cur, nxt, state, st);
        std::swap(cur, nxt);
    }
    const dim3 g2((n + 127) / 128, n);
    symmetrize_kernel<<<g2, 128, 0, st>>>(n, cur, P);
    PF_HIP(hipGetLastError());
}

}  // namespace pf
