// Batched SCF over many independent molecules: host driver and the batched
// dense kernels (per-problem MFMA GEMM tiles, DIIS, SP2, energy/convergence).
// Batch-level layout: molecule matrices (S, H, X, D_alpha + D_beta) and
// problem matrices (D, F, DIIS rings, SP2 iterates) each live in one pool,
// matrix of molecule/problem k at offset off[k] (n_k^2 doubles, row-major).
// Kernels take the problem (or molecule) from blockIdx.x and skip frozen
// (converged) molecules.
#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>

#include "batch.h"
#include "bdense.h"
#include "boys.h"
#include "cart.h"
#include "hostpar.h"
#include "linalg.h"

namespace pf {

size_t batch_digest_lds(int nmax);
void batch_digest_prepare(int nmax);
// Full FP64 builds of the molecules with mprec != 0; with G (mixed form) they
// also record G and the density built from, and the molecules with mprec == 0
// get FP32 increments on dD = D - Db (kets cut at Q Q max|dD| < tau).
void launch_batch_digest(const BatchDev &b, const std::array<BatchClass, 6> &cls, const BatchItem *items,
                         int nitems, const int *msplit, const double *D, const double *Dtot, const double *H,
                         double *F, double *gacc, int waves, hipStream_t st, double *G, double *Db, float *dD32,
                         float *dDt32, float *dmax, double tau, double tau_inc, const std::array<int, 6> &nbra);

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
__host__ __device__ inline int ntile(int n) { return ((n + 31) / 32) * ((n + 31) / 32); }
long long round64(long long n) { return (n + 63) / 64 * 64; }
size_t sp2_size(int n) { return 8 + 6 * (size_t)ntile(n) + 2 * (size_t)n; }
constexpr int SP2_MAX_STEPS = 100;  // purification steps per density at most (converged ones stop early)

// C_p = alpha a_p op(A_p) op(B_p) + beta Cin_p for each problem p (a_p = ascale[p], or 1).
struct BGemm {
    const double *A, *B, *Cin;
    double *C;
    const long long *oA, *oB, *oCin, *oC;
    const double *ascale;
    int ta, tb;
    double alpha, beta;
};
__global__ void __launch_bounds__(256) bgemm_kernel(BatchDev b, BGemm g) {
    __shared__ double red[4][16][64];
    const int p = blockIdx.x, n = b.pn[p];
    if (!b.mactive[b.pmol[p]]) return;
    const int nt = (n + 31) / 32, t = blockIdx.y;
    if (t >= nt * nt) return;
    const int r0 = (t / nt) * 32, c0 = (t % nt) * 32;
    mfma_tile(n, n, n, g.ta, g.tb, g.A + g.oA[p], n, g.B + g.oB[p], n, r0, c0, red);
    const double al = g.ascale ? g.alpha * g.ascale[p] : g.alpha;
    double *C = g.C + g.oC[p];
    const double *Cin = g.Cin ? g.Cin + g.oCin[p] : nullptr;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int idx = threadIdx.x + 256 * q, reg = idx >> 6, ln = idx & 63;
        int gr, gc;
        tile_pos(idx, r0, c0, gr, gc);
        if (gr < n && gc < n) {
            const double v = red[0][reg][ln] + red[1][reg][ln] + red[2][reg][ln] + red[3][reg][ln];
            const size_t o = (size_t)gr * n + gc;
            C[o] = Cin ? al * v + g.beta * Cin[o] : al * v;
        }
    }
}

// grid (P, chunks): element loop over problem p's n^2 entries
#define PROBLEM_LOOP(b)                                   \
    const int p = blockIdx.x, n = b.pn[p];                \
    if (!b.mactive[b.pmol[p]]) return;                    \
    for (int e = blockIdx.y * 256 + threadIdx.x; e < n * n; e += gridDim.y * 256)

// dst_p = src_p (src offsets per problem: os)
__global__ void __launch_bounds__(256) bcopy_kernel(BatchDev b, const double *src, const long long *os, double *dst) {
    PROBLEM_LOOP(b) dst[b.poff[p] + e] = src[os[p] + e];
}
// DIIS insert at the molecule's ring head: E = W - W^T (W = F D S), and F
__global__ void __launch_bounds__(256) bdiis_insert_kernel(BatchDev b, const double *W, const double *F, double *dE,
                                                           double *dF) {
    PROBLEM_LOOP(b) {
        const int r = e / n, c = e - r * n;
        const size_t slot = (size_t)b.mhead[b.pmol[p]] * n * n;
        dE[b.pdiis[p] + slot + e] = W[b.poff[p] + e] - W[b.poff[p] + (size_t)c * n + r];
        dF[b.pdiis[p] + slot + e] = F[b.poff[p] + e];
    }
}
// B[m][slot][k] = sum over the molecule's spins of E_slot . E_k, k < new ring size
__global__ void __launch_bounds__(256) bdiis_dots_kernel(BatchDev b, const double *dE) {
    __shared__ double sh[4];
    const int m = blockIdx.x, k = blockIdx.y;
    if (!b.mactive[m]) return;
    const int slot = b.mhead[m], ns = min(b.mnstore[m] + 1, b.ndiis);
    if (k >= ns) return;
    double s = 0.0;
    const int nd = b.muhf[m] ? 2 : 1;
    for (int sp = 0; sp < nd; ++sp) {
        const int p = b.mprob[m] + sp, n = b.pn[p];
        const double *es = dE + b.pdiis[p] + (size_t)slot * n * n, *ek = dE + b.pdiis[p] + (size_t)k * n * n;
        for (int e = threadIdx.x; e < n * n; e += 256) s += es[e] * ek[e];
    }
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) {
        b.mB[((size_t)m * b.ndiis + slot) * b.ndiis + k] = s;
        b.mB[((size_t)m * b.ndiis + k) * b.ndiis + slot] = s;
    }
}
// One thread per molecule: solve [B -1; -1 0][c; l] = [0; -1] over the ring
// (Gaussian elimination with partial pivoting), advance the ring.
__global__ void bdiis_solve_kernel(BatchDev b) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= b.M || !b.mactive[m]) return;
    constexpr int NMAX = 9;  // ring size <= 8
    const int nd = b.ndiis, ns = min(b.mnstore[m] + 1, nd), N = ns + 1;
    double A[NMAX][NMAX + 1];
    int ok = ns >= 2;
    if (ok) {
        for (int r = 0; r < ns; ++r) {
            for (int c = 0; c < ns; ++c) A[r][c] = b.mB[((size_t)m * nd + r) * nd + c];
            A[r][ns] = -1.0;
            A[r][N] = 0.0;
        }
        for (int c = 0; c < ns; ++c) A[ns][c] = -1.0;
        A[ns][ns] = 0.0;
        A[ns][N] = -1.0;
        for (int col = 0; col < N && ok; ++col) {
            int piv = col;
            for (int r = col + 1; r < N; ++r)
                if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
            if (fabs(A[piv][col]) < 1e-300) { ok = 0; break; }
            if (piv != col)
                for (int k = 0; k <= N; ++k) { const double t = A[col][k]; A[col][k] = A[piv][k]; A[piv][k] = t; }
            for (int r = 0; r < N; ++r) {
                if (r == col) continue;
                const double f = A[r][col] / A[col][col];
                for (int k = col; k <= N; ++k) A[r][k] -= f * A[col][k];
            }
        }
        if (ok)
            for (int r = 0; r < ns; ++r) b.mcoef[(size_t)m * nd + r] = A[r][N] / A[r][r];
    }
    b.muse[m] = ok;
    b.mhead[m] = (b.mhead[m] + 1) % nd;
    b.mnstore[m] = ns;
}
// Fx_p = sum_k c_k F_k (DIIS) or F_p
__global__ void __launch_bounds__(256) bdiis_extrap_kernel(BatchDev b, const double *F, const double *dF,
                                                           double *Fx) {
    PROBLEM_LOOP(b) {
        const int m = b.pmol[p];
        double v;
        if (b.muse[m]) {
            v = 0.0;
            for (int k = 0; k < b.mnstore[m]; ++k)
                v += b.mcoef[(size_t)m * b.ndiis + k] * dF[b.pdiis[p] + (size_t)k * n * n + e];
        } else {
            v = F[b.poff[p] + e];
        }
        Fx[b.poff[p] + e] = v;
    }
}
// D_alpha + D_beta of UHF molecules
__global__ void __launch_bounds__(256) bdtot_kernel(BatchDev b, const double *D, double *Dtot) {
    PROBLEM_LOOP(b) {
        const int m = b.pmol[p];
        if (!b.muhf[m] || p != b.mprob[m]) return;
        Dtot[b.moff[m] + e] = D[b.poff[p] + e] + D[b.poff[p + 1] + e];
    }
}

// Superposition of atomic densities: D_p = f x the block-diagonal of the
// atoms' densities (f = 1 RHF, 1/2 per spin UHF).  Basis function i of
// molecule m has bfi[3 (mbf[m] + i) ..] = (atom, the row's start in tab, its
// index on the isolated atom).
__global__ void __launch_bounds__(256) bsad_kernel(BatchDev b, const int *bfi, const int *mbf, const double *tab,
                                                   double *D) {
    PROBLEM_LOOP(b) {
        const int m = b.pmol[p], r = e / n, c = e - r * n;
        const int *ir = bfi + 3 * (mbf[m] + r), *ic = bfi + 3 * (mbf[m] + c);
        const double f = b.muhf[m] ? 0.5 : 1.0;
        D[b.poff[p] + e] = ir[0] == ic[0] ? f * tab[ir[1] + ic[2]] : 0.0;
    }
}

// ---- SP2 per problem: one workgroup runs the whole purification of its
// problem (Gershgorin bounds, X0 = (hi I - F')/(hi - lo), then X <- X^2 or
// 2X - X^2 chosen from tr X and tr X^2 = sum X_ij^2 until the idempotency
// error |tr X - tr X^2| is below tol or stops shrinking below 1e-8), as
// sp2_density does for one matrix with one launch per step.  The iterates
// ping-pong between Xa and Xb; the result, symmetrised, goes to P.
__device__ __forceinline__ double block_min256(double v, double *sh) {
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = fmin(fmin(sh[0], sh[1]), fmin(sh[2], sh[3]));
    __syncthreads();
    return t;
}
__global__ void __launch_bounds__(256) bsp2_kernel(BatchDev b, const double *Fp, double *Xa, double *Xb, double *Pout,
                                                    int max_steps) {
    __shared__ double red[4][16][64];
    __shared__ double sh[4];
    const int p = blockIdx.x, n = b.pn[p];
    if (!b.mactive[b.pmol[p]]) return;
    const size_t o = b.poff[p];
    const double *F = Fp + o;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // Gershgorin discs, one wavefront per row
    double lo = 1e300, hi = -1e300;
    for (int i = w; i < n; i += 4) {
        double r = 0.0;
        for (int j = lane; j < n; j += 64)
            if (j != i) r += fabs(F[(size_t)i * n + j]);
        for (int s = 32; s >= 1; s >>= 1) r += __shfl_xor(r, s, 64);
        const double d = F[(size_t)i * n + i];
        lo = fmin(lo, d - r);
        hi = fmax(hi, d + r);
    }
    lo = block_min256(lo, sh);
    hi = -block_min256(-hi, sh);
    const double sc = 1.0 / (hi - lo);
    double *cur = Xa + o, *nxt = Xb + o;
    double t1 = 0.0, t2 = 0.0;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const int r = e / n, c = e - r * n;
        const double v = ((r == c ? hi : 0.0) - F[e]) * sc;
        cur[e] = v;
        t2 += v * v;
        if (r == c) t1 += v;
    }
    t1 = block_sum256(t1, sh);
    t2 = block_sum256(t2, sh);
    const double nocc = b.pnocc[p], tol = 1e-13 * fmax(1.0, nocc);
    const int nt = (n + 31) / 32;
    double prev = 1e300;
    for (int k = 0; k < max_steps; ++k) {
        const double err = fabs(t1 - t2);
        if (k > 2 && (err < tol || (err < 1e-8 && err >= prev))) break;
        prev = err;
        const bool square = fabs(t2 - nocc) < fabs(2.0 * t1 - t2 - nocc);
        double u1 = 0.0, u2 = 0.0;
        for (int tile = 0; tile < nt * nt; ++tile) {
            const int r0 = (tile / nt) * 32, c0 = (tile % nt) * 32;
            mfma_tile(n, n, n, 0, 0, cur, n, cur, n, r0, c0, red);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int idx = threadIdx.x + 256 * q, reg = idx >> 6, ln = idx & 63;
                int gr, gc;
                tile_pos(idx, r0, c0, gr, gc);
                if (gr < n && gc < n) {
                    const size_t e = (size_t)gr * n + gc;
                    const double x2 = red[0][reg][ln] + red[1][reg][ln] + red[2][reg][ln] + red[3][reg][ln];
                    const double v = square ? x2 : 2.0 * cur[e] - x2;
                    nxt[e] = v;
                    u2 += v * v;
                    if (gr == gc) u1 += v;
                }
            }
            __syncthreads();  // red is rewritten by the next tile
        }
        t1 = block_sum256(u1, sh);  // also orders this step's writes before the next step's reads
        t2 = block_sum256(u2, sh);
        double *t = cur;
        cur = nxt;
        nxt = t;
    }
    // Not a projector: the occupied/virtual boundary falls inside a degenerate
    // level (an open shell from a spin-symmetric start), where purification
    // leaves fractional occupations; bjacobi_density_kernel then forms this
    // problem's density from eigenvectors (whole orbitals occupied).
    if (threadIdx.x == 0) b.psp2fail[p] = fabs(t1 - t2) >= 1e-8 ? 1 : 0;
    double *P = Pout + o;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const int r = e / n, c = e - r * n;
        P[e] = 0.5 * (cur[e] + cur[(size_t)c * n + r]);
    }
}

// Eigenvector fallback of the purification, one workgroup per problem whose
// bsp2_kernel flagged it: cyclic two-sided Jacobi on sym(F') with the
// round-robin pairing of jacobi_eigh (N/2 disjoint rotations per round, each
// 2x2 block of A and each (row, pair) of V updated in place by one thread),
// sweeps until the off-diagonal norm is below 1e-15 of the total, then
// P = sum over the nocc lowest eigenvalues of v v^T (ties: lower index
// first), the aufbau density of the reference (uscf.f90:177-184).  A and V
// live in dynamic LDS when 2 nmax^2 doubles fit (lds_mats), else in the
// problem's slots of Ag and Vg; the rotations, eigenvalues and occupations
// always in LDS.
__device__ __forceinline__ void brr_pair(int N, int r, int k, int &p, int &q) {
    if (k == 0) { p = r; q = N - 1; }
    else { p = (r + k) % (N - 1); q = (r - k + (N - 1)) % (N - 1); }
    if (p > q) { const int t = p; p = q; q = t; }
}
__global__ void __launch_bounds__(256) bjacobi_density_kernel(BatchDev b, const double *Fp, double *Ag, double *Vg,
                                                              double *Pout, int lds_mats) {
    extern __shared__ double jl[];
    __shared__ double sh[4];
    const int p = blockIdx.x, n = b.pn[p];
    if (!b.mactive[b.pmol[p]] || !b.psp2fail[p]) return;
    const size_t o = b.poff[p];
    const int N = n + (n & 1), m = N / 2, Nmax = b.nmax + (b.nmax & 1);
    const size_t nn = (size_t)b.nmax * b.nmax;
    double *A = lds_mats ? jl : Ag + o, *V = lds_mats ? jl + nn : Vg + o;
    double *cs = jl + (lds_mats ? 2 * nn : 0), *w = cs + Nmax, *occ = w + Nmax;
    const double *F = Fp + o;
    double tot = 0.0;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const int r = e / n, c = e - r * n;
        const double v = 0.5 * (F[e] + F[(size_t)c * n + r]);
        A[e] = v;
        V[e] = r == c ? 1.0 : 0.0;
        tot += v * v;
    }
    tot = block_sum256(tot, sh);  // also orders the writes above before the reads below
    for (int sweep = 0; sweep < 60 && n > 1; ++sweep) {
        double off = 0.0;
        for (int e = threadIdx.x; e < n * n; e += 256)
            if (e / n != e % n) off += A[e] * A[e];
        off = block_sum256(off, sh);
        if (off <= 1e-30 * tot) break;
        for (int r = 0; r < N - 1; ++r) {
            for (int k = threadIdx.x; k < m; k += 256) {
                int pp, qq;
                brr_pair(N, r, k, pp, qq);
                double c = 1.0, s = 0.0;
                if (qq < n) {
                    const double apq = A[(size_t)pp * n + qq];
                    if (apq != 0.0) {
                        const double th = (A[(size_t)qq * n + qq] - A[(size_t)pp * n + pp]) / (2.0 * apq);
                        const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                        c = 1.0 / sqrt(t * t + 1.0);
                        s = t * c;
                    }
                }
                cs[k] = c;
                cs[m + k] = s;
            }
            __syncthreads();
            for (int t = threadIdx.x; t < m * m; t += 256) {  // A <- G^T A G, block (k1, k2)
                const int k1 = t / m, k2 = t - k1 * m;
                int p1, q1, p2, q2;
                brr_pair(N, r, k1, p1, q1);
                brr_pair(N, r, k2, p2, q2);
                const double c1 = cs[k1], s1 = cs[m + k1], c2 = cs[k2], s2 = cs[m + k2];
                const bool q1ok = q1 < n, q2ok = q2 < n;
                const double a_pp = A[(size_t)p1 * n + p2];
                const double a_pq = q2ok ? A[(size_t)p1 * n + q2] : 0.0;
                const double a_qp = q1ok ? A[(size_t)q1 * n + p2] : 0.0;
                const double a_qq = (q1ok && q2ok) ? A[(size_t)q1 * n + q2] : 0.0;
                const double b_pp = c2 * a_pp - s2 * a_pq, b_pq = s2 * a_pp + c2 * a_pq;
                const double b_qp = c2 * a_qp - s2 * a_qq, b_qq = s2 * a_qp + c2 * a_qq;
                double o_pp = c1 * b_pp - s1 * b_qp, o_pq = c1 * b_pq - s1 * b_qq;
                double o_qp = s1 * b_pp + c1 * b_qp, o_qq = s1 * b_pq + c1 * b_qq;
                if (k1 == k2) { o_pq = 0.0; o_qp = 0.0; }
                A[(size_t)p1 * n + p2] = o_pp;
                if (q2ok) A[(size_t)p1 * n + q2] = o_pq;
                if (q1ok) A[(size_t)q1 * n + p2] = o_qp;
                if (q1ok && q2ok) A[(size_t)q1 * n + q2] = o_qq;
            }
            for (int t = threadIdx.x; t < n * m; t += 256) {  // V <- V G, (row, pair)
                const int row = t / m, k = t - row * m;
                int pp, qq;
                brr_pair(N, r, k, pp, qq);
                const double c = cs[k], s = cs[m + k];
                const double vp = V[(size_t)row * n + pp], vq = qq < n ? V[(size_t)row * n + qq] : 0.0;
                V[(size_t)row * n + pp] = c * vp - s * vq;
                if (qq < n) V[(size_t)row * n + qq] = s * vp + c * vq;
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += 256) w[i] = A[(size_t)i * n + i];
    __syncthreads();
    const double nocc = b.pnocc[p];
    for (int k = threadIdx.x; k < n; k += 256) {
        int rank = 0;
        for (int j = 0; j < n; ++j) rank += (w[j] < w[k] || (w[j] == w[k] && j < k)) ? 1 : 0;
        occ[k] = rank < nocc ? 1.0 : 0.0;
    }
    __syncthreads();
    double *P = Pout + o;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const int r = e / n, c = e - r * n;
        double v = 0.0;
        for (int k = 0; k < n; ++k) v += occ[k] * V[(size_t)r * n + k] * V[(size_t)c * n + k];
        P[e] = v;
    }
    if (threadIdx.x == 0) atomicAdd(b.nfallback, 1);
}

// Gershgorin discs of problem matrices (the orthogonaliser's molecule view):
// rows and bounds into the per-problem state at psp2.
__global__ void __launch_bounds__(256) bgersh_rows_kernel(BatchDev b, const double *A, double *st) {
    const int p = blockIdx.x, n = b.pn[p];
    if (!b.mactive[b.pmol[p]]) return;
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    const double *Ap = A + b.poff[p];
    double r = 0.0;
    for (int j = lane; j < n; j += 64)
        if (j != i) r += fabs(Ap[(size_t)i * n + j]);
    for (int o = 32; o >= 1; o >>= 1) r += __shfl_xor(r, o, 64);
    if (lane == 0) {
        double *rows = st + b.psp2[p] + 8 + 6 * (size_t)ntile(n);
        const double d = Ap[(size_t)i * n + i];
        rows[i] = d - r;
        rows[n + i] = d + r;
    }
}
__global__ void __launch_bounds__(256) bgersh_bounds_kernel(BatchDev b, double *st) {
    __shared__ double lo_s[4], hi_s[4];
    const int p = blockIdx.x, n = b.pn[p];
    if (!b.mactive[b.pmol[p]]) return;
    double *s = st + b.psp2[p];
    const double *rows = s + 8 + 6 * (size_t)ntile(n);
    double lo = 1e300, hi = -1e300;
    for (int i = threadIdx.x; i < n; i += 256) { lo = fmin(lo, rows[i]); hi = fmax(hi, rows[n + i]); }
    for (int o = 32; o >= 1; o >>= 1) { lo = fmin(lo, __shfl_xor(lo, o, 64)); hi = fmax(hi, __shfl_xor(hi, o, 64)); }
    if ((threadIdx.x & 63) == 0) { lo_s[threadIdx.x >> 6] = lo; hi_s[threadIdx.x >> 6] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < 4; ++w) { lo = fmin(lo, lo_s[w]); hi = fmax(hi, hi_s[w]); }
        s[0] = lo;
        s[1] = hi;
    }
}

// Energy, convergence and freezing, one workgroup per molecule (rscf.f90:142-156
// stopping rule: |dE| < eps and ||dD||_F < eps).  Mixed form (mprec): FP64
// opening (2) until ||dD||_F < mixed_open, FP32 increments (0) until
// ||dD||_F < mixed_switch, then one full FP64 build (1) and, with finish32,
// FP32 increments again (3), else full FP64 builds; only the phase after the
// switch may converge.  restart: DIIS starts a new subspace at the switch.
struct ConvArgs {
    const double *D, *Dold, *F, *H, *enuc;
    double *E, *Eold, *iters, *conv;
    int *count;
    double eps, mixed_switch, mixed_open;
    int max_iter, finish32, restart;
};
__global__ void __launch_bounds__(256) benergy_kernel(BatchDev b, ConvArgs a) {
    __shared__ double sh[4];
    const int m = blockIdx.x;
    if (!b.mactive[m]) return;
    const int n = b.mn[m], uhf = b.muhf[m], p0 = b.mprob[m];
    const double *H = a.H + b.moff[m];
    double e = 0.0, d2 = 0.0;
    for (int sp = 0; sp <= uhf; ++sp) {
        const size_t o = b.poff[p0 + sp];
        for (int i = threadIdx.x; i < n * n; i += 256) {
            const double d = a.D[o + i], dd = d - a.Dold[o + i];
            e += uhf ? 0.5 * d * (H[i] + a.F[o + i]) : d * (H[i] + a.F[o + i]);
            d2 += dd * dd;
        }
    }
    e = block_sum256(e, sh);
    d2 = block_sum256(d2, sh);
    if (threadIdx.x == 0) {
        const double eel = uhf ? e : 0.5 * e;
        const double dE = fabs(eel - a.Eold[m]), errD = sqrt(d2);
        const double it = a.iters[m] + 1.0;
        a.iters[m] = it;
        a.Eold[m] = eel;
        a.E[m] = eel + a.enuc[m];
        bool active = true;
        if (b.mprec[m] == 2 && errD < a.mixed_open) b.mprec[m] = 0;  // FP64 opening done: FP32 phase
        if (b.mprec[m] == 0) {
            if (errD < a.mixed_switch) {
                b.mprec[m] = 1;
                if (a.restart) { b.mnstore[m] = 0; b.mhead[m] = 0; }
            }
        } else if (dE < a.eps && errD < a.eps) {
            a.conv[m] = 1.0;
            active = false;
        } else if (b.mprec[m] == 1 && a.finish32) {
            b.mprec[m] = 3;  // after the switch's full FP64 build
        }
        if (it >= a.max_iter) active = false;
        if (!active) b.mactive[m] = 0;
        else atomicAdd(a.count, 1);
    }
}

// ---- molecule-level helpers for the orthogonaliser (problems = molecules)
__global__ void __launch_bounds__(256) ns_init_batch(BatchDev b, const double *S, const double *st, double *Y,
                                                     double *Z, double *I) {
    PROBLEM_LOOP(b) {
        const int r = e / n, c = e - r * n;
        const size_t o = b.poff[p] + e;
        Y[o] = S[o] / st[b.psp2[p] + 1];
        Z[o] = I[o] = r == c ? 1.0 : 0.0;
    }
}
__global__ void __launch_bounds__(256) ns_finish_batch(BatchDev b, const double *Z, const double *st, double *X) {
    PROBLEM_LOOP(b) X[b.poff[p] + e] = Z[b.poff[p] + e] / sqrt(st[b.psp2[p] + 1]);
}
__global__ void __launch_bounds__(256) resid_batch(BatchDev b, const double *T, double *out) {
    __shared__ double sh[4];
    const int p = blockIdx.x, n = b.pn[p];
    double mx = 0.0;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const int r = e / n, c = e - r * n;
        mx = fmax(mx, fabs(T[b.poff[p] + e] - (r == c ? 1.0 : 0.0)));
    }
    for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) out[p] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}
__global__ void fill_int_kernel(int n, int v, int *x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

dim3 pgrid(int P, int nmax) { return dim3(P, std::max(1, std::min(64, (nmax * nmax + 255) / 256))); }

// ---- autotune cache: {"key": value, ...} in a small JSON file
std::string tune_path() {
    if (const char *e = std::getenv("PYFACC_TUNE_CACHE")) return e;
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&tune_path), &info) && info.dli_fname) {
        std::string f = info.dli_fname;
        const size_t s = f.find_last_of('/');
        return (s == std::string::npos ? std::string(".") : f.substr(0, s)) + "/autotune.json";
    }
    return "autotune.json";
}
std::map<std::string, int> tune_read() {
    std::map<std::string, int> m;
    std::ifstream in(tune_path());
    if (!in) return m;
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string t = ss.str();
    size_t i = 0;
    while ((i = t.find('"', i)) != std::string::npos) {
        const size_t j = t.find('"', i + 1);
        if (j == std::string::npos) break;
        const size_t c = t.find(':', j);
        if (c == std::string::npos) break;
        m[t.substr(i + 1, j - i - 1)] = std::atoi(t.c_str() + c + 1);
        i = t.find_first_of(",}", c);
        if (i == std::string::npos) break;
    }
    return m;
}
void tune_write(const std::map<std::string, int> &m) {
    std::ofstream out(tune_path());
    if (!out) return;  // read-only location: the choice is simply not cached
    out << "{";
    bool first = true;
    for (auto &kv : m) {
        out << (first ? "" : ", ") << "\"" << kv.first << "\": " << kv.second;
        first = false;
    }
    out << "}\n";
}

}  // namespace

BatchEngine::BatchEngine(std::vector<BatchMolecule> mols, const std::string &basis_path, bool renormalise,
                         double schwarz)
    : mols_(std::move(mols)), basis_path_(basis_path), renorm_(renormalise), schwarz_(schwarz) {
    PF_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
}

BatchEngine::~BatchEngine() {
    if (st_) {
        (void)hipStreamSynchronize(st_);  // members go back to the block cache: nothing may be pending
        (void)hipStreamDestroy(st_);
    }
}

void BatchEngine::setup() {
    double tp = now();
    auto phase = [&](const char *what) {  // verbose: host-side phase times
        if (!verbose_) return;
        PF_HIP(hipStreamSynchronize(st_));
        const double t = now();
        std::fprintf(stderr, "batch setup: %-28s %8.3f s\n", what, t - tp);
        tp = t;
    };
    M_ = (int)mols_.size();
    if (M_ == 0) throw Error("batch: no molecules");
    const auto blocks_p = basis_blocks_cached(basis_path_);
    const BasisBlocks &blocks = *blocks_p;
    std::vector<int> shoff(M_ + 1, 0), aoff(M_ + 1, 0);
    std::vector<double> axyz;
    std::vector<int> Z;
    // each molecule's basis and nuclear repulsion on host threads
    std::vector<Basis> bases(M_);
    enuc_.resize(M_);
    parallel_ranges(M_, host_threads(M_, 128), [&](int, int m0, int m1) {
        for (int m = m0; m < m1; ++m) {
            bases[m] = build_basis_blocks(blocks, mols_[m].species, mols_[m].mol, renorm_);
            enuc_[m] = nuclear_repulsion(mols_[m].mol);
        }
    });
    for (int m = 0; m < M_; ++m) {
        const auto &bm = mols_[m];
        const Basis &bs = bases[m];
        for (const auto &sh : bs.shells)
            if (sh.l >= 2) throw Error("batch: molecule " + std::to_string(m) + ": s and p shells only (d shells: pf_scf)");
        for (auto &s : bs.shells) shells_.push_back(s);
        shoff[m + 1] = (int)shells_.size();
        aoff[m + 1] = aoff[m] + bm.mol.natoms;
        axyz.insert(axyz.end(), bm.mol.xyz.begin(), bm.mol.xyz.end());
        Z.insert(Z.end(), bm.mol.Z.begin(), bm.mol.Z.end());
        const int n = bs.nbf, uhf = bm.nalpha != bm.nbeta;
        if (bm.nalpha < 1 || bm.nbeta < 1 || bm.nalpha > n || bm.nbeta > n)
            throw Error("batch: molecule " + std::to_string(m) +
                        ": each spin needs between 1 and nbf electrons");
        mn_.push_back(n);
        muhf_.push_back(uhf);
        mprob_.push_back((int)pn_.size());
        nmax_ = std::max(nmax_, n);
        for (int s = 0; s <= uhf; ++s) { pn_.push_back(n); pmol_.push_back(m); }
    }
    shoff_ = shoff;
    P_ = (int)pn_.size();
    moff_.resize(M_);
    poff_.resize(P_);
    for (int m = 0; m < M_; ++m) { moff_[m] = mtot_; mtot_ += (long long)mn_[m] * mn_[m]; }
    std::vector<long long> pmoff(P_), pdiis(P_);
    std::vector<double> poccf(P_), pnocc(P_);
    long long dtot = 0;
    for (int p = 0; p < P_; ++p) {
        const int m = pmol_[p], n = pn_[p], spin = p - mprob_[m];
        poff_[p] = ptot_;
        ptot_ += (long long)n * n;
        pmoff[p] = moff_[m];
        pdiis[p] = dtot;
        dtot += (long long)ndiis_ * n * n;
        poccf[p] = muhf_[m] ? 1.0 : 2.0;
        pnocc[p] = spin == 0 ? mols_[m].nalpha : mols_[m].nbeta;
    }
    // device state
    const std::vector<double> tab = boys_table_host();
    tab_.upload(tab.data(), tab.size(), st_);
    d_mn_.upload(mn_.data(), M_, st_); d_muhf_.upload(muhf_.data(), M_, st_); d_mprob_.upload(mprob_.data(), M_, st_);
    d_pn_.upload(pn_.data(), P_, st_); d_pmol_.upload(pmol_.data(), P_, st_);
    d_moff_.upload(moff_.data(), M_, st_); d_poff_.upload(poff_.data(), P_, st_);
    d_pmoff_.upload(pmoff.data(), P_, st_); d_pdiis_.upload(pdiis.data(), P_, st_);
    d_poccf_.upload(poccf.data(), P_, st_); d_pnocc_.upload(pnocc.data(), P_, st_);
    d_enuc_.upload(enuc_.data(), M_, st_);
    d_mactive_.alloc(M_); d_mprec_.alloc(M_); d_mhead_.alloc(M_); d_mnstore_.alloc(M_); d_muse_.alloc(M_);
    d_count_.alloc(1);
    d_psp2fail_.alloc(P_);
    d_nfallback_.alloc(1);
    fill_int_kernel<<<(P_ + 255) / 256, 256, 0, st_>>>(P_, 0, d_psp2fail_.p);
    fill_int_kernel<<<1, 64, 0, st_>>>(1, 0, d_nfallback_.p);
    d_E_.alloc(M_); d_Eold_.alloc(M_); d_iters_.alloc(M_); d_conv_.alloc(M_);
    d_coef_.alloc((size_t)M_ * ndiis_); d_B_.alloc((size_t)M_ * ndiis_ * ndiis_);
    S_.alloc(mtot_); H_.alloc(mtot_); X_.alloc(mtot_); Dtot_.alloc(mtot_);
    for (auto *b : {&D_, &Dold_, &F_, &Fx_, &Fp_, &W1_, &W2_, &Pr_}) b->alloc(ptot_);
    dF_.alloc(std::max(dtot, 1LL)); dE_.alloc(std::max(dtot, 1LL));
    dev_.M = M_; dev_.P = P_; dev_.nmax = nmax_; dev_.ndiis = ndiis_;
    dev_.mn = d_mn_.p; dev_.muhf = d_muhf_.p; dev_.mprob = d_mprob_.p; dev_.moff = d_moff_.p;
    dev_.pn = d_pn_.p; dev_.pmol = d_pmol_.p; dev_.poccf = d_poccf_.p; dev_.pnocc = d_pnocc_.p;
    dev_.poff = d_poff_.p; dev_.pmoff = d_pmoff_.p; dev_.pdiis = d_pdiis_.p;
    dev_.mactive = d_mactive_.p; dev_.mprec = d_mprec_.p; dev_.mhead = d_mhead_.p; dev_.mnstore = d_mnstore_.p;
    dev_.muse = d_muse_.p; dev_.mcoef = d_coef_.p; dev_.mB = d_B_.p;
    dev_.psp2fail = d_psp2fail_.p; dev_.nfallback = d_nfallback_.p;
    {
        // the fallback's matrices in LDS when they fit beside its small arrays
        const int Nmax = nmax_ + (nmax_ & 1);
        const size_t small = sizeof(double) * 3 * (size_t)Nmax;
        const size_t mats = sizeof(double) * 2 * (size_t)nmax_ * nmax_;
        int maxlds = 0, dev = 0;
        PF_HIP(hipGetDevice(&dev));
        PF_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const size_t budget = (size_t)maxlds - 1024;  // static sh[4] and alignment
        jacobi_mats_in_lds_ = small + mats <= budget;
        const size_t dyn = jacobi_mats_in_lds_ ? small + mats : small;
        if (dyn > budget) throw Error("batch: molecules too large for the purification fallback");
        jacobi_lds_ = dyn;
        if (dyn > 65536)
            PF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(bjacobi_density_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, maxlds - 1024));
    }
    fill_int_kernel<<<(M_ + 255) / 256, 256, 0, st_>>>(M_, 1, d_mactive_.p);

    phase("basis sets and offsets");
    // one-electron integrals of every molecule in one launch
    d_shells_.upload(shells_.data(), shells_.size(), st_);
    {
        DevBuf<double> dxyz, T(mtot_), V(mtot_);
        DevBuf<int> dZ, dsh, dao;
        dxyz.upload(axyz.data(), axyz.size(), st_);
        dZ.upload(Z.data(), Z.size(), st_);
        dsh.upload(shoff.data(), shoff.size(), st_);
        dao.upload(aoff.data(), aoff.size(), st_);
        int max_nsh = 0;
        for (int m = 0; m < M_; ++m) max_nsh = std::max(max_nsh, shoff[m + 1] - shoff[m]);
        launch_onee_batch(M_, max_nsh, d_shells_.p, dsh.p, dao.p, d_mn_.p, d_moff_.p, dxyz.p, dZ.p, tab_.p, S_.p, T.p,
                          V.p, st_);
        dev_axpby(mtot_, 1.0, T.p, 1.0, V.p, H_.p, st_);
        PF_HIP(hipStreamSynchronize(st_));
    }

    phase("device state, 1e integrals");
    // shell pairs per class, all molecules concatenated; pass 1 gives Schwarz bounds
    std::array<std::vector<std::pair<int, int>>, 3> all;
    std::array<std::vector<int>, 3> beg1;
    for (int c = 0; c < 3; ++c) beg1[c].assign(1, 0);
    for (int m = 0; m < M_; ++m) {
        for (int a = shoff[m]; a < shoff[m + 1]; ++a)
            for (int b = shoff[m]; b <= a; ++b) {
                int f = a, s = b;
                if (shells_[b].l > shells_[a].l) std::swap(f, s);
                all[shells_[f].l + shells_[s].l].push_back({f, s});
            }
        for (int c = 0; c < 3; ++c) beg1[c].push_back((int)all[c].size());
    }
    const int cls_l[3][2] = {{0, 0}, {1, 0}, {1, 1}};
    double bound_max = 0.0;
    std::array<std::vector<double>, 3> Q1;
    for (int c = 0; c < 3; ++c) {
        PairSet &ps = pairs_[c];
        ps.la = cls_l[c][0]; ps.lb = cls_l[c][1];
        fill_pairset(ps, shells_, d_shells_.p, all[c], -1.0, bound_max, false, st_);
        ps.d_Q.alloc(std::max(ps.n, 1));
        if (ps.n) launch_schwarz<double>(ps.la, ps.lb, ps.view64(), tab_.p, ps.d_Q.p, st_);
        Q1[c].resize(ps.n);
        ps.d_Q.download(Q1[c].data(), ps.n, st_);
    }
    PF_HIP(hipStreamSynchronize(st_));
    phase("pairs pass 1 + Schwarz");
    // pass 2: per molecule keep Q * Qmax(molecule) >= tau, sorted by Q
    // descending; molecules in contiguous ranges on host threads, the ranges'
    // lists concatenated in molecule order
    std::array<std::vector<double>, 3> Q;
    std::array<std::vector<std::pair<int, int>>, 3> kept;
    {
        const int T = host_threads(M_, 128);
        std::vector<std::array<std::vector<double>, 3>> Qt(T);
        std::vector<std::array<std::vector<std::pair<int, int>>, 3>> kt(T);
        std::vector<std::array<std::vector<int>, 3>> nt(T);  // pairs kept per molecule
        parallel_ranges(M_, T, [&](int t, int m0, int m1) {
            std::vector<int> idx;
            for (int m = m0; m < m1; ++m) {
                double qmax = 0.0;
                for (int c = 0; c < 3; ++c)
                    for (int i = beg1[c][m]; i < beg1[c][m + 1]; ++i) qmax = std::max(qmax, Q1[c][i]);
                for (int c = 0; c < 3; ++c) {
                    idx.clear();
                    for (int i = beg1[c][m]; i < beg1[c][m + 1]; ++i)
                        if (Q1[c][i] * qmax >= schwarz_) idx.push_back(i);
                    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return Q1[c][x] > Q1[c][y]; });
                    for (int i : idx) { kt[t][c].push_back(all[c][i]); Qt[t][c].push_back(Q1[c][i]); }
                    nt[t][c].push_back((int)idx.size());
                }
            }
        });
        for (int c = 0; c < 3; ++c) {
            pbeg_[c].assign(1, 0);
            for (int t = 0; t < T; ++t) {
                kept[c].insert(kept[c].end(), kt[t][c].begin(), kt[t][c].end());
                Q[c].insert(Q[c].end(), Qt[t][c].begin(), Qt[t][c].end());
                for (int k : nt[t][c]) pbeg_[c].push_back(pbeg_[c].back() + k);
            }
        }
    }
    phase("pairs pass 2, host selection");
    for (int c = 0; c < 3; ++c) {
        fill_pairset(pairs_[c], shells_, d_shells_.p, kept[c], bound_max, bound_max, false, st_);
        pairs_[c].d_Q.upload(Q[c].data(), Q[c].size(), st_);
    }

    phase("pairs pass 2");
    // quartet plan per class (bra class x >= ket class y) and the integral store
    PF_HIP(hipStreamSynchronize(st_));
    const double t0 = now();
    auto &work = work_;  // stored values per bra pair
    auto &brq = brq_;    // per class: bra range of each molecule
    double t_plan = 0.0;
    int q = 0;
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y <= x; ++y, ++q) {
            const PairSet &B = pairs_[x], &K = pairs_[y];
            const bool same = x == y;
            const int nint = nc(B.la) * nc(B.lb) * nc(K.la) * nc(K.lb);
            const double tq = now();
            std::vector<long long> off(std::max(B.n, 1), 0);
            work[q].assign(std::max(B.n, 1), 0.0);
            std::vector<int> nk(std::max(B.n, 1), 0), kb(std::max(B.n, 1), 0), br(2 * M_);
            // molecules in contiguous ranges on host threads: offsets within
            // the range first, then shifted by the ranges before it
            const int T = host_threads(M_, 128);
            std::vector<long long> tot_t(T + 1, 0);
            std::vector<double> vals_t(T, 0.0);
            parallel_ranges(M_, T, [&](int t, int m0, int m1) {
                long long tot = 0;
                for (int m = m0; m < m1; ++m) {
                    br[2 * m] = pbeg_[x][m];
                    br[2 * m + 1] = pbeg_[x][m + 1];
                    const int k0 = pbeg_[y][m], k1 = pbeg_[y][m + 1];
                    for (int i = pbeg_[x][m]; i < pbeg_[x][m + 1]; ++i) {
                        const double thr = schwarz_ / Q[x][i];
                        int lo = k0, hi = k1;  // kets sorted by Q descending: prefix with Q >= thr
                        while (lo < hi) {
                            const int mid = (lo + hi) / 2;
                            if (Q[y][mid] >= thr) lo = mid + 1; else hi = mid;
                        }
                        int n = lo - k0;
                        if (same) n = std::min(n, i - k0 + 1);
                        nk[i] = n;
                        kb[i] = k0;
                        off[i] = tot;
                        tot += (long long)round64(n) * nint;
                        vals_t[t] += (double)n * nint;
                        work[q][i] = (double)round64(n) * nint;
                    }
                }
                tot_t[t + 1] = tot;
            });
            for (int t = 0; t < T; ++t) {
                tot_t[t + 1] += tot_t[t];
                eri_values_ += vals_t[t];
            }
            parallel_ranges(M_, T, [&](int t, int m0, int m1) {
                if (m0 == m1 || t == 0) return;
                for (int i = pbeg_[x][m0]; i < pbeg_[x][m1]; ++i) off[i] += tot_t[t];
            });
            const long long tot = tot_t[T];
            t_plan += now() - tq;
            s64_[q].alloc(std::max(tot, 1LL));
            s32_[q].alloc(std::max(tot, 1LL));
            soff_[q].upload(off.data(), off.size(), st_);
            snket_[q].upload(nk.data(), nk.size(), st_);
            skbase_[q].upload(kb.data(), kb.size(), st_);
            sbr_[q].upload(br.data(), br.size(), st_);
            brq[q] = br;
            if (B.n) {
                JKArgs<double> a;
                a.bra = B.view64(); a.ket = K.view64(); a.same = same; a.tau = schwarz_;
                a.nbf = nmax_; a.nsh = 0; a.tab = tab_.p;
                a.store = s64_[q].p; a.store32 = s32_[q].p;
                a.store_off = soff_[q].p; a.store_nket = snket_[q].p; a.store_kbase = skbase_[q].p;
                launch_jk_class<double>(B.la, B.lb, K.la, K.lb, 1, MODE_STORE, a, st_);
            }
            BatchClass &bc = cls_[q];
            bc.la = B.la; bc.lb = B.lb; bc.lc = K.la; bc.ld = K.lb;
            bc.sa = B.d_sa.p; bc.sb = B.d_sb.p; bc.oa = B.d_oa.p; bc.ob = B.d_ob.p;
            bc.ksa = K.d_sa.p; bc.ksb = K.d_sb.p; bc.koa = K.d_oa.p; bc.kob = K.d_ob.p;
            bc.off = soff_[q].p; bc.nket = snket_[q].p; bc.kbase = skbase_[q].p;
            bc.v64 = s64_[q].p; bc.v32 = s32_[q].p; bc.brange = sbr_[q].p; bc.same = same;
            bc.bQ = B.d_Q.p; bc.kQ = K.d_Q.p;
            nkeff_[q].alloc(std::max(B.n, 1));
            bc.nkeff = nkeff_[q].p;
        }
    // each pair's molecule, per pair class (the FP32 increments' ket cut)
    for (int c = 0; c < 3; ++c) {
        std::vector<int> pm(std::max(pairs_[c].n, 1), 0);
        for (int m = 0; m < M_; ++m)
            for (int i = pbeg_[c][m]; i < pbeg_[c][m + 1]; ++i) pm[i] = m;
        d_pairmol_[c].upload(pm.data(), pm.size(), st_);
    }
    q = 0;
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y <= x; ++y, ++q) cls_[q].bmol = d_pairmol_[x].p;
    PF_HIP(hipStreamSynchronize(st_));
    t_eri_ = now() - t0;
    if (verbose_) std::fprintf(stderr, "batch setup: of which host quartet plan  %8.3f s\n", t_plan);
    for (int s = 0; s < 2; ++s) {  // the autotuner may rebuild them with another shape
        build_items(s, 262144.0);
        sets_[s].waves = 8;
    }
    phase("quartet plan + stored ERIs");
    batch_digest_prepare(nmax_);
    setup_orthogonalizer();
    phase("orthogonalisers");
}

// X = S^{-1/2} per molecule by the coupled Newton-Schulz iteration on batched
// MFMA GEMMs (as inv_sqrt_spd); a molecule that does not converge falls back
// to the Jacobi solver's canonical U s^{-1/2}.
void BatchEngine::setup_orthogonalizer() {
    // a molecule-level view: problems = molecules
    std::vector<int> iota(M_);
    std::iota(iota.begin(), iota.end(), 0);
    std::vector<long long> st_off(M_);
    long long stot = 0;
    for (int m = 0; m < M_; ++m) { st_off[m] = stot; stot += (long long)sp2_size(mn_[m]); }
    DevBuf<int> dio, ones;
    DevBuf<long long> dst;
    dio.upload(iota.data(), M_, st_);
    dst.upload(st_off.data(), M_, st_);
    ones.alloc(M_);
    fill_int_kernel<<<(M_ + 255) / 256, 256, 0, st_>>>(M_, 1, ones.p);
    BatchDev v = dev_;
    v.P = M_; v.pn = d_mn_.p; v.pmol = dio.p; v.poff = d_moff_.p; v.pmoff = d_moff_.p; v.psp2 = dst.p;
    v.mactive = ones.p;
    DevBuf<double> Y(mtot_), Z(mtot_), T(mtot_), Y2(mtot_), Z2(mtot_), I(mtot_), stb(stot), res(M_);
    const dim3 g = pgrid(M_, nmax_), gt(M_, ntile(nmax_)), gr(M_, (nmax_ + 3) / 4);
    bgersh_rows_kernel<<<gr, 256, 0, st_>>>(v, S_.p, stb.p);
    bgersh_bounds_kernel<<<M_, 256, 0, st_>>>(v, stb.p);
    ns_init_batch<<<g, 256, 0, st_>>>(v, S_.p, stb.p, Y.p, Z.p, I.p);
    auto gemm = [&](const double *A, const double *B, double *C, double alpha, double beta, const double *Cin) {
        BGemm b{A, B, Cin, C, d_moff_.p, d_moff_.p, d_moff_.p, d_moff_.p, nullptr, 0, 0, alpha, beta};
        bgemm_kernel<<<gt, 256, 0, st_>>>(v, b);
    };
    double *y = Y.p, *z = Z.p, *y2 = Y2.p, *z2 = Z2.p;
    std::vector<double> hr(M_);
    bool ok = false;
    for (int round = 0; round < 6 && !ok; ++round) {
        for (int k = 0; k < 8; ++k) {
            gemm(z, y, T.p, -0.5, 1.5, I.p);  // (3I - ZY)/2
            gemm(y, T.p, y2, 1.0, 0.0, nullptr);
            gemm(T.p, z, z2, 1.0, 0.0, nullptr);
            std::swap(y, y2);
            std::swap(z, z2);
        }
        gemm(z, y, T.p, 1.0, 0.0, nullptr);
        resid_batch<<<M_, 256, 0, st_>>>(v, T.p, res.p);
        res.download(hr.data(), M_, st_);
        PF_HIP(hipStreamSynchronize(st_));
        ok = true;
        for (double r : hr) ok = ok && std::isfinite(r) && r < 1e-11;
    }
    ns_finish_batch<<<g, 256, 0, st_>>>(v, z, stb.p, X_.p);
    PF_HIP(hipStreamSynchronize(st_));
    for (int m = 0; m < M_; ++m) {
        if (std::isfinite(hr[m]) && hr[m] < 1e-11) continue;
        const int n = mn_[m];
        DevBuf<double> s(n), U((size_t)n * n);
        jacobi_eigh(n, S_.p + moff_[m], s.p, U.p, st_);
        std::vector<double> hs(n), hU((size_t)n * n);
        s.download(hs.data(), n, st_);
        U.download(hU.data(), hU.size(), st_);
        PF_HIP(hipStreamSynchronize(st_));
        if (hs[0] <= 0.0) throw Error("batch: overlap matrix of molecule " + std::to_string(m) + " is not positive definite");
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) hU[(size_t)r * n + c] /= std::sqrt(hs[c]);
        PF_HIP(hipMemcpyAsync(X_.p + moff_[m], hU.data(), sizeof(double) * hU.size(), hipMemcpyHostToDevice, st_));
        PF_HIP(hipStreamSynchronize(st_));
    }
}

// D_p = occ_p X P X^T with P the projector of F'_p = X^T F_p X (SP2), and D_alpha + D_beta.
void BatchEngine::density_step(const double *F) {
    const dim3 g = pgrid(P_, nmax_), gt(P_, ntile(nmax_));
    auto gemm = [&](const double *A, const long long *oA, int ta, const double *B, const long long *oB, int tb,
                    double *C, const double *ascale) {
        BGemm b{A, B, nullptr, C, oA, oB, d_poff_.p, d_poff_.p, ascale, ta, tb, 1.0, 0.0};
        bgemm_kernel<<<gt, 256, 0, st_>>>(dev_, b);
    };
    gemm(F, d_poff_.p, 0, X_.p, d_pmoff_.p, 0, W1_.p, nullptr);      // F X
    gemm(X_.p, d_pmoff_.p, 1, W1_.p, d_poff_.p, 0, Fp_.p, nullptr);  // X^T F X
    bsp2_kernel<<<P_, 256, 0, st_>>>(dev_, Fp_.p, Pr_.p, W2_.p, W1_.p, SP2_MAX_STEPS);  // P into W1
    // problems whose purification failed: aufbau density from eigenvectors, into W1
    bjacobi_density_kernel<<<P_, 256, jacobi_lds_, st_>>>(dev_, Fp_.p, Pr_.p, W2_.p, W1_.p,
                                                          jacobi_mats_in_lds_ ? 1 : 0);
    gemm(X_.p, d_pmoff_.p, 0, W1_.p, d_poff_.p, 0, W2_.p, nullptr);              // X P
    gemm(W2_.p, d_poff_.p, 0, X_.p, d_pmoff_.p, 1, D_.p, d_poccf_.p);            // occ X P X^T
    bdtot_kernel<<<g, 256, 0, st_>>>(dev_, D_.p, Dtot_.p);
}

void BatchEngine::build_items(int set, double item_values) {
    ItemSet &S = sets_[set];
    if (item_values == S.values) return;
    S.values = item_values;
    const auto &work = work_;
    const auto &brq = brq_;
    std::vector<BatchItem> items;
    std::vector<double> iwork;
    std::vector<int> msplit(M_, 0);
    std::vector<long long> goff(M_);
    long long gtot = 0;
    for (int m = 0; m < M_; ++m) {
        goff[m] = gtot;
        gtot += 3LL * mn_[m] * (mn_[m] + 1) / 2;
        double wm = 0.0;
        for (int c = 0; c < 6; ++c)
            for (int i = brq[c][2 * m]; i < brq[c][2 * m + 1]; ++i) wm += work[c][i];
        const int G = std::max(1, (int)std::ceil(wm / item_values));
        msplit[m] = G > 1;
        int sq = 0, si = brq[0][2 * m], g = 0;
        double cum = 0.0, icum = 0.0;
        for (int c = 0; c < 6 && G > 1; ++c)
            for (int i = brq[c][2 * m]; i < brq[c][2 * m + 1]; ++i) {
                cum += work[c][i];
                icum += work[c][i];
                if (g < G - 1 && cum >= wm * (g + 1) / G) {
                    items.push_back({m, sq, si, c, i + 1, 1});
                    iwork.push_back(icum);
                    sq = c; si = i + 1; icum = 0.0; ++g;
                }
            }
        items.push_back({m, sq, si, 5, brq[5][2 * m + 1], G > 1 ? 1 : 0});
        iwork.push_back(G > 1 ? icum : wm);
    }
    std::vector<int> order(items.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return iwork[x] > iwork[y]; });
    std::vector<BatchItem> sorted;
    for (int k : order) sorted.push_back(items[k]);
    S.n = (int)sorted.size();
    S.items.upload(sorted.data(), sorted.size(), st_);
    S.msplit.upload(msplit.data(), M_, st_);
    if (!gacc_.p) {
        d_gacc_off_.upload(goff.data(), M_, st_);
        gacc_.alloc(std::max(gtot, 1LL));
        dev_fill(gacc_.n, 0.0, gacc_.p, st_);
        dev_.gacc_off = d_gacc_off_.p;
    }
    PF_HIP(hipStreamSynchronize(st_));
}

void BatchEngine::digest(int set) {
    const ItemSet &S = sets_[set];
    std::array<int, 6> nbra{};
    int q = 0;
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y <= x; ++y, ++q) nbra[q] = pairs_[x].n;
    launch_batch_digest(dev_, cls_, S.items.p, S.n, S.msplit.p, D_.p, Dtot_.p, H_.p, F_.p, gacc_.p, S.waves, st_,
                        mixed_ ? G_.p : nullptr, Db_.p, dD32_.p, dDt32_.p, dmax_.p, schwarz_, schwarz_inc_, nbra);
}

// One SCF iteration for every active molecule (no host decision inside).
void BatchEngine::iteration(const pf_options &opt) {
    const dim3 g = pgrid(P_, nmax_), gt(P_, ntile(nmax_));
    fill_int_kernel<<<1, 64, 0, st_>>>(1, 0, d_count_.p);
    digest(cur_set_);
    bcopy_kernel<<<g, 256, 0, st_>>>(dev_, D_.p, d_poff_.p, Dold_.p);
    if (opt.diis > 1) {
        BGemm a{F_.p, D_.p, nullptr, W1_.p, d_poff_.p, d_poff_.p, d_poff_.p, d_poff_.p, nullptr, 0, 0, 1.0, 0.0};
        bgemm_kernel<<<gt, 256, 0, st_>>>(dev_, a);  // F D
        BGemm s{W1_.p, S_.p, nullptr, W2_.p, d_poff_.p, d_pmoff_.p, d_poff_.p, d_poff_.p, nullptr, 0, 0, 1.0, 0.0};
        bgemm_kernel<<<gt, 256, 0, st_>>>(dev_, s);  // F D S
        bdiis_insert_kernel<<<g, 256, 0, st_>>>(dev_, W2_.p, F_.p, dE_.p, dF_.p);
        bdiis_dots_kernel<<<dim3(M_, ndiis_), 256, 0, st_>>>(dev_, dE_.p);
        bdiis_solve_kernel<<<(M_ + 63) / 64, 64, 0, st_>>>(dev_);
        bdiis_extrap_kernel<<<g, 256, 0, st_>>>(dev_, F_.p, dF_.p, Fx_.p);
        density_step(Fx_.p);
    } else {
        density_step(F_.p);
    }
    ConvArgs c{D_.p, Dold_.p, F_.p, H_.p, d_enuc_.p, d_E_.p, d_Eold_.p, d_iters_.p, d_conv_.p, d_count_.p,
               opt.eps, opt.mixed_switch, opt.mixed_anchor, opt.max_iter,
               opt.precision == PF_PREC_MIXED && opt.mixed_finish == PF_FINISH_FP32 ? 1 : 0,
               1};
    benergy_kernel<<<M_, 256, 0, st_>>>(dev_, c);
}

void BatchEngine::run(const pf_options &opt, bool use_graph, double *energy, int *iterations, int *converged,
                      BatchStats *stats) {
    const double t0 = now();
    if (opt.diis > 8) throw Error("batch: DIIS subspace of at most 8");
    ndiis_ = std::max(opt.diis, 2);
    verbose_ = opt.verbose != 0;
    setup();
    const double t1 = now();
    // state: frozen flags, precision, DIIS rings, energies
    // mixed form: FP64 opening (2) until ||dD||_F < mixed_anchor (none if 0), FP32 (0), FP64 (1)
    const int prec0 = opt.precision != PF_PREC_MIXED ? 1 : (opt.mixed_anchor > 0.0 ? 2 : 0);
    mixed_ = opt.precision == PF_PREC_MIXED;
    schwarz_inc_ = opt.schwarz_inc > 0.0 ? opt.schwarz_inc : schwarz_;
    if (mixed_) {
        // G and the built density start at zero: an FP32 phase from the first
        // iteration (mixed_anchor 0) builds its first increment on all of D
        G_.alloc(ptot_); Db_.alloc(ptot_); dD32_.alloc(ptot_); dDt32_.alloc(mtot_); dmax_.alloc(M_);
        dev_fill(ptot_, 0.0, G_.p, st_);
        dev_fill(ptot_, 0.0, Db_.p, st_);
    }
    fill_int_kernel<<<(M_ + 255) / 256, 256, 0, st_>>>(M_, prec0, d_mprec_.p);
    fill_int_kernel<<<(M_ + 255) / 256, 256, 0, st_>>>(M_, 0, d_mhead_.p);
    fill_int_kernel<<<(M_ + 255) / 256, 256, 0, st_>>>(M_, 0, d_mnstore_.p);
    fill_int_kernel<<<(M_ + 255) / 256, 256, 0, st_>>>(M_, 0, d_muse_.p);
    for (auto *b : {&d_E_, &d_Eold_, &d_iters_, &d_conv_}) dev_fill(M_, 0.0, b->p, st_);
    if (opt.guess == PF_GUESS_SAD) {
        // superposition of atomic densities (as Engine::scf and sad_density):
        // RHF the sum, UHF half per spin; each element's atom and its place
        // on the isolated atom are tabulated here and the blocks laid out on
        // the device
        std::map<std::string, std::pair<int, int>> spec;  // element -> (offset in tab, nbf of the atom)
        std::vector<double> tab;
        std::vector<int> mbf(M_ + 1, 0), bfi;
        for (int m = 0; m < M_; ++m) {
            const auto &bm = mols_[m];
            const int na = bm.mol.natoms, s0 = shoff_[m], s1 = shoff_[m + 1];
            std::vector<int> loc(na, 0);
            mbf[m + 1] = mbf[m] + mn_[m];
            bfi.resize(3 * (size_t)mbf[m + 1]);
            for (int s = s0; s < s1; ++s) {  // an atom's shells in the order they have on the isolated atom
                const Shell &sh = shells_[s];
                const int a = sh.atom;
                auto it = spec.find(bm.species[a]);
                if (it == spec.end()) {
                    const std::vector<double> Da = atomic_density(basis_path_, renorm_, bm.species[a], bm.mol.Z[a]);
                    const int nat = (int)std::lround(std::sqrt((double)Da.size()));
                    it = spec.emplace(bm.species[a], std::make_pair((int)tab.size(), nat)).first;
                    tab.insert(tab.end(), Da.begin(), Da.end());
                }
                const auto [toff, nat] = it->second;
                for (int k = 0; k < ncart(sh.l); ++k, ++loc[a]) {
                    if (loc[a] >= nat) throw Error("SAD guess: atomic basis mismatch for " + bm.species[a]);
                    int *f = &bfi[3 * (size_t)(mbf[m] + sh.bf + k)];
                    f[0] = a;
                    f[1] = toff + loc[a] * nat;
                    f[2] = loc[a];
                }
            }
            for (int a = 0; a < na; ++a)
                if (loc[a] != spec.at(bm.species[a]).second)
                    throw Error("SAD guess: atomic basis mismatch for " + bm.species[a]);
        }
        DevBuf<int> d_bfi, d_mbf;
        DevBuf<double> d_tab;
        d_bfi.upload(bfi.data(), bfi.size(), st_);
        d_mbf.upload(mbf.data(), mbf.size(), st_);
        d_tab.upload(tab.data(), tab.size(), st_);
        bsad_kernel<<<pgrid(P_, nmax_), 256, 0, st_>>>(dev_, d_bfi.p, d_mbf.p, d_tab.p, D_.p);
        PF_HIP(hipStreamSynchronize(st_));  // before the tables go
        bdtot_kernel<<<pgrid(P_, nmax_), 256, 0, st_>>>(dev_, D_.p, Dtot_.p);
    } else {
        // core guess: D from H (rscf.f90:113-114; uscf.f90:140-144)
        bcopy_kernel<<<pgrid(P_, nmax_), 256, 0, st_>>>(dev_, H_.p, d_pmoff_.p, Fx_.p);
        density_step(Fx_.p);
    }
    if (verbose_) {
        PF_HIP(hipStreamSynchronize(st_));
        std::fprintf(stderr, "batch start: state and guess density     %8.3f s\n", now() - t1);
    }
    // autotune the digest's two shapes -- wavefronts per workgroup x stored
    // integrals per work item -- on the real batch; cached per device and
    // batch shape.  Shape 0 is timed with every molecule active (the first
    // iterations); shape 1 with only the 2 % of molecules with the most
    // stored integrals active, the state after most molecules have
    // converged, where one large molecule in one workgroup sets the time.
    hipDeviceProp_t prop;
    int dev = 0;
    PF_HIP(hipGetDevice(&dev));
    PF_HIP(hipGetDeviceProperties(&prop, dev));
    const std::string key = std::string("batch_digest_shapes:") + prop.gcnArchName + ":" + std::to_string(M_) + ":" +
                            std::to_string(nmax_);
    constexpr int NWAVES = 2, NITEM = 4, NSHAPE = NWAVES * NITEM;
    const int waves_c[NWAVES] = {4, 8};
    const double item_c[NITEM] = {65536.0, 262144.0, 1048576.0, 1e18};  // 1e18: no molecule split
    auto set_shape = [&](int set, int c) {
        sets_[set].waves = waves_c[c / NITEM];
        build_items(set, item_c[c % NITEM]);
    };
    auto cache = tune_read();
    const int code = cache.count(key) ? cache[key] : -1;
    if (code >= 0 && code < NSHAPE * NSHAPE) {
        set_shape(0, code / NSHAPE);
        set_shape(1, code % NSHAPE);
    } else {
        hipEvent_t e0, e1;
        PF_HIP(hipEventCreate(&e0));
        PF_HIP(hipEventCreate(&e1));
        std::vector<double> wm(M_, 0.0);
        for (int c = 0; c < 6; ++c)
            for (int m = 0; m < M_; ++m)
                for (int i = brq_[c][2 * m]; i < brq_[c][2 * m + 1]; ++i) wm[m] += work_[c][i];
        std::vector<double> ws = wm;
        std::sort(ws.begin(), ws.end());
        const double wcut = ws[(size_t)(0.98 * (M_ - 1))];
        std::vector<int> tail(M_);
        for (int m = 0; m < M_; ++m) tail[m] = wm[m] >= wcut ? 1 : 0;
        DevBuf<int> d_tail;
        d_tail.upload(tail.data(), M_, st_);
        auto time_digest = [&]() {
            PF_HIP(hipEventRecord(e0, st_));
            digest(0);
            PF_HIP(hipEventRecord(e1, st_));
            PF_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            PF_HIP(hipEventElapsedTime(&ms, e0, e1));
            return ms;
        };
        // three rounds over all shapes, each shape's best round counts
        std::vector<float> t_all(NSHAPE, 1e30f), t_tail(NSHAPE, 1e30f);
        for (int round = 0; round < 3; ++round)
            for (int c = 0; c < NSHAPE; ++c) {
                set_shape(0, c);
                if (round == 0) digest(0);  // warm
                t_all[c] = std::min(t_all[c], time_digest());
                PF_HIP(hipMemcpyAsync(d_mactive_.p, d_tail.p, sizeof(int) * M_, hipMemcpyDeviceToDevice, st_));
                t_tail[c] = std::min(t_tail[c], time_digest());
                fill_int_kernel<<<(M_ + 255) / 256, 256, 0, st_>>>(M_, 1, d_mactive_.p);
            }
        PF_HIP(hipStreamSynchronize(st_));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        int best0 = 0, best1 = 0;
        for (int c = 0; c < NSHAPE; ++c) {
            if (verbose_)
                std::fprintf(stderr, "batch digest: %d waves, items of %.0f values: %.3f ms all, %.3f ms tail\n",
                             waves_c[c / NITEM], item_c[c % NITEM], t_all[c], t_tail[c]);
            if (t_all[c] < t_all[best0]) best0 = c;
            if (t_tail[c] < t_tail[best1]) best1 = c;
        }
        set_shape(0, best0);
        set_shape(1, best1);
        cache[key] = best0 * NSHAPE + best1;
        tune_write(cache);
    }
    PF_HIP(hipStreamSynchronize(st_));
    const double t2 = now();
    // one graph per digest shape (one if both shapes are the same)
    const bool two = sets_[0].waves != sets_[1].waves || sets_[0].values != sets_[1].values;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    if (use_graph)
        for (int s = 0; s < (two ? 2 : 1); ++s) {
            cur_set_ = s;
            PF_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
            iteration(opt);
            PF_HIP(hipStreamEndCapture(st_, &graph[s]));
            PF_HIP(hipGraphInstantiate(&exec[s], graph[s], nullptr, nullptr, 0));
        }
    cur_set_ = 0;
    if (verbose_) std::fprintf(stderr, "batch start: tuning and graph capture   %8.3f s\n", now() - t2);
    int it = 0;
    for (; it < opt.max_iter; ++it) {
        if (use_graph) PF_HIP(hipGraphLaunch(exec[two ? cur_set_ : 0], st_));
        else iteration(opt);
        int active = 0;
        PF_HIP(hipMemcpyAsync(&active, d_count_.p, sizeof(int), hipMemcpyDeviceToHost, st_));
        PF_HIP(hipStreamSynchronize(st_));
        if (active == 0) { ++it; break; }
        if (2 * active <= M_) cur_set_ = 1;
    }
    for (int s = 0; s < 2; ++s) {
        if (exec[s]) (void)hipGraphExecDestroy(exec[s]);
        if (graph[s]) (void)hipGraphDestroy(graph[s]);
    }
    std::vector<double> E(M_), iters(M_), conv(M_);
    int nfallback = 0;
    d_nfallback_.download(&nfallback, 1, st_);
    d_E_.download(E.data(), M_, st_);
    d_iters_.download(iters.data(), M_, st_);
    d_conv_.download(conv.data(), M_, st_);
    PF_HIP(hipStreamSynchronize(st_));
    for (int m = 0; m < M_; ++m) {
        if (energy) energy[m] = E[m];
        if (iterations) iterations[m] = (int)iters[m];
        if (converged) converged[m] = conv[m] != 0.0;
    }
    const double t3 = now();
    if (stats) {
        stats->t_total = t3 - t0;
        stats->t_setup = t1 - t0;
        stats->t_eri = t_eri_;
        stats->t_scf = t3 - t2;
        stats->eri_values = eri_values_;
        double bytes = 0.0;
        for (int q = 0; q < 6; ++q) bytes += (double)s64_[q].n * 12.0;
        stats->store_bytes = bytes;
        stats->iterations = it;
        stats->graph = use_graph;
        stats->digest_waves = sets_[0].waves;
        stats->digest_item_values = sets_[0].values;
        stats->digest_tail_waves = sets_[1].waves;
        stats->digest_tail_item_values = sets_[1].values;
        stats->nproblems = P_;
        stats->sp2_fallbacks = nfallback;
    }
}

}  // namespace pf
