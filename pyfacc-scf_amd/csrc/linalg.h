// Dense FP64 helpers (linalg.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pf {

// Reductions to device scalars; part is scratch of RED_BLOCKS doubles per
// result (the sums are deterministic: fixed partition and order).
constexpr int RED_BLOCKS = 256;
struct DotPtrs {
    static constexpr int MAX = 8;
    const double *p[MAX];
};
// out[y] = a . b.p[y], y < m <= 8
void dev_dots(int64_t n, const double *a, const DotPtrs &b, int m, double *out, double *part, hipStream_t st);
void dev_diff_sq(int64_t n, const double *a, const double *b, double *out, double *part, hipStream_t st);
// Pulay DIIS on the device (no host round trip).  dev_diis_solve sets
// B[slot][k] = B[k][slot] = dots[k] (k < nstore; B is ndiis x ndiis) and
// solves [B -1; -1 0] [c; l] = [0; -1] over the nstore stored vectors by
// Gaussian elimination with partial pivoting in one workgroup: coef[k], and
// ok[0] = 0 when nstore < 2 or a pivot vanishes.  dev_diis_extrap writes
// Fx = sum_k coef[k] v.p[k] when ok[0], else Fx = F.
struct DiisPtrs {
    static constexpr int MAX = 32;
    const double *p[MAX];
};
void dev_diis_solve(int ndiis, int nstore, int slot, const double *dots, double *B, double *coef, int *ok,
                    hipStream_t st);
void dev_diis_extrap(int64_t n, int nstore, const double *coef, const int *ok, const DiisPtrs &v, const double *F,
                     double *Fx, hipStream_t st);
void dev_energy_rhf(int n, const double *D, const double *H, const double *F, double *out, double *part,
                    hipStream_t st);
void dev_energy_uhf(int n, const double *Da, const double *Db, const double *H, const double *Fa, const double *Fb,
                    double *out, double *part, hipStream_t st);
void dev_fill(int64_t n, double v, double *x, hipStream_t st);
// Deterministic J/K accumulation (JKArgs::fx): *fx = 2^e with e the largest
// exponent for which 2 qmax^2 (sum |D0| + sum |D1|) 2^e < 2^61, a bound on
// every partial sum of a J', K' accumulator (|(ab|cd)| <= Q_ab Q_cd <= qmax^2),
// so the int64 sums cannot overflow; D1 may be null.  One workgroup, fixed
// summation order.
void dev_fx_scale(int64_t n, const double *D0, const double *D1, double qmax, double *fx, hipStream_t st);
// x[i] (int64 fixed point with scale *fx, in place) -> double
void dev_fx_to_double(int64_t n, const double *fx, double *x, hipStream_t st);
void dev_fillf(int64_t n, float v, float *x, hipStream_t st);
template <typename R>
void dev_assemble_g(int n, const R *J, const R *K0, const R *K1, double *G0, double *G1, bool uhf, bool accumulate,
                    hipStream_t st);
void dev_axpby(int64_t n, double a, const double *x, double b, const double *y, double *z, hipStream_t st);
void dev_to_float(int64_t n, double scale, const double *x, const double *y, float *z, hipStream_t st);
void dev_antisym(int n, const double *M, double *E, hipStream_t st);
template <typename R>
void dev_shell_dmax(int nsh, int nbf, const int *sh_bf, const int *sh_nf, const double *D0, const double *D1, R *Dsh,
                    R *dmax, hipStream_t st, const double *D2 = nullptr);
// Symmetric eigensolver: A (n x n, row-major) is read only; w ascending, V
// columns are eigenvectors.  Synchronises the stream.
void jacobi_eigh(int n, const double *A, double *w, double *V, hipStream_t st, int max_sweeps = 30,
                 double tol = 1e-15);
// S^{-1/2} of a symmetric positive definite S (coupled Newton-Schulz on FP64
// MFMA GEMMs) and an estimate of lambda_min(S) (relative error < 1e-6).
// Returns false, leaving X undefined, when the iteration has not converged
// (S singular or nearly so); synchronises the stream.
bool inv_sqrt_spd(int n, const double *S, double *X, double *lam_min, hipStream_t st);
// Spectral projector onto the nocc lowest eigenvectors of symmetric Fp;
// X receives the projector, work is n x n scratch, state holds
// sp2_state_size(n) doubles.  Steps whose predecessor converged do no work,
// so the host launches first_check steps before it first reads the
// convergence flag (one stream synchronisation), then checks every 4 steps.
// Returns the number of purification steps the projector needed (a good
// first_check for the next, similar matrix is that plus one).
// Tiles of the purification step: 16 x 16 up to n = 1024 (four times as
// many workgroups where one step is latency-bound: 406 at n = 448), 32 x 32
// above.  The state holds per-tile partial sums of the upper-triangle tiles.
constexpr int sp2_tile_edge(int n) { return n <= 1024 ? 16 : 32; }
constexpr size_t sp2_tiles(int n) {
    return (size_t)((n + sp2_tile_edge(n) - 1) / sp2_tile_edge(n)) *
           ((n + sp2_tile_edge(n) - 1) / sp2_tile_edge(n) + 1) / 2;
}
constexpr size_t sp2_state_size(int n) { return 8 + 6 * sp2_tiles(n) + 2 * (size_t)n; }
int sp2_density(int n, const double *Fp, double nocc, double *X, double *work, double *state, hipStream_t st,
                int max_iter = 100, int first_check = 8);
// The same projector into P with exactly max_steps step launches, no host
// round trip and no allocation (X, work: n x n; state: sp2_state_size(n)).
void sp2_project_fixed(int n, const double *Fp, double nocc, double *P, double *X, double *work, double *state,
                       int max_steps, hipStream_t st);

}  // namespace pf
