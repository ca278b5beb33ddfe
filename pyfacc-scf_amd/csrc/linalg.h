// Dense FP64 helpers (linalg.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pf {

void dev_dot(int64_t n, const double *a, const double *b, double *out, hipStream_t st);
void dev_diff_sq(int64_t n, const double *a, const double *b, double *out, hipStream_t st);
void dev_trace(int n, const double *a, double *out, hipStream_t st);
void dev_energy_rhf(int n, const double *D, const double *H, const double *F, double *out, hipStream_t st);
void dev_energy_uhf(int n, const double *Da, const double *Db, const double *H, const double *Fa, const double *Fb,
                    double *out, hipStream_t st);
void dev_fill(int64_t n, double v, double *x, hipStream_t st);
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
// Spectral projector onto the nocc lowest eigenvectors of symmetric Fp;
// X receives the projector, work is n x n scratch, state holds 6 doubles.
// Returns the number of iterations used.
int sp2_density(int n, const double *Fp, double nocc, double *X, double *work, double *state, hipStream_t st,
                int max_iter = 100);

}  // namespace pf
