/*
 * C-ABI of the MI355X Hartree-Fock engine (libpyfacc.so).
 *
 * Drop-in boundary: the first two entry points have exactly the signatures
 * of the reference's shared library, so the reference's Python driver
 * (src/python/helper_functions.py:90-149 via ctypes) and its example script
 * (single_file.py:21-39) can load this library in place of libscf.so.
 * The pf_* entry points are additions for tests, benchmarks and callers that
 * want status codes, options and intermediate results.  All buffers are
 * caller-owned host memory, borrowed for the duration of the call; nothing
 * outlives a call.  Every call is synchronous and owns its GPU state
 * (no module-global SCF state, unlike libint_integrals.f90:15).
 *
 * Conventions (the reference's): coordinates in bohr, per-atom interleaved
 * x y z (Fortran (3, natoms), main.f90:21 / molecule.f90:16); species as
 * 2*natoms space-padded bytes without NUL (helper_functions.py:105-107);
 * matrices nbf x nbf row-major (symmetric ones are the same in either order);
 * basis-function order = the reference's shell order (basis.f90:44-85) x
 * Cartesian x, y, z (cartesian.f90:7-23).
 */
#ifndef PYFACC_H
#define PYFACC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces run_scf_c (src/c/scf_interface.c:12-17): total energy
 * E_elec + E_nuc written to *final_energy.  RHF when nalpha == nbeta, else UHF
 * (the dispatch main.f90:57-61 intends).  Plain Roothaan iterations with the
 * reference's stopping rule (|dE| < eps and ||dD||_F < eps, rscf.f90:152-155),
 * at most max_iter iterations, unconverged result returned silently as the
 * reference does.  Basis file ./data/basis_set/<basis_set_c>.txt relative to
 * the working directory (main.f90:33), then $PYFACC_DATA/basis_set, then the
 * data directory shipped next to this library.  On error *final_energy is NaN
 * and pf_last_error() describes it. */
void run_scf_c(int natoms, int nelectrons, int charge, int nalpha, int nbeta, const char *species, double *coords,
               double eps_scf, int max_iter, double *final_energy, const char *basis_set_c, int32_t *z_num);

/* Replaces the Fortran bind(C) entry run_scf_interface_py (main.f90:15-64;
 * declared at src/c/scf_interface.c:7-9): every argument by reference, the
 * basis name of length *len_basis_set, not necessarily NUL-terminated. */
void run_scf_interface_py(int32_t *natoms, int32_t *nelectrons, int32_t *charge, int32_t *nalpha, int32_t *nbeta,
                          const char *species, double *coords, double *eps_scf, int32_t *max_iter,
                          double *final_energy, const char *basis_set_c, int32_t *len_basis_set, int32_t *z_num);

/* ---------------------------------------------------------------- pf_* */

enum { PF_OK = 0, PF_ERR_INPUT = 1, PF_ERR_BASIS = 2, PF_ERR_DEVICE = 3, PF_ERR_INTERNAL = 4 };
enum { PF_PREC_FP64 = 0, PF_PREC_MIXED = 1 };
enum { PF_ERI_DIRECT = 0, PF_ERI_STORED = 1 };
enum { PF_DIAG_SP2 = 0, PF_DIAG_JACOBI = 1 };
enum { PF_GUESS_CORE = 0, PF_GUESS_SAD = 1 };
enum { PF_DIIS_AO = 0, PF_DIIS_ORTHO = 1 };
enum { PF_FINISH_FP32 = 0, PF_FINISH_FP64 = 1 };

typedef struct pf_options {
    double eps;           /* convergence threshold on |dE| and ||dD||_F (rscf.f90:152-155) */
    int max_iter;
    int precision;        /* PF_PREC_FP64 or PF_PREC_MIXED (FP32 incremental J/K, FP64 finish) */
    int incremental;      /* 1: F_n = F_{n-1} + G[D_n - D_{n-1}] (FP64 path) */
    int eri_mode;         /* PF_ERI_DIRECT or PF_ERI_STORED (FP64 only) */
    int diis;             /* 0: plain Roothaan (reference); n > 1: Pulay DIIS over n vectors */
    int diag;             /* PF_DIAG_SP2 (GEMM spectral projector) or PF_DIAG_JACOBI */
    int renormalize;      /* 1: contracted functions normalised (Libint); 0: primitive norms only (reference) */
    double schwarz;       /* screening threshold on Q_ab Q_cd max|D| */
    double mixed_switch;  /* mixed precision: ||dD||_F below which the FP64 finish starts */
    int verbose;          /* 1: one line per iteration on stderr */
    double schwarz_inc;   /* screening threshold of FP64 incremental builds (on Q Q max|dD|) */
    int rebuild_every;    /* incremental FP64 builds: full rebuild after this many increments (0: never) */
    double schwarz_fp32;  /* mixed precision: screening threshold (on Q Q max|dD|; 0: schwarz) of the
                             FP32 incremental builds before the full FP64 rebuild that replaces their sum */
    int guess;            /* PF_GUESS_CORE: core Hamiltonian (rscf.f90:113-114, the reference's); PF_GUESS_SAD:
                             superposition of spherically averaged atomic UHF densities */
    double level_shift;   /* Eh added to the virtual space, F + b (S - S D S / occ), after DIIS extrapolation,
                             while ||dD||_F of the previous iteration is >= level_shift_until (0: off;
                             pf_scf and pf_scf_gradient; pf_batch_scf rejects a nonzero value) */
    double level_shift_until;
    int diis_basis;       /* DIIS error vector: PF_DIIS_AO, F D S - S D F; PF_DIIS_ORTHO, X^T (F D S - S D F) X
                             with X the orthogonaliser (pf_scf and pf_scf_gradient; pf_batch_scf: AO only) */
    double mixed_anchor;  /* mixed precision: the full FP64 rebuild inside the FP32 phase comes when ||dD||_F
                             first falls below mixed_anchor (0: none; pf_scf and pf_scf_gradient); pf_batch_scf:
                             full FP64 builds until ||dD||_F first falls below mixed_anchor, FP32 increments after
                             (0: FP32 increments from the start) */
    int profile;          /* 1: time each J/K class kernel with HIP events around it (pf_scf_profile's
                             milliseconds; the events add short gaps to the build), 0: no events */
    int jk_fallback;      /* testing: direct J/K builds take the paths that large systems need, at any
                             size: bit 0, exchange rows added with device atomics (rows too long for LDS);
                             bit 1, J'_cd added with device atomics (ket lists beyond the J'_cd buffer).
                             0: chosen by size (default) */
    int deterministic;    /* 1: every J', K' accumulation of a direct build is an int64 fixed-point add
                             (scale 2^e with B 2^e < 2^61 for the bound B = 2 Qmax^2 sum|D| on the sums; the
                             resolution 2^-e is below 1e-12 while B < 2^21), so repeated SCFs give bit-identical
                             energies and densities; s and p shells, direct builds only (pf_scf,
                             pf_scf_gradient's SCF; not pf_batch_scf).  0: FP64 atomics (repeated runs differ
                             by ~1e-11) */
    int mixed_finish;     /* mixed precision, after the switch's full FP64 build: PF_FINISH_FP32, the increments
                             that follow run in FP32 (screened at schwarz_inc; their rounding scales with
                             ||dD||_F < mixed_switch), accumulated in FP64; PF_FINISH_FP64, they run in FP64
                             (pf_batch_scf: full FP64 builds).  pf_scf, pf_scf_gradient and pf_batch_scf */
} pf_options;

typedef struct pf_result {
    double energy;        /* E_elec + E_nuc (Eh) */
    double e_elec;
    double e_nuc;
    double last_dE;
    double last_dD;
    int iterations;
    int converged;
    int fock_builds_fp64; /* full or incremental FP64 J/K builds */
    int fock_builds_fp32; /* FP32 incremental builds (mixed precision) */
    int nbf;
    int nshell;
    double t_total;       /* seconds, wall clock of the SCF call */
    double t_setup;       /* one-electron integrals, S^-1/2, pair lists, Schwarz, stored ERIs */
    double t_fock;        /* J/K builds (pf_options.profile = 1; 0 otherwise) */
    double t_diag;        /* density from Fock (orthogonalise + SP2/Jacobi + back-transform; profile = 1) */
    double quartets;      /* shell quartets evaluated over all builds */
    double min_overlap_eig; /* smallest eigenvalue of S (la.f90:40-44 warns below 1e-8) */
} pf_result;

/* Defaults: eps 1e-8, max_iter 100, FP64, full builds, direct, DIIS 8, SP2,
 * renormalised, Schwarz 1e-10, mixed_switch 1e-5, schwarz_inc 1e-12,
 * rebuild_every 0, schwarz_fp32 1e-6, core guess, no level shift (level_shift_until 0.1), AO DIIS error,
 * mixed_anchor 0.1, profile 0, jk_fallback 0, deterministic 0, mixed_finish PF_FINISH_FP32. */
void pf_default_options(pf_options *opt);

/* Message for the last failed call on this thread ("" if none). */
const char *pf_last_error(void);

/* Number of visible GPUs (0 if none). */
int pf_device_count(void);

/* Basis size and shells.  Outputs may be NULL; shell arrays hold nshell
 * entries (exp/coef: nshell*8, coefficients include the primitive norm). */
int pf_basis_info(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
                  int renormalize, int *nbf, int *nshell, int *shell_l, int *shell_atom, int *shell_nprim,
                  double *shell_exp, double *shell_coef);

/* S, T, V (nbf x nbf) on the GPU (overlap.f90:240, kinetic.f90:117, nuclear.f90:202). */
int pf_one_electron(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
                    int renormalize, double *S, double *T, double *V);

/* All unique ERIs in the canonical packed order of rscf.f90:67-88 (0-based:
 * ij = i(i+1)/2 + j, i >= j; index = ij(ij+1)/2 + kl, ij >= kl), unscreened;
 * packed holds nbf(nbf+1)/2 * (nbf(nbf+1)/2 + 1)/2 doubles. */
int pf_eri_packed(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
                  int renormalize, double *packed);

/* ERI blocks of shell quartets in Libint's layout: quartet q is
 * (quartets[4q] quartets[4q+1] | quartets[4q+2] quartets[4q+3]), 0-based shell
 * indices in the order of pf_basis_info, any orientation; its block holds
 * n1 n2 n3 n4 values (n = 1 for s, 3 for p), index ((i1 n2 + i2) n3 + i3) n4 + i4
 * over the shells' Cartesian components (x, y, z), over contracted functions
 * normalised when renormalize = 1 (Libint's convention) or with primitive
 * norms only (0).  Blocks are concatenated in list order; offsets (may be
 * NULL) receives nquart + 1 entries, block q at out + offsets[q].  Replaces
 * what compute_ERI_libint receives from Libint per quartet before scattering
 * it (libint_integrals.f90:111-139); every primitive quartet is included. */
int pf_eri_shell_quartets(int natoms, const char *species, const int32_t *z, const double *coords,
                          const char *basis_name, int renormalize, int nquart, const int32_t *quartets, double *out,
                          int64_t *offsets);

/* Schwarz bounds Q_ab = sqrt(max |(ab|ab)|) over the Cartesian components of
 * every shell pair, Q[nshell * nshell] symmetric, shells in pf_basis_info's
 * order; the values the direct J/K builds screen with (Q_ab Q_cd max|D| <
 * schwarz).  The reference leaves schwarz_matrix empty (schwarz.f90:13-21). */
int pf_schwarz(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
               int renormalize, double *Q);

/* J[D] (summed over densities) and K[D_k] per density, for ndens = 1 or 2
 * host densities (ndens x nbf x nbf).  eri_mode as in pf_options; precision
 * PF_PREC_MIXED evaluates in FP32.  schwarz <= 0 selects 1e-14. */
int pf_jk(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
          int renormalize, int ndens, const double *D, double *J, double *K, int eri_mode, int precision,
          double schwarz);

/* SCF with options.  D_out (ndens x nbf^2; ndens = 1 RHF total density, 2 UHF
 * alpha, beta) and eps_out (ndens x nbf orbital energies, Jacobi at the end)
 * may be NULL. */
int pf_scf(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
           int nalpha, int nbeta, const pf_options *opt, pf_result *res, double *D_out, double *eps_out);

/* SCF (as pf_scf) followed by the analytic nuclear gradient dE/dR at the
 * converged density: grad[natoms*3] in Eh/bohr (one-electron, two-electron
 * and nuclear-repulsion terms; W = 1/2 D F D, RHF, or sum_s D_s F_s D_s, UHF).
 * The reference has no gradients; this extension is bound to torch.autograd
 * by pyfacc.grad.  D_out may be NULL;
 * terms (may be NULL) receives 5 x natoms x 3 contributions: nuclear
 * repulsion, kinetic (D dT), nuclear attraction (D dV), overlap (-W dS),
 * two-electron. */
int pf_scf_gradient(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
                    int nalpha, int nbeta, const pf_options *opt, pf_result *res, double *grad, double *D_out,
                    double *terms);

/* Fock-build profile of the last pf_scf call on this thread, summed over its
 * J/K builds, 12 entries indexed precision*6 + class (precision 0 FP64,
 * 1 FP32; classes (ss|ss) (ps|ss) (ps|ps) (pp|ss) (pp|ps) (pp|pp)):
 * quartets that passed screening, milliseconds of the class's main ERI
 * kernel (HIP events on the engine's stream right around it, recorded only
 * when pf_options.profile = 1, else 0: the screening
 * count/scan and J'_cd reduction helpers are not included), kernel launches
 * and primitive quartets evaluated (summed over quartets);
 * class_l[24] gives per class the angular momenta (la lb | lc ld) in the
 * orientation launched.  Any output may be NULL. */
int pf_scf_profile(unsigned long long *quartets, double *ms, int *launches, int *class_l,
                   unsigned long long *prims);

/* Statistics of one pf_batch_scf call. */
typedef struct {
    double t_total;     /* seconds, whole call */
    double t_setup;     /* integrals, pairs, stored ERIs and orthogonalisers */
    double t_eri;       /* of which: computing the stored integrals */
    double t_scf;       /* the SCF iterations */
    double eri_values;  /* unique integrals stored (each kept in FP64 and FP32) */
    double store_bytes; /* device bytes of the stored integrals */
    int iterations;     /* batched iterations run (the slowest molecule's count) */
    int graph;          /* 1: iterations replayed from one captured hipGraph */
    int digest_waves;   /* autotuned wavefronts per workgroup in the J/K digest while more than half the
                           molecules iterate */
    int nproblems;      /* spin densities in the batch (1 per RHF, 2 per UHF molecule) */
    int sp2_fallbacks;  /* densities formed from eigenvectors (aufbau) because the purification did not
                           converge: the occupied/virtual boundary inside a degenerate level */
    int digest_tail_waves; /* the same after at least half the molecules have converged */
    double digest_item_values; /* autotuned work item of the digest while more than half the molecules
                                  iterate: stored integrals per item (a larger molecule is split over
                                  several workgroups; 1e18: none is) */
    double digest_tail_item_values; /* the same after at least half the molecules have converged */
} pf_batch_stats;

/* Batched SCF over many independent molecules in one set of launches.  No
 * counterpart in the reference: it is for screening workloads (thousands of
 * small water clusters) where one molecule cannot fill the GPU.  Each
 * molecule follows the reference's RHF/UHF iteration and stopping rule
 * (rscf.f90:49-164, uscf.f90:65-213; UHF when nalpha != nbeta, core guess),
 * with DIIS and integrals computed once and stored.  Molecule m has natoms[m]
 * atoms; species (2 chars per atom, as pf_scf), z and coords (bohr, xyz per
 * atom) are concatenated in molecule order.  opt: eps, max_iter, diis (2..8;
 * 0 or 1 = plain Roothaan), schwarz, renormalize, precision (FP64, or mixed:
 * FP32 stored integrals until ||dD||_F < mixed_switch, then FP64 to
 * convergence).  use_graph: replay the iteration from a captured hipGraph.
 * Outputs per molecule (may be NULL): total energy (Eh), iterations,
 * converged flag. */
int pf_batch_scf(int nmol, const int32_t *natoms, const char *species, const int32_t *z, const double *coords,
                 const char *basis_name, const int32_t *nalpha, const int32_t *nbeta, const pf_options *opt,
                 int use_graph, double *energy, int32_t *iterations, int32_t *converged, pf_batch_stats *stats);

/* The projector of pf_projector for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid hipMalloc).
 * No counterpart in the reference.  F (n x n, symmetric, row-major) and P (the
 * projector onto the nocc lowest eigenvectors of F) are caller-owned DEVICE
 * buffers; stream is a hipStream_t (NULL: the null stream).  The call only
 * enqueues work: exactly max_steps (1..200) purification steps (steps after
 * convergence do nothing; 100 covers any gap of practical interest), no host
 * wait, no hipMalloc/hipFree, no memset.  Its working memory (2 n^2 + O(n)
 * doubles) is a fixed device array of the library, one per GPU, sized for
 * n <= PF_PROJECTOR_NOALLOC_NMAX (about 16 MiB), held until the process ends;
 * larger n is rejected (PF_ERR_INPUT).  Calls that overlap on one GPU (two
 * streams, two threads) share that array and race: serialise them.  The first
 * call on each GPU resolves the array's address (no allocation); make it
 * before a graph capture. */
#define PF_PROJECTOR_NOALLOC_NMAX 1024
int pf_projector_noalloc(int n, const double *F, double nocc, double *P, int max_steps, void *stream);

/* Device memory kept between calls.  Every call drains its own stream, so the
 * library caches released device blocks per calling thread for its next call
 * (up to 8 GiB per thread); a thread that ends hands its blocks to a
 * process-wide idle list (up to 8 GiB, reused by any thread, the rest freed).
 * pf_release_device_cache frees the calling thread's cache, that idle list and
 * the idle scratch blocks of finished SCFs; no library call may be running on
 * another thread meanwhile.  pf_cached_device_bytes: bytes held by the calling
 * thread's cache plus the idle list.  No counterpart in the reference (which
 * allocates and frees per call); for long-running services. */
int pf_release_device_cache(void);
unsigned long long pf_cached_device_bytes(void);

/* Device kernels exposed for testing on host buffers. */
int pf_boys(int nmax, int npts, const double *x, double *F);               /* F[npts*(nmax+1)], nmax <= 6 */
int pf_eigh(int n, const double *A, double *w, double *V);                  /* Jacobi: w ascending, V columns */
int pf_projector(int n, const double *F, double nocc, double *P);           /* SP2 spectral projector */
int pf_gemm(int m, int n, int k, int ta, int tb, double alpha, const double *A, const double *B, double beta,
            double *C);                                                     /* row-major, FP64 MFMA */
/* Plain device-to-device copy of `bytes` (kernel ms in *ms): the reference
 * kernel for calibrating HBM byte counters. */
int pf_device_copy(unsigned long long bytes, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* PYFACC_H */
