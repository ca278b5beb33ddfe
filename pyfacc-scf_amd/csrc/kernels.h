// Internal kernel launchers and device views shared by the host engine.
#pragma once
#include <hip/hip_runtime.h>

#include "pf_common.h"

namespace pf {

// ---------------------------------------------------------------- 1e
void launch_onee(const Shell *d_shells, int nsh, int nbf, int natoms, const double *d_axyz, const int *d_Z,
                 const double *d_tab, double *S, double *T, double *V, hipStream_t st);
// Many molecules at once (device arrays: shells from shoff[m], atoms from
// aoff[m], both with nmol + 1 entries; matrices of nbf mnbf[m] at moff[m]).
void launch_onee_batch(int nmol, int max_nsh, const Shell *d_shells, const int *shoff, const int *aoff,
                       const int *mnbf, const long long *moff, const double *d_axyz, const int *d_Z,
                       const double *d_tab, double *S, double *T, double *V, hipStream_t st);

// ---------------------------------------------------------------- pairs
// Per primitive pair data, in this field order, stored structure-of-arrays.
enum PrimField { PF_P = 0, PF_PX, PF_PY, PF_PZ, PF_AX, PF_AY, PF_AZ, PF_K, PF_O2P, PF_BND, PF_NFIELD };
// stride of one primitive in the pair-major copy (a bra pair's primitive
// fields are read by wavefront-uniform scalar loads, contiguous there)
constexpr int PF_AOS = 16;
// fields of the gradient's bra records ([pair][primitive][PF_AOS], shell
// order): p, 1/2p, P, P - A, K, 2 alpha K, 2 beta K, P - B
enum GradRecField { GR_P = 0, GR_O2P, GR_PX, GR_PY, GR_PZ, GR_AX, GR_AY, GR_AZ, GR_K, GR_KA, GR_KB, GR_BX, GR_BY, GR_BZ };
// A ket primitive pair is visited for a bra pair only when its bound PF_BND =
// |K|/sqrt(q) times the bra's largest primitive bound reaches PRIM_EPS
// (|(ab|cd)_prim| is at most the product of the two bounds times the
// polynomial factors of the p-type recurrences).
constexpr double PRIM_EPS = 1e-16;
// PF_AX.. hold P - A (A = first shell of the pair); PF_K holds
// sqrt(2) pi^{5/4} c_a c_b exp(-ab/p |AB|^2) / p, so that the quartet
// prefactor is K_bra K_ket / sqrt(p + q); PF_O2P holds 1/(2p).

// Device view of one shell-pair class (LA >= LB), pairs sorted by Schwarz
// bound Q descending.  Arrays are structure-of-arrays with stride npad.
template <typename R>
struct PairView {
    int n = 0, npad = 0, kp = 0;
    const int *sa = nullptr, *sb = nullptr;  // shell indices
    const int *oa = nullptr, *ob = nullptr;  // first basis function of each shell
    const R *Q = nullptr;                    // [n] Schwarz bound
    const R *AB = nullptr;                   // [3][npad]  A - B
    const R *prim = nullptr;                 // [kp][PF_NFIELD][npad], per pair by bound |K|/sqrt(p) descending
    const R *aos = nullptr;                  // the same, [npad][kp][PF_AOS]: one pair's primitives contiguous
    const int *kpn = nullptr;                // [npad] primitive pairs that the ERI kernels visit
    R peps = R(0);                           // ket-primitive cut against the bra's bound (0: none)
};

// Work description shared by all ERI/J/K kernels of one launch.
template <typename R>
struct JKArgs {
    PairView<R> bra, ket;
    int same = 0;          // bra and ket are the same class: only ket j <= bra i
    R tau = 0;             // screening threshold on Q_i Q_j dmax
    int nbf = 0, nsh = 0;
    const R *Dj = nullptr;   // density for J (total density)
    const R *Dk0 = nullptr;  // densities for K
    const R *Dk1 = nullptr;
    const R *Dsh = nullptr;  // [nsh*nsh] max |D| per shell block (screening), nullptr = Schwarz only
    const R *dmax = nullptr; // device scalar: 2 * max(Dsh)
    // non-symmetric accumulators J', K': FP64 in every precision (FP32 builds
    // form each quartet's contribution in FP32 and add it in FP64)
    double *J = nullptr, *K0 = nullptr, *K1 = nullptr;
    const R *tab = nullptr;  // Boys table, stride BOYS_MMAX+1
    // stored-integral mode
    R *store = nullptr;
    const long long *store_off = nullptr;  // per bra: first element in store
    const int *store_nket = nullptr;       // per bra: number of stored kets
    const int *store_kbase = nullptr;      // per bra: first ket (batched stores), nullptr = 0
    float *store32 = nullptr;              // optional FP32 copy of the store (same layout)
    // packed canonical export
    double *packed = nullptr;
    // statistics (quartets that passed screening), may be nullptr
    unsigned long long *count = nullptr;
    unsigned long long *pcount = nullptr;  // primitive quartets evaluated, may be nullptr
    // direct-mode work plan (device): per bra ket-prefix length [bra.n],
    // first work item [bra.n + 1], queue head, item total and J'_cd-buffer
    // flag [3]; kets per item
    int *nket = nullptr, *item_off = nullptr, *work = nullptr;
    int chunk = 0;
    // J'_cd of every evaluated quartet, stored per (bra, ket of its prefix) at
    // (joff[bra] + ket) * NC*ND and summed per ket afterwards; jcap doubles /
    // (NC*ND) entries at most (else device atomics)
    double *jbuf = nullptr;
    long long *joff = nullptr;
    int *jhi = nullptr;  // per ket: end of the bra range holding it [ket.n]
    long long jcap = 0;
    int krows = 0;  // direct mode: bra exchange rows summed in LDS
    int katomic = 0;  // direct mode: force exchange rows to device atomics (the no-LDS-rows path)
    int planned = 0;  // direct mode: the work plan and the J'_cd sums are launched for all classes at once
    // direct mode, deterministic accumulation: device scalar 2^s.  Every
    // addition into J', K' (LDS rows and device memory) is then the integer
    // add of round(v 2^s) to the accumulator's 8 bytes read as a two's-
    // complement int64, so the sums do not depend on the order of the adds.
    // nullptr: FP64 atomics.
    const double *fx = nullptr;
    // host side only: when set, recorded on the stream right before and after
    // the class's main ERI kernel (its duration without the helper kernels)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

enum JKMode { MODE_DIRECT = 0, MODE_STORE = 1, MODE_LOAD = 2, MODE_PACKED = 3 };

// Direct-mode work plan of one Fock build for all its quartet classes at once:
// one launch finds every class's ket prefixes (Q_i Q_j dmax >= tau) and item
// counts, one scans them (a workgroup per class); after the direct kernels
// one launch finds, per ket, the bras whose prefix holds it, and one sums the
// J'_cd buffer's columns into J'.  Six launches per build instead of 24.
template <typename R>
struct PlanClass {
    PairView<R> bra, ket;
    int same = 0, chunk = 0, ncd = 0;  // ncd: components of a ket pair (J'_cd values per slot)
    int *nket = nullptr, *item_off = nullptr, *work = nullptr, *jhi = nullptr;
    long long *joff = nullptr;
    long long jcap = -1;               // J'_cd buffer capacity in slots (-1: no buffer)
    const double *jbuf = nullptr;
    int cb0 = 0, cb1 = 0;              // this class's blocks of the count launch
    int rb0 = 0, rb1 = 0;              // ... of the J'_cd range launch
    int sb0 = 0, sb1 = 0;              // ... of the J'_cd sum launch (x dimension)
};
template <typename R>
struct PlanArgs {
    PlanClass<R> c[6];
    int n = 0;                         // classes
    R tau = 0;
    const R *dmax = nullptr;
    double *J = nullptr;
    int nbf = 0;
    const double *fx = nullptr;        // deterministic accumulation scale (nullptr: FP64 atomics)
};
// plan (count + scan) before the direct kernels; J'_cd ranges and sums after them
template <typename R>
void launch_plan(PlanArgs<R> &p, hipStream_t st);
template <typename R>
void launch_jcd_sums(PlanArgs<R> &p, hipStream_t st);
// kets per direct-mode work item of a class (the direct kernel's chunk)
template <typename R>
int direct_chunk(int la, int lb, int lc, int ld, int scale);

// Launch one quartet class (bra class la,lb | ket class lc,ld).
template <typename R>
void launch_jk_class(int la, int lb, int lc, int ld, int nk, int mode, const JKArgs<R> &args, hipStream_t st);

// Schwarz bounds for every pair of a class: Q[i] = sqrt(max |(ab|ab)|).
template <typename R>
void launch_schwarz(int la, int lb, const PairView<R> &pv, const R *tab, R *Q, hipStream_t st);

// ---------------------------------------------------------------- gradient
struct GradArgs;
void launch_grad1e(int la, int lb, const Shell *shells, const int *pa, const int *pb, int npairs, int nbf, int natoms,
                   const double *axyz, const int *Z, const double *tab, const double *D, const double *W,
                   double *gT, double *gV, double *gS, hipStream_t st);
// queue: a device int the persistent derivative kernel takes its items from
void launch_grad2e_class(int la, int lb, int lc, int ld, int nk, const GradArgs &g, int *work, int cap, hipStream_t st);

// ---------------------------------------------------------------- dense FP64
// C = alpha op(A) op(B) + beta C, row-major n x n (lda = ldb = ldc = n), MFMA f64.
void gemm_f64(int n, bool ta, bool tb, double alpha, const double *A, const double *B, double beta, double *C,
              hipStream_t st);
// General variant: C = alpha op(A) op(B) + beta Cin; alpha_p / beta_p (device
// pointers) override alpha / beta when non-null.
void gemm_f64_ex(int m, int n, int k, bool ta, bool tb, double alpha, const double *alpha_p, const double *A, int lda,
                 const double *B, int ldb, double beta, const double *beta_p, const double *Cin, double *C, int ldc,
                 hipStream_t st);
// General m x n x k row-major variant.
void gemm_f64_mnk(int m, int n, int k, bool ta, bool tb, double alpha, const double *A, int lda,
                  const double *B, int ldb, double beta, double *C, int ldc, hipStream_t st);

}  // namespace pf
