// Batched SCF over many independent small molecules on one GPU.
//
// Every molecule's unique shell quartets (Schwarz-screened) are computed once
// into stored FP64 and FP32 integral lists; each SCF iteration then runs a
// fixed sequence of batched kernels over all molecules (spin densities are
// "problems": one per RHF molecule, two per UHF one):
//   digest (one workgroup per molecule, J/K in LDS) -> F = H + G,
//   DIIS (error FDS - SDF, B matrix, per-molecule solve, extrapolation),
//   F' = X^T F X, SP2 purification (one workgroup per problem, steps until
//   its own convergence), D = occ X P X^T, energy and convergence.
// Converged molecules are frozen.  The sequence contains no host decision, so
// it is captured once into a hipGraph and replayed every iteration; the host
// reads one counter (molecules still running) per iteration.
#pragma once
#include <string>
#include <vector>

#include "engine.h"
#include "kernels.h"
#include "pf_common.h"

namespace pf {

struct BatchMolecule {
    Molecule mol;
    std::vector<std::string> species;
    int nalpha = 0, nbeta = 0;
};

// Device view of the stored integrals of one quartet class for the digest.
struct BatchClass {
    int la, lb, lc, ld;
    const int *sa = nullptr, *sb = nullptr, *oa = nullptr, *ob = nullptr;  // bra pairs
    const int *ksa = nullptr, *ksb = nullptr, *koa = nullptr, *kob = nullptr;  // ket pairs
    const long long *off = nullptr;  // per bra: first stored value
    const int *nket = nullptr, *kbase = nullptr;
    const double *v64 = nullptr;
    const float *v32 = nullptr;
    const int *brange = nullptr;  // per molecule: bra pairs brange[2m] .. brange[2m+1]-1
    int same = 0;
    // FP32 increments: Schwarz bounds of the bra and ket pairs (kets of a
    // molecule sorted by Q descending), each bra's molecule, and per bra the
    // ket count that passes Q_i Q_j max|dD| >= tau this iteration
    const double *bQ = nullptr, *kQ = nullptr;
    const int *bmol = nullptr;
    int *nkeff = nullptr;
};

// A digest work item: molecule m's bra pairs from (class q0, pair i0) to
// (class q1, pair i1) exclusive; split = the molecule has several items.
struct BatchItem {
    int m, q0, i0, q1, i1, split;
};

// Per-molecule and per-problem state shared by the batched kernels.
struct BatchDev {
    int M = 0, P = 0, nmax = 0, ndiis = 0;
    const int *mn = nullptr, *muhf = nullptr, *mprob = nullptr;  // nbf, UHF flag, first problem
    const long long *moff = nullptr;                              // molecule matrix offsets
    const long long *gacc_off = nullptr;                          // split molecules' J/K sums
    const int *pn = nullptr, *pmol = nullptr;
    const double *poccf = nullptr, *pnocc = nullptr;              // occupation factor, occupied count
    const long long *poff = nullptr, *pmoff = nullptr;            // problem / its molecule's matrix offsets
    const long long *pdiis = nullptr;                             // DIIS ring offsets
    const long long *psp2 = nullptr;  // Gershgorin scratch offsets (orthogonaliser's molecule view)
    int *mactive = nullptr;   // 1 while the molecule iterates
    int *mprec = nullptr;     // 1 = full FP64 builds; mixed form: 2 = FP64 opening, 0 = FP32 increments, then 1
    int *mhead = nullptr, *mnstore = nullptr, *muse = nullptr;  // DIIS ring head / size / extrapolate flag
    double *mcoef = nullptr;  // DIIS coefficients [M][ndiis]
    double *mB = nullptr;     // DIIS B matrices [M][ndiis][ndiis]
    int *psp2fail = nullptr;  // per problem: 1 when the last purification did not converge
    int *nfallback = nullptr; // densities formed from eigenvectors instead (whole run)
};

struct BatchStats {
    double t_total = 0, t_setup = 0, t_eri = 0, t_scf = 0;
    double eri_values = 0, store_bytes = 0;
    int iterations = 0, graph = 0, digest_waves = 0, nproblems = 0;
    int sp2_fallbacks = 0, digest_tail_waves = 0;
    double digest_item_values = 0, digest_tail_item_values = 0;
};

class BatchEngine {
   public:
    BatchEngine(std::vector<BatchMolecule> mols, const std::string &basis_path, bool renormalise, double schwarz);
    ~BatchEngine();
    BatchEngine(const BatchEngine &) = delete;
    BatchEngine &operator=(const BatchEngine &) = delete;
    // energy/iterations/converged per molecule; use_graph: replay iterations from a hipGraph
    void run(const pf_options &opt, bool use_graph, double *energy, int *iterations, int *converged,
             BatchStats *stats);

   private:
    void setup();
    void setup_orthogonalizer();
    void density_step(const double *F);  // D from F (per problem) for the active molecules
    void iteration(const pf_options &opt);
    void digest(int set);
    // work items of digest shape `set`: molecules whose stored integrals
    // exceed item_values split into about equal parts (whole bra pairs),
    // largest items first
    void build_items(int set, double item_values);

    hipStream_t st_ = nullptr;
    std::vector<BatchMolecule> mols_;
    std::string basis_path_;
    bool renorm_;
    double schwarz_;
    double t_eri_ = 0.0;
    bool verbose_ = false;
    int M_ = 0, P_ = 0, nmax_ = 0, ndiis_ = 8;
    std::vector<int> mn_, muhf_, mprob_, pn_, pmol_;
    std::vector<double> enuc_;
    std::vector<long long> moff_, poff_;
    long long mtot_ = 0, ptot_ = 0;
    std::vector<Shell> shells_;
    std::vector<int> shoff_;  // molecule m's shells: shells_[shoff_[m] .. shoff_[m+1])
    DevBuf<Shell> d_shells_;
    // device
    DevBuf<double> tab_;
    DevBuf<int> d_mn_, d_muhf_, d_mprob_, d_pn_, d_pmol_, d_mactive_, d_mprec_, d_mhead_, d_mnstore_, d_muse_;
    DevBuf<int> d_count_;
    DevBuf<int> d_psp2fail_, d_nfallback_;
    size_t jacobi_lds_ = 0;            // dynamic LDS of the eigenvector fallback
    bool jacobi_mats_in_lds_ = false;  // its A and V in LDS (else in device memory)
    DevBuf<double> d_poccf_, d_pnocc_, d_enuc_, d_E_, d_Eold_, d_iters_, d_conv_, d_coef_, d_B_;
    DevBuf<long long> d_moff_, d_poff_, d_pmoff_, d_pdiis_;
    DevBuf<double> S_, H_, X_, Dtot_;                    // molecule pools
    DevBuf<double> D_, Dold_, F_, Fx_, Fp_, W1_, W2_, Pr_;  // problem pools
    DevBuf<double> dF_, dE_;                             // DIIS rings
    // mixed form: the two-electron part G (F = H + G) and the density it was
    // built from (problem pools), the FP32 difference densities of the
    // increments (problem pool; UHF molecules' spin sum in a molecule pool) and
    // each molecule's max|dD|
    bool mixed_ = false;
    double schwarz_inc_ = 1e-12;  // ket cut of the increments after the switch (Q Q max|dD|)
    DevBuf<double> G_, Db_;
    DevBuf<float> dD32_, dDt32_, dmax_;
    std::array<DevBuf<int>, 3> d_pairmol_;  // per pair class: each pair's molecule
    std::array<DevBuf<int>, 6> nkeff_;      // per quartet class: kets of each bra this iteration
    std::array<PairSet, 3> pairs_;
    std::array<std::vector<int>, 3> pbeg_;  // per class: molecule m's pairs pbeg[m] .. pbeg[m+1]-1
    // stored integrals per quartet class
    std::array<DevBuf<double>, 6> s64_;
    std::array<DevBuf<float>, 6> s32_;
    std::array<DevBuf<long long>, 6> soff_;
    std::array<DevBuf<int>, 6> snket_, skbase_, sbr_;
    std::array<BatchClass, 6> cls_;
    // two digest shapes (work items + wavefronts per workgroup): [0] while
    // more than half the molecules iterate, [1] for the rest of the run
    struct ItemSet {
        DevBuf<BatchItem> items;
        DevBuf<int> msplit;  // per molecule: 1 if split over several items
        int n = 0, waves = 0;
        double values = 0.0;  // work-item size the items were built with
    };
    std::array<ItemSet, 2> sets_;
    int cur_set_ = 0;
    DevBuf<long long> d_gacc_off_;
    DevBuf<double> gacc_;
    std::array<std::vector<double>, 6> work_;  // per class: stored values per bra pair
    std::array<std::vector<int>, 6> brq_;      // per class: bra range of each molecule
    BatchDev dev_;
    double eri_values_ = 0;
};

}  // namespace pf
