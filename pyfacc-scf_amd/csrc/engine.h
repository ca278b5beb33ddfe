// SCF engine: one molecule on one GPU.
#pragma once
#include <array>
#include <memory>
#include <string>
#include <vector>

#include "generic.h"
#include "kernels.h"
#include "pf_common.h"

extern "C" {
#include "../../include/pyfacc.h"
}

namespace pf {

// One shell-pair class on the device, in both precisions.
struct PairSet {
    int la = 0, lb = 0;
    int n = 0, npad = 0, kp = 0;
    double peps = 0.0;         // ket-primitive cut against the bra's bound (0: none)
    std::vector<int> sa, sb;   // host copies, sorted by Q descending
    std::vector<double> Q;
    DevBuf<int> d_sa, d_sb, d_oa, d_ob, d_kpn;
    DevBuf<double> d_Q, d_AB, d_prim, d_aos;
    DevBuf<float> f_Q, f_AB, f_prim, f_aos;
    PairView<double> view64() const;
    PairView<float> view32() const;
};

// Fill ps from a list of (first, second) shell indices into sh (host copy) /
// d_shells (device copy), built on the device (pairs.hip): SoA fields in FP64
// (and FP32 when want32), primitive pairs sorted by bound with the significant
// count kpn against kref (kref < 0 keeps all); bound_max is raised to the
// largest primitive bound.  Synchronises the stream.
void fill_pairset(PairSet &ps, const std::vector<Shell> &sh, const Shell *d_shells,
                  const std::vector<std::pair<int, int>> &list, double kref, double &bound_max, bool want32,
                  hipStream_t st);

// The gradient's primitive-pair data of ps, primitives in shell order: the
// structure-of-arrays sets w (P - A, K 2 alpha) and r (P - A, K), each
// kp * PF_NFIELD * npad doubles, and the records rec (npad * kp * PF_AOS, GR_
// fields).
void fill_grad_pairset(const PairSet &ps, const Shell *d_shells, double *w, double *r, double *rec, hipStream_t st);

struct QuartetClass {
    int bra, ket;  // indices into Engine::pairs_
    int slot;      // unordered class: (ss|ss) (ps|ss) (ps|ps) (pp|ss) (pp|ps) (pp|pp)
};

struct Timing {
    double ms = 0.0;
    int calls = 0;
};

// Spin-summed density of the isolated atom (element label, atomic number Z)
// in the basis at path, spherically averaged per shell block; cached.
std::vector<double> atomic_density(const std::string &path, bool renorm, const std::string &el, int Z);
// Superposition of atomic densities over shells sh[0..nsh) (atom indices and
// first basis functions local to the molecule): nbf^2, row-major, spin-summed.
std::vector<double> sad_density(const Shell *sh, int nsh, int nbf, const std::vector<std::string> &species,
                                const std::vector<int> &Z, const std::string &path, bool renorm);

class Engine {
   public:
    Engine(const Molecule &mol, const std::vector<std::string> &species, const std::string &basis_path,
           bool renormalise, double schwarz);
    ~Engine();

    int nbf() const { return basis_.nbf; }
    int nshell() const { return (int)basis_.shells.size(); }
    const Basis &basis() const { return basis_; }
    hipStream_t stream() const { return st_; }

    void one_electron(double *S, double *T, double *V);  // host outputs
    // J/K over host densities (ndens = 1 or 2); host outputs J (total) and K per density.
    void jk_host(int ndens, const double *D, double *J, double *K, int mode, int precision);
    void eri_packed_host(double *packed);
    long long schwarz_pairs() const;
    // Q_ab of every shell pair (nshell^2, symmetric), host output.
    void schwarz_host(double *Q);

    // Full SCF.  D_out (ndens x nbf^2) and eps_out (ndens x nbf) may be null.
    void scf(int nalpha, int nbeta, const pf_options &opt, pf_result &res, double *D_out, double *eps_out);
    // Nuclear gradient (natoms x 3, Eh/bohr) at the density of the last scf()
    // call; two-electron quartets screened at tau on Q Q max|Gamma|.
    // terms (optional, 5 x natoms x 3): nuclear repulsion, kinetic (D dT),
    // nuclear attraction (D dV), overlap (-W dS), two-electron.
    void gradient(double *grad_out, double tau, double *terms = nullptr);
    // ERI blocks of shell quartets quartets[4q..4q+3] (shell indices) in
    // Libint's layout, concatenated into out (host); offsets[q] (host, nq + 1
    // entries, may be null) is where block q starts.
    void eri_quartets(int nq, const int *quartets, double *out, long long *offsets);

    // Fock-build statistics accumulated over an SCF, index precision*6 + class
    // (precision 0 = FP64, 1 = FP32): quartets, milliseconds of the main ERI
    // kernel (HIP events on the engine stream right around it, filled when
    // profile_classes is set) and launches.
    std::array<unsigned long long, 12> prof_quartets{};
    std::array<unsigned long long, 12> prof_prims{};  // primitive quartets evaluated
    std::array<double, 12> prof_ms{};
    std::array<int, 12> prof_launches{};
    std::array<int, 24> prof_class_l{};  // per slot: la lb lc ld as launched (bra | ket)
    bool profile_classes = false;
    int jk_fallback = 0;  // pf_options::jk_fallback: bit 0 K rows, bit 1 J'_cd through device atomics
    bool deterministic = false;  // pf_options::deterministic: order-independent (int64 fixed-point) J/K sums

   private:
    void setup_one_electron();
    void setup_pairs();
    void setup_store();
    void setup_grad_pairs();
    // G (FP64) += or = G[D] in precision R.  D0/D1 device FP64 densities (D1 null for RHF).
    // tau > 0 overrides the Schwarz threshold (used for incremental builds).
    // Nothing returns to the host: quartet counts accumulate on the device and
    // the class timings in recorded events, both read by collect_profile.
    template <typename R>
    void build_g(const double *D0, const double *D1, double *G0, double *G1, bool accumulate, int mode,
                 double tau = 0.0);
    // Synchronises the stream, adds the builds since the last call to the
    // prof_* statistics and returns their shell-quartet count.
    unsigned long long collect_profile();
    // D from F (eps_dev: orbital energies by Jacobi as well).  With defer_slot
    // >= 0 and a step count from an earlier projector, SP2 runs a fixed number
    // of steps with no host round trip and returns false: its convergence
    // record is copied to sp2_rec_[defer_slot] for the caller to read with the
    // iteration's energy (and to redo this density if it is unset).
    bool density_from_fock(const double *F, int nocc, double occ, double *D, const pf_options &opt,
                           double *eps_dev, int defer_slot = -1);
    // Superposition of atomic densities (spin-summed, host, nbf^2 row-major).
    std::vector<double> sad_density();

    Molecule mol_;
    std::vector<std::string> species_;
    std::string basis_path_;
    bool renorm_ = true;
    Basis basis_;
    double schwarz_;
    hipStream_t st_ = nullptr;
    DevBuf<Shell> d_shells_;
    DevBuf<double> d_axyz_;
    DevBuf<int> d_Z_;
    DevBuf<double> d_tab_;
    DevBuf<float> f_tab_;
    DevBuf<double> S_, T_, V_, H_, X_;
    DevBuf<int> d_sh_bf_, d_sh_nf_;
    std::array<PairSet, 3> pairs_;
    std::vector<QuartetClass> classes_;
    bool pairs_ready_ = false;
    bool one_e_ready_ = false;
    // Shell pairs with a d shell (generic McMurchie-Davidson path): the kept
    // ones (Q * Qmax >= schwarz), first shell the higher l, sorted by Q, and
    // the ket list of their quartets (these pairs first, then every kept s/p pair).
    bool has_d_ = false;
    int ngen_ = 0, nket_gen_ = 0;
    DevBuf<int> d_gsa_, d_gsb_, d_ksa_, d_ksb_;
    DevBuf<double> d_gQ_, d_kQ_;
    GenQuartets gen_quartets() const;
    // J/K accumulators
    DevBuf<double> J64_, K64a_, K64b_, Dsh64_, dmax64_;
    DevBuf<double> fx_;   // deterministic builds: the fixed-point scale (device scalar)
    double qmax_ = 0.0;   // largest Schwarz bound of any shell pair
    DevBuf<float> Dsh32_, dmax32_, Dj32_, Dk32a_, Dk32b_;
    DevBuf<double> Dj64_;
    DevBuf<unsigned long long> count_;  // [precision][quartets 6, primitive quartets 6]
    struct PendingBuild {
        int pbase;
        hipEvent_t ev[12];  // per class: before and after its main ERI kernel
    };
    std::vector<PendingBuild> pending_;  // builds whose class timings are not yet read
    std::array<DevBuf<int>, 6> plan_nket_, plan_off_, plan_work_;  // direct-mode work plans per class
    std::array<DevBuf<long long>, 6> plan_joff_;
    std::array<DevBuf<int>, 6> plan_jhi_;
    std::array<PooledBlock, 6> jbuf_;  // per class: J'_cd of each evaluated quartet (pooled)
    std::array<long long, 6> jcap_{};
    static constexpr long long JBUF_MAX_VALUES = 1LL << 27;  // per class (1 GiB), beyond: device atomics
    // stored integrals
    bool store_ready_ = false;
    std::array<DevBuf<double>, 6> store_;
    std::array<DevBuf<long long>, 6> store_off_;
    std::array<DevBuf<int>, 6> store_nket_;
    // work
    DevBuf<double> w1_, w2_, w3_, w4_, state_;
    DevBuf<double> sp2_rec_;      // deferred purifications: convergence record per spin
    int sp2_hint_ = 7;            // purification steps the last projector needed
    bool sp2_hint_set_ = false;   // sp2_hint_ comes from a finished projector
    int sp2_defer_steps_ = 0;     // steps of the last deferred purification
    // converged densities of the last SCF (for the gradient)
    DevBuf<double> Dfin_;
    bool have_fin_ = false, uhf_fin_ = false;
    // gradient pair data: 2 alpha weighted, reversed 2 beta weighted, reversed
    bool grad_ready_ = false;
    std::array<DevBuf<double>, 3> gw_, gr_, grec_;
    // one-electron derivative pairs: ordered shell pairs by (la, lb) = (0,0), (0,1), (1,0), (1,1)
    DevBuf<int> g1pairs_;                // the four classes' (a, b) lists, a's then b's per class
    std::array<int, 4> g1n_{}, g1off_{};
};

// Thread-local error text for the C-ABI.
void set_last_error(const std::string &msg);

}  // namespace pf
