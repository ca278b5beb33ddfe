// SCF engine: integrals, Fock builds and the Roothaan/DIIS loop on one GPU.
//
// Control flow follows run_rscf (rscf.f90:49-164) and run_uscf
// (uscf.f90:65-213): core-Hamiltonian guess, then per iteration
// G[D] -> F = H + G -> D_old = D -> D from F -> E = 1/2 sum D (H + F)
// (UHF: 1/2 sum (Da+Db) H + 1/2 sum (Da Fa + Db Fb)) -> stop when
// |dE| < eps and ||dD||_F < eps.  Everything between the one-time setup and
// the two scalars read back per iteration stays on the device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <thread>

#include "boys.h"
#include "engine.h"
#include "grad_kernel.h"
#include "linalg.h"

namespace pf {

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int round64(int n) { return ((n + 63) / 64) * 64; }

__global__ void scale_cols_kernel(int n, const double *U, const double *s, double *X) {
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) X[(size_t)i * n + j] = U[(size_t)i * n + j] / sqrt(s[j]);
}

}  // namespace

PairView<double> PairSet::view64() const {
    PairView<double> v;
    v.n = n; v.npad = npad; v.kp = kp;
    v.sa = d_sa.p; v.sb = d_sb.p; v.oa = d_oa.p; v.ob = d_ob.p;
    v.Q = d_Q.p; v.AB = d_AB.p; v.prim = d_prim.p; v.aos = d_aos.p;
    v.kpn = d_kpn.p;
    v.peps = peps;
    return v;
}
PairView<float> PairSet::view32() const {
    PairView<float> v;
    v.n = n; v.npad = npad; v.kp = kp;
    v.sa = d_sa.p; v.sb = d_sb.p; v.oa = d_oa.p; v.ob = d_ob.p;
    v.Q = f_Q.p; v.AB = f_AB.p; v.prim = f_prim.p; v.aos = f_aos.p;
    v.kpn = d_kpn.p;
    v.peps = float(peps);
    return v;
}

Engine::Engine(const Molecule &mol, const std::vector<std::string> &species, const std::string &basis_path,
               bool renormalise, double schwarz)
    : mol_(mol), species_(species), basis_path_(basis_path), renorm_(renormalise), schwarz_(schwarz) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error("no GPU device visible");
    basis_ = build_basis(basis_path, species, mol_, renormalise);
    for (const auto &sh : basis_.shells) has_d_ = has_d_ || sh.l >= 2;
    PF_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    const int nsh = nshell();
    d_shells_.upload(basis_.shells.data(), nsh, st_);
    d_axyz_.upload(mol_.xyz.data(), mol_.xyz.size(), st_);
    d_Z_.upload(mol_.Z.data(), mol_.Z.size(), st_);
    static const std::vector<double> tab = boys_table_host();  // computed once per process
    static const std::vector<float> ftab(tab.begin(), tab.end());
    d_tab_.upload(tab.data(), tab.size(), st_);
    f_tab_.upload(ftab.data(), ftab.size(), st_);
    std::vector<int> bf(nsh), nf(nsh);
    for (int s = 0; s < nsh; ++s) { bf[s] = basis_.shells[s].bf; nf[s] = ncart(basis_.shells[s].l); }
    d_sh_bf_.upload(bf.data(), nsh, st_);
    d_sh_nf_.upload(nf.data(), nsh, st_);
    PF_HIP(hipStreamSynchronize(st_));
}

Engine::~Engine() {
    if (st_) {
        (void)hipStreamSynchronize(st_);
        (void)hipStreamDestroy(st_);
    }
    for (auto &b : jbuf_) pool_release(b);  // the stream is idle: nothing pending uses them
    for (auto &b : pending_)                  // class timings of builds never collected (an SCF that threw)
        for (auto &e : b.ev) event_release(e);
}

void Engine::setup_one_electron() {
    if (one_e_ready_) return;
    const size_t n2 = (size_t)nbf() * nbf();
    S_.alloc(n2); T_.alloc(n2); V_.alloc(n2); H_.alloc(n2);
    launch_onee(d_shells_.p, nshell(), nbf(), mol_.natoms, d_axyz_.p, d_Z_.p, d_tab_.p, S_.p, T_.p, V_.p, st_);
    if (has_d_) {  // pairs with a d shell: generic kernel
        std::vector<int> pa, pb;
        const auto &sh = basis_.shells;
        for (int a = 0; a < nshell(); ++a)
            for (int b = 0; b <= a; ++b)
                if (sh[a].l >= 2 || sh[b].l >= 2) { pa.push_back(a); pb.push_back(b); }
        DevBuf<int> dpa, dpb;
        dpa.upload(pa.data(), pa.size(), st_);
        dpb.upload(pb.data(), pb.size(), st_);
        launch_onee_generic(d_shells_.p, dpa.p, dpb.p, (int)pa.size(), nbf(), mol_.natoms, d_axyz_.p, d_Z_.p,
                            d_tab_.p, S_.p, T_.p, V_.p, st_);
        PF_HIP(hipStreamSynchronize(st_));  // dpa, dpb go back to the block cache
    }
    dev_axpby(n2, 1.0, T_.p, 1.0, V_.p, H_.p, st_);
    one_e_ready_ = true;
}

void Engine::one_electron(double *S, double *T, double *V) {
    setup_one_electron();
    const size_t n2 = (size_t)nbf() * nbf();
    if (S) S_.download(S, n2, st_);
    if (T) T_.download(T, n2, st_);
    if (V) V_.download(V, n2, st_);
    PF_HIP(hipStreamSynchronize(st_));
}

// Indices of q (all >= 0) in the order of q descending, ties by index: the
// order of a stable sort, by an LSD radix sort on the IEEE bit patterns
// (monotonic for non-negative doubles; complemented for descending order),
// 8 bits per pass, passes whose digit is the same for every key skipped.
// (std::sort with a comparator took ~0.8 ms for the 33k (s,s) pairs of
// (H2O)64 on the host, the largest host gap of an SCF's set-up.)
static std::vector<int> order_by_q_desc(const std::vector<double> &q) {
    const size_t n = q.size();
    std::vector<uint64_t> key(n), key2(n);
    std::vector<int> idx(n), idx2(n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t b;
        std::memcpy(&b, &q[i], sizeof(b));
        key[i] = ~b;
        idx[i] = (int)i;
    }
    for (int shift = 0; shift < 64; shift += 8) {
        size_t count[257] = {0};
        for (size_t i = 0; i < n; ++i) ++count[((key[i] >> shift) & 0xFF) + 1];
        bool trivial = false;
        for (int d = 1; d <= 256; ++d)
            if (count[d] == n) trivial = true;
        if (trivial) continue;
        for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
        for (size_t i = 0; i < n; ++i) {
            const size_t o = count[(key[i] >> shift) & 0xFF]++;
            key2[o] = key[i];
            idx2[o] = idx[i];
        }
        key.swap(key2);
        idx.swap(idx2);
    }
    return idx;
}

// Shell pairs: class 0 = (s,s), 1 = (p,s), 2 = (p,p); the first shell of a
// pair has the higher angular momentum (ties: the higher shell index).
void Engine::setup_pairs() {
    if (pairs_ready_) return;
    const int nsh = nshell();
    const auto &sh = basis_.shells;
    std::array<std::vector<std::pair<int, int>>, 3> all;
    std::vector<std::pair<int, int>> gall;  // pairs with a d shell
    for (int a = 0; a < nsh; ++a)
        for (int b = 0; b <= a; ++b) {
            int f = a, s = b;
            if (sh[b].l > sh[a].l) std::swap(f, s);
            if (sh[f].l >= 2) gall.push_back({f, s});
            else all[sh[f].l + sh[s].l].push_back({f, s});
        }
    double bound_max = 0.0;
    auto fill = [&](PairSet &ps, const std::vector<std::pair<int, int>> &list, double kref, bool want32) {
        fill_pairset(ps, sh, d_shells_.p, list, kref, bound_max, want32, st_);
    };
    std::vector<double> gQall(gall.size());
    if (!gall.empty()) {
        std::vector<int> ga(gall.size()), gb(gall.size());
        for (size_t k = 0; k < gall.size(); ++k) { ga[k] = gall[k].first; gb[k] = gall[k].second; }
        DevBuf<int> dga, dgb;
        DevBuf<double> dq(gall.size());
        dga.upload(ga.data(), ga.size(), st_);
        dgb.upload(gb.data(), gb.size(), st_);
        launch_schwarz_generic(d_shells_.p, dga.p, dgb.p, (int)gall.size(), d_tab_.p, dq.p, st_);
        dq.download(gQall.data(), gall.size(), st_);
        PF_HIP(hipStreamSynchronize(st_));
    }
    const int cls_l[3][2] = {{0, 0}, {1, 0}, {1, 1}};
    // pass 1: all pairs, Schwarz bounds
    std::array<std::vector<double>, 3> Qall;
    for (int c = 0; c < 3; ++c) {
        PairSet &ps = pairs_[c];
        ps.la = cls_l[c][0]; ps.lb = cls_l[c][1];
        fill(ps, all[c], -1.0, false);  // FP64 only: the bounds need no FP32 copy
        ps.d_Q.alloc(std::max(ps.n, 1));
        launch_schwarz<double>(ps.la, ps.lb, ps.view64(), d_tab_.p, ps.d_Q.p, st_);
        Qall[c].resize(ps.n);
        ps.d_Q.download(Qall[c].data(), ps.n, st_);
    }
    PF_HIP(hipStreamSynchronize(st_));
    double qmax = 0.0;
    for (auto &q : Qall) for (double v : q) qmax = std::max(qmax, v);
    for (double v : gQall) qmax = std::max(qmax, v);
    qmax_ = qmax;
    // pass 2: keep pairs that can pass Q_ab Q_max >= tau, sorted by Q descending
    for (int c = 0; c < 3; ++c) {
        // pairs by Q descending, ties by index; those that can pass Q_ab Q_max >= tau
        std::vector<std::pair<int, int>> kept;
        std::vector<double> q;
        for (int i : order_by_q_desc(Qall[c])) {
            if (Qall[c][i] * qmax < schwarz_) break;  // descending: the rest fail too
            kept.push_back(all[c][i]);
            q.push_back(Qall[c][i]);
        }
        PairSet &ps = pairs_[c];
        fill(ps, kept, bound_max, true);
        ps.Q = q;
        ps.d_Q.upload(q.data(), q.size(), st_);
        std::vector<float> fq(q.begin(), q.end());
        // FP32 bound rounded up so FP32 screening never drops what FP64 keeps
        for (auto &v : fq) v = std::nextafter(v, 1e30f);
        ps.f_Q.upload(fq.data(), fq.size(), st_);
    }
    if (!gall.empty()) {
        // kept d pairs sorted by Q descending; kets: those pairs, then every kept s/p pair
        std::vector<int> ksa, ksb;
        std::vector<double> kq;
        for (int i : order_by_q_desc(gQall)) {
            if (gQall[i] * qmax < schwarz_) break;
            ksa.push_back(gall[i].first);
            ksb.push_back(gall[i].second);
            kq.push_back(gQall[i]);
        }
        ngen_ = (int)ksa.size();
        d_gsa_.upload(ksa.data(), ksa.size(), st_);
        d_gsb_.upload(ksb.data(), ksb.size(), st_);
        d_gQ_.upload(kq.data(), kq.size(), st_);
        for (int c = 0; c < 3; ++c) {
            ksa.insert(ksa.end(), pairs_[c].sa.begin(), pairs_[c].sa.end());
            ksb.insert(ksb.end(), pairs_[c].sb.begin(), pairs_[c].sb.end());
            kq.insert(kq.end(), pairs_[c].Q.begin(), pairs_[c].Q.end());
        }
        nket_gen_ = (int)ksa.size();
        d_ksa_.upload(ksa.data(), ksa.size(), st_);
        d_ksb_.upload(ksb.data(), ksb.size(), st_);
        d_kQ_.upload(kq.data(), kq.size(), st_);
    }
    // One launch per unordered class.
    classes_.clear();
    int slot = 0;
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y <= x; ++y, ++slot) {
            // bra = the pair class of higher angular momentum: its exchange
            // rows are summed on chip, only J'_cd of the ket goes to memory
            QuartetClass qc{x, y, slot};
            classes_.push_back(qc);
            const int l4[4] = {pairs_[qc.bra].la, pairs_[qc.bra].lb, pairs_[qc.ket].la, pairs_[qc.ket].lb};
            for (int k = 0; k < 4; ++k) prof_class_l[slot * 4 + k] = l4[k];
        }
    const size_t n2 = (size_t)nbf() * nbf();
    J64_.alloc(n2); K64a_.alloc(n2); K64b_.alloc(n2); Dj64_.alloc(n2);
    Dsh64_.alloc((size_t)nshell() * nshell()); dmax64_.alloc(1);
    count_.alloc(24);
    PF_HIP(hipMemsetAsync(count_.p, 0, 24 * sizeof(unsigned long long), st_));
    for (size_t q = 0; q < classes_.size(); ++q) {  // direct-mode work plans
        const PairSet &B = pairs_[classes_[q].bra], &K = pairs_[classes_[q].ket];
        const int nb = std::max(B.n, 1);
        plan_nket_[q].alloc(nb); plan_off_[q].alloc(nb + 1); plan_work_[q].alloc(3); plan_joff_[q].alloc(nb);
        // J'_cd buffer: every (bra, ket) slot the largest possible prefixes can reach, capped
        const bool same = classes_[q].bra == classes_[q].ket;
        const long long full = same ? (long long)B.n * (B.n + 1) / 2 : (long long)B.n * K.n;
        const long long ncd = (long long)ncart(K.la) * ncart(K.lb);
        jcap_[q] = std::min<long long>(full, JBUF_MAX_VALUES / ncd);
        jbuf_[q] = pool_acquire(sizeof(double) * (size_t)std::max<long long>(jcap_[q] * ncd, 1));
        plan_jhi_[q].alloc(std::max(K.n, 1));
    }
    PF_HIP(hipStreamSynchronize(st_));
    pairs_ready_ = true;
}

void Engine::schwarz_host(double *Q) {
    const int nsh = nshell();
    const auto &sh = basis_.shells;
    std::array<std::vector<std::pair<int, int>>, 3> all;
    for (int a = 0; a < nsh; ++a)
        for (int b = 0; b <= a; ++b) {
            int f = a, s = b;
            if (sh[b].l > sh[a].l) std::swap(f, s);
            all[sh[f].l + sh[s].l].push_back({f, s});
        }
    const int cls_l[3][2] = {{0, 0}, {1, 0}, {1, 1}};
    std::fill(Q, Q + (size_t)nsh * nsh, 0.0);
    for (int c = 0; c < 3; ++c) {
        PairSet ps;
        ps.la = cls_l[c][0]; ps.lb = cls_l[c][1];
        double bound_max = 0.0;
        fill_pairset(ps, sh, d_shells_.p, all[c], -1.0, bound_max, false, st_);
        if (ps.n == 0) continue;
        ps.d_Q.alloc(ps.n);
        launch_schwarz<double>(ps.la, ps.lb, ps.view64(), d_tab_.p, ps.d_Q.p, st_);
        std::vector<double> q(ps.n);
        ps.d_Q.download(q.data(), ps.n, st_);
        PF_HIP(hipStreamSynchronize(st_));
        for (int i = 0; i < ps.n; ++i) {
            Q[(size_t)ps.sa[i] * nsh + ps.sb[i]] = q[i];
            Q[(size_t)ps.sb[i] * nsh + ps.sa[i]] = q[i];
        }
    }
}

long long Engine::schwarz_pairs() const {
    long long n = 0;
    for (auto &p : pairs_) n += p.n;
    return n;
}

GenQuartets Engine::gen_quartets() const {
    GenQuartets q;
    q.shells = d_shells_.p;
    q.bsa = d_gsa_.p; q.bsb = d_gsb_.p; q.bQ = d_gQ_.p; q.nbra = ngen_;
    q.ksa = d_ksa_.p; q.ksb = d_ksb_.p; q.kQ = d_kQ_.p; q.nket = nket_gen_; q.nsame = ngen_;
    q.tab = d_tab_.p;
    return q;
}

void Engine::setup_store() {
    if (store_ready_) return;
    if (has_d_) throw Error("stored ERIs: s and p shells only (d shells run direct)");
    setup_pairs();
    const int ints[3] = {1, 3, 9};
    for (size_t q = 0; q < classes_.size(); ++q) {
        const PairSet &B = pairs_[classes_[q].bra], &K = pairs_[classes_[q].ket];
        const bool same = classes_[q].bra == classes_[q].ket;
        const int nint = ints[classes_[q].bra] * ints[classes_[q].ket];
        std::vector<long long> off(std::max(B.n, 1), 0);
        std::vector<int> nk(std::max(B.n, 1), 0);
        long long tot = 0;
        for (int i = 0; i < B.n; ++i) {
            const double thr = schwarz_ / B.Q[i];
            int lo = 0, hi = K.n;
            while (lo < hi) {
                const int mid = (lo + hi) / 2;
                if (K.Q[mid] >= thr) lo = mid + 1; else hi = mid;
            }
            nk[i] = same ? std::min(lo, i + 1) : lo;
            off[i] = tot;
            tot += (long long)round64(nk[i]) * nint;
        }
        store_[q].alloc(std::max<long long>(tot, 1));
        store_off_[q].upload(off.data(), off.size(), st_);
        store_nket_[q].upload(nk.data(), nk.size(), st_);
        JKArgs<double> a;
        a.bra = B.view64(); a.ket = K.view64(); a.same = same; a.tau = schwarz_;
        a.nbf = nbf(); a.nsh = nshell(); a.tab = d_tab_.p;
        a.store = store_[q].p; a.store_off = store_off_[q].p; a.store_nket = store_nket_[q].p;
        launch_jk_class<double>(pairs_[classes_[q].bra].la, pairs_[classes_[q].bra].lb, pairs_[classes_[q].ket].la,
                                pairs_[classes_[q].ket].lb, 1, MODE_STORE, a, st_);
    }
    PF_HIP(hipStreamSynchronize(st_));
    store_ready_ = true;
}

template <typename R>
void Engine::build_g(const double *D0, const double *D1, double *G0, double *G1, bool accumulate,
                                   int mode, double tau) {
    setup_pairs();
    const int n = nbf(), nsh = nshell();
    const size_t n2 = (size_t)n * n;
    const bool uhf = D1 != nullptr;
    JKArgs<R> a;
    a.nbf = n; a.nsh = nsh; a.tau = R(tau > 0 ? tau : schwarz_);
    double *J = J64_.p, *Ka = K64a_.p, *Kb = K64b_.p;
    dev_fill(n2, 0.0, J, st_); dev_fill(n2, 0.0, Ka, st_);
    if (uhf) dev_fill(n2, 0.0, Kb, st_);
    if constexpr (std::is_same<R, double>::value) {
        if (uhf) {
            dev_axpby(n2, 1.0, D0, 1.0, D1, Dj64_.p, st_);
            a.Dj = Dj64_.p; a.Dk0 = D0; a.Dk1 = D1;
        } else {
            a.Dj = D0; a.Dk0 = D0;
        }
        a.tab = d_tab_.p;
        dev_shell_dmax<double>(nsh, n, d_sh_bf_.p, d_sh_nf_.p, D0, D1, Dsh64_.p, dmax64_.p, st_);
        a.Dsh = Dsh64_.p; a.dmax = dmax64_.p;
    } else {
        if (!Dj32_.p) {
            Dj32_.alloc(n2); Dk32a_.alloc(n2); Dk32b_.alloc(n2);
            Dsh32_.alloc((size_t)nsh * nsh); dmax32_.alloc(1);
        }
        if (uhf) {
            dev_to_float(n2, 1.0, D0, D1, Dj32_.p, st_);
            dev_to_float(n2, 1.0, D0, nullptr, Dk32a_.p, st_);
            dev_to_float(n2, 1.0, D1, nullptr, Dk32b_.p, st_);
            a.Dj = Dj32_.p; a.Dk0 = Dk32a_.p; a.Dk1 = Dk32b_.p;
        } else {
            dev_to_float(n2, 1.0, D0, nullptr, Dj32_.p, st_);
            a.Dj = Dj32_.p; a.Dk0 = Dj32_.p;
        }
        a.tab = f_tab_.p;
        dev_shell_dmax<float>(nsh, n, d_sh_bf_.p, d_sh_nf_.p, D0, D1, Dsh32_.p, dmax32_.p, st_);
        a.Dsh = Dsh32_.p; a.dmax = dmax32_.p;
    }
    a.J = J; a.K0 = Ka; a.K1 = uhf ? Kb : nullptr;
    if (deterministic) {
        if (mode != MODE_DIRECT) throw Error("deterministic J/K: direct builds only (eri_mode 0)");
        if (ngen_ > 0) throw Error("deterministic J/K: s and p shells only");
        if (!fx_.p) fx_.alloc(1);
        dev_fx_scale((int64_t)n2, D0, D1, qmax_, fx_.p, st_);
        a.fx = fx_.p;
    }
    const int pbase = std::is_same<R, double>::value ? 0 : 6;
    PendingBuild pb{pbase, {}};
    hipEvent_t *ev = pb.ev;
    if (profile_classes)
        for (int e = 0; e < 12; ++e) ev[e] = event_acquire();
    // direct builds: every class's work plan in two launches before the
    // direct kernels, the J'_cd sums in two after them
    const bool planned = mode == MODE_DIRECT;
    PlanArgs<R> plan;
    // Work items grow with the basis: an item's fixed cost is the flush of its
    // exchange rows, nbf long, while short items balance the end of a launch.
    // Measured (mixed / FP64 SCF): (H2O)64, 448 bf, 49.2 / 112.8 ms at scale 1,
    // 49.9 / 115.9 at 2; gly60, 1387 bf, 508 / 983 ms at 1, 474 / 944 at 3,
    // 461 / 938 at 4; rna4, 1943 bf, 2937 / 7391 ms at 1, 2369 / 6537 at 4,
    // 2388 / 6769 at 8.
    const int chunk_scale = std::min(4, std::max(1, n / 448));
    if (planned) {
        plan.n = (int)classes_.size();
        plan.tau = a.tau;
        plan.dmax = a.dmax;
        plan.J = J;
        plan.nbf = n;
        plan.fx = a.fx;
        for (size_t q = 0; q < classes_.size(); ++q) {
            const PairSet &B = pairs_[classes_[q].bra], &K = pairs_[classes_[q].ket];
            PlanClass<R> &c = plan.c[q];
            if constexpr (std::is_same<R, double>::value) { c.bra = B.view64(); c.ket = K.view64(); }
            else { c.bra = B.view32(); c.ket = K.view32(); }
            c.same = classes_[q].bra == classes_[q].ket;
            c.chunk = direct_chunk<R>(B.la, B.lb, K.la, K.lb, chunk_scale);
            c.ncd = ncart(K.la) * ncart(K.lb);
            c.nket = plan_nket_[q].p; c.item_off = plan_off_[q].p; c.work = plan_work_[q].p;
            c.jhi = plan_jhi_[q].p; c.joff = plan_joff_[q].p;
            c.jbuf = static_cast<const double *>(jbuf_[q].p);
            c.jcap = c.jbuf ? ((jk_fallback & 2) ? 0 : jcap_[q]) : -1;
        }
        launch_plan<R>(plan, st_);
    }
    for (size_t q = 0; q < classes_.size(); ++q) {
        const PairSet &B = pairs_[classes_[q].bra], &K = pairs_[classes_[q].ket];
        if constexpr (std::is_same<R, double>::value) { a.bra = B.view64(); a.ket = K.view64(); }
        else { a.bra = B.view32(); a.ket = K.view32(); }
        a.same = classes_[q].bra == classes_[q].ket;
        if (mode == MODE_LOAD) {
            if constexpr (std::is_same<R, double>::value) {
                a.store = store_[q].p; a.store_off = store_off_[q].p; a.store_nket = store_nket_[q].p;
            }
        }
        a.count = count_.p + 2 * pbase + classes_[q].slot;
        a.pcount = count_.p + 2 * pbase + 6 + classes_[q].slot;
        a.nket = plan_nket_[q].p; a.item_off = plan_off_[q].p; a.work = plan_work_[q].p;
        a.joff = plan_joff_[q].p; a.jbuf = static_cast<double *>(jbuf_[q].p); a.jhi = plan_jhi_[q].p;
        a.jcap = (jk_fallback & 2) ? 0 : jcap_[q];  // 0: no ket list fits, J'_cd through device atomics
        a.katomic = (jk_fallback & 1) ? 1 : 0;
        a.planned = planned ? 1 : 0;
        a.chunk = planned ? plan.c[q].chunk : 0;
        a.ev0 = profile_classes ? ev[2 * q] : nullptr;
        a.ev1 = profile_classes ? ev[2 * q + 1] : nullptr;
        launch_jk_class<R>(B.la, B.lb, K.la, K.lb, uhf ? 2 : 1, mode, a, st_);
    }
    if (planned) launch_jcd_sums<R>(plan, st_);
    if (ngen_ > 0) {
        // quartets with a d shell, in FP64 in either precision (FP64 densities)
        const double *Dj = D0;
        if (uhf) {
            dev_axpby(n2, 1.0, D0, 1.0, D1, Dj64_.p, st_);
            Dj = Dj64_.p;
        }
        if constexpr (!std::is_same<R, double>::value)
            dev_shell_dmax<double>(nsh, n, d_sh_bf_.p, d_sh_nf_.p, D0, D1, Dsh64_.p, dmax64_.p, st_);
        launch_jk_generic(gen_quartets(), n, nsh, Dj, D0, uhf ? D1 : nullptr, Dsh64_.p, tau > 0 ? tau : schwarz_, J,
                          Ka, uhf ? Kb : nullptr, st_);
    }
    if (deterministic) {
        dev_fx_to_double((int64_t)n2, fx_.p, J, st_);
        dev_fx_to_double((int64_t)n2, fx_.p, Ka, st_);
        if (uhf) dev_fx_to_double((int64_t)n2, fx_.p, Kb, st_);
    }
    dev_assemble_g<double>(n, J, Ka, uhf ? Kb : nullptr, G0, G1, uhf, accumulate, st_);
    for (int q = 0; q < 6; ++q) prof_launches[pbase + q] += 1;
    if (profile_classes) pending_.push_back(pb);
}

unsigned long long Engine::collect_profile() {
    if (!count_.p) return 0;
    unsigned long long c[24];
    PF_HIP(hipMemcpyAsync(c, count_.p, sizeof(c), hipMemcpyDeviceToHost, st_));
    PF_HIP(hipStreamSynchronize(st_));
    PF_HIP(hipMemsetAsync(count_.p, 0, sizeof(c), st_));
    unsigned long long total = 0;
    for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 6; ++q) {
            prof_quartets[6 * p + q] += c[12 * p + q];
            prof_prims[6 * p + q] += c[12 * p + 6 + q];
            total += c[12 * p + q];
        }
    // taken out of pending_ first: a failed read below leaves its events
    // unreleased rather than released twice (by the destructor as well)
    std::vector<PendingBuild> done;
    done.swap(pending_);
    for (auto &b : done) {
        for (int q = 0; q < 6; ++q) {
            float ms = 0.f;
            PF_HIP(hipEventElapsedTime(&ms, b.ev[2 * q], b.ev[2 * q + 1]));
            prof_ms[b.pbase + classes_[q].slot] += ms;
        }
        for (auto &e : b.ev) event_release(e);
    }
    return total;
}

void Engine::jk_host(int ndens, const double *D, double *J, double *K, int mode, int precision) {
    setup_pairs();
    const int n = nbf();
    const size_t n2 = (size_t)n * n;
    DevBuf<double> dD(n2 * ndens), G0(n2), G1(n2);
    dD.upload(D, n2 * ndens, st_);
    int kmode = MODE_DIRECT;
    if (mode == PF_ERI_STORED) {
        if (precision != PF_PREC_FP64) throw Error("stored ERIs are FP64 only");
        setup_store();
        kmode = MODE_LOAD;
    }
    // J and K separately: run with accumulate=false, then read the raw accumulators
    const double *D1 = ndens == 2 ? dD.p + n2 : nullptr;
    if (precision == PF_PREC_FP64) build_g<double>(dD.p, D1, G0.p, G1.p, false, kmode);
    else build_g<float>(dD.p, D1, G0.p, G1.p, false, kmode);
    // symmetrised J', K' -> host
    std::vector<double> j(n2), k0(n2), k1(ndens == 2 ? n2 : 0);
    J64_.download(j.data(), n2, st_);
    K64a_.download(k0.data(), n2, st_);
    if (ndens == 2) K64b_.download(k1.data(), n2, st_);
    PF_HIP(hipStreamSynchronize(st_));
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            J[(size_t)r * n + c] = j[(size_t)r * n + c] + j[(size_t)c * n + r];
            K[(size_t)r * n + c] = k0[(size_t)r * n + c] + k0[(size_t)c * n + r];
            if (ndens == 2) K[n2 + (size_t)r * n + c] = k1[(size_t)r * n + c] + k1[(size_t)c * n + r];
        }
}

void Engine::eri_packed_host(double *packed) {
    setup_pairs();
    const long long np = (long long)nbf() * (nbf() + 1) / 2;
    const long long total = np * (np + 1) / 2;
    DevBuf<double> d(total);
    dev_fill(total, 0.0, d.p, st_);
    for (size_t q = 0; q < classes_.size(); ++q) {
        const PairSet &B = pairs_[classes_[q].bra], &K = pairs_[classes_[q].ket];
        JKArgs<double> a;
        a.bra = B.view64(); a.ket = K.view64(); a.same = classes_[q].bra == classes_[q].ket;
        a.nbf = nbf(); a.nsh = nshell(); a.tab = d_tab_.p; a.packed = d.p;
        launch_jk_class<double>(B.la, B.lb, K.la, K.lb, 1, MODE_PACKED, a, st_);
    }
    if (ngen_ > 0) launch_packed_generic(gen_quartets(), d.p, st_);
    d.download(packed, total, st_);
    PF_HIP(hipStreamSynchronize(st_));
}

// D = occ * X P X^T with P the projector onto the nocc lowest eigenvectors
// of F' = X^T F X (la.f90:52-54 and rscf.f90:140 / uscf.f90:183-184).
bool Engine::density_from_fock(const double *F, int nocc, double occ, double *D, const pf_options &opt,
                               double *eps_dev, int defer_slot) {
    const int n = nbf();
    gemm_f64(n, false, false, 1.0, F, X_.p, 0.0, w1_.p, st_);   // F X
    gemm_f64(n, true, false, 1.0, X_.p, w1_.p, 0.0, w2_.p, st_);  // X^T F X
    if (nocc == 0) {
        dev_fill((size_t)n * n, 0.0, D, st_);
        if (eps_dev) {
            DevBuf<double> C(size_t(n) * n);
            jacobi_eigh(n, w2_.p, eps_dev, C.p, st_);
        }
        return true;
    }
    auto by_eigenvectors = [&]() {
        DevBuf<double> w(n), C(size_t(n) * n);
        jacobi_eigh(n, w2_.p, eps_dev ? eps_dev : w.p, C.p, st_);
        gemm_f64(n, false, false, 1.0, X_.p, C.p, 0.0, w1_.p, st_);  // C = X C'
        gemm_f64_mnk(n, n, nocc, false, true, occ, w1_.p, n, w1_.p, n, 0.0, D, n, st_);
    };
    if (opt.diag == PF_DIAG_JACOBI || eps_dev) {
        by_eigenvectors();
        return true;
    }
    constexpr int SP2_MAX = 100, SP2_SLACK = 3;
    if (defer_slot >= 0 && sp2_hint_set_) {
        // the last projector's step count plus a margin, no host round trip
        const int nsteps = std::min(SP2_MAX, sp2_hint_ + SP2_SLACK);
        sp2_project_fixed(n, w2_.p, double(nocc), w3_.p, w1_.p, w4_.p, state_.p, nsteps, st_);
        // record of the last step: steps since convergence (0: not converged)
        PF_HIP(hipMemcpyAsync(sp2_rec_.p + defer_slot, state_.p + 3 + 2 * (nsteps % 3), sizeof(double),
                              hipMemcpyDeviceToDevice, st_));
        sp2_defer_steps_ = nsteps;
        gemm_f64(n, false, false, 1.0, X_.p, w3_.p, 0.0, w1_.p, st_);  // X P
        gemm_f64(n, false, true, occ, w1_.p, X_.p, 0.0, D, st_);       // occ X P X^T
        return false;
    }
    const int used = sp2_density(n, w2_.p, double(nocc), w3_.p, w1_.p, state_.p, st_, SP2_MAX, sp2_hint_ + 1);
    if (used >= SP2_MAX) {
        // No projector: the occupied/virtual boundary falls inside a degenerate
        // level (an open shell from a spin-symmetric start, such as the pi pair
        // of OH), where purification leaves fractional occupations and only
        // summation-order noise would break the symmetry.  The eigenvector
        // route occupies whole orbitals.
        by_eigenvectors();
        return true;
    }
    sp2_hint_ = used;
    sp2_hint_set_ = true;
    gemm_f64(n, false, false, 1.0, X_.p, w3_.p, 0.0, w1_.p, st_);       // X P
    gemm_f64(n, false, true, occ, w1_.p, X_.p, 0.0, D, st_);            // occ X P X^T
    return true;
}

// Spin-summed density of the isolated atom of element `el` (atomic number Z)
// in the basis at `path`: UHF with n_alpha = ceil(Z/2), n_beta = floor(Z/2)
// (Jacobi densities: robust for one-function atoms), then spherically
// averaged shell block by shell block -- s-p blocks vanish, a p-p block keeps
// a third of its trace on the diagonal -- so the guess has the atom's
// symmetry whatever orbitals the open shell settled in.  Cached per (basis
// file, normalisation, element); the atom's basis functions in shell order.
std::vector<double> atomic_density(const std::string &path, bool renorm, const std::string &el, int Z) {
    static std::mutex mu;
    static std::map<std::tuple<std::string, bool, std::string>, std::vector<double>> cache;
    const auto key = std::make_tuple(path, renorm, el);
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    Molecule atom;
    atom.natoms = 1;
    atom.Z = {Z};
    atom.xyz = {0.0, 0.0, 0.0};
    Engine eng(atom, {el}, path, renorm, 1e-14);
    pf_options o;
    pf_default_options(&o);
    o.eps = 1e-9;
    o.max_iter = 200;
    o.diis = 8;
    o.diag = PF_DIAG_JACOBI;
    const int na = (Z + 1) / 2, nb = Z / 2, n = eng.nbf();
    const size_t n2 = (size_t)n * n;
    std::vector<double> Ds(2 * n2, 0.0), D(n2, 0.0);
    pf_result r;
    eng.scf(na, nb, o, r, Ds.data(), nullptr);
    const bool uhf = na != nb;
    for (size_t e = 0; e < n2; ++e) D[e] = uhf ? Ds[e] + Ds[n2 + e] : Ds[e];
    const auto &sh = eng.basis().shells;
    for (size_t i = 0; i < sh.size(); ++i)
        for (size_t j = 0; j < sh.size(); ++j) {
            const int ni = ncart(sh[i].l), nj = ncart(sh[j].l);
            if (sh[i].l == 0 && sh[j].l == 0) continue;
            double tr = 0.0;
            if (sh[i].l == sh[j].l)
                for (int k = 0; k < ni; ++k) tr += D[(size_t)(sh[i].bf + k) * n + sh[j].bf + k];
            for (int a = 0; a < ni; ++a)
                for (int b = 0; b < nj; ++b)
                    D[(size_t)(sh[i].bf + a) * n + sh[j].bf + b] =
                        (sh[i].l == sh[j].l && a == b) ? tr / ni : 0.0;
        }
    std::lock_guard<std::mutex> lock(mu);
    cache[key] = D;
    return D;
}

std::vector<double> sad_density(const Shell *sh, int nsh, int nbf, const std::vector<std::string> &species,
                                const std::vector<int> &Z, const std::string &path, bool renorm) {
    const int n = nbf;
    std::vector<double> D((size_t)n * n, 0.0);
    for (int a = 0; a < (int)species.size(); ++a) {
        const std::vector<double> Da = atomic_density(path, renorm, species[a], Z[a]);
        // the atom's shells in the molecule, in the order they have on the isolated atom
        std::vector<int> mine;
        for (int s = 0; s < nsh; ++s)
            if (sh[s].atom == a) mine.push_back(s);
        int na = 0;
        std::vector<int> off;
        for (int s : mine) { off.push_back(na); na += ncart(sh[s].l); }
        if ((size_t)na * na != Da.size()) throw Error("SAD guess: atomic basis mismatch for " + species[a]);
        for (size_t i = 0; i < mine.size(); ++i)
            for (size_t j = 0; j < mine.size(); ++j)
                for (int p = 0; p < ncart(sh[mine[i]].l); ++p)
                    for (int q = 0; q < ncart(sh[mine[j]].l); ++q)
                        D[(size_t)(sh[mine[i]].bf + p) * n + sh[mine[j]].bf + q] =
                            Da[(size_t)(off[i] + p) * na + off[j] + q];
    }
    return D;
}

std::vector<double> Engine::sad_density() {
    return pf::sad_density(basis_.shells.data(), nshell(), nbf(), species_, mol_.Z, basis_path_, renorm_);
}

void Engine::scf(int nalpha, int nbeta, const pf_options &opt, pf_result &res, double *D_out, double *eps_out) {
    const double t0 = now();
    res = pf_result{};
    const int n = nbf();
    const size_t n2 = (size_t)n * n;
    const bool uhf = nalpha != nbeta;
    const int nd = uhf ? 2 : 1;
    res.nbf = n;
    res.nshell = nshell();
    res.e_nuc = nuclear_repulsion(mol_);
    if (nalpha < 0 || nbeta < 0 || nalpha > n || nbeta > n)
        throw Error("electron count does not fit the basis (nalpha=" + std::to_string(nalpha) +
                    ", nbeta=" + std::to_string(nbeta) + ", nbf=" + std::to_string(n) + ")");
    if (opt.eri_mode == PF_ERI_STORED && opt.precision != PF_PREC_FP64)
        throw Error("stored ERIs are FP64 only");
    if (opt.jk_fallback < 0 || opt.jk_fallback > 3) throw Error("jk_fallback must be 0..3");
    jk_fallback = opt.jk_fallback;
    deterministic = opt.deterministic != 0;

    setup_one_electron();
    w1_.alloc(n2); w2_.alloc(n2); w3_.alloc(n2); w4_.alloc(n2); state_.alloc(sp2_state_size(n));
    sp2_rec_.alloc(2);
    // orthogonaliser with X X^T = S^{-1}, computed once: S^{-1/2} by Newton-Schulz,
    // or the reference's X = U s^{-1/2} (la.f90:36-50) from the Jacobi solver
    // when that iteration does not converge
    X_.alloc(n2);
    if (!inv_sqrt_spd(n, S_.p, X_.p, &res.min_overlap_eig, st_)) {
        DevBuf<double> s(n), U(n2);
        jacobi_eigh(n, S_.p, s.p, U.p, st_);
        std::vector<double> hs(n);
        s.download(hs.data(), n, st_);
        PF_HIP(hipStreamSynchronize(st_));
        res.min_overlap_eig = n ? hs[0] : 0.0;
        if (n && hs[0] <= 0.0) throw Error("overlap matrix is not positive definite (coincident atoms?)");
        dim3 g2((n + 127) / 128, n);
        scale_cols_kernel<<<g2, 128, 0, st_>>>(n, U.p, s.p, X_.p);
        PF_HIP(hipGetLastError());
    }
    setup_pairs();
    if (opt.eri_mode == PF_ERI_STORED) setup_store();
    PF_HIP(hipStreamSynchronize(st_));
    res.t_setup = now() - t0;

    const int nocc[2] = {uhf ? nalpha : nalpha, nbeta};
    const double occ = uhf ? 1.0 : 2.0;
    DevBuf<double> D(n2 * nd), Dold(n2 * nd), Dbuilt(n2 * nd), dD(n2 * nd), G(n2 * nd), F(n2 * nd), red(4);
    DevBuf<double> part((size_t)RED_BLOCKS * DotPtrs::MAX);  // reduction partials
    if (opt.guess == PF_GUESS_SAD) {
        // superposition of atomic densities: RHF D = the sum (occupation 2 folded
        // in), UHF D_alpha = D_beta = half of it
        const std::vector<double> sad = sad_density();
        D.upload(sad.data(), n2, st_);
        if (uhf) {
            dev_axpby(n2, 0.5, D.p, 0.0, D.p, D.p + n2, st_);
            dev_axpby(n2, 0.5, D.p, 0.0, D.p, D.p, st_);
        }
    } else {
        // core guess (rscf.f90:113-114; uscf.f90:140-144, C_alpha = C_beta)
        const double td = now();
        for (int s = 0; s < nd; ++s) density_from_fock(H_.p, nocc[s], occ, D.p + s * n2, opt, nullptr);
        if (opt.profile) {
            PF_HIP(hipStreamSynchronize(st_));
            res.t_diag += now() - td;
        }
    }
    dev_fill(n2 * nd, 0.0, Dbuilt.p, st_);
    dev_fill(n2 * nd, 0.0, G.p, st_);

    // DIIS storage
    const int ndiis = opt.diis > 1 ? std::min(opt.diis, 32) : 0;
    std::vector<DevBuf<double>> dF(ndiis), dE(ndiis);
    for (int k = 0; k < ndiis; ++k) { dF[k].alloc(n2 * nd); dE[k].alloc(n2 * nd); }
    DevBuf<double> dots(std::max(ndiis, 1)), Fx(ndiis ? n2 * nd : 1), Bdev(std::max(ndiis * ndiis, 1)),
        coef(std::max(ndiis, 1));
    DevBuf<int> dok(1);
    DevBuf<double> Fls(opt.level_shift > 0.0 ? n2 * nd : 1);  // level-shifted Fock
    DevBuf<double> dots_all(opt.precision == PF_PREC_MIXED && ndiis ? (size_t)ndiis * ndiis : 1);
    int nstore = 0, head = 0;

    bool fp32_phase = opt.precision == PF_PREC_MIXED;
    const bool incremental = opt.incremental || opt.precision == PF_PREC_MIXED;
    const int gmode = opt.eri_mode == PF_ERI_STORED ? MODE_LOAD : MODE_DIRECT;
    double Eold = 0.0, Eel = 0.0, last_dD = 1e300;
    int it = 0, since_full = 0;
    bool anchored = false;
    bool conv = false;
    // per iteration, when profiling: events at the start of the Fock build, its
    // end, and the end of the density/energy step (t_fock and t_diag without
    // host round trips)
    std::vector<std::array<hipEvent_t, 3>> tev;
    struct ReleaseEvents {  // back to the pool on every exit, also when the SCF throws
        std::vector<std::array<hipEvent_t, 3>> &v;
        ~ReleaseEvents() {
            for (auto &a : v)
                for (auto e : a) event_release(e);
        }
    } release_tev{tev};
    // DIIS error in the orthonormal basis on request, e' = X^T e X (in place; w1_
    // as scratch): the AO commutator weights each element by the overlap metric;
    // on large molecules with small gaps the orthonormal form can take fewer
    // steps (either may end on another stationary point of UHF)
    auto to_orth = [&](double *e) {
        if (opt.diis_basis != PF_DIIS_ORTHO) return;
        gemm_f64(n, true, false, 1.0, X_.p, e, 0.0, w1_.p, st_);
        gemm_f64(n, false, false, 1.0, w1_.p, X_.p, 0.0, e, st_);
    };
    auto mark = [&](int k) {
        // per-iteration events only when profiling: reading 3 per iteration back
        // costs ~0.8 ms of host time at the end of a 16-iteration SCF
        if (!opt.profile) return;
        if (k == 0) {
            tev.emplace_back();
            for (auto &e : tev.back()) e = event_acquire();
        }
        PF_HIP(hipEventRecord(tev.back()[k], st_));
    };
    for (it = 1; it <= opt.max_iter; ++it) {
        // ---- Fock build
        mark(0);
        const double *Dcur0 = D.p, *Dcur1 = uhf ? D.p + n2 : nullptr;
        double *G0 = G.p, *G1 = uhf ? G.p + n2 : nullptr;
        // Mixed precision: the FP32 rounding of an incremental build scales with
        // |dD|, so the first builds (dD as large as D itself) leave almost all
        // of the accumulated error.  One full FP64 rebuild when ||dD|| first
        // falls below mixed_anchor clears it; the switch to FP64
        // at mixed_switch then moves the density by little more than a normal
        // step.
        const bool switch_now = fp32_phase && last_dD < opt.mixed_switch;
        const bool anchor_now = fp32_phase && !switch_now && !anchored && last_dD < opt.mixed_anchor;
        const bool rebuild = !fp32_phase && incremental && opt.rebuild_every > 0 &&
                             since_full >= opt.rebuild_every;
        if (switch_now || anchor_now || rebuild || !incremental) {
            // full FP64 build: the anchor of the FP32 phase, or the switch to FP64 (the
            // last full build of a mixed SCF: every FP32 increment drops out of G, and
            // only FP64 increments follow it)
            DevBuf<double> Gs;
            // DIIS keeps its subspace: at the anchor the error vectors (~|dD|) dwarf
            // the offset, and those entries leave the subspace before it matters;
            // at a switch with no anchor before it, the offset is removed below
            const bool correct = switch_now && !anchored;
            if (anchor_now) anchored = true;
            if (switch_now) fp32_phase = false;
            if (correct) {
                if (nstore > 0) {
                    // the FP32 phase's G at the current density (one more incremental
                    // FP32 build), to measure the offset its accumulated rounding left
                    Gs.alloc(n2 * nd);
                    PF_HIP(hipMemcpyAsync(Gs.p, G.p, sizeof(double) * n2 * nd, hipMemcpyDeviceToDevice, st_));
                    dev_axpby(n2 * nd, 1.0, D.p, -1.0, Dbuilt.p, dD.p, st_);
                    build_g<float>(dD.p, uhf ? dD.p + n2 : nullptr, Gs.p, uhf ? Gs.p + n2 : nullptr, true,
                                   MODE_DIRECT);
                    ++res.fock_builds_fp32;
                }
            }
            build_g<double>(Dcur0, Dcur1, G0, G1, false, gmode);
            ++res.fock_builds_fp64;
            since_full = 0;
            if (Gs.p) {
                // The stored DIIS Fock matrices share (nearly) this offset
                // d = G_fp32 - G_fp64, accumulated mostly by the first builds:
                // remove it from each F_k
                // and its error vector, e_k -= d D S - S D d (or X^T . X; with the current D for
                // D_k, an error of |d| |D_k - D|), and rebuild the B matrix.  The
                // subspace then continues instead of restarting.
                dev_axpby(n2 * nd, 1.0, Gs.p, -1.0, G.p, Gs.p, st_);
                for (int s = 0; s < nd; ++s) {
                    gemm_f64(n, false, false, 1.0, Gs.p + s * n2, D.p + s * n2, 0.0, w1_.p, st_);
                    gemm_f64(n, false, false, 1.0, w1_.p, S_.p, 0.0, w2_.p, st_);
                    dev_antisym(n, w2_.p, dD.p + s * n2, st_);
                    to_orth(dD.p + s * n2);
                }
                for (int k = 0; k < nstore; ++k) {
                    dev_axpby(n2 * nd, 1.0, dF[k].p, -1.0, Gs.p, dF[k].p, st_);
                    dev_axpby(n2 * nd, 1.0, dE[k].p, -1.0, dD.p, dE[k].p, st_);
                }
                for (int r = 0; r < nstore; ++r)
                    for (int k0 = 0; k0 < nstore; k0 += DotPtrs::MAX) {
                        const int m = std::min(DotPtrs::MAX, nstore - k0);
                        DotPtrs vecs{};
                        for (int k = 0; k < m; ++k) vecs.p[k] = dE[k0 + k].p;
                        dev_dots(n2 * nd, dE[r].p, vecs, m, dots_all.p + (size_t)r * ndiis + k0, part.p, st_);
                    }
                PF_HIP(hipMemcpyAsync(Bdev.p, dots_all.p, sizeof(double) * (size_t)nstore * ndiis,
                                      hipMemcpyDeviceToDevice, st_));
            }
        } else {
            dev_axpby(n2 * nd, 1.0, D.p, -1.0, Dbuilt.p, dD.p, st_);
            const double *d0 = dD.p, *d1 = uhf ? dD.p + n2 : nullptr;
            if (fp32_phase) {
                // before the anchor rebuild (which replaces G) screening may be looser
                build_g<float>(d0, d1, G0, G1, true, MODE_DIRECT, anchored ? 0.0 : opt.schwarz_fp32);
                ++res.fock_builds_fp32;
            } else if (opt.precision == PF_PREC_MIXED && opt.mixed_finish == PF_FINISH_FP32) {
                // after the switch's full FP64 build: the increments' FP32 rounding
                // is ~1e-7 of G[dD] with ||dD||_F < mixed_switch, far below eps
                build_g<float>(d0, d1, G0, G1, true, MODE_DIRECT, opt.schwarz_inc);
                ++res.fock_builds_fp32;
            } else {
                build_g<double>(d0, d1, G0, G1, true, gmode, opt.schwarz_inc);
                ++res.fock_builds_fp64;
                ++since_full;
            }
        }
        if (incremental) PF_HIP(hipMemcpyAsync(Dbuilt.p, D.p, sizeof(double) * n2 * nd, hipMemcpyDeviceToDevice, st_));
        for (int s = 0; s < nd; ++s) dev_axpby(n2, 1.0, H_.p, 1.0, G.p + s * n2, F.p + s * n2, st_);
        mark(1);

        PF_HIP(hipMemcpyAsync(Dold.p, D.p, sizeof(double) * n2 * nd, hipMemcpyDeviceToDevice, st_));
        const double *Fuse = F.p;
        if (ndiis) {
            // error e = F D S - S D F per spin (or X^T e X), with the density F was built from
            const int slot = head;
            for (int s = 0; s < nd; ++s) {
                gemm_f64(n, false, false, 1.0, F.p + s * n2, D.p + s * n2, 0.0, w1_.p, st_);
                gemm_f64(n, false, false, 1.0, w1_.p, S_.p, 0.0, w2_.p, st_);
                dev_antisym(n, w2_.p, dE[slot].p + s * n2, st_);
                to_orth(dE[slot].p + s * n2);
            }
            PF_HIP(hipMemcpyAsync(dF[slot].p, F.p, sizeof(double) * n2 * nd, hipMemcpyDeviceToDevice, st_));
            head = (head + 1) % ndiis;
            if (nstore < ndiis) ++nstore;
            for (int k0 = 0; k0 < nstore; k0 += DotPtrs::MAX) {  // B-matrix row, up to 8 dots per launch
                const int m = std::min(DotPtrs::MAX, nstore - k0);
                DotPtrs vecs{};
                for (int k = 0; k < m; ++k) vecs.p[k] = dE[k0 + k].p;
                dev_dots(n2 * nd, dE[slot].p, vecs, m, dots.p + k0, part.p, st_);
            }
            if (nstore >= 2) {
                // B row/column of this slot, the Pulay system and the extrapolation, all on the device
                dev_diis_solve(ndiis, nstore, slot, dots.p, Bdev.p, coef.p, dok.p, st_);
                DiisPtrs vs{};
                for (int k = 0; k < nstore; ++k) vs.p[k] = dF[k].p;
                dev_diis_extrap(n2 * nd, nstore, coef.p, dok.p, vs, F.p, Fx.p, st_);
                Fuse = Fx.p;
            } else {
                dev_diis_solve(ndiis, nstore, slot, dots.p, Bdev.p, coef.p, dok.p, st_);  // records the B entries
            }
        }
        if (opt.level_shift > 0.0 && last_dD >= opt.level_shift_until) {
            // virtual-space level shift: F + b (S - S D S / occ), with the density
            // F was built from (S D S / occ is the occupied projector in the AO
            // metric); the converged solution, where D is that projector of F,
            // does not depend on b
            const double b = opt.level_shift;
            for (int s = 0; s < nd; ++s) {
                gemm_f64(n, false, false, 1.0, S_.p, D.p + s * n2, 0.0, w1_.p, st_);
                gemm_f64(n, false, false, 1.0, w1_.p, S_.p, 0.0, w2_.p, st_);
                dev_axpby(n2, 1.0, Fuse + s * n2, b, S_.p, Fls.p + s * n2, st_);
                dev_axpby(n2, 1.0, Fls.p + s * n2, -b / occ, w2_.p, Fls.p + s * n2, st_);
            }
            Fuse = Fls.p;
        }
        // energy with the new D and the F built this iteration (rscf.f90:144-151; uscf.f90:186-193)
        const auto energy_and_dd = [&]() {
            if (!uhf) dev_energy_rhf(n, D.p, H_.p, F.p, red.p, part.p, st_);
            else dev_energy_uhf(n, D.p, D.p + n2, H_.p, F.p, F.p + n2, red.p, part.p, st_);
            dev_diff_sq(n2 * nd, D.p, Dold.p, red.p + 1, part.p, st_);
        };
        bool deferred = false, dfr[2] = {false, false};
        for (int s = 0; s < nd; ++s) {
            dfr[s] = !density_from_fock(Fuse + s * n2, nocc[s], occ, D.p + s * n2, opt, nullptr, s);
            deferred = deferred || dfr[s];
        }
        energy_and_dd();
        mark(2);
        double hv[2], rec[2] = {1.0, 1.0};
        PF_HIP(hipMemcpyAsync(hv, red.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st_));
        if (deferred) PF_HIP(hipMemcpyAsync(rec, sp2_rec_.p, sizeof(double) * nd, hipMemcpyDeviceToHost, st_));
        PF_HIP(hipStreamSynchronize(st_));
        if (deferred) {
            // a purification that had not converged in its fixed step count: this
            // iteration's densities again with the checked purification (and its
            // eigenvector fallback), then the energy and ||dD|| again
            bool redo = false;
            int used = 0;
            for (int s = 0; s < nd; ++s) {
                if (!dfr[s]) continue;
                if (rec[s] == 0.0) redo = true;
                else used = std::max(used, sp2_defer_steps_ - (int)rec[s]);
            }
            if (redo) {
                for (int s = 0; s < nd; ++s) density_from_fock(Fuse + s * n2, nocc[s], occ, D.p + s * n2, opt, nullptr);
                energy_and_dd();
                PF_HIP(hipMemcpyAsync(hv, red.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st_));
                PF_HIP(hipStreamSynchronize(st_));
            } else {
                sp2_hint_ = std::max(used, 1);
            }
        }
        Eel = uhf ? hv[0] : 0.5 * hv[0];
        const double dE_ = std::fabs(Eel - Eold), errD = std::sqrt(hv[1]);
        last_dD = errD;
        res.last_dE = dE_;
        res.last_dD = errD;
        if (opt.verbose)
            std::fprintf(stderr, "iter %3d  E = %.12f  dE = %.3e  dD = %.3e  %s\n", it, Eel + res.e_nuc, dE_, errD,
                         fp32_phase ? "fp32" : "fp64");
        if (!std::isfinite(Eel)) throw Error("SCF produced a non-finite energy");
        if (dE_ < opt.eps && errD < opt.eps && !fp32_phase) { conv = true; break; }
        Eold = Eel;
    }
    // Only a converged density is kept for the gradient: W = 1/2 D F D gives
    // dE/dR only at a stationary point.
    have_fin_ = conv;
    uhf_fin_ = uhf;
    if (conv) {
        Dfin_.alloc(n2 * nd);
        PF_HIP(hipMemcpyAsync(Dfin_.p, D.p, sizeof(double) * n2 * nd, hipMemcpyDeviceToDevice, st_));
    }
    res.quartets = (double)collect_profile();  // also synchronises the stream
    for (auto &e3 : tev) {
        float a = 0.f, b = 0.f;
        PF_HIP(hipEventElapsedTime(&a, e3[0], e3[1]));
        PF_HIP(hipEventElapsedTime(&b, e3[1], e3[2]));
        res.t_fock += 1e-3 * a;
        res.t_diag += 1e-3 * b;
    }
    res.iterations = std::min(it, opt.max_iter);
    res.converged = conv ? 1 : 0;
    res.e_elec = Eel;
    res.energy = Eel + res.e_nuc;
    if (D_out) { D.download(D_out, n2 * nd, st_); PF_HIP(hipStreamSynchronize(st_)); }
    if (eps_out) {
        DevBuf<double> e(n), Dtmp(n2);
        for (int s = 0; s < nd; ++s) {
            density_from_fock(F.p + s * n2, nocc[s], occ, Dtmp.p, opt, e.p);
            e.download(eps_out + (size_t)s * n, n, st_);
        }
        PF_HIP(hipStreamSynchronize(st_));
    }
    res.t_total = now() - t0;
}

// Primitive-pair data variants for the derivative kernels, same layout as
// PairSet::d_prim: variant 0 weights K by 2 alpha of the first shell; variant
// 1 is the reversed pair (P - B, K weighted by 2 beta of the second shell);
// variant 2 the reversed pair with plain K.
void Engine::setup_grad_pairs() {
    if (grad_ready_) return;
    setup_pairs();
    const auto &sh = basis_.shells;
    for (int c = 0; c < 3; ++c) {
        const PairSet &ps = pairs_[c];
        const size_t sz = (size_t)ps.kp * PF_NFIELD * ps.npad;
        gw_[c].alloc(std::max<size_t>(sz, 1));
        gr_[c].alloc(std::max<size_t>(sz, 1));
        grec_[c].alloc(std::max<size_t>((size_t)ps.npad * ps.kp * PF_AOS, 1));
        fill_grad_pairset(ps, d_shells_.p, gw_[c].p, gr_[c].p, grec_[c].p, st_);
    }
    const int nsh = nshell();
    std::vector<int> flat;
    for (int la = 0; la <= 1; ++la)
        for (int lb = 0; lb <= 1; ++lb) {
            std::vector<int> pa, pb;
            for (int a = 0; a < nsh; ++a)
                for (int b = 0; b < nsh; ++b)
                    if (sh[a].l == la && sh[b].l == lb) { pa.push_back(a); pb.push_back(b); }
            const int k = 2 * la + lb;
            g1n_[k] = (int)pa.size();
            g1off_[k] = (int)flat.size();
            flat.insert(flat.end(), pa.begin(), pa.end());
            flat.insert(flat.end(), pb.begin(), pb.end());
        }
    g1pairs_.upload(flat.data(), flat.size(), st_);
    grad_ready_ = true;
}

void Engine::gradient(double *grad_out, double tau, double *terms) {
    if (!have_fin_) throw Error("gradient: no converged density (the SCF did not converge, or was not run)");
    if (has_d_) throw Error("gradient: s and p shells only (this basis has d shells)");
    setup_grad_pairs();
    const int n = nbf(), nsh = nshell(), na = mol_.natoms;
    const size_t n2 = (size_t)n * n;
    const bool uhf = uhf_fin_;
    const int nd = uhf ? 2 : 1;
    const double *D0 = Dfin_.p, *D1 = uhf ? Dfin_.p + n2 : nullptr;
    // F at the final density (full FP64 build, tight screening)
    DevBuf<double> G(n2 * nd), F(n2 * nd), W(n2), Dt(n2), tmp(n2), grad(9 * (size_t)na);
    build_g<double>(D0, D1, G.p, uhf ? G.p + n2 : nullptr, false, MODE_DIRECT, 1e-14);
    for (int s = 0; s < nd; ++s) dev_axpby(n2, 1.0, H_.p, 1.0, G.p + s * n2, F.p + s * n2, st_);
    // W = 1/2 D F D (RHF) or sum_s D_s F_s D_s (UHF)
    dev_fill(n2, 0.0, W.p, st_);
    for (int s = 0; s < nd; ++s) {
        gemm_f64(n, false, false, 1.0, F.p + s * n2, Dfin_.p + s * n2, 0.0, tmp.p, st_);
        gemm_f64(n, false, false, uhf ? 1.0 : 0.5, Dfin_.p + s * n2, tmp.p, 1.0, W.p, st_);
    }
    if (uhf) dev_axpby(n2, 1.0, D0, 1.0, D1, Dt.p, st_);
    else PF_HIP(hipMemcpyAsync(Dt.p, D0, sizeof(double) * n2, hipMemcpyDeviceToDevice, st_));
    dev_fill(9 * (size_t)na, 0.0, grad.p, st_);  // kinetic, nuclear attraction, overlap
    // ---- one-electron part over ordered shell pairs, by (la, lb)
    for (int la = 0; la <= 1; ++la)
        for (int lb = 0; lb <= 1; ++lb) {
            const int k = 2 * la + lb;
            if (g1n_[k] == 0) continue;
            const int *pa = g1pairs_.p + g1off_[k];
            launch_grad1e(la, lb, d_shells_.p, pa, pa + g1n_[k], g1n_[k], n, na, d_axyz_.p, d_Z_.p, d_tab_.p, Dt.p, W.p,
                          grad.p, grad.p + 3 * na, grad.p + 6 * na, st_);
        }
    // ---- two-electron part
    dev_shell_dmax<double>(nsh, n, d_sh_bf_.p, d_sh_nf_.p, Dt.p, D0 != Dt.p ? D0 : nullptr, Dsh64_.p, dmax64_.p,
                           st_, D1);
    std::array<DevBuf<double>, 3> gp;
    int cap = 2;
    for (int c = 0; c < 3; ++c) cap = std::max(cap, pairs_[c].n + 2);
    DevBuf<int> work(cap);  // queue head and work plan of each class in turn
    for (int c = 0; c < 3; ++c) {
        gp[c].alloc(std::max<size_t>(6 * (size_t)pairs_[c].n, 1));
        dev_fill(6 * (size_t)pairs_[c].n, 0.0, gp[c].p, st_);
    }
    for (size_t q = 0; q < classes_.size(); ++q) {
        const int bc = classes_[q].bra, kc = classes_[q].ket;
        GradArgs g;
        g.bra = pairs_[bc].view64();
        g.ket = pairs_[kc].view64();
        g.bra_rec = grec_[bc].p; g.ket_w = gw_[kc].p; g.ket_r = gr_[kc].p;
        g.same = bc == kc;
        g.tau = tau;
        g.nbf = n; g.nsh = nsh;
        g.Dt = Dt.p;
        g.Dk0 = uhf ? D0 : Dt.p;
        g.Dk1 = D1;
        g.Dsh = Dsh64_.p; g.dmax = dmax64_.p;
        g.gbra = gp[bc].p; g.gket = gp[kc].p;
        g.tab = d_tab_.p;
        launch_grad2e_class(pairs_[bc].la, pairs_[bc].lb, pairs_[kc].la, pairs_[kc].lb, uhf ? 2 : 1, g, work.p, cap, st_);
    }
    std::vector<double> h1(9 * (size_t)na), hg(3 * (size_t)na, 0.0), h2(3 * (size_t)na, 0.0),
        hn(3 * (size_t)na, 0.0);
    grad.download(h1.data(), h1.size(), st_);
    std::array<std::vector<double>, 3> hp;
    for (int c = 0; c < 3; ++c) {
        hp[c].resize(6 * (size_t)pairs_[c].n);
        gp[c].download(hp[c].data(), hp[c].size(), st_);
    }
    PF_HIP(hipStreamSynchronize(st_));
    const auto &sh = basis_.shells;
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < pairs_[c].n; ++i)
            for (int x = 0; x < 3; ++x) {
                h2[3 * sh[pairs_[c].sa[i]].atom + x] += hp[c][6 * (size_t)i + x];
                h2[3 * sh[pairs_[c].sb[i]].atom + x] += hp[c][6 * (size_t)i + 3 + x];
            }
    // nuclear repulsion: d/dR_A sum_{B != A} Z_A Z_B / R_AB
    for (int A = 0; A < na; ++A)
        for (int B = 0; B < na; ++B) {
            if (A == B) continue;
            double d[3], r2 = 0.0;
            for (int x = 0; x < 3; ++x) { d[x] = mol_.xyz[3 * A + x] - mol_.xyz[3 * B + x]; r2 += d[x] * d[x]; }
            const double r = std::sqrt(r2);
            if (r < 1e-10) continue;
            for (int x = 0; x < 3; ++x) hn[3 * A + x] -= double(mol_.Z[A]) * mol_.Z[B] * d[x] / (r2 * r);
        }
    for (size_t k = 0; k < hg.size(); ++k) hg[k] = hn[k] + h1[k] + h1[3 * na + k] + h1[6 * na + k] + h2[k];
    std::copy(hg.begin(), hg.end(), grad_out);
    if (terms) {
        const size_t m = 3 * (size_t)na;
        std::copy(hn.begin(), hn.end(), terms);
        std::copy(h1.begin(), h1.end(), terms + m);  // kinetic, nuclear attraction, overlap
        std::copy(h2.begin(), h2.end(), terms + 4 * m);
    }
}

template void Engine::build_g<double>(const double *, const double *, double *, double *, bool, int,
                                                    double);
template void Engine::build_g<float>(const double *, const double *, double *, double *, bool, int,
                                                   double);

}  // namespace pf
