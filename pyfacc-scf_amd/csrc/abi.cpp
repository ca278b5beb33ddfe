// C-ABI entry points of libpyfacc.so (declared in include/pyfacc.h).
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "batch.h"
#include "engine.h"
#include "linalg.h"

namespace pf {
void boys_device(int nmax, int npts, const double *x, double *F, hipStream_t st);
double device_copy(size_t bytes, hipStream_t st);

namespace {
thread_local std::string g_last_error;
thread_local std::array<unsigned long long, 12> g_prof_q{};
thread_local std::array<unsigned long long, 12> g_prof_p{};
thread_local std::array<double, 12> g_prof_ms{};
thread_local std::array<int, 12> g_prof_n{};
thread_local std::array<int, 24> g_prof_l{};

struct Input {
    Molecule mol;
    std::vector<std::string> species;
};

Input make_input(int natoms, const char *species, const int32_t *z, const double *coords) {
    if (natoms <= 0) throw Error("natoms must be positive");
    if (!species || !z || !coords) throw Error("null species, z or coords");
    Input in;
    in.mol.natoms = natoms;
    in.mol.Z.assign(z, z + natoms);
    in.mol.xyz.assign(coords, coords + 3 * (size_t)natoms);
    for (int a = 0; a < natoms; ++a) {
        std::string s(species + 2 * a, species + 2 * a + 2);
        while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
        while (!s.empty() && s.front() == ' ') s.erase(s.begin());
        if (s.empty()) throw Error("empty species label for atom " + std::to_string(a));
        in.species.push_back(s);
        if (in.mol.Z[a] <= 0) throw Error("non-positive atomic number for atom " + std::to_string(a));
    }
    for (double c : in.mol.xyz)
        if (!std::isfinite(c)) throw Error("non-finite coordinate");
    return in;
}

template <typename Fn>
int guarded(Fn &&fn) {
    // A call that failed part-way may have released buffers (into the block
    // cache) with work on them still queued: drain the device before returning.
    struct DrainOnError {
        bool failed = true;
        ~DrainOnError() {
            if (failed) (void)hipDeviceSynchronize();
        }
    } drain;
    try {
        g_last_error.clear();
        fn();
        drain.failed = false;
        return PF_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        const std::string w = e.what();
        if (w.find("basis") != std::string::npos) return PF_ERR_BASIS;
        if (w.find("HIP") != std::string::npos || w.find("GPU") != std::string::npos) return PF_ERR_DEVICE;
        return PF_ERR_INPUT;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return PF_ERR_INTERNAL;
    } catch (...) {
        g_last_error = "unknown error";
        return PF_ERR_INTERNAL;
    }
}

}  // namespace

void set_last_error(const std::string &msg) { g_last_error = msg; }

}  // namespace pf

using namespace pf;

extern "C" {

void pf_default_options(pf_options *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->eps = 1e-8;
    o->max_iter = 100;
    o->precision = PF_PREC_FP64;
    o->incremental = 0;
    o->eri_mode = PF_ERI_DIRECT;
    o->diis = 8;
    o->diag = PF_DIAG_SP2;
    o->renormalize = 1;
    o->schwarz = 1e-10;
    o->mixed_switch = 1e-5;
    o->verbose = 0;
    o->schwarz_inc = 1e-12;
    o->rebuild_every = 0;
    o->schwarz_fp32 = 1e-6;
    o->guess = PF_GUESS_CORE;
    o->level_shift = 0.0;
    o->level_shift_until = 0.1;
    o->diis_basis = PF_DIIS_AO;
    o->mixed_anchor = 0.1;
    o->profile = 0;
    o->jk_fallback = 0;
    o->deterministic = 0;
    o->mixed_finish = PF_FINISH_FP32;
}

const char *pf_last_error(void) { return g_last_error.c_str(); }

int pf_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int pf_basis_info(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
                  int renormalize, int *nbf, int *nshell, int *shell_l, int *shell_atom, int *shell_nprim,
                  double *shell_exp, double *shell_coef) {
    return guarded([&] {
        Input in = make_input(natoms, species, z, coords);
        Basis b = build_basis(resolve_basis_path(basis_name ? basis_name : ""), in.species, in.mol, renormalize != 0);
        if (nbf) *nbf = b.nbf;
        if (nshell) *nshell = (int)b.shells.size();
        for (size_t s = 0; s < b.shells.size(); ++s) {
            if (shell_l) shell_l[s] = b.shells[s].l;
            if (shell_atom) shell_atom[s] = b.shells[s].atom;
            if (shell_nprim) shell_nprim[s] = b.shells[s].np;
            for (int k = 0; k < KMAX; ++k) {
                if (shell_exp) shell_exp[s * KMAX + k] = k < b.shells[s].np ? b.shells[s].exp[k] : 0.0;
                if (shell_coef) shell_coef[s * KMAX + k] = k < b.shells[s].np ? b.shells[s].coef[k] : 0.0;
            }
        }
    });
}

int pf_one_electron(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
                    int renormalize, double *S, double *T, double *V) {
    return guarded([&] {
        Input in = make_input(natoms, species, z, coords);
        Engine eng(in.mol, in.species, resolve_basis_path(basis_name ? basis_name : ""), renormalize != 0, 1e-14);
        eng.one_electron(S, T, V);
    });
}

int pf_eri_packed(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
                  int renormalize, double *packed) {
    return guarded([&] {
        Input in = make_input(natoms, species, z, coords);
        // no screening: every unique quartet is written
        Engine eng(in.mol, in.species, resolve_basis_path(basis_name ? basis_name : ""), renormalize != 0, 0.0);
        eng.eri_packed_host(packed);
    });
}

int pf_eri_shell_quartets(int natoms, const char *species, const int32_t *z, const double *coords,
                          const char *basis_name, int renormalize, int nquart, const int32_t *quartets, double *out,
                          int64_t *offsets) {
    return guarded([&] {
        if (nquart < 0) throw Error("nquart must be >= 0");
        if (nquart > 0 && (!quartets || !out)) throw Error("null quartets or out");
        Input in = make_input(natoms, species, z, coords);
        Engine eng(in.mol, in.species, resolve_basis_path(basis_name ? basis_name : ""), renormalize != 0, 0.0);
        std::vector<long long> off((size_t)nquart + 1);
        eng.eri_quartets(nquart, quartets, out, off.data());
        if (offsets)
            for (int q = 0; q <= nquart; ++q) offsets[q] = off[q];
    });
}

int pf_jk(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
          int renormalize, int ndens, const double *D, double *J, double *K, int eri_mode, int precision,
          double schwarz) {
    return guarded([&] {
        if (ndens != 1 && ndens != 2) throw Error("ndens must be 1 or 2");
        Input in = make_input(natoms, species, z, coords);
        Engine eng(in.mol, in.species, resolve_basis_path(basis_name ? basis_name : ""), renormalize != 0,
                   schwarz > 0 ? schwarz : 1e-14);
        eng.jk_host(ndens, D, J, K, eri_mode, precision);
    });
}

int pf_schwarz(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
               int renormalize, double *Q) {
    return guarded([&] {
        if (!Q) throw Error("null output");
        Input in = make_input(natoms, species, z, coords);
        Engine eng(in.mol, in.species, resolve_basis_path(basis_name ? basis_name : ""), renormalize != 0, 1e-14);
        eng.schwarz_host(Q);
    });
}

int pf_scf(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
           int nalpha, int nbeta, const pf_options *opt, pf_result *res, double *D_out, double *eps_out) {
    return guarded([&] {
        pf_options o;
        if (opt) o = *opt;
        else pf_default_options(&o);
        if (o.max_iter < 1) throw Error("max_iter must be >= 1");
        if (!(o.level_shift >= 0.0)) throw Error("level_shift must be >= 0");
        if (o.diis_basis != PF_DIIS_AO && o.diis_basis != PF_DIIS_ORTHO) throw Error("diis_basis must be 0 or 1");
        if (o.mixed_finish != PF_FINISH_FP32 && o.mixed_finish != PF_FINISH_FP64) throw Error("mixed_finish must be 0 or 1");
        Input in = make_input(natoms, species, z, coords);
        Engine eng(in.mol, in.species, resolve_basis_path(basis_name ? basis_name : ""), o.renormalize != 0,
                   o.schwarz > 0 ? o.schwarz : 1e-14);
        eng.profile_classes = o.profile != 0;
        pf_result r;
        eng.scf(nalpha, nbeta, o, r, D_out, eps_out);
        g_prof_q = eng.prof_quartets;
        g_prof_p = eng.prof_prims;
        g_prof_ms = eng.prof_ms;
        g_prof_n = eng.prof_launches;
        g_prof_l = eng.prof_class_l;
        if (res) *res = r;
    });
}

int pf_batch_scf(int nmol, const int32_t *natoms, const char *species, const int32_t *z, const double *coords,
                 const char *basis_name, const int32_t *nalpha, const int32_t *nbeta, const pf_options *opt,
                 int use_graph, double *energy, int32_t *iterations, int32_t *converged, pf_batch_stats *stats) {
    return guarded([&] {
        pf_options o;
        if (opt) o = *opt;
        else pf_default_options(&o);
        if (o.max_iter < 1) throw Error("max_iter must be >= 1");
        if (nmol <= 0) throw Error("nmol must be positive");
        if (o.level_shift != 0.0 || o.diis_basis != PF_DIIS_AO)
            throw Error("level_shift and diis_basis are not available in the batched engine (pf_batch_scf)");
        if (o.deterministic) throw Error("deterministic is not available in the batched engine (pf_batch_scf)");
        if (!natoms || !nalpha || !nbeta) throw Error("null natoms, nalpha or nbeta");
        std::vector<BatchMolecule> mols(nmol);
        size_t a0 = 0;
        for (int m = 0; m < nmol; ++m) {
            if (natoms[m] <= 0) throw Error("batch: molecule " + std::to_string(m) + " has no atoms");
            Input in = make_input(natoms[m], species + 2 * a0, z + a0, coords + 3 * a0);
            mols[m].mol = in.mol;
            mols[m].species = in.species;
            mols[m].nalpha = nalpha[m];
            mols[m].nbeta = nbeta[m];
            a0 += natoms[m];
        }
        BatchEngine be(std::move(mols), resolve_basis_path(basis_name ? basis_name : ""), o.renormalize != 0,
                       o.schwarz > 0 ? o.schwarz : 1e-14);
        BatchStats bs;
        be.run(o, use_graph != 0, energy, iterations, converged, &bs);
        if (stats) {
            stats->t_total = bs.t_total; stats->t_setup = bs.t_setup; stats->t_eri = bs.t_eri;
            stats->t_scf = bs.t_scf; stats->eri_values = bs.eri_values; stats->store_bytes = bs.store_bytes;
            stats->iterations = bs.iterations; stats->graph = bs.graph; stats->digest_waves = bs.digest_waves;
            stats->nproblems = bs.nproblems;
            stats->sp2_fallbacks = bs.sp2_fallbacks;
            stats->digest_tail_waves = bs.digest_tail_waves;
            stats->digest_item_values = bs.digest_item_values;
            stats->digest_tail_item_values = bs.digest_tail_item_values;
        }
    });
}

int pf_scf_gradient(int natoms, const char *species, const int32_t *z, const double *coords, const char *basis_name,
                    int nalpha, int nbeta, const pf_options *opt, pf_result *res, double *grad, double *D_out,
                    double *terms) {
    return guarded([&] {
        pf_options o;
        if (opt) o = *opt;
        else pf_default_options(&o);
        if (o.max_iter < 1) throw Error("max_iter must be >= 1");
        if (!(o.level_shift >= 0.0)) throw Error("level_shift must be >= 0");
        if (o.diis_basis != PF_DIIS_AO && o.diis_basis != PF_DIIS_ORTHO) throw Error("diis_basis must be 0 or 1");
        if (o.mixed_finish != PF_FINISH_FP32 && o.mixed_finish != PF_FINISH_FP64) throw Error("mixed_finish must be 0 or 1");
        if (!grad) throw Error("null gradient output");
        Input in = make_input(natoms, species, z, coords);
        Engine eng(in.mol, in.species, resolve_basis_path(basis_name ? basis_name : ""), o.renormalize != 0,
                   o.schwarz > 0 ? o.schwarz : 1e-14);
        pf_result r;
        eng.scf(nalpha, nbeta, o, r, D_out, nullptr);
        if (res) *res = r;
        if (!r.converged)
            throw Error("gradient: the SCF did not converge in " + std::to_string(r.iterations) +
                        " iterations; the analytic gradient holds only at a converged density");
        eng.gradient(grad, 1e-13, terms);
    });
}

int pf_scf_profile(unsigned long long *quartets, double *ms, int *launches, int *class_l,
                   unsigned long long *prims) {
    for (int q = 0; q < 12; ++q) {
        if (quartets) quartets[q] = g_prof_q[q];
        if (prims) prims[q] = g_prof_p[q];
        if (ms) ms[q] = g_prof_ms[q];
        if (launches) launches[q] = g_prof_n[q];
    }
    for (int q = 0; q < 24; ++q)
        if (class_l) class_l[q] = g_prof_l[q];
    return PF_OK;
}

int pf_boys(int nmax, int npts, const double *x, double *F) {
    return guarded([&] {
        if (pf_device_count() == 0) throw Error("no GPU device visible");
        hipStream_t st;
        PF_HIP(hipStreamCreate(&st));
        boys_device(nmax, npts, x, F, st);
        PF_HIP(hipStreamDestroy(st));
    });
}

int pf_eigh(int n, const double *A, double *w, double *V) {
    return guarded([&] {
        if (n <= 0) throw Error("n must be positive");
        if (pf_device_count() == 0) throw Error("no GPU device visible");
        hipStream_t st;
        PF_HIP(hipStreamCreate(&st));
        {
            DevBuf<double> dA((size_t)n * n), dw(n), dV((size_t)n * n);
            dA.upload(A, (size_t)n * n, st);
            jacobi_eigh(n, dA.p, dw.p, dV.p, st);
            dw.download(w, n, st);
            dV.download(V, (size_t)n * n, st);
            PF_HIP(hipStreamSynchronize(st));
        }
        PF_HIP(hipStreamDestroy(st));
    });
}

int pf_projector(int n, const double *F, double nocc, double *P) {
    return guarded([&] {
        if (n <= 0) throw Error("n must be positive");
        if (pf_device_count() == 0) throw Error("no GPU device visible");
        hipStream_t st;
        PF_HIP(hipStreamCreate(&st));
        {
            DevBuf<double> dF((size_t)n * n), dP((size_t)n * n), w((size_t)n * n), state(sp2_state_size(n));
            dF.upload(F, (size_t)n * n, st);
            sp2_density(n, dF.p, nocc, dP.p, w.p, state.p, st);
            dP.download(P, (size_t)n * n, st);
            PF_HIP(hipStreamSynchronize(st));
        }
        PF_HIP(hipStreamDestroy(st));
    });
}

int pf_gemm(int m, int n, int k, int ta, int tb, double alpha, const double *A, const double *B, double beta,
            double *C) {
    return guarded([&] {
        if (m <= 0 || n <= 0 || k <= 0) throw Error("gemm sizes must be positive");
        if (pf_device_count() == 0) throw Error("no GPU device visible");
        hipStream_t st;
        PF_HIP(hipStreamCreate(&st));
        {
            DevBuf<double> dA((size_t)m * k), dB((size_t)k * n), dC((size_t)m * n);
            dA.upload(A, (size_t)m * k, st);
            dB.upload(B, (size_t)k * n, st);
            dC.upload(C, (size_t)m * n, st);
            gemm_f64_mnk(m, n, k, ta != 0, tb != 0, alpha, dA.p, ta ? m : k, dB.p, tb ? k : n, beta, dC.p, n, st);
            dC.download(C, (size_t)m * n, st);
            PF_HIP(hipStreamSynchronize(st));
        }
        PF_HIP(hipStreamDestroy(st));
    });
}

int pf_device_copy(unsigned long long bytes, double *ms) {
    return guarded([&] {
        if (pf_device_count() == 0) throw Error("no GPU device visible");
        hipStream_t st;
        PF_HIP(hipStreamCreate(&st));
        const double t = device_copy((size_t)bytes, st);
        if (ms) *ms = t;
        PF_HIP(hipStreamDestroy(st));
    });
}

int pf_release_device_cache(void) {
    return guarded([&] { release_device_cache(); });
}

unsigned long long pf_cached_device_bytes(void) { return (unsigned long long)cached_device_bytes(); }

void run_scf_c(int natoms, int nelectrons, int charge, int nalpha, int nbeta, const char *species, double *coords,
               double eps_scf, int max_iter, double *final_energy, const char *basis_set_c, int32_t *z_num) {
    (void)charge;
    const int rc = guarded([&] {
        if (nelectrons != nalpha + nbeta) throw Error("nelectrons != nalpha + nbeta");
        pf_options o;
        pf_default_options(&o);
        o.eps = eps_scf;
        o.max_iter = max_iter;
        o.diis = 0;          // plain Roothaan iterations as rscf.f90:118-157
        o.schwarz = 1e-14;   // the reference keeps every integral
        Input in = make_input(natoms, species, z_num, coords);
        Engine eng(in.mol, in.species, resolve_basis_path(basis_set_c ? basis_set_c : ""), true, o.schwarz);
        pf_result r;
        eng.scf(nalpha, nbeta, o, r, nullptr, nullptr);
        if (final_energy) *final_energy = r.energy;
    });
    if (rc != PF_OK && final_energy) *final_energy = std::numeric_limits<double>::quiet_NaN();
}

void run_scf_interface_py(int32_t *natoms, int32_t *nelectrons, int32_t *charge, int32_t *nalpha, int32_t *nbeta,
                          const char *species, double *coords, double *eps_scf, int32_t *max_iter,
                          double *final_energy, const char *basis_set_c, int32_t *len_basis_set, int32_t *z_num) {
    if (!natoms || !nelectrons || !charge || !nalpha || !nbeta || !eps_scf || !max_iter || !len_basis_set ||
        !basis_set_c) {
        g_last_error = "null argument";
        if (final_energy) *final_energy = std::numeric_limits<double>::quiet_NaN();
        return;
    }
    const std::string name(basis_set_c, basis_set_c + std::max(0, *len_basis_set));
    run_scf_c(*natoms, *nelectrons, *charge, *nalpha, *nbeta, species, coords, *eps_scf, *max_iter, final_energy,
              name.c_str(), z_num);
}

}  // extern "C"
