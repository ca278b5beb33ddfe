// Shell-quartet ERI kernels with the fused J/K digest.
//
// One launch per quartet class (LA LB|LC LD); the bra pair class is the one
// of higher angular momentum.  Pairs of each class are sorted by Schwarz
// bound Q descending, so the kets that can pass Q_i Q_j dmax >= tau form a
// prefix of the ket list.  The bra pair's data is uniform across a
// wavefront (scalar loads); ket data is structure-of-arrays so the 64 lanes
// (one ket each) read consecutive addresses.
//
// J/K digest for a unique shell quartet with degeneracy factor
// deg = (1/2)^{[a=b] + [c=d] + [ab=cd]}, into non-symmetric accumulators
// (symmetrised afterwards: J = J' + J'^T, K = K' + K'^T):
//   J'_ab += 2 deg (ab|cd) D_cd      (reduced over the kets of the bra)
//   J'_cd += 2 deg (ab|cd) D_ab      (device atomics)
//   K'_ac += deg (ab|cd) D_bd,  K'_ad += deg (ab|cd) D_bc,
//   K'_bc += deg (ab|cd) D_ad,  K'_bd += deg (ab|cd) D_ac
// Every exchange term lands on a row of the bra (a or b).  Direct builds
// (jk_direct_kernel) give each workgroup one bra pair and a block of its kets
// at a time and sum those rows, and J'_ab, in LDS, where FP64 atomics run at
// LDS rate (scattered FP64 atomics to device memory sustain only ~24e9/s);
// the rows go to device memory once per block.
// This replaces the reference's G loop (rscf.f90:121-134, uscf.f90:149-175),
// which gathers nbf^4 values from a stored list and drops every
// non-canonical permutation.
#pragma once
#include <algorithm>

#include "eri_core.h"
#include "kernels.h"

namespace pf {

namespace {

template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ long long pair_index(long long i, long long j) {
    return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}
// Canonical packed position of (ij|kl) over basis functions: ij = i(i+1)/2 + j
// (i >= j), index = ij(ij+1)/2 + kl (ij >= kl); the set enumerated by
// rscf.f90:67-88, here 0-based.
__device__ __forceinline__ long long quad_index(int i, int j, int k, int l) {
    const long long ij = pair_index(i, j), kl = pair_index(k, l);
    return ij >= kl ? ij * (ij + 1) / 2 + kl : kl * (kl + 1) / 2 + ij;
}

// One addition into a J'/K' accumulator (LDS or device memory): an FP64
// atomic, or (DET) the int64 atomic add of round(v s) on the same 8 bytes.
template <bool DET>
__device__ __forceinline__ void acc_add(double *p, double v, double s) {
    if constexpr (DET)
        atomicAdd(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(__double2ll_rn(v * s)));
    else
        atomicAdd(p, v);
}

}  // namespace

// Wavefronts per workgroup: eight (two per SIMD, sharing one LDS Boys table)
// where the class fits in 256 registers, four for the register-heavy classes
// of total angular momentum >= 3.
template <int LA, int LB, int LC, int LD>
constexpr int jk_waves() { return LA + LB + LC + LD >= 3 ? 4 : 8; }
// Direct-mode workgroup shape per class and precision (measured on (H2O)64
// STO-3G): FP32 (ps|ss) runs twelve wavefronts per workgroup held to 80
// registers (two workgroups: six wavefronts per SIMD; 13 % faster than
// eight wavefronts at five), the others keep jk_waves and the compiler's
// register choice.
template <int LA, int LB, int LC, int LD, typename R>
constexpr int jk_direct_waves() {
    return LA + LB + LC + LD == 1 && std::is_same<R, float>::value ? 12 : jk_waves<LA, LB, LC, LD>();
}
// minimum wavefronts per SIMD the register allocation must allow (1: no limit)
template <int LA, int LB, int LC, int LD, typename R>
constexpr int jk_direct_occupancy() {
    return LA + LB + LC + LD == 1 && std::is_same<R, float>::value ? 6 : 1;
}
// Ket-primitive prefetch in eri_block: 2-4 % faster for every class but FP64
// (ss|ss), where its registers cost a wavefront per SIMD (16 % slower).
template <int LA, int LB, int LC, int LD, typename R>
constexpr bool jk_prefetch() { return LA + LB + LC + LD != 0 || !std::is_same<R, double>::value; }
// FP32 builds (the mixed path's increments): a ket primitive whose bound
// times the bra's largest bound times the quartet's density factor dq is
// under JK_PCUT_REL tau ends the lane's primitive loop (primitives are sorted
// by bound), so each primitive quartet left out is below JK_PCUT_REL tau / dq
// in the integral and below JK_PCUT_REL tau in J and K (up to the p-function
// polynomial factors), against tau for a whole quartet left out by the
// screen.  FP64 builds keep the fixed cut PRIM_EPS.  Measured on (H2O)64
// (mixed SCF): 49.1 -> 47.4 ms with 1e-2, the energy moving 3e-10 Eh; 3e-2
// gives the same time and moves it 2.6e-9 Eh.
constexpr double JK_PCUT_REL = 1e-2;

// Kets per direct-mode work item (one workgroup): passes of every wavefront,
// few enough that the last items of a launch do not leave the GPU waiting on
// one workgroup (the FP32 builds are mostly increments, whose surviving
// quartets crowd into the items of the largest bras) and enough that the
// per-item cost (queue atomic, row flush) stays small beside the item's
// integrals.  Measured on (H2O)64 (mixed / FP64 SCF, ms): 8-4-2 passes for
// L <= 1, <= 3, 4 in both precisions 52.1 / 116.6; 4-2-1: 49.5 / 113.1;
// FP64 4-2-1 with FP32 2-1-1: 49.1 / 113.1 (used); 2-1-1 in both
// 50.2 / 122.1; one pass everywhere 52.7 / 134.7.
template <int LA, int LB, int LC, int LD, typename R>
constexpr int jk_chunk() {
    constexpr int L = LA + LB + LC + LD;
    constexpr int passes = std::is_same<R, float>::value ? (L <= 1 ? 2 : 1) : (L <= 1 ? 4 : (L <= 3 ? 2 : 1));
    return 64 * jk_direct_waves<LA, LB, LC, LD, R>() * passes;
}

// Contracted (ab|cd) block, eri[((ia*NB+ib)*NC+ic)*ND+id], of bra pair i
// (uniform; its primitive data through bprim) and ket pair j (this lane).
template <int LA, int LB, int LC, int LD, typename R, bool PREFETCH = true>
__device__ __forceinline__ int eri_block(const PairView<R> &B, const R *__restrict__ bprim, int i, R ABx, R ABy,
                                          R ABz, const PairView<R> &K, int j, const R *lds_boys, R *eri,
                                          R pcut = R(0)) {
    using E = Eri<LA, LB, LC, LD, R>;
    R acc[E::NACC];
#pragma unroll
    for (int k = 0; k < E::NACC; ++k) acc[k] = R(0);
    // significant primitive pairs only (bra: uniform count; ket: this lane's,
    // and of those only the ones whose bound times the bra's largest reaches
    // the cut: ket primitives are sorted by bound, so the loop stops there)
    const int kpk = K.kpn[j], kpb = __builtin_amdgcn_readfirstlane(B.kpn[i]);
    // bra primitive 0 has the largest bound; pcut: this quartet's own cut (0: none)
    const R *bpair = bprim + (size_t)i * B.kp * PF_AOS;  // pair-major copy: pair i's primitives contiguous
    const R bmax = bpair[PF_BND];
    R cut = B.peps;
    if constexpr (std::is_same<R, float>::value) cut = max(cut, pcut);
    // this lane's ket primitive fields, the next primitive's loaded while the
    // current one runs through the bra primitives (hides the load latency)
    struct KetPrim {
        R bnd, q, Qx, Qy, Qz, QCx, QCy, QCz, Kk, o2q;
    };
    const auto load = [&](int u) {
        const R *kq = K.prim + (size_t)u * PF_NFIELD * K.npad + j;
        return KetPrim{kq[PF_BND * K.npad], kq[PF_P * K.npad],  kq[PF_PX * K.npad], kq[PF_PY * K.npad],
                       kq[PF_PZ * K.npad],  kq[PF_AX * K.npad], kq[PF_AY * K.npad], kq[PF_AZ * K.npad],
                       kq[PF_K * K.npad],   kq[PF_O2P * K.npad]};
    };
    int nu = 0;
    KetPrim nx = (PREFETCH && kpk > 0) ? load(0) : KetPrim{};
    for (int u = 0; u < kpk; ++u, ++nu) {
        const KetPrim c = PREFETCH ? nx : load(u);
        if (c.bnd * bmax < cut) break;
        if (PREFETCH && u + 1 < kpk) nx = load(u + 1);
        for (int v = 0; v < kpb; ++v) {
            // bra data: read-only in these kernels and wavefront-uniform -> scalar loads
            const R *bp = bpair + (size_t)v * PF_AOS;
            E::template prim<true>(bp[PF_P], bp[PF_O2P], bp[PF_PX], bp[PF_PY], bp[PF_PZ], bp[PF_AX], bp[PF_AY],
                                   bp[PF_AZ], bp[PF_K], c.q, c.o2q, c.Qx, c.Qy, c.Qz, c.QCx, c.QCy, c.QCz, c.Kk,
                                   lds_boys, acc);
        }
    }
    const R CDx = K.AB[j], CDy = K.AB[K.npad + j], CDz = K.AB[2 * K.npad + j];
    E::hrr(acc, ABx, ABy, ABz, CDx, CDy, CDz, eri);
    return nu * kpb;  // primitive quartets evaluated
}

// Densities a digest reads: Dj for J (the total density), Dk0/Dk1 for K.
template <typename R>
struct DensView {
    const R *Dj, *Dk0, *Dk1;
    int nbf;
};
template <typename R>
__device__ __forceinline__ DensView<R> dens_view(const JKArgs<R> &g) {
    return DensView<R>{g.Dj, g.Dk0, g.Dk1, g.nbf};
}

// J/K digest of one quartet (bra functions oa.., ob.., ket functions oc..,
// od..).  jab[NA*NB] accumulates J'_ab (reduced by the caller);
// jadd(row, col, v) adds v to J'(row, col) for the ket block;
// kadd(kk, r, col, v) adds v to K'(row r, col) of density kk, bra row r < NA
// being function oa + r and r >= NA function ob + r - NA.
template <int LA, int LB, int LC, int LD, typename R, int NK, typename JAdd, typename KAdd>
__device__ __forceinline__ void digest(const DensView<R> &g, const R *eri, int oa, int ob, int oc, int od, R deg,
                                       const R *dab, R *jab, JAdd &&jadd, KAdd &&kadd) {
    using E = Eri<LA, LB, LC, LD, R>;
    constexpr int NA = E::NA, NB = E::NB, NC = E::NC, ND = E::ND;
    const int nbf = g.nbf;
    const R deg2 = deg + deg;
    // ---- Coulomb
    R dcd[NC * ND];
#pragma unroll
    for (int ic = 0; ic < NC; ++ic)
#pragma unroll
        for (int id = 0; id < ND; ++id) dcd[ic * ND + id] = g.Dj[(size_t)(oc + ic) * nbf + od + id];
#pragma unroll
    for (int ab = 0; ab < NA * NB; ++ab) {
        R s = R(0);
#pragma unroll
        for (int cd = 0; cd < NC * ND; ++cd) s += eri[ab * NC * ND + cd] * dcd[cd];
        jab[ab] += deg2 * s;
    }
#pragma unroll
    for (int ic = 0; ic < NC; ++ic)
#pragma unroll
        for (int id = 0; id < ND; ++id) {
            R s = R(0);
#pragma unroll
            for (int ab = 0; ab < NA * NB; ++ab) s += eri[ab * NC * ND + ic * ND + id] * dab[ab];
            jadd(oc + ic, od + id, deg2 * s);
        }
    // ---- exchange, one pass per density
#pragma unroll
    for (int kk = 0; kk < NK; ++kk) {
        const R *Dk = kk == 0 ? g.Dk0 : g.Dk1;
        R dbd[NB * ND], dbc[NB * NC], dad[NA * ND], dac[NA * NC];
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
#pragma unroll
            for (int id = 0; id < ND; ++id) dbd[ib * ND + id] = Dk[(size_t)(ob + ib) * nbf + od + id];
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) dbc[ib * NC + ic] = Dk[(size_t)(ob + ib) * nbf + oc + ic];
        }
#pragma unroll
        for (int ia = 0; ia < NA; ++ia) {
#pragma unroll
            for (int id = 0; id < ND; ++id) dad[ia * ND + id] = Dk[(size_t)(oa + ia) * nbf + od + id];
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) dac[ia * NC + ic] = Dk[(size_t)(oa + ia) * nbf + oc + ic];
        }
#define ERI_(ia, ib, ic, id) eri[(((ia) * NB + (ib)) * NC + (ic)) * ND + (id)]
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) {
                R s = R(0);
#pragma unroll
                for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                    for (int id = 0; id < ND; ++id) s += ERI_(ia, ib, ic, id) * dbd[ib * ND + id];
                kadd(kk, ia, oc + ic, deg * s);  // K'_ac
            }
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                R s = R(0);
#pragma unroll
                for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                    for (int ic = 0; ic < NC; ++ic) s += ERI_(ia, ib, ic, id) * dbc[ib * NC + ic];
                kadd(kk, ia, od + id, deg * s);  // K'_ad
            }
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) {
                R s = R(0);
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                    for (int id = 0; id < ND; ++id) s += ERI_(ia, ib, ic, id) * dad[ia * ND + id];
                kadd(kk, NA + ib, oc + ic, deg * s);  // K'_bc
            }
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                R s = R(0);
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                    for (int ic = 0; ic < NC; ++ic) s += ERI_(ia, ib, ic, id) * dac[ia * NC + ic];
                kadd(kk, NA + ib, od + id, deg * s);  // K'_bd
            }
#undef ERI_
    }
}

// Bra-major kernel for the stored-integral and export modes: one wavefront
// per bra pair, lanes over its kets (store layout: per bra, blocks of 64
// kets, integral k of lane l at k*64 + l).
//   MODE_STORE:  compute and store the quartets that pass Schwarz (FP64)
//   MODE_LOAD:   read them back and digest into J', K' (device atomics)
//   MODE_PACKED: compute every unique quartet into the canonical packed list
template <int LA, int LB, int LC, int LD, typename R, int NK, int MODE>
__global__ void __launch_bounds__((64 * jk_waves<LA, LB, LC, LD>())) jk_kernel(JKArgs<R> g, const R *__restrict__ bprim) {
    constexpr int W = jk_waves<LA, LB, LC, LD>();
    using E = Eri<LA, LB, LC, LD, R>;
    using BT = BoysLds<E::L, R>;
    constexpr int NA = E::NA, NB = E::NB, NC = E::NC, ND = E::ND, NINT = E::NINT;
    constexpr bool COMPUTE = MODE == MODE_STORE || MODE == MODE_PACKED;
    __shared__ R lds_boys[COMPUTE ? BT::SIZE : 1];
    if constexpr (COMPUTE) {
        BT::fill(lds_boys, g.tab, threadIdx.x, 64 * W);
        __syncthreads();
    }
    const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * W + (threadIdx.x >> 6));
    const PairView<R> &B = g.bra;
    const PairView<R> &K = g.ket;
    if (i >= B.n) return;
    const int lane = threadIdx.x & 63;
    const int nbf = g.nbf;
    const int nket = MODE == MODE_PACKED ? (g.same ? i + 1 : K.n) : g.store_nket[i];
    // batched stores: bra i's kets start at store_kbase[i] of the ket list
    const int kbase = (MODE == MODE_STORE && g.store_kbase) ? g.store_kbase[i] : 0;
    const int a = B.sa[i], b = B.sb[i], oa = B.oa[i], ob = B.ob[i];
    const R ABx = B.AB[i], ABy = B.AB[B.npad + i], ABz = B.AB[2 * B.npad + i];
    const R deg_ab = (a == b) ? R(0.5) : R(1);
    R dab[NA * NB], jab[NA * NB];
    if constexpr (MODE == MODE_LOAD) {
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
                dab[ia * NB + ib] = g.Dj[(size_t)(oa + ia) * nbf + ob + ib];
                jab[ia * NB + ib] = R(0);
            }
    }
    unsigned long long counted = 0, pcounted = 0;
    for (int j0 = 0; j0 < nket; j0 += 64) {
        if (j0 + lane >= nket) break;
        const int j = kbase + j0 + lane;
        ++counted;
        R eri[NINT];
        if constexpr (MODE == MODE_LOAD) {
            const R *src = g.store + g.store_off[i] + (size_t)j0 * NINT + lane;
#pragma unroll
            for (int k = 0; k < NINT; ++k) eri[k] = src[k * 64];
        } else {
            pcounted += eri_block<LA, LB, LC, LD, R>(B, bprim, i, ABx, ABy, ABz, K, j, lds_boys, eri);
        }
        if constexpr (MODE == MODE_STORE) {
            R *dst = g.store + g.store_off[i] + (size_t)j0 * NINT + lane;
#pragma unroll
            for (int k = 0; k < NINT; ++k) dst[k * 64] = eri[k];
            if (g.store32) {
                float *d32 = g.store32 + g.store_off[i] + (size_t)j0 * NINT + lane;
#pragma unroll
                for (int k = 0; k < NINT; ++k) d32[k * 64] = float(eri[k]);
            }
        } else if constexpr (MODE == MODE_PACKED) {
            const int oc = K.oa[j], od = K.ob[j];
#pragma unroll
            for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                    for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                        for (int id = 0; id < ND; ++id)
                            g.packed[quad_index(oa + ia, ob + ib, oc + ic, od + id)] =
                                double(eri[((ia * NB + ib) * NC + ic) * ND + id]);
        } else {
            const int c = K.sa[j], d = K.sb[j], oc = K.oa[j], od = K.ob[j];
            const R deg = deg_ab * ((c == d) ? R(0.5) : R(1)) * ((g.same && i == j) ? R(0.5) : R(1));
            digest<LA, LB, LC, LD, R, NK>(
                dens_view(g), eri, oa, ob, oc, od, deg, dab, jab,
                [&](int row, int col, R v) { atomicAdd(&g.J[(size_t)row * nbf + col], double(v)); },
                [&](int kk, int r, int col, R v) {
                    double *Ko = kk == 0 ? g.K0 : g.K1;
                    const int row = r < NA ? oa + r : ob + r - NA;
                    atomicAdd(&Ko[(size_t)row * nbf + col], double(v));
                });
        }
    }
    if constexpr (MODE == MODE_LOAD) {
#pragma unroll
        for (int ab = 0; ab < NA * NB; ++ab) {
            const double s = wave_sum(double(jab[ab]));
            if (lane == 0) atomicAdd(&g.J[(size_t)(oa + ab / NB) * nbf + ob + ab % NB], s);
        }
    }
    if (g.count) {
        unsigned long long tot = counted, ptot = pcounted;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            tot += __shfl_xor(tot, o, 64);
            ptot += __shfl_xor(ptot, o, 64);
        }
        if (lane == 0) {
            atomicAdd(g.count, tot);
            if (g.pcount) atomicAdd(g.pcount, ptot);
        }
    }
}

// Direct-mode kernel: persistent workgroups take items (bra pair i, block of
// up to g.chunk of its kets) from a queue laid out by jk_count_kernel /
// jk_scan_kernel (bra i owns items item_off[i] .. item_off[i+1]-1).  The
// workgroup's wavefronts split the block; each screens its candidates 64 at
// a time with the density-weighted test and queues the survivors in LDS, so
// passes run with full wavefronts.  The bra's exchange rows K'(a|b, :) of
// every density and its J'_ab are summed in LDS (dynamic shared memory, when
// g.krows) and flushed to device memory after the item; with g.krows = 0
// (rows too long for LDS) the exchange terms go straight to device memory.
template <int LA, int LB, int LC, int LD, typename R, int NK, bool DET>
__global__ void __launch_bounds__((64 * jk_direct_waves<LA, LB, LC, LD, R>()), (jk_direct_occupancy<LA, LB, LC, LD, R>()))
    jk_direct_kernel(JKArgs<R> g, const R *__restrict__ bprim, const int *__restrict__ item_off) {
    constexpr int W = jk_direct_waves<LA, LB, LC, LD, R>(), BLOCK = 64 * W;
    using E = Eri<LA, LB, LC, LD, R>;
    using BT = BoysLds<E::L, R>;
    constexpr int NA = E::NA, NB = E::NB, NINT = E::NINT, NAB = NA + NB;
    __shared__ R lds_boys[BT::SIZE];
    __shared__ int lds_list[W][128];
    // FP32 builds: each queued ket's density factor, for its primitive cut
    constexpr bool PCUT = std::is_same<R, float>::value;
    __shared__ float lds_dq[PCUT ? W : 1][PCUT ? 128 : 1];
    __shared__ int s_item;
    extern __shared__ double lds_acc[];  // [NK][NAB][nbf] exchange rows if g.krows, then [NA*NB] J'_ab
    BT::fill(lds_boys, g.tab, threadIdx.x, BLOCK);
    const PairView<R> &B = g.bra;
    const PairView<R> &K = g.ket;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nbf = g.nbf;
    const int nrow = g.krows ? NK * NAB * nbf : 0;
    double *ljab = lds_acc + nrow;
    for (int e = threadIdx.x; e < nrow + NA * NB; e += BLOCK) lds_acc[e] = 0.0;
    int *list = lds_list[wave];
    float *dql = lds_dq[PCUT ? wave : 0];
    const bool compact = g.Dsh != nullptr;
    const int total = g.work[1];
    constexpr int NCD = E::NC * E::ND;
    const bool jstore = g.jbuf != nullptr && g.work[2] != 0;
    const double fxs = DET ? *g.fx : 0.0;
    unsigned long long counted = 0, pcounted = 0;
    __syncthreads();
    while (true) {
        if (threadIdx.x == 0) s_item = atomicAdd(&g.work[0], 1);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_item);
        if (t >= total) break;
        int i = 0, hi = B.n - 1;  // largest i with item_off[i] <= t
        while (i < hi) {
            const int mid = (i + hi + 1) >> 1;
            if (item_off[mid] <= t) i = mid;
            else hi = mid - 1;
        }
        const int jbeg = (t - item_off[i]) * g.chunk, jend = min(g.nket[i], jbeg + g.chunk);
        // this wavefront's share of the block, a multiple of 64 kets
        const int per = (((jend - jbeg) + W - 1) / W + 63) & ~63;
        const int wb = min(jend, jbeg + wave * per), we = min(jend, wb + per);

        const R Qi = B.Q[i];
        const int a = B.sa[i], b = B.sb[i], oa = B.oa[i], ob = B.ob[i];
        const R ABx = B.AB[i], ABy = B.AB[B.npad + i], ABz = B.AB[2 * B.npad + i];
        const R deg_ab = (a == b) ? R(0.5) : R(1);
        R dab[NA * NB], jab[NA * NB];
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
                dab[ia * NB + ib] = g.Dj[(size_t)(oa + ia) * nbf + ob + ib];
                jab[ia * NB + ib] = R(0);
            }

        int next = wb, queued = 0;  // wavefront-uniform
        while (true) {
            int j;
            R jdq = R(0);  // FP32: the queued ket's density factor
            if (compact) {
                const int ns = g.nsh;
                while (queued < 64 && next < we) {
                    const int jj = next + lane;
                    bool keep = false;
                    R dqs = R(0);
                    if (jj < we) {
                        const int c = K.sa[jj], d = K.sb[jj];
                        R dq = R(2) * max(g.Dsh[a * ns + b], g.Dsh[c * ns + d]);
                        dq = max(dq, max(max(g.Dsh[b * ns + d], g.Dsh[b * ns + c]),
                                         max(g.Dsh[a * ns + d], g.Dsh[a * ns + c])));
                        keep = Qi * K.Q[jj] * dq >= g.tau;
                        if constexpr (PCUT) dqs = dq;
                        if (!keep && jstore) {  // a screened ket adds nothing to its J'_cd column
                            double *jb = g.jbuf + (g.joff[i] + jj) * NCD;
#pragma unroll
                            for (int q = 0; q < NCD; ++q) jb[q] = 0.0;
                        }
                    }
                    const unsigned long long m = __ballot(keep);
                    if (keep) {
                        const int pos = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                        list[queued + pos] = jj;
                        if constexpr (PCUT) dql[queued + pos] = dqs;
                    }
                    queued += __popcll(m);
                    next += 64;
                }
                // the queue is read across lanes: keep its LDS writes ahead of the reads
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if (queued == 0) break;
                const int take = min(queued, 64);
                j = lane < take ? list[lane] : -1;
                if constexpr (PCUT) jdq = lane < take ? dql[lane] : R(0);
                // entries 64 .. queued-1 move down; lane l alone reads and writes slot l
                if (lane + 64 < queued) list[lane] = list[lane + 64];
                if constexpr (PCUT)
                    if (lane + 64 < queued) dql[lane] = dql[lane + 64];
                queued -= take;
            } else {
                if (next >= we) break;
                j = next + lane < we ? next + lane : -1;
                next += 64;
            }
            if (j < 0) continue;
            const int c = K.sa[j], d = K.sb[j], oc = K.oa[j], od = K.ob[j];
            ++counted;
            R eri[NINT];
            // FP32: primitive quartets whose bound times the quartet's density
            // factor is under JK_PCUT_REL tau are left out
            const R pcut = (PCUT && jdq > R(0)) ? g.tau * R(JK_PCUT_REL) / jdq : R(0);
            pcounted += eri_block<LA, LB, LC, LD, R, jk_prefetch<LA, LB, LC, LD, R>()>(B, bprim, i, ABx, ABy, ABz, K, j,
                                                                                   lds_boys, eri, pcut);
            const R deg = deg_ab * ((c == d) ? R(0.5) : R(1)) * ((g.same && i == j) ? R(0.5) : R(1));
            double jcd[NCD];  // this quartet's J'_cd block
            const auto jadd = [&](int row, int col, R v) { jcd[(row - oc) * E::ND + (col - od)] = double(v); };
            if (g.krows) {
                digest<LA, LB, LC, LD, R, NK>(dens_view(g), eri, oa, ob, oc, od, deg, dab, jab, jadd,
                                              [&](int kk, int r, int col, R v) {
                                                  acc_add<DET>(&lds_acc[(kk * NAB + r) * nbf + col], double(v), fxs);
                                              });
            } else {
                digest<LA, LB, LC, LD, R, NK>(dens_view(g), eri, oa, ob, oc, od, deg, dab, jab, jadd,
                                              [&](int kk, int r, int col, R v) {
                                                  double *Ko = kk == 0 ? g.K0 : g.K1;
                                                  const int row = r < NA ? oa + r : ob + r - NA;
                                                  acc_add<DET>(&Ko[(size_t)row * nbf + col], double(v), fxs);
                                              });
            }
            if (jstore) {
                double *jb = g.jbuf + (g.joff[i] + j) * NCD;
#pragma unroll
                for (int q = 0; q < NCD; ++q) jb[q] = jcd[q];
            } else {
#pragma unroll
                for (int q = 0; q < NCD; ++q)
                    acc_add<DET>(&g.J[(size_t)(oc + q / E::ND) * nbf + od + q % E::ND], jcd[q], fxs);
            }
        }
#pragma unroll
        for (int ab = 0; ab < NA * NB; ++ab) {
            const double s = wave_sum(double(jab[ab]));
            if (lane == 0) acc_add<DET>(&ljab[ab], s, fxs);
        }
        __syncthreads();
        // flush the item's rows and J'_ab (partial sums: doubles, or with DET
        // int64 in the same bytes); entries are left zero for the next item
        for (int e = threadIdx.x; e < nrow + NA * NB; e += BLOCK) {
            double *dst;
            if (e < nrow) {
                const int kk = e / (NAB * nbf), rc = e - kk * NAB * nbf, r = rc / nbf, col = rc - r * nbf;
                const int row = r < NA ? oa + r : ob + r - NA;
                dst = &(kk == 0 ? g.K0 : g.K1)[(size_t)row * nbf + col];
            } else {
                const int ab = e - nrow;
                dst = &g.J[(size_t)(oa + ab / NB) * nbf + ob + ab % NB];
            }
            if constexpr (DET) {
                unsigned long long *li = reinterpret_cast<unsigned long long *>(lds_acc);
                const unsigned long long v = li[e];
                if (v == 0) continue;
                li[e] = 0;
                atomicAdd(reinterpret_cast<unsigned long long *>(dst), v);
            } else {
                const double v = lds_acc[e];
                if (v == 0.0) continue;
                lds_acc[e] = 0.0;
                atomicAdd(dst, v);
            }
        }
        __syncthreads();
    }
    if (g.count) {
        unsigned long long tot = counted, ptot = pcounted;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            tot += __shfl_xor(tot, o, 64);
            ptot += __shfl_xor(ptot, o, 64);
        }
        if (lane == 0) {
            atomicAdd(g.count, tot);
            if (g.pcount) atomicAdd(g.pcount, ptot);
        }
    }
}

template <int LA, int LB, typename R>
__global__ void __launch_bounds__(64) schwarz_kernel(PairView<R> pv, const R *__restrict__ tab, R *Q) {
    using E = Eri<LA, LB, LA, LB, R>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pv.n) return;
    R acc[E::NACC];
#pragma unroll
    for (int k = 0; k < E::NACC; ++k) acc[k] = R(0);
    for (int u = 0; u < pv.kp; ++u) {
        const R *kq = pv.prim + (size_t)u * PF_NFIELD * pv.npad + i;
        for (int v = 0; v < pv.kp; ++v) {
            const R *bp = pv.prim + (size_t)v * PF_NFIELD * pv.npad + i;
            E::prim(bp[PF_P * pv.npad], bp[PF_O2P * pv.npad], bp[PF_PX * pv.npad], bp[PF_PY * pv.npad],
                    bp[PF_PZ * pv.npad], bp[PF_AX * pv.npad], bp[PF_AY * pv.npad], bp[PF_AZ * pv.npad],
                    bp[PF_K * pv.npad], kq[PF_P * pv.npad], kq[PF_O2P * pv.npad], kq[PF_PX * pv.npad],
                    kq[PF_PY * pv.npad], kq[PF_PZ * pv.npad], kq[PF_AX * pv.npad], kq[PF_AY * pv.npad],
                    kq[PF_AZ * pv.npad], kq[PF_K * pv.npad], tab, acc);
        }
    }
    R eri[E::NINT];
    const R ABx = pv.AB[i], ABy = pv.AB[pv.npad + i], ABz = pv.AB[2 * pv.npad + i];
    E::hrr(acc, ABx, ABy, ABz, ABx, ABy, ABz, eri);
    constexpr int NAB = E::NA * E::NB;
    R m = R(0);
#pragma unroll
    for (int ab = 0; ab < NAB; ++ab) m = max(m, fabs(eri[ab * NAB + ab]));
    Q[i] = sqrt(m);
}

// Direct-mode work plan, step 1 (one thread per bra pair i): the ket prefix
// that can pass Q_i Q_j dmax >= tau (kets sorted by Q descending, binary
// search; same-class launches stop at j = i) and its number of items.
template <typename R>
__global__ void __launch_bounds__(256) jk_count_kernel(JKArgs<R> g) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= g.bra.n) return;
    const R qthr = g.tau / (g.bra.Q[i] * (g.dmax ? *g.dmax : R(1)));
    int lo = 0, hi = g.ket.n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (g.ket.Q[mid] >= qthr) lo = mid + 1;
        else hi = mid;
    }
    const int nk = g.same ? min(lo, i + 1) : lo;
    g.nket[i] = nk;
    g.item_off[i] = (nk + g.chunk - 1) / g.chunk;
}

// Step 2 (one workgroup): exclusive prefix sum of the item counts in place,
// item_off[n] = total, and the queue head reset (work[0] = 0, work[1] = total).
__global__ void __launch_bounds__(1024) jk_scan_kernel(int n, int *item_off, const int *nket, long long *joff,
                                                       long long jcap, int *work);

// Step 4 (after the direct kernel, when the J'_cd buffer was used): column
// sums.  The bras whose prefix holds ket j form a contiguous range [lo, hi):
// prefixes shrink as Q_bra falls (same-class lists start at bra = ket), so
// jk_jrange_kernel finds hi by binary search.  jcd_reduce_kernel then takes
// JRED_TILE bras per blockIdx.y and one ket-pair component per thread; the
// tile's sum goes to J' with one device atomic (one per component and tile
// instead of one per quartet).
template <typename R>
__global__ void __launch_bounds__(256) jk_jrange_kernel(JKArgs<R> g, int *jhi) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= g.ket.n || !g.work[2]) return;
    int lo = g.same ? j : 0, hi = g.bra.n;  // first bra with nket <= j
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (g.nket[mid] > j) lo = mid + 1;
        else hi = mid;
    }
    jhi[j] = lo;
}
constexpr int JRED_TILE = 256;
template <int NC, int ND, typename R, bool DET>
__global__ void __launch_bounds__(256) jcd_reduce_kernel(JKArgs<R> g, const int *jhi) {
    constexpr int NCD = NC * ND;
    if (!g.work[2]) return;
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    const int j = (int)(t / NCD), k = (int)(t - (long long)j * NCD);
    if (j >= g.ket.n) return;
    const int i0 = max((int)blockIdx.y * JRED_TILE, g.same ? j : 0);
    const int i1 = min(jhi[j], (int)(blockIdx.y + 1) * JRED_TILE);
    double s = 0.0;
#pragma unroll 8
    for (int i = i0; i < i1; ++i) s += g.jbuf[(g.joff[i] + j) * NCD + k];
    if (s != 0.0) acc_add<DET>(&g.J[(size_t)(g.ket.oa[j] + k / ND) * g.nbf + g.ket.ob[j] + k % ND], s, DET ? *g.fx : 0.0);
}

template <int LA, int LB, int LC, int LD, typename R, int NK, bool DET>
void launch_direct(const JKArgs<R> &args, hipStream_t st) {
    constexpr int W = jk_direct_waves<LA, LB, LC, LD, R>();
    constexpr int NA = nc(LA), NB = nc(LB);
    if (!args.work || !args.item_off || !args.nket || !args.joff)
        throw Error("jk: direct mode needs its work-plan buffers");
    JKArgs<R> a = args;
    // a planned build counted its items with the plan's chunk
    if (!a.planned || a.chunk <= 0) a.chunk = jk_chunk<LA, LB, LC, LD, R>();
    const auto kern = jk_direct_kernel<LA, LB, LC, LD, R, NK, DET>;
    // launch shape per (device, nbf), worked out once per thread: the
    // attribute and occupancy queries cost host time on every build otherwise
    struct Shape {
        int dev = -1, nbf = -1, grid = 0;
        bool krows = false;
        size_t dyn = 0;
    };
    static thread_local Shape sh;
    int dev = 0;
    PF_HIP(hipGetDevice(&dev));
    if (sh.dev != dev || sh.nbf != a.nbf) {
        int ncu = 0, maxlds = 0, per = 0;
        PF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        PF_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        hipFuncAttributes fa;
        PF_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(kern)));
        const size_t rows = sizeof(double) * (size_t)NK * (NA + NB) * a.nbf, jab = sizeof(double) * NA * NB;
        const bool krows = fa.sharedSizeBytes + rows + jab <= (size_t)maxlds;
        const size_t dyn = (krows ? rows : 0) + jab;
        // the limit is process-wide state: always the device maximum, so every
        // thread's cached shape stays launchable whatever another thread set
        if (fa.sharedSizeBytes + dyn > 65536)
            PF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       maxlds - (int)fa.sharedSizeBytes));
        PF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 64 * W, dyn));
        sh = Shape{dev, a.nbf, std::max(1, per) * std::max(1, ncu), krows, dyn};  // persistent: all that fit
    }
    a.krows = sh.krows && !args.katomic;
    const size_t dyn = a.krows ? sh.dyn : sizeof(double) * NA * NB;
    const int grid = sh.grid;
    if (!a.planned) {
        jk_count_kernel<R><<<(a.bra.n + 255) / 256, 256, 0, st>>>(a);
        jk_scan_kernel<<<1, 1024, 0, st>>>(a.bra.n, a.item_off, a.nket, a.joff, a.jbuf ? a.jcap : -1, a.work);
    }
    if (a.ev0) PF_HIP(hipEventRecord(a.ev0, st));
    kern<<<grid, 64 * W, dyn, st>>>(a, a.bra.aos, a.item_off);
    if (a.ev1) PF_HIP(hipEventRecord(a.ev1, st));
    if (a.jbuf && !a.planned) {
        constexpr int NCD = nc(LC) * nc(LD);
        const long long nt = (long long)a.ket.n * NCD;
        jk_jrange_kernel<R><<<(a.ket.n + 255) / 256, 256, 0, st>>>(a, a.jhi);
        const dim3 grid((unsigned)((nt + 255) / 256), (a.bra.n + JRED_TILE - 1) / JRED_TILE);
        jcd_reduce_kernel<nc(LC), nc(LD), R, DET><<<grid, 256, 0, st>>>(a, a.jhi);
    }
}

template <int LA, int LB, int LC, int LD, typename R>
void launch_nk_stored(int nk, int mode, const JKArgs<R> &args, dim3 grid, dim3 block, hipStream_t st) {
    if constexpr (std::is_same<R, double>::value) {
        if (mode == MODE_STORE) { jk_kernel<LA, LB, LC, LD, R, 1, MODE_STORE><<<grid, block, 0, st>>>(args, args.bra.aos); return; }
        if (mode == MODE_PACKED) { jk_kernel<LA, LB, LC, LD, R, 1, MODE_PACKED><<<grid, block, 0, st>>>(args, args.bra.aos); return; }
        if (mode == MODE_LOAD) {
            if (nk == 1) jk_kernel<LA, LB, LC, LD, R, 1, MODE_LOAD><<<grid, block, 0, st>>>(args, args.bra.aos);
            else jk_kernel<LA, LB, LC, LD, R, 2, MODE_LOAD><<<grid, block, 0, st>>>(args, args.bra.aos);
            return;
        }
    }
    throw Error("jk: mode not available in this precision");
}


template <int LA, int LB, int LC, int LD, typename R>
void launch_nk_mode(int nk, int mode, const JKArgs<R> &args, hipStream_t st) {
    if (args.bra.n == 0) {  // nothing to launch; the events still bracket an (empty) interval
        if (args.ev0) PF_HIP(hipEventRecord(args.ev0, st));
        if (args.ev1) PF_HIP(hipEventRecord(args.ev1, st));
        return;
    }
    constexpr int W = jk_waves<LA, LB, LC, LD>();
    const dim3 grid((args.bra.n + W - 1) / W), block(64 * W);
    if (mode != MODE_DIRECT) {  // the stored-integral kernels are one launch: time it whole
        if (args.ev0) PF_HIP(hipEventRecord(args.ev0, st));
        launch_nk_stored<LA, LB, LC, LD, R>(nk, mode, args, grid, block, st);
        if (args.ev1) PF_HIP(hipEventRecord(args.ev1, st));
        return;
    }
    if (args.fx) {
        if (nk == 1) launch_direct<LA, LB, LC, LD, R, 1, true>(args, st);
        else launch_direct<LA, LB, LC, LD, R, 2, true>(args, st);
    } else {
        if (nk == 1) launch_direct<LA, LB, LC, LD, R, 1, false>(args, st);
        else launch_direct<LA, LB, LC, LD, R, 2, false>(args, st);
    }
}

}  // namespace pf
