// Shell-quartet ERI kernels with the fused J/K digest.
//
// Work decomposition (one launch per quartet class (LA LB|LC LD), bra class
// >= ket class): one 64-lane wavefront per bra shell pair, lanes stride over
// that pair's ket list.  Pairs of each class are sorted by Schwarz bound Q
// descending, so the kets that survive Q_i Q_j dmax >= tau form a prefix found
// by binary search; a per-lane density-weighted test then removes the rest.
// The bra pair's primitive data is wavefront-uniform (scalar loads); ket data
// is structure-of-arrays so the 64 lanes read consecutive addresses.
//
// J/K digest for a unique shell quartet with degeneracy factor
// deg = (1/2)^{[a=b] + [c=d] + [ab=cd]}, into non-symmetric accumulators
// (symmetrised afterwards: J = J' + J'^T, K = K' + K'^T):
//   J'_ab += 2 deg (ab|cd) D_cd      (wavefront-reduced, one atomic per element)
//   J'_cd += 2 deg (ab|cd) D_ab
//   K'_ac += deg (ab|cd) D_bd,  K'_ad += deg (ab|cd) D_bc,
//   K'_bc += deg (ab|cd) D_ad,  K'_bd += deg (ab|cd) D_ac
// This replaces the reference's G loop (rscf.f90:121-134, uscf.f90:149-175),
// which gathers nbf^4 values from a stored list and drops every
// non-canonical permutation.
#pragma once
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

}  // namespace

// Four wavefronts per workgroup, one bra pair each.
constexpr int JK_BLOCK = 256;

template <int LA, int LB, int LC, int LD, typename R, int NK, int MODE>
__global__ void __launch_bounds__(JK_BLOCK) jk_kernel(JKArgs<R> g) {
    using E = Eri<LA, LB, LC, LD, R>;
    constexpr int NA = E::NA, NB = E::NB, NC = E::NC, ND = E::ND, NINT = E::NINT;
    constexpr int NT = BoysTraits<R>::terms;
    // Boys table orders L .. L+NT-1, shared by the workgroup's wavefronts
    __shared__ R lds_boys[BOYS_NGRID * NT];
    if constexpr (MODE == MODE_DIRECT || MODE == MODE_STORE || MODE == MODE_PACKED) {
        for (int k = threadIdx.x; k < BOYS_NGRID * NT; k += JK_BLOCK) {
            const int gpt = k / NT, m = k - gpt * NT;
            lds_boys[k] = g.tab[gpt * (BOYS_MMAX + 1) + E::L + m];
        }
        __syncthreads();
    }
    const int i = blockIdx.x * (JK_BLOCK / 64) + (threadIdx.x >> 6);
    const PairView<R> &B = g.bra;
    const PairView<R> &K = g.ket;
    if (i >= B.n) return;
    const int lane = threadIdx.x & 63;
    const R Qi = B.Q[i];

    int nket;
    if constexpr (MODE == MODE_LOAD || MODE == MODE_STORE) {
        nket = g.store_nket[i];
    } else if constexpr (MODE == MODE_PACKED) {
        nket = g.same ? i + 1 : K.n;
    } else {
        const R qthr = g.tau / (Qi * (g.dmax ? *g.dmax : R(1)));
        int lo = 0, hi = K.n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (K.Q[mid] >= qthr) lo = mid + 1;
            else hi = mid;
        }
        nket = g.same ? min(lo, i + 1) : lo;
    }

    const int a = B.sa[i], b = B.sb[i], oa = B.oa[i], ob = B.ob[i];
    const R ABx = B.AB[i], ABy = B.AB[B.npad + i], ABz = B.AB[2 * B.npad + i];
    const R deg_ab = (a == b) ? R(0.5) : R(1);
    const int nbf = g.nbf;

    R dab[NA * NB];
    R jab[NA * NB];
    if constexpr (MODE == MODE_DIRECT || MODE == MODE_LOAD) {
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
                dab[ia * NB + ib] = g.Dj[(size_t)(oa + ia) * nbf + ob + ib];
                jab[ia * NB + ib] = R(0);
            }
    }
    unsigned long long counted = 0;

    for (int j0 = 0; j0 < nket; j0 += 64) {
        const int j = j0 + lane;
        if (j >= nket) break;
        const int c = K.sa[j], d = K.sb[j], oc = K.oa[j], od = K.ob[j];
        if constexpr (MODE == MODE_DIRECT) {
            if (g.Dsh) {
                const int ns = g.nsh;
                R dq = R(2) * max(g.Dsh[a * ns + b], g.Dsh[c * ns + d]);
                dq = max(dq, max(max(g.Dsh[b * ns + d], g.Dsh[b * ns + c]), max(g.Dsh[a * ns + d], g.Dsh[a * ns + c])));
                if (Qi * K.Q[j] * dq < g.tau) continue;
            }
        }
        ++counted;
        R eri[NINT];
        if constexpr (MODE == MODE_LOAD) {
            const R *src = g.store + g.store_off[i] + (size_t)(j0 >> 6) * 64 * NINT + lane;
#pragma unroll
            for (int k = 0; k < NINT; ++k) eri[k] = src[k * 64];
        } else {
            R acc[E::NACC];
#pragma unroll
            for (int k = 0; k < E::NACC; ++k) acc[k] = R(0);
            const int kpk = K.kp, kpb = B.kp;
            for (int u = 0; u < kpk; ++u) {
                const R *kq = K.prim + (size_t)u * PF_NFIELD * K.npad + j;
                const R q = kq[PF_P * K.npad], Qx = kq[PF_PX * K.npad], Qy = kq[PF_PY * K.npad],
                        Qz = kq[PF_PZ * K.npad];
                const R QCx = kq[PF_AX * K.npad], QCy = kq[PF_AY * K.npad], QCz = kq[PF_AZ * K.npad],
                        Kk = kq[PF_K * K.npad];
                for (int v = 0; v < kpb; ++v) {
                    const R *bp = B.prim + (size_t)v * PF_NFIELD * B.npad + i;
                    E::template prim<true>(bp[PF_P * B.npad], bp[PF_PX * B.npad], bp[PF_PY * B.npad], bp[PF_PZ * B.npad],
                            bp[PF_AX * B.npad], bp[PF_AY * B.npad], bp[PF_AZ * B.npad], bp[PF_K * B.npad], q, Qx,
                            Qy, Qz, QCx, QCy, QCz, Kk, lds_boys, acc);
                }
            }
            const R CDx = K.AB[j], CDy = K.AB[K.npad + j], CDz = K.AB[2 * K.npad + j];
            E::hrr(acc, ABx, ABy, ABz, CDx, CDy, CDz, eri);
        }
        if constexpr (MODE == MODE_STORE) {
            R *dst = g.store + g.store_off[i] + (size_t)(j0 >> 6) * 64 * NINT + lane;
#pragma unroll
            for (int k = 0; k < NINT; ++k) dst[k * 64] = eri[k];
            continue;
        }
        if constexpr (MODE == MODE_PACKED) {
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
            continue;
        }
        if constexpr (MODE == MODE_DIRECT || MODE == MODE_LOAD) {
            const R deg = deg_ab * ((c == d) ? R(0.5) : R(1)) * ((g.same && i == j) ? R(0.5) : R(1));
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
                    atomicAdd(&g.J[(size_t)(oc + ic) * nbf + od + id], double(deg2 * s));
                }
            // ---- exchange, one pass per density
#pragma unroll
            for (int kk = 0; kk < NK; ++kk) {
                const R *Dk = kk == 0 ? g.Dk0 : g.Dk1;
                double *Kout = kk == 0 ? g.K0 : g.K1;
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
                        atomicAdd(&Kout[(size_t)(oa + ia) * nbf + oc + ic], double(deg * s));
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
                        atomicAdd(&Kout[(size_t)(oa + ia) * nbf + od + id], double(deg * s));
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
                        atomicAdd(&Kout[(size_t)(ob + ib) * nbf + oc + ic], double(deg * s));
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
                        atomicAdd(&Kout[(size_t)(ob + ib) * nbf + od + id], double(deg * s));
                    }
#undef ERI_
            }
        }
    }
    if constexpr (MODE == MODE_DIRECT || MODE == MODE_LOAD) {
#pragma unroll
        for (int ab = 0; ab < NA * NB; ++ab) {
            const double s = wave_sum(double(jab[ab]));
            if (lane == 0) atomicAdd(&g.J[(size_t)(oa + ab / NB) * nbf + ob + ab % NB], s);
        }
    }
    if (g.count) {
        unsigned long long tot = counted;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) tot += __shfl_xor(tot, o, 64);
        if (lane == 0) atomicAdd(g.count, tot);
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
            E::prim(bp[PF_P * pv.npad], bp[PF_PX * pv.npad], bp[PF_PY * pv.npad], bp[PF_PZ * pv.npad],
                    bp[PF_AX * pv.npad], bp[PF_AY * pv.npad], bp[PF_AZ * pv.npad], bp[PF_K * pv.npad],
                    kq[PF_P * pv.npad], kq[PF_PX * pv.npad], kq[PF_PY * pv.npad], kq[PF_PZ * pv.npad],
                    kq[PF_AX * pv.npad], kq[PF_AY * pv.npad], kq[PF_AZ * pv.npad], kq[PF_K * pv.npad], tab, acc);
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

template <int LA, int LB, int LC, int LD, typename R>
void launch_nk_mode(int nk, int mode, const JKArgs<R> &args, hipStream_t st) {
    const dim3 grid((args.bra.n + JK_BLOCK / 64 - 1) / (JK_BLOCK / 64)), block(JK_BLOCK);
    if (args.bra.n == 0) return;
    if constexpr (std::is_same<R, double>::value) {
        if (mode == MODE_STORE) { jk_kernel<LA, LB, LC, LD, R, 1, MODE_STORE><<<grid, block, 0, st>>>(args); return; }
        if (mode == MODE_PACKED) { jk_kernel<LA, LB, LC, LD, R, 1, MODE_PACKED><<<grid, block, 0, st>>>(args); return; }
        if (mode == MODE_LOAD) {
            if (nk == 1) jk_kernel<LA, LB, LC, LD, R, 1, MODE_LOAD><<<grid, block, 0, st>>>(args);
            else jk_kernel<LA, LB, LC, LD, R, 2, MODE_LOAD><<<grid, block, 0, st>>>(args);
            return;
        }
    }
    if (mode != MODE_DIRECT) throw Error("jk: mode not available in this precision");
    if (nk == 1) jk_kernel<LA, LB, LC, LD, R, 1, MODE_DIRECT><<<grid, block, 0, st>>>(args);
    else jk_kernel<LA, LB, LC, LD, R, 2, MODE_DIRECT><<<grid, block, 0, st>>>(args);
}

}  // namespace pf
