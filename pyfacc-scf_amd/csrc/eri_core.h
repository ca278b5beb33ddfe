// Electron-repulsion integrals over s/p(/d) Cartesian Gaussians:
// Obara-Saika vertical recurrence per primitive quartet, contraction, then the
// Head-Gordon-Pople horizontal recurrence on the contracted [e0|f0].
//
// Replaces the reference's Libint call path (libint_integrals.f90:46-148,
// libint_f.f90:303-625), whose per-primitive prefactors are
//   P, Q, W, PA, QC, WP, WQ, 1/2zeta, 1/2eta, 1/2(zeta+eta), rho/zeta, rho/eta
// (libint_f.f90:339-461) and (ss|ss)^(m) = 2 pi^{5/2} K1 K2 / (zeta eta sqrt(zeta+eta)) F_m(T)
// (libint_f.f90:514-604), with T = rho |PQ|^2 evaluated per primitive quartet
// (the reference's compute_boys_x averages T over the quartet, boys.f90:85-138).
//
// VRR (Obara & Saika 1986; Head-Gordon & Pople 1988):
//   [e+1_i|f]^m = PA_i [e|f]^m + WP_i [e|f]^{m+1}
//                 + e_i/2p ([e-1_i|f]^m - rho/p [e-1_i|f]^{m+1}) + f_i/2(p+q) [e|f-1_i]^{m+1}
//   [e|f+1_i]^m = QC_i [e|f]^m + WQ_i [e|f]^{m+1}
//                 + f_i/2q ([e|f-1_i]^m - rho/q [e|f-1_i]^{m+1}) + e_i/2(p+q) [e-1_i|f]^{m+1}
// HRR: (a, b+1_i| = (a+1_i, b| + AB_i (a, b|   (and the same on the ket with CD).
//
// All loops have compile-time trip counts and are fully unrolled so the
// intermediate arrays live in registers.
#pragma once
#include <type_traits>

#include "boys.h"
#include "cart.h"
#include "kernels.h"

namespace pf {

// Compile-time loop: f(std::integral_constant<int, i>) for i in [B, E).
template <int B, int E, typename Fn>
__device__ __forceinline__ void static_for(Fn &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

template <int LA, int LB, int LC, int LD, typename R>
struct Eri {
    static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
    static constexpr int NA = nc(LA), NB = nc(LB), NC = nc(LC), ND = nc(LD);
    static constexpr int E0 = coff(LA), NE = coff(LAB + 1) - coff(LA);
    static constexpr int F0 = coff(LC), NF = coff(LCD + 1) - coff(LC);
    static constexpr int GE = coff(LAB + 1), GF = coff(LCD + 1);
    static constexpr int NINT = NA * NB * NC * ND;
    static constexpr int NACC = NE * NF;

    // Compact storage of [g|h]^m: only m <= L - L(g) - L(h) is ever needed.
    static constexpr int nm(int g, int h) { return L - cL(g) - cL(h) + 1; }
    // Offset of [g|h]^0 in the compact array: entries are laid out g-major,
    // h-minor, each (g, h) holding nm(g, h) orders.  Closed form (no loops),
    // so it folds to a constant inside the unrolled recurrences:
    //   row(l)    = sum_{Lh<=LCD} nc(Lh) (L - l - Lh + 1) = (L - l + 1) coff(LCD+1) - S1(LCD+1)
    //   S1(n)     = sum_{l<n} l nc(l) = (n-1) n (n+1) (n+2) / 8
    //   base(g,h) = sum_{l<Lg} nc(l) row(l) + (g - coff(Lg)) row(Lg)
    //             + (L - Lg + 1) coff(Lh) - S1(Lh) + (h - coff(Lh)) (L - Lg - Lh + 1)
    static constexpr int S1(int n) { return (n - 1) * n * (n + 1) * (n + 2) / 8; }
    static constexpr int row(int l) { return (L - l + 1) * coff(LCD + 1) - S1(LCD + 1); }
    static constexpr int base(int g, int h) {
        const int Lg = cL(g), Lh = cL(h);
        return coff(Lg) * ((L + 1) * coff(LCD + 1) - S1(LCD + 1)) - coff(LCD + 1) * S1(Lg) +
               (g - coff(Lg)) * row(Lg) + (L - Lg + 1) * coff(Lh) - S1(Lh) + (h - coff(Lh)) * (L - Lg - Lh + 1);
    }
    static constexpr int NTH = base(GE - 1, GF - 1) + nm(GE - 1, GF - 1);

    // Add one primitive quartet to acc[NE*NF] (the contracted [e0|f0]).
    // o2p = 1/(2p), o2q = 1/(2q).  LDS: tab is a workgroup table laid out by
    // BoysLds<MT, R> (MT >= L: a table of a higher order serves too),
    // otherwise the full table (orders 0 .. BOYS_MMAX).
    template <bool LDS = false, int MT = L>
    __device__ __forceinline__ static void prim(R p, R o2p, R Px, R Py, R Pz, R PAx, R PAy, R PAz, R Kb, R q, R o2q,
                                                R Qx, R Qy, R Qz, R QCx, R QCy, R QCz, R Kk,
                                                const R *__restrict__ tab, R *acc) {
        // r = 1/sqrt(p+q) gives both 1/(p+q) and the prefactor's square root
        const R r = rsqrt_nr(p + q), ipq = r * r;
        const R rop = q * ipq, roq = p * ipq;  // rho/p, rho/q
        const R rho = p * rop;
        const R PQ[3] = {Px - Qx, Py - Qy, Pz - Qz};
        const R T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
        static_assert(MT >= L, "Boys table of too low an order");
        R F[(LDS ? MT : L) + 1];
        if constexpr (LDS) BoysLds<MT, R>::eval(T, tab, F);
        else boys_eval<L, R>(T, tab, F);
        const R pref = Kb * Kk * r;
        if constexpr (L == 0) {
            acc[0] = fma(pref, F[0], acc[0]);  // (ss|ss)
            return;
        } else if constexpr (LA == 1 && LB == 0 && LCD == 0) {
            // (ps|ss): [p_i|s] = pref (PA_i F_0 + WP_i F_1), WP_i = -(rho/p) PQ_i,
            // accumulated as two FMAs per component (acc holds [p|s] only; for
            // (sp|ss), E0 = 0, acc holds [s|s] and [p|s] and takes the general path)
            const R a = pref * F[0], b = (pref * rop) * F[1];
            acc[0] = fma(PAx, a, acc[0]);
            acc[1] = fma(PAy, a, acc[1]);
            acc[2] = fma(PAz, a, acc[2]);
            acc[0] = fma(-PQ[0], b, acc[0]);
            acc[1] = fma(-PQ[1], b, acc[1]);
            acc[2] = fma(-PQ[2], b, acc[2]);
            return;
        }
        // the recurrences are linear in [s|s]^m: run them on F_m and scale the
        // results by pref as they are accumulated (one FMA each)
        R th[NTH];
#pragma unroll
        for (int m = 0; m <= L; ++m) th[m] = F[m];
        if constexpr (LAB > 0) {
            const R PA[3] = {PAx, PAy, PAz};
            const R WP[3] = {-rop * PQ[0], -rop * PQ[1], -rop * PQ[2]};
#pragma unroll
            for (int g = 1; g < GE; ++g) {
                const int d = cdir(g), g1 = cminus(g, d), em = ccomp(g, d) - 1;
                const int bg = base(g, 0), bg1 = base(g1, 0);
#pragma unroll
                for (int m = 0; m < nm(g, 0); ++m) {
                    R v = PA[d] * th[bg1 + m] + WP[d] * th[bg1 + m + 1];
                    if (em > 0) {
                        const int bg2 = base(cminus(g1, d), 0);
                        v += R(em) * o2p * (th[bg2 + m] - rop * th[bg2 + m + 1]);
                    }
                    th[bg + m] = v;
                }
            }
        }
        if constexpr (LCD > 0) {
            const R QC[3] = {QCx, QCy, QCz};
            const R WQ[3] = {roq * PQ[0], roq * PQ[1], roq * PQ[2]};
            const R o2pq = R(0.5) * ipq;
#pragma unroll
            for (int h = 1; h < GF; ++h) {
                const int d = cdir(h), h1 = cminus(h, d), fm = ccomp(h, d) - 1;
#pragma unroll
                for (int g = 0; g < GE; ++g) {
                    const int eg = ccomp(g, d);
                    const int bgh = base(g, h), bgh1 = base(g, h1);
#pragma unroll
                    for (int m = 0; m < nm(g, h); ++m) {
                        R v = QC[d] * th[bgh1 + m] + WQ[d] * th[bgh1 + m + 1];
                        if (fm > 0) {
                            const int bgh2 = base(g, cminus(h1, d));
                            v += R(fm) * o2q * (th[bgh2 + m] - roq * th[bgh2 + m + 1]);
                        }
                        if (eg > 0) v += R(eg) * o2pq * th[base(cminus(g, d), h1) + m + 1];
                        th[bgh + m] = v;
                    }
                }
            }
        }
#pragma unroll
        for (int e = 0; e < NE; ++e)
#pragma unroll
            for (int f = 0; f < NF; ++f) acc[e * NF + f] = fma(pref, th[base(E0 + e, F0 + f)], acc[e * NF + f]);
    }

    // Contracted [e0|f0] -> (ab|cd), out[((ia*NB+ib)*NC+ic)*ND+id].
    __device__ __forceinline__ static void hrr(const R *acc, R ABx, R ABy, R ABz, R CDx, R CDy, R CDz, R *out) {
        static_assert(LB <= 1 && LD <= 1, "HRR implemented for b, d in {s, p}");
        const R AB[3] = {ABx, ABy, ABz}, CD[3] = {CDx, CDy, CDz};
        R bra[NA * NB][NF];
#pragma unroll
        for (int ia = 0; ia < NA; ++ia) {
            const int ga = coff(LA) + ia;
#pragma unroll
            for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    if constexpr (LB == 0) {
                        bra[ia * NB + ib][f] = acc[(ga - E0) * NF + f];
                    } else {
                        bra[ia * NB + ib][f] = acc[(cplus(ga, ib) - E0) * NF + f] + AB[ib] * acc[(ga - E0) * NF + f];
                    }
                }
        }
#pragma unroll
        for (int ab = 0; ab < NA * NB; ++ab)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic) {
                const int gc = coff(LC) + ic;
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    R v;
                    if constexpr (LD == 0) {
                        v = bra[ab][gc - F0];
                    } else {
                        v = bra[ab][cplus(gc, id) - F0] + CD[id] * bra[ab][gc - F0];
                    }
                    out[(ab * NC + ic) * ND + id] = v;
                }
            }
    }
};

}  // namespace pf

namespace pf {
static_assert(Eri<1, 1, 1, 1, double>::base(0, 0) == 0 && Eri<1, 1, 1, 1, double>::base(0, 1) == 5 &&
                  Eri<1, 1, 1, 1, double>::base(1, 0) == 5 + 3 * 4 + 6 * 3,
              "compact VRR layout");
static_assert(Eri<1, 1, 1, 1, double>::NTH == 200, "compact VRR size");
static_assert(Eri<1, 0, 1, 1, double>::NTH == 70 && Eri<0, 0, 0, 0, double>::NTH == 1, "compact VRR size");
static_assert(Eri<1, 1, 1, 1, double>::base(9, 9) == 199 && Eri<1, 0, 1, 1, double>::base(3, 9) == 69, "last entry");
}  // namespace pf
