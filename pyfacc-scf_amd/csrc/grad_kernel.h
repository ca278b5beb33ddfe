// Nuclear gradient kernels of the SCF energy (FP64).
//
//   dE/dR = sum D dH/dR - sum W dS/dR + sum_unique deg Gamma d(ab|cd)/dR + dE_nuc/dR
//   W = 1/2 D F D (RHF, D with the factor 2), W = sum_s D_s F_s D_s (UHF)
//   Gamma_abcd = 4 Dt_ab Dt_cd - f sum_s (Ds_ac Ds_bd + Ds_ad Ds_bc)
//   (RHF: f = 1, Ds = Dt; UHF: f = 2 over the two spins), deg as in the J/K digest.
//
// Derivative of a primitive Cartesian Gaussian with respect to its centre:
//   d/dA_x [a] = 2 alpha [a + 1_x] - a_x [a - 1_x]
// so the derivative block of a contracted quartet is an ordinary contracted
// ERI of one higher angular momentum on the differentiated centre, computed
// with pair coefficients weighted by 2 alpha, minus the lower block; centre D
// follows from translational invariance, dD = -(dA + dB + dC).  Classes with
// b and d s-type get all three blocks from one recurrence (DFused); the
// others take centre B from the reversed pair (b a| and centre C from the
// swapped quartet (cd|ab).  The reference has no gradients; this extends the
// ERI path of libint_f.f90:303-625 / rscf.f90:118-137.
#pragma once
#include "eri_core.h"
#include "kernels.h"

namespace pf {

struct GradArgs {
    PairView<double> bra, ket;        // normal pair data (Q, shells, AB, prim)
    const double *bra_rec = nullptr;  // bra records [npad][kp][PF_AOS] (GR_ fields): one wide scalar load each
    const double *ket_w = nullptr;    // ket prim data in shell order, K weighted by 2 gamma (first shell)
    const double *ket_r = nullptr;    // ket prim data in shell order, K unweighted
    int same = 0;
    double tau = 0;
    int nbf = 0, nsh = 0;
    const double *Dt = nullptr, *Dk0 = nullptr, *Dk1 = nullptr;
    const double *Dsh = nullptr, *dmax = nullptr;
    double *gbra = nullptr;  // [bra.n][6]: dA xyz, dB xyz
    double *gket = nullptr;  // [ket.n][6]: dC xyz, dD xyz
    const double *tab = nullptr;
    unsigned long long *count = nullptr;
};

// Contracted block of Eri<LA,LB,LC,LD> for bra pair i (prim data bp) and ket
// pair j (prim data kp); AB / CD the pair separations; tab the workgroup's
// Boys table of order MT (BoysLds<MT>).
template <int LA, int LB, int LC, int LD, int MT>
__device__ __forceinline__ void contracted_block(const double *__restrict__ bp, int bnpad, int bi, int kpb,
                                                 const double *__restrict__ kq, int knpad, int kj, int kpk,
                                                 double ABx, double ABy, double ABz, double CDx, double CDy,
                                                 double CDz, const double *__restrict__ tab, double *out) {
    using E = Eri<LA, LB, LC, LD, double>;
    double acc[E::NACC];
#pragma unroll
    for (int k = 0; k < E::NACC; ++k) acc[k] = 0.0;
    for (int u = 0; u < kpk; ++u) {
        const double *q = kq + (size_t)u * PF_NFIELD * knpad + kj;
        const double qq = q[PF_P * knpad], Qx = q[PF_PX * knpad], Qy = q[PF_PY * knpad], Qz = q[PF_PZ * knpad];
        const double QCx = q[PF_AX * knpad], QCy = q[PF_AY * knpad], QCz = q[PF_AZ * knpad], Kk = q[PF_K * knpad];
        const double o2q = q[PF_O2P * knpad];
        for (int v = 0; v < kpb; ++v) {
            const double *b = bp + (size_t)v * PF_NFIELD * bnpad + bi;
            E::template prim<true, MT>(b[PF_P * bnpad], b[PF_O2P * bnpad], b[PF_PX * bnpad], b[PF_PY * bnpad],
                                       b[PF_PZ * bnpad], b[PF_AX * bnpad], b[PF_AY * bnpad], b[PF_AZ * bnpad],
                                       b[PF_K * bnpad], qq, o2q, Qx, Qy, Qz, QCx, QCy, QCz, Kk, tab, acc);
        }
    }
    E::hrr(acc, ABx, ABy, ABz, CDx, CDy, CDz, out);
}

// The same with the bra pair's primitives from its records (bi uniform: one
// wide scalar load per primitive), K field KF and displacement fields DX
// (GR_AX: P - A, GR_BX: P - B).
template <int LA, int LB, int LC, int LD, int MT, int KF, int DX>
__device__ __forceinline__ void contracted_block_rec(const double *__restrict__ brec, int bi, int kpb,
                                                     const double *__restrict__ kq, int knpad, int kj, int kpk,
                                                     double ABx, double ABy, double ABz, double CDx, double CDy,
                                                     double CDz, const double *__restrict__ tab, double *out) {
    using E = Eri<LA, LB, LC, LD, double>;
    double acc[E::NACC];
#pragma unroll
    for (int k = 0; k < E::NACC; ++k) acc[k] = 0.0;
    const double *bpair = brec + (size_t)bi * kpb * PF_AOS;
    for (int u = 0; u < kpk; ++u) {
        const double *q = kq + (size_t)u * PF_NFIELD * knpad + kj;
        const double qq = q[PF_P * knpad], Qx = q[PF_PX * knpad], Qy = q[PF_PY * knpad], Qz = q[PF_PZ * knpad];
        const double QCx = q[PF_AX * knpad], QCy = q[PF_AY * knpad], QCz = q[PF_AZ * knpad], Kk = q[PF_K * knpad];
        const double o2q = q[PF_O2P * knpad];
        for (int v = 0; v < kpb; ++v) {
            const double *b = bpair + (size_t)v * PF_AOS;
            E::template prim<true, MT>(b[GR_P], b[GR_O2P], b[GR_PX], b[GR_PY], b[GR_PZ], b[DX], b[DX + 1], b[DX + 2],
                                       b[KF], qq, o2q, Qx, Qy, Qz, QCx, QCy, QCz, Kk, tab, acc);
        }
    }
    E::hrr(acc, ABx, ABy, ABz, CDx, CDy, CDz, out);
}

template <typename R>
__device__ __forceinline__ R gwave_sum(R v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---- Fused derivative blocks of the classes (a s|c s), a, c in {s, p}.
// One primitive loop serves all three centres: the recurrences run once to
// total order L + 1 and give [a+1|c], [a|c+1] and the lower [a|c], [a-1|c],
// [a|c-1]; the accumulations differ only in their weights (2 alpha K for A,
// 2 beta K for B, 2 gamma K for C, K for the lower blocks).  Centre B's block
// follows from the HRR (a, b+1_x| = (a+1_x, b| + AB_x (a, b| with b = s.

// Compact [g|h]^(m) for Lg <= LE, Lh <= LF, m <= T - Lg - Lh, g-major; entries
// with Lg + Lh > T are absent.
template <int LE, int LF, int T>
struct DVrr {
    static constexpr int GE = coff(LE + 1), GF = coff(LF + 1);
    static constexpr int nm(int g, int h) { return T - cL(g) - cL(h) + 1 > 0 ? T - cL(g) - cL(h) + 1 : 0; }
    static constexpr int base(int g, int h) {
        int o = 0;
        for (int gg = 0; gg < GE; ++gg)
            for (int hh = 0; hh < GF; ++hh) {
                if (gg == g && hh == h) return o;
                o += nm(gg, hh);
            }
        return o;
    }
    static constexpr int NTH = base(GE, 0);
};

template <int LA, int LC>
struct DFused {
    static constexpr int T = LA + LC + 1;
    using V = DVrr<LA + 1, LC + 1, T>;
    static constexpr int NA = nc(LA), NC = nc(LC), NA1 = nc(LA + 1), NC1 = nc(LC + 1);
    // accumulators: A1 = [a+1|c] (2 alpha K), B1 = [a+1|c] and B0 = [a|c] (2 beta K),
    // C1 = [a|c+1] (2 gamma K), A0 = [s|c] (K, p on A), C0 = [a|s] (K, p on C)
    static constexpr int OA1 = 0, OB1 = OA1 + NA1 * NC, OB0 = OB1 + NA1 * NC, OC1 = OB0 + NA * NC;
    static constexpr int OA0 = OC1 + NA * NC1, OC0 = OA0 + (LA ? NC : 0), NACC = OC0 + (LC ? NA : 0);

    // bra primitive: p, 1/2p, P, P - A and the weights K, 2 alpha K, 2 beta K;
    // ket primitive: q, 1/2q, Q, Q - C and the weights K, 2 gamma K.
    __device__ __forceinline__ static void prim(double p, double o2p, double Px, double Py, double Pz, double PAx,
                                                double PAy, double PAz, double Kb, double Ka, double Kbb, double q,
                                                double o2q, double Qx, double Qy, double Qz, double QCx, double QCy,
                                                double QCz, double Kk, double Kg, const double *tab, double *acc) {
        const double r = rsqrt_nr(p + q), ipq = r * r;
        const double rop = q * ipq, roq = p * ipq, rho = p * rop;
        const double PQ[3] = {Px - Qx, Py - Qy, Pz - Qz};
        const double X = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
        double th[V::NTH];
        static_assert(V::base(0, 0) == 0 && V::nm(0, 0) == T + 1, "[s|s] orders first");
        BoysLds<T, double>::eval(X, tab, th);
        const double PA[3] = {PAx, PAy, PAz}, QC[3] = {QCx, QCy, QCz};
        const double WP[3] = {-rop * PQ[0], -rop * PQ[1], -rop * PQ[2]};
        const double WQ[3] = {roq * PQ[0], roq * PQ[1], roq * PQ[2]};
        const double o2pq = 0.5 * ipq;
        static_for<1, V::GE>([&](auto gi) {
            constexpr int g = decltype(gi)::value;
            constexpr int d = cdir(g), g1 = cminus(g, d), em = ccomp(g, d) - 1;
            constexpr int bg = V::base(g, 0), bg1 = V::base(g1, 0);
            static_for<0, V::nm(g, 0)>([&](auto mi) {
                constexpr int m = decltype(mi)::value;
                double v = PA[d] * th[bg1 + m] + WP[d] * th[bg1 + m + 1];
                if constexpr (em > 0) {
                    constexpr int bg2 = V::base(cminus(g1, d), 0);
                    v += double(em) * o2p * (th[bg2 + m] - rop * th[bg2 + m + 1]);
                }
                th[bg + m] = v;
            });
        });
        static_for<1, V::GF>([&](auto hi) {
            constexpr int h = decltype(hi)::value;
            constexpr int d = cdir(h), h1 = cminus(h, d), fm = ccomp(h, d) - 1;
            static_for<0, V::GE>([&](auto gi) {
                constexpr int g = decltype(gi)::value;
                constexpr int eg = ccomp(g, d);
                constexpr int bgh = V::base(g, h), bgh1 = V::base(g, h1);
                static_for<0, V::nm(g, h)>([&](auto mi) {
                    constexpr int m = decltype(mi)::value;
                    double v = QC[d] * th[bgh1 + m] + WQ[d] * th[bgh1 + m + 1];
                    if constexpr (fm > 0) {
                        constexpr int bgh2 = V::base(g, cminus(h1, d));
                        v += double(fm) * o2q * (th[bgh2 + m] - roq * th[bgh2 + m + 1]);
                    }
                    if constexpr (eg > 0) v += double(eg) * o2pq * th[V::base(cminus(g, d), h1) + m + 1];
                    th[bgh + m] = v;
                });
            });
        });
        const double rk = r * Kk;
        const double wA = Ka * rk, wB = Kbb * rk, w0 = Kb * rk, wC = Kb * Kg * r;
        static_for<0, NC>([&](auto ci) {
            constexpr int c = decltype(ci)::value;
            static_for<0, NA1>([&](auto ai) {
                constexpr int a1 = decltype(ai)::value;
                constexpr int b = V::base(coff(LA + 1) + a1, coff(LC) + c);
                acc[OA1 + a1 * NC + c] = fma(wA, th[b], acc[OA1 + a1 * NC + c]);
                acc[OB1 + a1 * NC + c] = fma(wB, th[b], acc[OB1 + a1 * NC + c]);
            });
            static_for<0, NA>([&](auto ai) {
                constexpr int a = decltype(ai)::value;
                constexpr int b = V::base(coff(LA) + a, coff(LC) + c);
                acc[OB0 + a * NC + c] = fma(wB, th[b], acc[OB0 + a * NC + c]);
            });
            if constexpr (LA == 1) {
                constexpr int b = V::base(0, coff(LC) + c);
                acc[OA0 + c] = fma(w0, th[b], acc[OA0 + c]);
            }
        });
        static_for<0, NA>([&](auto ai) {
            constexpr int a = decltype(ai)::value;
            static_for<0, NC1>([&](auto ci) {
                constexpr int c1 = decltype(ci)::value;
                constexpr int b = V::base(coff(LA) + a, coff(LC + 1) + c1);
                acc[OC1 + a * NC1 + c1] = fma(wC, th[b], acc[OC1 + a * NC1 + c1]);
            });
            if constexpr (LC == 1) {
                constexpr int b = V::base(coff(LA) + a, 0);
                acc[OC0 + a] = fma(w0, th[b], acc[OC0 + a]);
            }
        });
    }

    // d/dA, d/dB, d/dC of sum Gamma (a s|c s) from the contracted
    // accumulators; gam[a * NC + c]
    __device__ __forceinline__ static void contract(const double *acc, const double *gam, double ABx, double ABy,
                                                    double ABz, double *dA, double *dB, double *dC) {
        const double AB[3] = {ABx, ABy, ABz};
        static_for<0, NA>([&](auto ai) {
            constexpr int a = decltype(ai)::value;
            static_for<0, 3>([&](auto xi) {
                constexpr int x = decltype(xi)::value;
                constexpr int ap = cplus(coff(LA) + a, x) - coff(LA + 1);
                static_for<0, NC>([&](auto ci) {
                    constexpr int c = decltype(ci)::value;
                    constexpr int cp = cplus(coff(LC) + c, x) - coff(LC + 1);
                    const double gv = gam[a * NC + c];
                    dA[x] += gv * acc[OA1 + ap * NC + c];
                    dB[x] += gv * (acc[OB1 + ap * NC + c] + AB[x] * acc[OB0 + a * NC + c]);
                    dC[x] += gv * acc[OC1 + a * NC1 + cp];
                });
            });
        });
        if constexpr (LA == 1)
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
                for (int c = 0; c < NC; ++c) dA[x] -= gam[x * NC + c] * acc[OA0 + c];
        if constexpr (LC == 1)
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
                for (int a = 0; a < NA; ++a) dC[x] -= gam[a * NC + x] * acc[OC0 + a];
    }
};

// Wavefronts per workgroup of the derivative kernel: they share one LDS copy
// of the Boys table and take work items from a queue.
constexpr int GRAD_W = 4;

// Work plan of a derivative class (bra blocks of ib pairs, items of 64-ket
// tiles), written to work[0 .. nblk + 1]; grad.hip.
void launch_grad_plan(const GradArgs &g, int ib, int nblk, int *work, hipStream_t st);

// Persistent: every wavefront takes items (bra block, ket tile) from the
// queue until the plan's item count.  Lanes hold one ket pair each for the
// whole item, loop over the block's bra pairs, and keep the ket-side
// derivatives (C, D) in registers until the item ends; the bra side (A, B)
// is a wavefront sum per bra pair.  One device atomic per value per lane and
// item for the kets, per bra pair and item for the bras.
template <int LA, int LB, int LC, int LD, int NK>
__global__ void __launch_bounds__(64 * GRAD_W) grad2e_kernel(GradArgs g, int ib, int nblk, int *work) {
    constexpr int NA = nc(LA), NB = nc(LB), NC = nc(LC), ND = nc(LD);
    constexpr int NAp = nc(LA + 1), NBp = nc(LB + 1), NCp = nc(LC + 1);
    constexpr bool FUSED = LB == 0 && LD == 0;
    // one table of the derivative blocks' order L + 1 serves the L - 1 blocks too
    constexpr int MT = LA + LB + LC + LD + 1;
    using BT = BoysLds<MT, double>;
    __shared__ double lds_boys[BT::SIZE];
    BT::fill(lds_boys, g.tab, threadIdx.x, 64 * GRAD_W);
    __syncthreads();
    const PairView<double> &B = g.bra;
    const PairView<double> &K = g.ket;
    const int lane = threadIdx.x & 63;
    const int *off = work + 1;
    const int total = off[nblk];
    const double dm = g.dmax ? *g.dmax : 1.0;
    const int nbf = g.nbf, ns = g.nsh;
    const double kfac = NK == 1 ? 1.0 : 2.0;
    unsigned long long counted = 0;
    while (true) {
        int t = 0;
        if (lane == 0) t = atomicAdd(work, 1);
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= total) break;
        int blk = 0, bhi = nblk - 1;  // largest block whose first item is <= t
        while (blk < bhi) {
            const int mid = (blk + bhi + 1) >> 1;
            if (off[mid] <= t) blk = mid;
            else bhi = mid - 1;
        }
        const int j0 = (t - off[blk]) * 64, j = j0 + lane;
        const int i0 = blk * ib, i1 = min(B.n, i0 + ib);
        const bool jin = j < K.n;
        const double Qj0 = K.Q[j0];
        int c = 0, d = 0, oc = 0, od = 0;
        double Qj = 0.0, CDx = 0.0, CDy = 0.0, CDz = 0.0;
        if (jin) {
            c = K.sa[j]; d = K.sb[j]; oc = K.oa[j]; od = K.ob[j];
            Qj = K.Q[j];
            CDx = K.AB[j]; CDy = K.AB[K.npad + j]; CDz = K.AB[2 * K.npad + j];
        }
        double gC[3] = {0, 0, 0}, gD[3] = {0, 0, 0};
        bool touched = false;
        for (int i = i0; i < i1; ++i) {
            const double Qi = B.Q[i];
            const double qthr = g.tau / (Qi * dm * dm);
            if (Qj0 < qthr) break;  // Q_i falls along the block: no later bra reaches this tile
            if (g.same && j0 > i) continue;
            const int a = B.sa[i], b = B.sb[i], oa = B.oa[i], ob = B.ob[i];
            bool act = jin && Qj >= qthr && (!g.same || j <= i);
            if (act) {
                const double ab = g.Dsh[a * ns + b], cd = g.Dsh[c * ns + d];
                const double x = g.Dsh[a * ns + c] * g.Dsh[b * ns + d] + g.Dsh[a * ns + d] * g.Dsh[b * ns + c];
                const double dq = max(4.0 * ab * cd, kfac * x);
                act = Qi * Qj * dq >= g.tau;
            }
            if (__ballot(act) == 0) continue;
            const double ABx = B.AB[i], ABy = B.AB[B.npad + i], ABz = B.AB[2 * B.npad + i];
            double dA[3] = {0, 0, 0}, dB[3] = {0, 0, 0}, dC[3] = {0, 0, 0};
            if (act) {
                ++counted;
                const double deg = ((a == b) ? 0.5 : 1.0) * ((c == d) ? 0.5 : 1.0) * ((g.same && i == j) ? 0.5 : 1.0);
                // Gamma block
                double gam[NA * NB * NC * ND];
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                    for (int ib2 = 0; ib2 < NB; ++ib2)
#pragma unroll
                        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                            for (int id = 0; id < ND; ++id) {
                                const int mu = oa + ia, nu = ob + ib2, la = oc + ic, si = od + id;
                                double v = 4.0 * g.Dt[(size_t)mu * nbf + nu] * g.Dt[(size_t)la * nbf + si];
#pragma unroll
                                for (int kk = 0; kk < NK; ++kk) {
                                    const double *Dk = kk == 0 ? g.Dk0 : g.Dk1;
                                    v -= kfac * (Dk[(size_t)mu * nbf + la] * Dk[(size_t)nu * nbf + si] +
                                                 Dk[(size_t)mu * nbf + si] * Dk[(size_t)nu * nbf + la]);
                                }
                                gam[((ia * NB + ib2) * NC + ic) * ND + id] = deg * v;
                            }
                if constexpr (FUSED) {
                    using DF = DFused<LA, LC>;
                    double acc[DF::NACC];
#pragma unroll
                    for (int k = 0; k < DF::NACC; ++k) acc[k] = 0.0;
                    for (int u = 0; u < K.kp; ++u) {
                        const size_t ko = (size_t)u * PF_NFIELD * K.npad + j;
                        const double *q = g.ket_w + ko;
                        const double qq = q[PF_P * K.npad], o2q = q[PF_O2P * K.npad];
                        const double Qx = q[PF_PX * K.npad], Qy = q[PF_PY * K.npad], Qz = q[PF_PZ * K.npad];
                        const double QCx = q[PF_AX * K.npad], QCy = q[PF_AY * K.npad], QCz = q[PF_AZ * K.npad];
                        const double Kg = q[PF_K * K.npad], Kk = g.ket_r[ko + PF_K * K.npad];
                        const double *brec = g.bra_rec + (size_t)i * B.kp * PF_AOS;
                        for (int v = 0; v < B.kp; ++v) {
                            const double *w = brec + (size_t)v * PF_AOS;
                            DF::prim(w[GR_P], w[GR_O2P], w[GR_PX], w[GR_PY], w[GR_PZ], w[GR_AX], w[GR_AY], w[GR_AZ],
                                     w[GR_K], w[GR_KA], w[GR_KB], qq, o2q, Qx, Qy, Qz, QCx, QCy, QCz, Kk, Kg, lds_boys,
                                     acc);
                        }
                    }
                    DF::contract(acc, gam, ABx, ABy, ABz, dA, dB, dC);
                } else {
                    // ---- centre A: (a+1, b | c d) with 2 alpha weights, minus (a-1, b | c d)
                    {
                        double e1[NAp * NB * NC * ND];
                        contracted_block_rec<LA + 1, LB, LC, LD, MT, GR_KA, GR_AX>(g.bra_rec, i, B.kp, K.prim, K.npad,
                                                                                   j, K.kp, ABx, ABy, ABz, CDx, CDy,
                                                                                   CDz, lds_boys, e1);
#pragma unroll
                        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                            for (int x = 0; x < 3; ++x) {
                                const int ap = cplus(coff(LA) + ia, x) - coff(LA + 1);
#pragma unroll
                                for (int r = 0; r < NB * NC * ND; ++r)
                                    dA[x] += gam[ia * NB * NC * ND + r] * e1[ap * NB * NC * ND + r];
                            }
                        if constexpr (LA == 1) {
                            double e0[NB * NC * ND];
                            contracted_block_rec<0, LB, LC, LD, MT, GR_K, GR_AX>(g.bra_rec, i, B.kp, K.prim, K.npad, j,
                                                                                 K.kp, ABx, ABy, ABz, CDx, CDy, CDz,
                                                                                 lds_boys, e0);
#pragma unroll
                            for (int x = 0; x < 3; ++x)
#pragma unroll
                                for (int r = 0; r < NB * NC * ND; ++r) dA[x] -= gam[x * NB * NC * ND + r] * e0[r];
                        }
                    }
                    // ---- centre B: reversed pair (b+1, a | c d) with 2 beta weights, minus (b-1, a | c d)
                    {
                        double e1[NBp * NA * NC * ND];
                        contracted_block_rec<LB + 1, LA, LC, LD, MT, GR_KB, GR_BX>(g.bra_rec, i, B.kp, K.prim, K.npad,
                                                                                   j, K.kp, -ABx, -ABy, -ABz, CDx,
                                                                                   CDy, CDz, lds_boys, e1);
#pragma unroll
                        for (int ib2 = 0; ib2 < NB; ++ib2)
#pragma unroll
                            for (int x = 0; x < 3; ++x) {
                                const int bp = cplus(coff(LB) + ib2, x) - coff(LB + 1);
#pragma unroll
                                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                                    for (int r = 0; r < NC * ND; ++r)
                                        dB[x] += gam[(ia * NB + ib2) * NC * ND + r] * e1[(bp * NA + ia) * NC * ND + r];
                            }
                        if constexpr (LB == 1) {
                            double e0[NA * NC * ND];
                            contracted_block_rec<0, LA, LC, LD, MT, GR_K, GR_BX>(g.bra_rec, i, B.kp, K.prim, K.npad, j,
                                                                                 K.kp, -ABx, -ABy, -ABz, CDx, CDy, CDz,
                                                                                 lds_boys, e0);
#pragma unroll
                            for (int x = 0; x < 3; ++x)
#pragma unroll
                                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                                    for (int r = 0; r < NC * ND; ++r)
                                        dB[x] -= gam[(ia * NB + x) * NC * ND + r] * e0[ia * NC * ND + r];
                        }
                    }
                    // ---- centre C: swapped quartet (c+1, d | a b) with 2 gamma weights, minus (c-1, d | a b)
                    {
                        double e1[NCp * ND * NA * NB];
                        contracted_block<LC + 1, LD, LA, LB, MT>(g.ket_w, K.npad, j, K.kp, B.prim, B.npad, i, B.kp, CDx,
                                                                 CDy, CDz, ABx, ABy, ABz, lds_boys, e1);
#pragma unroll
                        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                            for (int x = 0; x < 3; ++x) {
                                const int cp = cplus(coff(LC) + ic, x) - coff(LC + 1);
#pragma unroll
                                for (int id = 0; id < ND; ++id)
#pragma unroll
                                    for (int ab2 = 0; ab2 < NA * NB; ++ab2)
                                        dC[x] += gam[(ab2 * NC + ic) * ND + id] * e1[(cp * ND + id) * NA * NB + ab2];
                            }
                        if constexpr (LC == 1) {
                            double e0[ND * NA * NB];
                            contracted_block<0, LD, LA, LB, MT>(K.prim, K.npad, j, K.kp, B.prim, B.npad, i, B.kp, CDx,
                                                                CDy, CDz, ABx, ABy, ABz, lds_boys, e0);
#pragma unroll
                            for (int x = 0; x < 3; ++x)
#pragma unroll
                                for (int id = 0; id < ND; ++id)
#pragma unroll
                                    for (int ab2 = 0; ab2 < NA * NB; ++ab2)
                                        dC[x] -= gam[(ab2 * NC + x) * ND + id] * e0[id * NA * NB + ab2];
                        }
                    }
                }
                touched = true;
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    gC[x] += dC[x];
                    gD[x] -= dA[x] + dB[x] + dC[x];
                }
            }
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                const double sa = gwave_sum(dA[x]), sb = gwave_sum(dB[x]);
                if (lane == 0) {
                    atomicAdd(&g.gbra[(size_t)i * 6 + x], sa);
                    atomicAdd(&g.gbra[(size_t)i * 6 + 3 + x], sb);
                }
            }
        }
        if (touched) {
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                atomicAdd(&g.gket[(size_t)j * 6 + x], gC[x]);
                atomicAdd(&g.gket[(size_t)j * 6 + 3 + x], gD[x]);
            }
        }
    }
    if (g.count) {
        unsigned long long tot = counted;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) tot += __shfl_xor(tot, o, 64);
        if (lane == 0) atomicAdd(g.count, tot);
    }
}

// work: device ints, at least bra.n + 2 of them (queue head and item plan).
template <int LA, int LB, int LC, int LD>
void launch_grad2e(int nk, const GradArgs &g, int *work, int cap, hipStream_t st) {
    if (g.bra.n == 0 || g.ket.n == 0) return;
    const auto kern = nk == 1 ? grad2e_kernel<LA, LB, LC, LD, 1> : grad2e_kernel<LA, LB, LC, LD, 2>;
    // persistent grid: all the workgroups that fit, worked out once per thread and device
    struct Shape {
        int dev = -1, grid = 0;
    };
    static thread_local Shape sh[2];
    int dev = 0;
    PF_HIP(hipGetDevice(&dev));
    Shape &s = sh[nk == 1 ? 0 : 1];
    if (s.dev != dev) {
        int ncu = 0, per = 0;
        PF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        PF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 64 * GRAD_W, 0));
        s = Shape{dev, std::max(1, per) * std::max(1, ncu)};
    }
    // bra blocks as long as the class still makes a few items per wavefront
    // before screening (a same-class list is a triangle)
    const double tiles = (g.ket.n + 63) / 64 * (g.same ? 0.5 : 1.0);
    const double want = 4.0 * s.grid * GRAD_W;
    int ib = 16;
    while (ib > 1 && (double)((g.bra.n + ib - 1) / ib) * tiles < want) ib >>= 1;
    const int nblk = (g.bra.n + ib - 1) / ib;
    if (nblk + 2 > cap) throw Error("gradient: work plan buffer too small");
    launch_grad_plan(g, ib, nblk, work, st);
    kern<<<s.grid, 64 * GRAD_W, 0, st>>>(g, ib, nblk, work);
}

}  // namespace pf
