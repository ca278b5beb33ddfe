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
// with pair coefficients weighted by 2 alpha (a separate primitive-pair data
// set), minus the lower block.  Centre B uses the reversed pair (b a|, centre
// C the swapped quartet (cd|ab), centre D translational invariance
// dD = -(dA + dB + dC).  The reference has no gradients; this extends the
// ERI path of libint_f.f90:303-625 / rscf.f90:118-137.
#pragma once
#include "eri_core.h"
#include "kernels.h"

namespace pf {

struct GradArgs {
    PairView<double> bra, ket;        // normal pair data (Q, shells, AB, prim)
    const double *bra_w = nullptr;    // bra prim data, K weighted by 2 alpha(first shell)
    const double *bra_rw = nullptr;   // bra reversed (P - B, K 2 beta)
    const double *bra_r = nullptr;    // bra reversed (P - B, K)
    const double *ket_w = nullptr;    // ket prim data, K weighted by 2 gamma(first shell)
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

template <typename R>
__device__ __forceinline__ R gwave_sum(R v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Wavefronts per workgroup of the derivative kernel: they share one LDS copy
// of the Boys table and take (bra pair, ket slice) items from a queue.
constexpr int GRAD_W = 4;

// Persistent: every wavefront takes items t = (bra pair i, ket slice s) from
// the queue (*queue, cleared before the launch) until t >= B.n * nslice.
template <int LA, int LB, int LC, int LD, int NK>
__global__ void __launch_bounds__(64 * GRAD_W) grad2e_kernel(GradArgs g, int nslice, int *queue) {
    constexpr int NA = nc(LA), NB = nc(LB), NC = nc(LC), ND = nc(LD);
    constexpr int NAp = nc(LA + 1), NBp = nc(LB + 1), NCp = nc(LC + 1);
    // one table of the derivative blocks' order L + 1 serves the L - 1 blocks too
    constexpr int MT = LA + LB + LC + LD + 1;
    using BT = BoysLds<MT, double>;
    __shared__ double lds_boys[BT::SIZE];
    BT::fill(lds_boys, g.tab, threadIdx.x, 64 * GRAD_W);
    __syncthreads();
    const PairView<double> &B = g.bra;
    const PairView<double> &K = g.ket;
    const int lane = threadIdx.x & 63;
    const int total = B.n * nslice;
    // each item's slice of the ket list (slices of a multiple of 64 covering
    // all K.n kets), so a class with few bra pairs still fills the GPU;
    // same-class lists end at ket i, so later slices have nothing to do
    const int slice = ((K.n + nslice - 1) / nslice + 63) & ~63;
    while (true) {
    int t = 0;
    if (lane == 0) t = atomicAdd(queue, 1);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= total) break;
    const int i = t / nslice;
    const int jbeg = (t - i * nslice) * slice;
    if (jbeg >= K.n || (g.same && jbeg > i)) continue;
    const double Qi = B.Q[i];
    const double dm = g.dmax ? *g.dmax : 1.0;
    const double qthr = g.tau / (Qi * dm * dm);
    int lo = 0, hi = K.n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (K.Q[mid] >= qthr) lo = mid + 1;
        else hi = mid;
    }
    const int nket = g.same ? min(lo, i + 1) : lo;
    const int jend = min(nket, jbeg + slice);
    if (jbeg >= jend) continue;
    const int a = B.sa[i], b = B.sb[i], oa = B.oa[i], ob = B.ob[i];
    const double ABx = B.AB[i], ABy = B.AB[B.npad + i], ABz = B.AB[2 * B.npad + i];
    const double deg_ab = (a == b) ? 0.5 : 1.0;
    const int nbf = g.nbf, ns = g.nsh;
    const double kfac = NK == 1 ? 1.0 : 2.0;
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0};
    unsigned long long counted = 0;

    for (int j0 = jbeg; j0 < jend; j0 += 64) {
        const int j = j0 + lane;
        if (j >= jend) break;
        const int c = K.sa[j], d = K.sb[j], oc = K.oa[j], od = K.ob[j];
        {
            const double ab = g.Dsh[a * ns + b], cd = g.Dsh[c * ns + d];
            const double x = g.Dsh[a * ns + c] * g.Dsh[b * ns + d] + g.Dsh[a * ns + d] * g.Dsh[b * ns + c];
            const double dq = max(4.0 * ab * cd, kfac * x);
            if (Qi * K.Q[j] * dq < g.tau) continue;
        }
        ++counted;
        const double deg = deg_ab * ((c == d) ? 0.5 : 1.0) * ((g.same && i == j) ? 0.5 : 1.0);
        const double CDx = K.AB[j], CDy = K.AB[K.npad + j], CDz = K.AB[2 * K.npad + j];
        // Gamma block
        double gam[NA * NB * NC * ND];
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                    for (int id = 0; id < ND; ++id) {
                        const int mu = oa + ia, nu = ob + ib, la = oc + ic, si = od + id;
                        double v = 4.0 * g.Dt[(size_t)mu * nbf + nu] * g.Dt[(size_t)la * nbf + si];
#pragma unroll
                        for (int kk = 0; kk < NK; ++kk) {
                            const double *Dk = kk == 0 ? g.Dk0 : g.Dk1;
                            v -= kfac * (Dk[(size_t)mu * nbf + la] * Dk[(size_t)nu * nbf + si] +
                                         Dk[(size_t)mu * nbf + si] * Dk[(size_t)nu * nbf + la]);
                        }
                        gam[((ia * NB + ib) * NC + ic) * ND + id] = deg * v;
                    }
        double dA[3] = {0, 0, 0}, dB[3] = {0, 0, 0}, dC[3] = {0, 0, 0};
        // ---- centre A: (a+1, b | c d) with 2 alpha weights, minus (a-1, b | c d)
        {
            double e1[NAp * NB * NC * ND];
            contracted_block<LA + 1, LB, LC, LD, MT>(g.bra_w, B.npad, i, B.kp, K.prim, K.npad, j, K.kp, ABx, ABy, ABz,
                                                  CDx, CDy, CDz, lds_boys, e1);
#pragma unroll
            for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const int ap = cplus(coff(LA) + ia, x) - coff(LA + 1);
#pragma unroll
                    for (int r = 0; r < NB * NC * ND; ++r) dA[x] += gam[ia * NB * NC * ND + r] * e1[ap * NB * NC * ND + r];
                }
            if constexpr (LA == 1) {
                double e0[NB * NC * ND];
                contracted_block<0, LB, LC, LD, MT>(B.prim, B.npad, i, B.kp, K.prim, K.npad, j, K.kp, ABx, ABy, ABz, CDx,
                                                CDy, CDz, lds_boys, e0);
#pragma unroll
                for (int x = 0; x < 3; ++x)
#pragma unroll
                    for (int r = 0; r < NB * NC * ND; ++r) dA[x] -= gam[x * NB * NC * ND + r] * e0[r];
            }
        }
        // ---- centre B: reversed pair (b+1, a | c d) with 2 beta weights, minus (b-1, a | c d)
        {
            double e1[NBp * NA * NC * ND];
            contracted_block<LB + 1, LA, LC, LD, MT>(g.bra_rw, B.npad, i, B.kp, K.prim, K.npad, j, K.kp, -ABx, -ABy,
                                                  -ABz, CDx, CDy, CDz, lds_boys, e1);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const int bp = cplus(coff(LB) + ib, x) - coff(LB + 1);
#pragma unroll
                    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                        for (int r = 0; r < NC * ND; ++r)
                            dB[x] += gam[(ia * NB + ib) * NC * ND + r] * e1[(bp * NA + ia) * NC * ND + r];
                }
            if constexpr (LB == 1) {
                double e0[NA * NC * ND];
                contracted_block<0, LA, LC, LD, MT>(g.bra_r, B.npad, i, B.kp, K.prim, K.npad, j, K.kp, -ABx, -ABy, -ABz,
                                                CDx, CDy, CDz, lds_boys, e0);
#pragma unroll
                for (int x = 0; x < 3; ++x)
#pragma unroll
                    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                        for (int r = 0; r < NC * ND; ++r) dB[x] -= gam[(ia * NB + x) * NC * ND + r] * e0[ia * NC * ND + r];
            }
        }
        // ---- centre C: swapped quartet (c+1, d | a b) with 2 gamma weights, minus (c-1, d | a b)
        {
            double e1[NCp * ND * NA * NB];
            contracted_block<LC + 1, LD, LA, LB, MT>(g.ket_w, K.npad, j, K.kp, B.prim, B.npad, i, B.kp, CDx, CDy, CDz,
                                                  ABx, ABy, ABz, lds_boys, e1);
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const int cp = cplus(coff(LC) + ic, x) - coff(LC + 1);
#pragma unroll
                    for (int id = 0; id < ND; ++id)
#pragma unroll
                        for (int ab = 0; ab < NA * NB; ++ab)
                            dC[x] += gam[(ab * NC + ic) * ND + id] * e1[(cp * ND + id) * NA * NB + ab];
                }
            if constexpr (LC == 1) {
                double e0[ND * NA * NB];
                contracted_block<0, LD, LA, LB, MT>(K.prim, K.npad, j, K.kp, B.prim, B.npad, i, B.kp, CDx, CDy, CDz, ABx,
                                                ABy, ABz, lds_boys, e0);
#pragma unroll
                for (int x = 0; x < 3; ++x)
#pragma unroll
                    for (int id = 0; id < ND; ++id)
#pragma unroll
                        for (int ab = 0; ab < NA * NB; ++ab)
                            dC[x] -= gam[(ab * NC + x) * ND + id] * e0[id * NA * NB + ab];
            }
        }
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            gA[x] += dA[x];
            gB[x] += dB[x];
            atomicAdd(&g.gket[(size_t)j * 6 + x], dC[x]);
            atomicAdd(&g.gket[(size_t)j * 6 + 3 + x], -(dA[x] + dB[x] + dC[x]));
        }
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        const double sa = gwave_sum(gA[x]), sb = gwave_sum(gB[x]);
        if (lane == 0) {
            atomicAdd(&g.gbra[(size_t)i * 6 + x], sa);
            atomicAdd(&g.gbra[(size_t)i * 6 + 3 + x], sb);
        }
    }
    if (g.count) {
        unsigned long long tot = counted;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) tot += __shfl_xor(tot, o, 64);
        if (lane == 0) atomicAdd(g.count, tot);
    }
    }
}

// queue: a device int, cleared here (fill kernel) before the launch.
template <int LA, int LB, int LC, int LD>
void launch_grad2e(int nk, const GradArgs &g, int *queue, hipStream_t st) {
    if (g.bra.n == 0) return;
    // items (bra pair, ket slice): slices of at least 64 kets, at most 64 of them
    const int nslice = std::max(1, std::min(64, (g.ket.n + 63) / 64));
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
    dev_fill_int(1, 0, queue, st);
    kern<<<s.grid, 64 * GRAD_W, 0, st>>>(g, nslice, queue);
}

}  // namespace pf
