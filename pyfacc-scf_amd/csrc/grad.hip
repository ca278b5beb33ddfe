// One-electron part of the nuclear gradient and the host-side dispatch of
// the two-electron derivative kernels (grad_kernel.h, instantiated per class
// in grad_<class>.hip).
//
// For every ordered shell pair (a, b) and only the derivative of the first
// function (the ordered pair (b, a) supplies the other), with
//   d/dA_x [a] = 2 alpha [a + 1_x] - a_x [a - 1_x]:
//   g_A += 2 sum_{mu in a, nu in b} ( D_mu nu dT - W_mu nu dS + sum_C D_mu nu dV_C )
//   g_C -= 2 sum D_mu nu dV_C            (Hellmann-Feynman term by translational
//                                         invariance of each nuclear-attraction term)
#include "boys.h"
#include "cart.h"
#include "grad_kernel.h"
#include "kernels.h"

namespace pf {

constexpr int GRAD1E_ATOMS = 4;  // nuclei per workgroup of the attraction term

template <int LA, int LB>
__global__ void __launch_bounds__(64) grad1e_kernel(const Shell *__restrict__ shells, const int *__restrict__ pa,
                                                     const int *__restrict__ pb, int npairs, int nbf, int natoms,
                                                     const double *__restrict__ axyz, const int *__restrict__ Z,
                                                     const double *__restrict__ tab, const double *__restrict__ D,
                                                     const double *__restrict__ W, double *gT, double *gV,
                                                     double *gS) {
    constexpr int NA = nc(LA), NB = nc(LB);
    constexpr int LT = LA + LB + 1;
    const int pid = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = pid < npairs;
    const Shell &A = shells[active ? pa[pid] : pa[0]];
    const Shell &B = shells[active ? pb[pid] : pb[0]];
    double dab[NA * NB], wab[NA * NB];
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) {
            const size_t o = (size_t)(A.bf + ia) * nbf + B.bf + ib;
            dab[ia * NB + ib] = active ? D[o] : 0.0;
            wab[ia * NB + ib] = active ? W[o] : 0.0;
        }
    double AB[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) AB[d] = A.xyz[d] - B.xyz[d];
    const double r2ab = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    double gownT[3] = {0, 0, 0}, gownV[3] = {0, 0, 0}, gownS[3] = {0, 0, 0};
    // blockIdx.y takes GRAD1E_ATOMS nuclei of the attraction term; the overlap
    // and kinetic terms belong to the first chunk
    const int cbeg = (int)blockIdx.y * GRAD1E_ATOMS, cend = min(natoms, cbeg + GRAD1E_ATOMS);
    const bool first = blockIdx.y == 0;

    // ---- overlap and kinetic
    if (first)
    for (int i = 0; i < A.np; ++i)
        for (int j = 0; j < B.np; ++j) {
            const double a = A.exp[i], b = B.exp[j], p = a + b, ip = 1.0 / p, o2p = 0.5 * ip;
            const double pre = A.coef[i] * B.coef[j] * exp(-a * b * ip * r2ab) * ip * sqrt(M_PI * ip) * M_PI;
            double s1[3][LA + 2][LB + 3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double P = (a * A.xyz[d] + b * B.xyz[d]) * ip;
                const double PA = P - A.xyz[d], PB = P - B.xyz[d];
                s1[d][0][0] = 1.0;
#pragma unroll
                for (int jj = 0; jj < LB + 2; ++jj)
                    s1[d][0][jj + 1] = PB * s1[d][0][jj] + (jj > 0 ? jj * o2p * s1[d][0][jj - 1] : 0.0);
#pragma unroll
                for (int ii = 0; ii < LA + 1; ++ii)
#pragma unroll
                    for (int jj = 0; jj <= LB + 2; ++jj)
                        s1[d][ii + 1][jj] = PA * s1[d][ii][jj] + (ii > 0 ? ii * o2p * s1[d][ii - 1][jj] : 0.0) +
                                            (jj > 0 ? jj * o2p * s1[d][ii][jj - 1] : 0.0);
            }
            auto t1d = [&](int d, int ii, int jj) {
                return b * (2 * jj + 1) * s1[d][ii][jj] - 2.0 * b * b * s1[d][ii][jj + 2] -
                       (jj >= 2 ? 0.5 * jj * (jj - 1) * s1[d][ii][jj >= 2 ? jj - 2 : 0] : 0.0);
            };
#pragma unroll
            for (int ia = 0; ia < NA; ++ia) {
                const int ga = coff(LA) + ia;
                const int ac[3] = {cx(ga), cy(ga), cz(ga)};
#pragma unroll
                for (int ib = 0; ib < NB; ++ib) {
                    const int gb = coff(LB) + ib;
                    const int bc[3] = {cx(gb), cy(gb), cz(gb)};
#pragma unroll
                    for (int x = 0; x < 3; ++x) {
                        // S and T for a + 1_x and a - 1_x
                        double sp = 1.0, sm = 1.0, tp = 0.0, tm = 0.0;
                        double sd_p[3], sd_m[3], td_p[3], td_m[3];
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            const int ip_ = ac[d] + (d == x), im_ = ac[d] - (d == x);
                            sd_p[d] = s1[d][ip_][bc[d]];
                            td_p[d] = t1d(d, ip_, bc[d]);
                            sd_m[d] = im_ >= 0 ? s1[d][im_ >= 0 ? im_ : 0][bc[d]] : 0.0;
                            td_m[d] = im_ >= 0 ? t1d(d, im_ >= 0 ? im_ : 0, bc[d]) : 0.0;
                        }
                        sp = sd_p[0] * sd_p[1] * sd_p[2];
                        tp = td_p[0] * sd_p[1] * sd_p[2] + sd_p[0] * td_p[1] * sd_p[2] + sd_p[0] * sd_p[1] * td_p[2];
                        if (ac[x] > 0) {
                            sm = sd_m[0] * sd_m[1] * sd_m[2];
                            tm = td_m[0] * sd_m[1] * sd_m[2] + sd_m[0] * td_m[1] * sd_m[2] + sd_m[0] * sd_m[1] * td_m[2];
                        } else {
                            sm = 0.0;
                        }
                        const double dS = pre * (2.0 * a * sp - ac[x] * sm);
                        const double dT = pre * (2.0 * a * tp - ac[x] * tm);
                        gownT[x] += 2.0 * dab[ia * NB + ib] * dT;
                        gownS[x] -= 2.0 * wab[ia * NB + ib] * dS;
                    }
                }
            }
        }

    // ---- nuclear attraction, one nucleus at a time (wavefront-uniform loop)
    for (int c = cbeg; c < cend; ++c) {
        double val[3] = {0, 0, 0};
        const double zc = -double(Z[c]);
        for (int i = 0; i < A.np; ++i)
            for (int j = 0; j < B.np; ++j) {
                const double a = A.exp[i], b = B.exp[j], p = a + b, ip = 1.0 / p, o2p = 0.5 * ip;
                const double vpre = A.coef[i] * B.coef[j] * exp(-a * b * ip * r2ab) * 2.0 * M_PI * ip;
                double PA[3], PC[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double P = (a * A.xyz[d] + b * B.xyz[d]) * ip;
                    PA[d] = P - A.xyz[d];
                    PC[d] = P - axyz[3 * c + d];
                }
                const double xx = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                double F[LT + 1];
                boys_eval<LT, double>(xx, tab, F);
                double th[LT + 1][coff(LT + 1)];
#pragma unroll
                for (int m = 0; m <= LT; ++m) th[m][0] = vpre * F[m];
#pragma unroll
                for (int g = 1; g < coff(LT + 1); ++g) {
                    const int L = cL(g), d = cdir(g), g1 = cminus(g, d);
                    const int em = ccomp(g, d) - 1;
#pragma unroll
                    for (int m = 0; m <= LT - L; ++m) {
                        double v = PA[d] * th[m][g1] - PC[d] * th[m + 1][g1];
                        if (em > 0) {
                            const int g2 = cminus(g1, d);
                            v += em * o2p * (th[m][g2] - th[m + 1][g2]);
                        }
                        th[m][g] = v;
                    }
                }
                auto vint = [&](int gaa, int ib) {  // (a'|V_C|b), a' global index
                    if (LB == 0) return th[0][gaa];
                    return th[0][cplus(gaa, ib)] + AB[ib] * th[0][gaa];
                };
#pragma unroll
                for (int ia = 0; ia < NA; ++ia) {
                    const int ga = coff(LA) + ia;
#pragma unroll
                    for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                        for (int x = 0; x < 3; ++x) {
                            double dv = 2.0 * a * vint(cplus(ga, x), ib);
                            if (ccomp(ga, x) > 0) dv -= ccomp(ga, x) * vint(cminus(ga, x), ib);
                            val[x] += zc * dab[ia * NB + ib] * dv;
                        }
                }
            }
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            gownV[x] += 2.0 * val[x];
            const double s = gwave_sum(val[x]);
            if (threadIdx.x == 0 && s != 0.0) atomicAdd(&gV[3 * c + x], -2.0 * s);
        }
    }
    if (active)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            if (first) {
                atomicAdd(&gT[3 * A.atom + x], gownT[x]);
                atomicAdd(&gS[3 * A.atom + x], gownS[x]);
            }
            atomicAdd(&gV[3 * A.atom + x], gownV[x]);
        }
}

void launch_grad1e(int la, int lb, const Shell *shells, const int *pa, const int *pb, int npairs, int nbf, int natoms,
                   const double *axyz, const int *Z, const double *tab, const double *D, const double *W,
                   double *gT, double *gV, double *gS, hipStream_t st) {
    if (npairs == 0) return;
    // (pair block, chunk of nuclei): one thread per pair alone leaves a
    // small class at a few wavefronts for the whole chip
    const dim3 grid((npairs + 63) / 64, std::max(1, (natoms + GRAD1E_ATOMS - 1) / GRAD1E_ATOMS)), block(64);
    if (la == 0 && lb == 0) grad1e_kernel<0, 0><<<grid, block, 0, st>>>(shells, pa, pb, npairs, nbf, natoms, axyz, Z, tab, D, W, gT, gV, gS);
    else if (la == 0 && lb == 1) grad1e_kernel<0, 1><<<grid, block, 0, st>>>(shells, pa, pb, npairs, nbf, natoms, axyz, Z, tab, D, W, gT, gV, gS);
    else if (la == 1 && lb == 0) grad1e_kernel<1, 0><<<grid, block, 0, st>>>(shells, pa, pb, npairs, nbf, natoms, axyz, Z, tab, D, W, gT, gV, gS);
    else grad1e_kernel<1, 1><<<grid, block, 0, st>>>(shells, pa, pb, npairs, nbf, natoms, axyz, Z, tab, D, W, gT, gV, gS);
    PF_HIP(hipGetLastError());
}

// Work plan of a derivative class, one workgroup of 1024 threads: bra pairs
// in blocks of ib (pairs are sorted by Q descending, so a block's first pair
// has the longest Schwarz ket prefix), and an item is (bra block, tile of 64
// kets) for every tile of the first pair's prefix.  work[0] is the queue head
// (cleared here), work[1 + b] the first item of block b, work[1 + nblk] the
// number of items.
__global__ void __launch_bounds__(1024) grad_plan_kernel(GradArgs g, int ib, int nblk, int *work) {
    __shared__ int part[1024];
    const PairView<double> &B = g.bra;
    const PairView<double> &K = g.ket;
    const int tid = threadIdx.x;
    const int per = (nblk + 1023) / 1024;
    const int b0 = min(nblk, tid * per), b1 = min(nblk, b0 + per);
    const double dm = g.dmax ? *g.dmax : 1.0;
    int *off = work + 1;
    int sum = 0;
    for (int b = b0; b < b1; ++b) {
        const int i0 = b * ib;
        const double qthr = g.tau / (B.Q[i0] * dm * dm);
        int lo = 0, hi = K.n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (K.Q[mid] >= qthr) lo = mid + 1;
            else hi = mid;
        }
        // same-class lists stop at the block's last bra pair
        const int nket = g.same ? min(lo, min(B.n, i0 + ib)) : lo;
        off[b] = (nket + 63) / 64;
        sum += off[b];
    }
    part[tid] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - sum;  // exclusive prefix of this thread's blocks
    for (int b = b0; b < b1; ++b) {
        const int c = off[b];
        off[b] = run;
        run += c;
    }
    if (tid == 1023) {
        off[nblk] = part[1023];
        work[0] = 0;
    }
}

void launch_grad_plan(const GradArgs &g, int ib, int nblk, int *work, hipStream_t st) {
    grad_plan_kernel<<<1, 1024, 0, st>>>(g, ib, nblk, work);
    PF_HIP(hipGetLastError());
}

#define PF_GRAD_EXTERN(a, b, c, d) extern template void launch_grad2e<a, b, c, d>(int, const GradArgs &, int *, int, hipStream_t);
PF_GRAD_EXTERN(0, 0, 0, 0)
PF_GRAD_EXTERN(1, 0, 0, 0)
PF_GRAD_EXTERN(0, 0, 1, 0)
PF_GRAD_EXTERN(1, 0, 1, 0)
PF_GRAD_EXTERN(1, 1, 0, 0)
PF_GRAD_EXTERN(0, 0, 1, 1)
PF_GRAD_EXTERN(1, 1, 1, 0)
PF_GRAD_EXTERN(1, 0, 1, 1)
PF_GRAD_EXTERN(1, 1, 1, 1)
#undef PF_GRAD_EXTERN

void launch_grad2e_class(int la, int lb, int lc, int ld, int nk, const GradArgs &g, int *work, int cap, hipStream_t st) {
    const int key = ((la * 2 + lb) * 2 + lc) * 2 + ld;
    switch (key) {
        case 0: launch_grad2e<0, 0, 0, 0>(nk, g, work, cap, st); break;
        case 8: launch_grad2e<1, 0, 0, 0>(nk, g, work, cap, st); break;
        case 2: launch_grad2e<0, 0, 1, 0>(nk, g, work, cap, st); break;
        case 10: launch_grad2e<1, 0, 1, 0>(nk, g, work, cap, st); break;
        case 12: launch_grad2e<1, 1, 0, 0>(nk, g, work, cap, st); break;
        case 3: launch_grad2e<0, 0, 1, 1>(nk, g, work, cap, st); break;
        case 14: launch_grad2e<1, 1, 1, 0>(nk, g, work, cap, st); break;
        case 11: launch_grad2e<1, 0, 1, 1>(nk, g, work, cap, st); break;
        case 15: launch_grad2e<1, 1, 1, 1>(nk, g, work, cap, st); break;
        default: throw Error("gradient: unsupported quartet class");
    }
    PF_HIP(hipGetLastError());
}

}  // namespace pf
