// One-electron integrals S, T, V over contracted s/p shells (Obara-Saika).
//
// Replaces the reference's three OpenACC shell-pair loops S_overlap
// (overlap.f90:240-340), T_kinetic (kinetic.f90:117-214) and V_nuclear
// (nuclear.f90:202-328) with one fused kernel: per shell pair (a >= b) all
// primitive pairs in registers, the nuclei split over 16 lanes (one thread
// per pair in the batched form, whose many small molecules fill the chip), the
// Boys table staged once per workgroup in LDS.  No per-atom V scratch (the reference
// holds natoms x nbf^2 doubles) and no host-side summation.
//
//   overlap 1-D OS:  S_{i+1,j} = PA S_ij + (i S_{i-1,j} + j S_{i,j-1}) / 2p
//                    S_{i,j+1} = PB S_ij + (i S_{i-1,j} + j S_{i,j-1}) / 2p
//   kinetic:         T_x(i,j) = b(2j+1) S_ij - 2 b^2 S_{i,j+2} - j(j-1)/2 S_{i,j-2}
//   nuclear OS VRR:  [e+1_i]^m = PA_i [e]^m - PC_i [e]^{m+1}
//                                + e_i/2p ([e-1_i]^m - [e-1_i]^{m+1})
//                    [0]^m = 2 pi / p K_AB F_m(p |PC|^2),  HRR (a|b+1_i) = (a+1_i|b) + AB_i (a|b)
#include "boys.h"
#include "cart.h"
#include "pf_common.h"
#include "kernels.h"

namespace pf {

constexpr int ONEE_NORD = 2 + BOYS_TAYLOR;

template <int LA, int LB>
__device__ void onee_block(const Shell &A, const Shell &B, int natoms, const double *__restrict__ axyz,
                           const int *__restrict__ Z, const double *tab, double *S, double *T, double *V, int c0,
                           int cs) {
    constexpr int NA = nc(LA), NB = nc(LB);
    constexpr int LT = LA + LB;
    double s_acc[NA * NB], t_acc[NA * NB], v_acc[NA * NB];
#pragma unroll
    for (int k = 0; k < NA * NB; ++k) s_acc[k] = t_acc[k] = v_acc[k] = 0.0;
    double AB[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) AB[d] = A.xyz[d] - B.xyz[d];
    const double r2ab = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];

    for (int i = 0; i < A.np; ++i) {
        for (int j = 0; j < B.np; ++j) {
            const double a = A.exp[i], b = B.exp[j], p = a + b, ip = 1.0 / p;
            const double kab = exp(-a * b * ip * r2ab);
            const double cc = A.coef[i] * B.coef[j];
            double P[3], PA[3], PB[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                P[d] = (a * A.xyz[d] + b * B.xyz[d]) * ip;
                PA[d] = P[d] - A.xyz[d];
                PB[d] = P[d] - B.xyz[d];
            }
            const double o2p = 0.5 * ip;
            // 1-D overlaps, i <= LA, j <= LB + 2 (unit S_00; prefactor applied below)
            double s1[3][LA + 1][LB + 3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                s1[d][0][0] = 1.0;
#pragma unroll
                for (int jj = 0; jj < LB + 2; ++jj)
                    s1[d][0][jj + 1] = PB[d] * s1[d][0][jj] + (jj > 0 ? jj * o2p * s1[d][0][jj - 1] : 0.0);
#pragma unroll
                for (int ii = 0; ii < LA; ++ii)
#pragma unroll
                    for (int jj = 0; jj <= LB + 2; ++jj)
                        s1[d][ii + 1][jj] = PA[d] * s1[d][ii][jj] + (ii > 0 ? ii * o2p * s1[d][ii - 1][jj] : 0.0) +
                                            (jj > 0 ? jj * o2p * s1[d][ii][jj - 1] : 0.0);
            }
            const double pre = cc * kab * ip * sqrt(M_PI * ip) * M_PI;  // (pi/p)^{3/2} K_AB c_a c_b
#pragma unroll
            for (int ia = 0; ia < NA; ++ia) {
                constexpr int ga0 = coff(LA);
                const int ga = ga0 + ia;
                const int ax = cx(ga), ay = cy(ga), az = cz(ga);
#pragma unroll
                for (int ib = 0; ib < NB; ++ib) {
                    const int gb = coff(LB) + ib;
                    const int bx = cx(gb), by = cy(gb), bz = cz(gb);
                    const int ab[3] = {ax, ay, az}, bb[3] = {bx, by, bz};
                    double sd[3], td[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const int ii = ab[d], jj = bb[d];
                        sd[d] = s1[d][ii][jj];
                        td[d] = b * (2 * jj + 1) * s1[d][ii][jj] - 2.0 * b * b * s1[d][ii][jj + 2] -
                                (jj >= 2 ? 0.5 * jj * (jj - 1) * s1[d][ii][jj >= 2 ? jj - 2 : 0] : 0.0);
                    }
                    s_acc[ia * NB + ib] += pre * sd[0] * sd[1] * sd[2];
                    t_acc[ia * NB + ib] += pre * (td[0] * sd[1] * sd[2] + sd[0] * td[1] * sd[2] + sd[0] * sd[1] * td[2]);
                }
            }
            // nuclear attraction
            const double vpre = cc * kab * 2.0 * M_PI * ip;
            for (int c = c0; c < natoms; c += cs) {  // this thread's nuclei
                double PC[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) PC[d] = P[d] - axyz[3 * c + d];
                const double x = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                double F[LT + 1];
                boys_eval<LT, double, ONEE_NORD>(x, tab, F);
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
                            const int g2 = cminus(g1 >= 0 ? g1 : 0, d);
                            v += em * o2p * (th[m][g2] - th[m + 1][g2]);
                        }
                        th[m][g] = v;
                    }
                }
                const double zc = -double(Z[c]);
#pragma unroll
                for (int ia = 0; ia < NA; ++ia) {
                    const int ga = coff(LA) + ia;
#pragma unroll
                    for (int ib = 0; ib < NB; ++ib) {
                        double v;
                        if (LB == 0) {
                            v = th[0][ga];
                        } else {
                            // (a|p_j) = (a+1_j|s) + AB_j (a|s)
                            v = th[0][cplus(ga, ib)] + AB[ib] * th[0][ga];
                        }
                        v_acc[ia * NB + ib] += zc * v;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NA * NB; ++k) {
        S[k] = s_acc[k];
        T[k] = t_acc[k];
        V[k] = v_acc[k];
    }
}

// Shell pair pid (a >= b) -> (a, b).
__device__ __forceinline__ void onee_pair_index(int pid, int &a, int &b) {
    a = int((sqrt(8.0 * pid + 1.0) - 1.0) * 0.5);
    while (a * (a + 1) / 2 > pid) --a;
    while ((a + 1) * (a + 2) / 2 <= pid) ++a;
    b = pid - a * (a + 1) / 2;
}

// S, T and the nuclei c0, c0 + cs, ... of V for shell pair (A, B) (s[k], t[k],
// v[k], k = ia * nb + ib).
__device__ __forceinline__ void onee_pair_blocks(const Shell &A, const Shell &B, int natoms,
                                                 const double *__restrict__ axyz, const int *__restrict__ Z,
                                                 const double *lds_tab, double *s, double *t, double *v, int c0,
                                                 int cs) {
    const int la = A.l, lb = B.l;
    if (la == 0 && lb == 0) onee_block<0, 0>(A, B, natoms, axyz, Z, lds_tab, s, t, v, c0, cs);
    else if (la == 1 && lb == 0) onee_block<1, 0>(A, B, natoms, axyz, Z, lds_tab, s, t, v, c0, cs);
    else if (la == 0 && lb == 1) onee_block<0, 1>(A, B, natoms, axyz, Z, lds_tab, s, t, v, c0, cs);
    else onee_block<1, 1>(A, B, natoms, axyz, Z, lds_tab, s, t, v, c0, cs);
}

__device__ __forceinline__ void onee_store(const Shell &A, const Shell &B, int nbf, const double *s, const double *t,
                                           const double *v, double *Smat, double *Tmat, double *Vmat) {
    const int na = nc(A.l), nb = nc(B.l);
    for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib) {
            const size_t mu = A.bf + ia, nu = B.bf + ib, k = ia * nb + ib;
            Smat[mu * nbf + nu] = s[k]; Smat[nu * nbf + mu] = s[k];
            Tmat[mu * nbf + nu] = t[k]; Tmat[nu * nbf + mu] = t[k];
            Vmat[mu * nbf + nu] = v[k]; Vmat[nu * nbf + mu] = v[k];
        }
}

// One shell pair (a >= b, pair index pid) of one molecule by one thread: S, T,
// V blocks written to both triangles.
__device__ __forceinline__ void onee_pair(const Shell *__restrict__ shells, int pid, int nbf, int natoms,
                                          const double *__restrict__ axyz, const int *__restrict__ Z,
                                          const double *lds_tab, double *Smat, double *Tmat, double *Vmat) {
    int a, b;
    onee_pair_index(pid, a, b);
    double s[9], t[9], v[9];
    onee_pair_blocks(shells[a], shells[b], natoms, axyz, Z, lds_tab, s, t, v, 0, 1);
    onee_store(shells[a], shells[b], nbf, s, t, v, Smat, Tmat, Vmat);
}

__device__ __forceinline__ void onee_stage_table(const double *__restrict__ gtab, double *lds_tab) {
    // orders 0 .. 2 + 6 (enough for LA + LB <= 2) staged in LDS, stride ONEE_NORD
    for (int k = threadIdx.x; k < BOYS_NGRID * ONEE_NORD; k += blockDim.x) {
        const int g = k / ONEE_NORD, m = k - g * ONEE_NORD;
        lds_tab[k] = gtab[g * (BOYS_MMAX + 1) + m];
    }
    __syncthreads();
}

// One molecule: ONEE_LPP lanes per shell pair split the nuclear sum (V from a
// sub-wavefront shuffle reduction), so that a large molecule's tens of
// thousands of pairs fill the chip; S and T are computed alike by the lanes.
constexpr int ONEE_LPP = 16;
__global__ void __launch_bounds__(256) onee_kernel(const Shell *__restrict__ shells, int nsh, int nbf, int natoms,
                                                   const double *__restrict__ axyz, const int *__restrict__ Z,
                                                   const double *__restrict__ gtab, double *Smat, double *Tmat,
                                                   double *Vmat) {
    extern __shared__ double lds_tab[];
    onee_stage_table(gtab, lds_tab);
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int pid = g / ONEE_LPP, sub = g % ONEE_LPP;
    const bool valid = pid < nsh * (nsh + 1) / 2;
    double s[9], t[9], v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) s[k] = t[k] = v[k] = 0.0;
    int a = 0, b = 0;
    if (valid) onee_pair_index(pid, a, b);
    const bool sp = valid && shells[a].l <= 1 && shells[b].l <= 1;  // pairs with a d shell: generic kernel
    if (sp) onee_pair_blocks(shells[a], shells[b], natoms, axyz, Z, lds_tab, s, t, v, sub, ONEE_LPP);
    // every lane takes part (groups of ONEE_LPP lanes lie within one wavefront)
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
        for (int o = ONEE_LPP / 2; o >= 1; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    if (sp && sub == 0) onee_store(shells[a], shells[b], nbf, s, t, v, Smat, Tmat, Vmat);
}

// Many molecules in one launch: blockIdx.y = molecule m, whose shells start at
// shoff[m], atoms at aoff[m] and matrices at moff[m].
__global__ void __launch_bounds__(256) onee_batch_kernel(const Shell *__restrict__ shells, const int *shoff,
                                                         const int *aoff, const int *mnbf, const long long *moff,
                                                         const double *__restrict__ axyz, const int *__restrict__ Z,
                                                         const double *__restrict__ gtab, double *Smat, double *Tmat,
                                                         double *Vmat) {
    extern __shared__ double lds_tab[];
    const int m = blockIdx.y;
    const int nsh = shoff[m + 1] - shoff[m];
    if ((int)(blockIdx.x * blockDim.x) >= nsh * (nsh + 1) / 2) return;
    onee_stage_table(gtab, lds_tab);
    const int pid = blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= nsh * (nsh + 1) / 2) return;
    const size_t o = moff[m];
    onee_pair(shells + shoff[m], pid, mnbf[m], aoff[m + 1] - aoff[m], axyz + 3 * (size_t)aoff[m], Z + aoff[m], lds_tab,
              Smat + o, Tmat + o, Vmat + o);
}

void launch_onee_batch(int nmol, int max_nsh, const Shell *d_shells, const int *shoff, const int *aoff,
                       const int *mnbf, const long long *moff, const double *d_axyz, const int *d_Z,
                       const double *d_tab, double *S, double *T, double *V, hipStream_t st) {
    const int npair = max_nsh * (max_nsh + 1) / 2;
    const dim3 grid((npair + 255) / 256, nmol);
    const size_t lds = sizeof(double) * BOYS_NGRID * ONEE_NORD;
    onee_batch_kernel<<<grid, 256, lds, st>>>(d_shells, shoff, aoff, mnbf, moff, d_axyz, d_Z, d_tab, S, T, V);
    PF_HIP(hipGetLastError());
}

void launch_onee(const Shell *d_shells, int nsh, int nbf, int natoms, const double *d_axyz, const int *d_Z,
                 const double *d_tab, double *S, double *T, double *V, hipStream_t st) {
    const long long nthreads = (long long)nsh * (nsh + 1) / 2 * ONEE_LPP;
    const int blocks = (int)((nthreads + 255) / 256);
    const size_t lds = sizeof(double) * BOYS_NGRID * ONEE_NORD;
    onee_kernel<<<blocks, 256, lds, st>>>(d_shells, nsh, nbf, natoms, d_axyz, d_Z, d_tab, S, T, V);
    PF_HIP(hipGetLastError());
}

}  // namespace pf
