// Integrals over shell pairs and quartets that contain a d shell:
// McMurchie-Davidson with runtime angular momenta.
//
// The reference's Cartesian machinery is general in l (cartesian.f90:7-23,
// overlap.f90:195 sizes its recursion for l_max = 3) and its README names
// larger basis sets (README.md:60); its parser skips D lines
// (basis.f90:76-82).  Here a d shell gets its 6 Cartesian components in the
// reference's order (xx, xy, xz, yy, yz, zz) and each component the
// primitive norm of gaussian.f90:80-98, i.e. 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!)
// relative to the (l,0,0) norm folded into the shell coefficients.
//
// McMurchie & Davidson (J. Comput. Phys. 26, 218, 1978), as in Helgaker,
// Jorgensen & Olsen ch. 9:
//   Hermite expansion, per Cartesian direction (E^{00}_0 = 1, K_AB apart):
//     E^{i+1,j}_t = E^{ij}_{t-1} / 2p + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}
//     E^{i,j+1}_t = E^{ij}_{t-1} / 2p + X_PB E^{ij}_t + (t+1) E^{ij}_{t+1}
//   overlap S_ij = E^{ij}_0 sqrt(pi/p); kinetic per direction
//     T_ij = b (2j+1) S_ij - 2 b^2 S_{i,j+2} - j (j-1) / 2 S_{i,j-2}
//   Hermite Coulomb integrals R^n_{000} = (-2 alpha)^n F_n(alpha |R|^2),
//     R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{t,u,v}  (likewise u, v)
//   (a|1/r_C|b) = 2 pi / p K_AB sum_tuv E_t E_u E_v R_tuv(p, P - C)
//   (ab|cd) = 2 pi^{5/2} / (p q sqrt(p+q)) K_AB K_CD
//             sum_tuv E^{ab}_tuv sum_{tau nu phi} (-1)^{tau+nu+phi} E^{cd}_{tau nu phi} R_{t+tau,u+nu,v+phi}(alpha, P - Q)
// Work is spread one thread per shell pair (one-electron) or per (shell
// quartet, component of the ket pair): the few quartets that touch a d shell
// of a third-row atom, whose arrays live in per-lane scratch.
#include <cmath>

#include "boys.h"
#include "cart.h"
#include "generic.h"

namespace pf {

namespace {

constexpr int GL2 = 2 * GEN_LMAX;      // pair (4)
constexpr int GL4 = 4 * GEN_LMAX;      // quartet (8)
constexpr int NH2 = coff(GL2 + 1);     // Hermite (t,u,v), t+u+v <= 4: 35
constexpr int NH4 = coff(GL4 + 1);     // t+u+v <= 8: 165
constexpr int NCMAX = nc(GEN_LMAX);    // Cartesian components of one shell (6)
constexpr int NAB = NCMAX * NCMAX;     // of a pair (36)
constexpr int EJ = GEN_LMAX + 3;       // j <= lb + 2 (kinetic)
constexpr int ET = 2 * GEN_LMAX + 4;   // t <= la + lb + 2, plus the zero above
static_assert(BOYS_MMAX >= GL4 + BOYS_TAYLOR - 1, "Boys table too short for (dd|dd)");

__device__ double dfact_dev(int n) {
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}
// norm of Cartesian component (lx, ly, lz) relative to (l, 0, 0)
__device__ double comp_scale(int g) {
    const int lx = cx(g), ly = cy(g), lz = cz(g), l = lx + ly + lz;
    return sqrt(dfact_dev(2 * l - 1) / (dfact_dev(2 * lx - 1) * dfact_dev(2 * ly - 1) * dfact_dev(2 * lz - 1)));
}

// F_0 .. F_M at x, M <= BOYS_MMAX - BOYS_TAYLOR + 1: Taylor expansion of F_M
// from the full table, exp, downward recursion (as boys_eval, runtime M).
__device__ void boys_rt(int M, double x, const double *__restrict__ tab, double *F) {
    if (x < BOYS_XMAX) {
        const int k = int(x * (1.0 / BOYS_H) + 0.5);
        const double d = k * BOYS_H - x;
        const double *t = tab + k * (BOYS_MMAX + 1) + M;
        double acc = t[BOYS_TAYLOR - 1];
        for (int j = BOYS_TAYLOR - 1; j >= 1; --j) acc = t[j - 1] + acc * d * (1.0 / j);
        F[M] = acc;
        const double e = exp(-x);
        for (int m = M - 1; m >= 0; --m) F[m] = (2.0 * x * F[m + 1] + e) * (1.0 / (2 * m + 1));
    } else {
        const double e = exp(-x), o2x = 0.5 / x;
        F[0] = 0.5 * sqrt(M_PI / x);
        for (int m = 0; m < M; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * o2x;
    }
}

// 1-D Hermite coefficients E[i][j][t], i <= la, j <= lb (t beyond i + j is 0).
template <int JD, int TD>
__device__ void hermite_e(int la, int lb, double o2p, double PA, double PB, double (*E)[JD][TD]) {
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j)
            for (int t = 0; t < TD; ++t) E[i][j][t] = 0.0;
    E[0][0][0] = 1.0;
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j) {
            if (i == 0 && j == 0) continue;
            const double *src = j == 0 ? E[i - 1][0] : E[i][j - 1];
            const double X = j == 0 ? PA : PB;
            for (int t = 0; t <= i + j; ++t) {
                double v = X * src[t];
                if (t > 0) v += o2p * src[t - 1];
                if (t + 1 <= i + j - 1) v += (t + 1) * src[t + 1];
                E[i][j][t] = v;
            }
        }
}

// R[cidx(t,u,v)] = R^0_tuv(alpha, X, Y, Z) for t+u+v <= L, from F[0..L];
// the levels n = L .. 0 alternate between R and W (both NH4 long).
__device__ void hermite_r(int L, double alpha, double X, double Y, double Z, const double *F, double *R, double *W) {
    double *cur = (L % 2 == 0) ? R : W, *nxt = (L % 2 == 0) ? W : R;
    double fac = 1.0;
    for (int n = 0; n < L; ++n) fac *= -2.0 * alpha;  // (-2 alpha)^L
    cur[0] = fac * F[L];
    for (int n = L - 1; n >= 0; --n) {
        fac /= -2.0 * alpha;
        nxt[0] = fac * F[n];
        // (t, u, v) enumerated directly: the table-free decoders cx/cy/cz of
        // cart.h cover total angular momenta up to 5 only, and here N <= 8
        for (int N = 1; N <= L - n; ++N)
            for (int t = N; t >= 0; --t)
                for (int u = N - t; u >= 0; --u) {
                    const int v = N - t - u;
                    double r;
                    if (t > 0) {
                        r = X * cur[cidx(t - 1, u, v)];
                        if (t > 1) r += (t - 1) * cur[cidx(t - 2, u, v)];
                    } else if (u > 0) {
                        r = Y * cur[cidx(t, u - 1, v)];
                        if (u > 1) r += (u - 1) * cur[cidx(t, u - 2, v)];
                    } else {
                        r = Z * cur[cidx(t, u, v - 1)];
                        if (v > 1) r += (v - 1) * cur[cidx(t, u, v - 2)];
                    }
                    nxt[cidx(t, u, v)] = r;
                }
        double *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
}

// One-electron blocks of an ordered shell pair: s[ia*nb + ib], t, v.
__device__ void onee_generic_pair(const Shell &A, const Shell &B, int natoms, const double *__restrict__ axyz,
                                  const int *__restrict__ Z, const double *__restrict__ tab, double *s, double *t,
                                  double *v) {
    const int la = A.l, lb = B.l, na = nc(la), nb = nc(lb);
    for (int k = 0; k < na * nb; ++k) s[k] = t[k] = v[k] = 0.0;
    double AB[3], r2 = 0.0;
    for (int d = 0; d < 3; ++d) { AB[d] = A.xyz[d] - B.xyz[d]; r2 += AB[d] * AB[d]; }
    double E[3][GEN_LMAX + 1][EJ][ET];
    double F[GL2 + 1], R[NH4], W[NH4];
    for (int i = 0; i < A.np; ++i)
        for (int j = 0; j < B.np; ++j) {
            const double a = A.exp[i], b = B.exp[j], p = a + b, ip = 1.0 / p;
            const double cab = A.coef[i] * B.coef[j] * exp(-a * b * ip * r2);
            double P[3];
            for (int d = 0; d < 3; ++d) {
                P[d] = (a * A.xyz[d] + b * B.xyz[d]) * ip;
                hermite_e<EJ, ET>(la, lb + 2, 0.5 * ip, P[d] - A.xyz[d], P[d] - B.xyz[d], E[d]);
            }
            const double sp = cab * M_PI * ip * sqrt(M_PI * ip);  // K_AB (pi/p)^{3/2} c_a c_b
            for (int ia = 0; ia < na; ++ia) {
                const int ga = coff(la) + ia, ax[3] = {cx(ga), cy(ga), cz(ga)};
                for (int ib = 0; ib < nb; ++ib) {
                    const int gb = coff(lb) + ib, bx[3] = {cx(gb), cy(gb), cz(gb)};
                    double S1[3], T1[3];
                    for (int d = 0; d < 3; ++d) {
                        const int ii = ax[d], jj = bx[d];
                        S1[d] = E[d][ii][jj][0];
                        T1[d] = b * (2 * jj + 1) * E[d][ii][jj][0] - 2.0 * b * b * E[d][ii][jj + 2][0] -
                                (jj > 1 ? 0.5 * jj * (jj - 1) * E[d][ii][jj - 2][0] : 0.0);
                    }
                    s[ia * nb + ib] += sp * S1[0] * S1[1] * S1[2];
                    t[ia * nb + ib] += sp * (T1[0] * S1[1] * S1[2] + S1[0] * T1[1] * S1[2] + S1[0] * S1[1] * T1[2]);
                }
            }
            // nuclear attraction, nucleus by nucleus
            const double vp = -2.0 * M_PI * ip * cab;
            for (int c = 0; c < natoms; ++c) {
                const double PC[3] = {P[0] - axyz[3 * c], P[1] - axyz[3 * c + 1], P[2] - axyz[3 * c + 2]};
                boys_rt(la + lb, p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), tab, F);
                hermite_r(la + lb, p, PC[0], PC[1], PC[2], F, R, W);
                const double zc = vp * Z[c];
                for (int ia = 0; ia < na; ++ia) {
                    const int ga = coff(la) + ia, ax = cx(ga), ay = cy(ga), az = cz(ga);
                    for (int ib = 0; ib < nb; ++ib) {
                        const int gb = coff(lb) + ib, bx = cx(gb), by = cy(gb), bz = cz(gb);
                        double sum = 0.0;
                        for (int tt = 0; tt <= ax + bx; ++tt)
                            for (int uu = 0; uu <= ay + by; ++uu)
                                for (int vv = 0; vv <= az + bz; ++vv)
                                    sum += E[0][ax][bx][tt] * E[1][ay][by][uu] * E[2][az][bz][vv] *
                                           R[cidx(tt, uu, vv)];
                        v[ia * nb + ib] += zc * sum;
                    }
                }
            }
        }
    for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib) {
            const double sc = comp_scale(coff(la) + ia) * comp_scale(coff(lb) + ib);
            s[ia * nb + ib] *= sc;
            t[ia * nb + ib] *= sc;
            v[ia * nb + ib] *= sc;
        }
}

// (ab|cd) for every component pair of (A, B) at ket component kc = ic*nd + id
// of (C, D): out[ia*nb + ib], including the component norms.
__device__ void eri_generic(const Shell &A, const Shell &B, const Shell &C, const Shell &D, int kc,
                            const double *__restrict__ tab, double *out) {
    const int la = A.l, lb = B.l, lc = C.l, ld = D.l;
    const int na = nc(la), nb = nc(lb), nd = nc(ld);
    const int lab = la + lb, L = lab + lc + ld;
    const int ic = kc / nd, id = kc - ic * nd;
    const int gc = coff(lc) + ic, gd = coff(ld) + id;
    const int cxyz[3] = {cx(gc), cy(gc), cz(gc)}, dxyz[3] = {cx(gd), cy(gd), cz(gd)};
    for (int k = 0; k < na * nb; ++k) out[k] = 0.0;
    double AB[3], CD[3], rab = 0.0, rcd = 0.0;
    for (int d = 0; d < 3; ++d) {
        AB[d] = A.xyz[d] - B.xyz[d]; rab += AB[d] * AB[d];
        CD[d] = C.xyz[d] - D.xyz[d]; rcd += CD[d] * CD[d];
    }
    constexpr int QJ = GEN_LMAX + 1, QT = 2 * GEN_LMAX + 1;  // j <= l, t <= i + j
    double Eb[3][GEN_LMAX + 1][QJ][QT], Ek[3][GEN_LMAX + 1][QJ][QT];
    double F[GL4 + 1], R[NH4], W[NH4], G[NH2];
    for (int i = 0; i < A.np; ++i)
        for (int j = 0; j < B.np; ++j) {
            const double a = A.exp[i], b = B.exp[j], p = a + b, ip = 1.0 / p;
            const double cab = A.coef[i] * B.coef[j] * exp(-a * b * ip * rab);
            double P[3];
            for (int d = 0; d < 3; ++d) {
                P[d] = (a * A.xyz[d] + b * B.xyz[d]) * ip;
                hermite_e<QJ, QT>(la, lb, 0.5 * ip, P[d] - A.xyz[d], P[d] - B.xyz[d], Eb[d]);
            }
            for (int k = 0; k < C.np; ++k)
                for (int l = 0; l < D.np; ++l) {
                    const double c = C.exp[k], dd = D.exp[l], q = c + dd, iq = 1.0 / q;
                    const double ccd = C.coef[k] * D.coef[l] * exp(-c * dd * iq * rcd);
                    double Q[3];
                    for (int d = 0; d < 3; ++d) {
                        Q[d] = (c * C.xyz[d] + dd * D.xyz[d]) * iq;
                        hermite_e<QJ, QT>(cxyz[d], dxyz[d], 0.5 * iq, Q[d] - C.xyz[d], Q[d] - D.xyz[d], Ek[d]);
                    }
                    const double alpha = p * q / (p + q);
                    const double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
                    boys_rt(L, alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]), tab, F);
                    hermite_r(L, alpha, PQ[0], PQ[1], PQ[2], F, R, W);
                    const double pref = 2.0 * M_PI * M_PI * sqrt(M_PI) / (p * q * sqrt(p + q)) * cab * ccd;
                    // ket contraction at this component: G_tuv, t+u+v <= la+lb
                    const int tc = cxyz[0] + dxyz[0], uc = cxyz[1] + dxyz[1], vc = cxyz[2] + dxyz[2];
                    for (int g = 0; g < coff(lab + 1); ++g) {
                        const int t = cx(g), u = cy(g), v = cz(g);
                        double sum = 0.0;
                        for (int tau = 0; tau <= tc; ++tau)
                            for (int nu = 0; nu <= uc; ++nu)
                                for (int phi = 0; phi <= vc; ++phi) {
                                    const double e = Ek[0][cxyz[0]][dxyz[0]][tau] * Ek[1][cxyz[1]][dxyz[1]][nu] *
                                                     Ek[2][cxyz[2]][dxyz[2]][phi];
                                    const double r = R[cidx(t + tau, u + nu, v + phi)];
                                    sum += ((tau + nu + phi) & 1) ? -e * r : e * r;
                                }
                        G[g] = sum;
                    }
                    for (int ia = 0; ia < na; ++ia) {
                        const int ga = coff(la) + ia, ax = cx(ga), ay = cy(ga), az = cz(ga);
                        for (int ib = 0; ib < nb; ++ib) {
                            const int gb = coff(lb) + ib, bx = cx(gb), by = cy(gb), bz = cz(gb);
                            double sum = 0.0;
                            for (int t = 0; t <= ax + bx; ++t)
                                for (int u = 0; u <= ay + by; ++u)
                                    for (int v = 0; v <= az + bz; ++v)
                                        sum += Eb[0][ax][bx][t] * Eb[1][ay][by][u] * Eb[2][az][bz][v] * G[cidx(t, u, v)];
                            out[ia * nb + ib] += pref * sum;
                        }
                    }
                }
        }
    const double sk = comp_scale(gc) * comp_scale(gd);
    for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib) out[ia * nb + ib] *= sk * comp_scale(coff(la) + ia) * comp_scale(coff(lb) + ib);
}

__global__ void __launch_bounds__(64) onee_generic_kernel(const Shell *__restrict__ shells, const int *pa,
                                                          const int *pb, int npair, int nbf, int natoms,
                                                          const double *__restrict__ axyz, const int *__restrict__ Z,
                                                          const double *__restrict__ tab, double *Smat, double *Tmat,
                                                          double *Vmat) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npair) return;
    const Shell A = shells[pa[k]], B = shells[pb[k]];
    double s[NAB], t[NAB], v[NAB];
    onee_generic_pair(A, B, natoms, axyz, Z, tab, s, t, v);
    const int na = nc(A.l), nb = nc(B.l);
    for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib) {
            const size_t mu = A.bf + ia, nu = B.bf + ib, e = ia * nb + ib;
            Smat[mu * nbf + nu] = s[e]; Smat[nu * nbf + mu] = s[e];
            Tmat[mu * nbf + nu] = t[e]; Tmat[nu * nbf + mu] = t[e];
            Vmat[mu * nbf + nu] = v[e]; Vmat[nu * nbf + mu] = v[e];
        }
}

// diag[k * NAB + kc] = (ab|ab) at component kc of pair k (a's component kc/nb, b's kc%nb)
__global__ void __launch_bounds__(64) schwarz_generic_kernel(const Shell *__restrict__ shells, const int *pa,
                                                             const int *pb, int npair, const double *__restrict__ tab,
                                                             double *diag) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = g / NAB, kc = g - k * NAB;
    if (k >= npair) return;
    const Shell A = shells[pa[k]], B = shells[pb[k]];
    if (kc >= nc(A.l) * nc(B.l)) { diag[g] = 0.0; return; }
    double out[NAB];
    eri_generic(A, B, A, B, kc, tab, out);
    diag[g] = fabs(out[kc]);
}
__global__ void schwarz_generic_max_kernel(int npair, const double *diag, double *Q) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npair) return;
    double m = 0.0;
    for (int c = 0; c < NAB; ++c) m = fmax(m, diag[k * NAB + c]);
    Q[k] = sqrt(m);
}

__device__ __forceinline__ long long gen_pair_index(long long i, long long j) {
    return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

// One thread per (quartet, ket component); mode 0: J/K digest, 1: packed export.
template <int MODE>
__global__ void __launch_bounds__(64) quartet_generic_kernel(GenQuartets q, int nbf, int nsh, const double *Dj,
                                                             const double *Dk0, const double *Dk1, const double *Dsh,
                                                             double tau, double *J, double *K0, double *K1,
                                                             double *packed) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long qi = g / NAB;
    const int kc = (int)(g - qi * NAB);
    if (qi >= (long long)q.nbra * q.nket) return;
    const int k = (int)(qi / q.nket), j = (int)(qi - (long long)k * q.nket);
    if (j < q.nsame && j > k) return;  // (k|j) with j > k is (j|k): once
    const int a = q.bsa[k], b = q.bsb[k], c = q.ksa[j], d = q.ksb[j];
    const Shell A = q.shells[a], B = q.shells[b], C = q.shells[c], D = q.shells[d];
    const int nb = nc(B.l), ncd = nc(C.l) * nc(D.l), nab = nc(A.l) * nb;
    if (kc >= ncd) return;
    if (MODE == 0) {
        double dq = 2.0 * fmax(Dsh[a * nsh + b], Dsh[c * nsh + d]);
        dq = fmax(dq, fmax(fmax(Dsh[b * nsh + d], Dsh[b * nsh + c]), fmax(Dsh[a * nsh + d], Dsh[a * nsh + c])));
        if (q.bQ[k] * q.kQ[j] * dq < tau) return;
    }
    double v[NAB];
    eri_generic(A, B, C, D, kc, q.tab, v);
    const int nd = nc(D.l), ic = kc / nd, id = kc - ic * nd;
    const int fc = C.bf + ic, fd = D.bf + id;
    if (MODE == 1) {
        for (int ab = 0; ab < nab; ++ab) {
            const long long ij = gen_pair_index(A.bf + ab / nb, B.bf + ab % nb), kl = gen_pair_index(fc, fd);
            packed[ij >= kl ? ij * (ij + 1) / 2 + kl : kl * (kl + 1) / 2 + ij] = v[ab];
        }
        return;
    }
    // degeneracy of the unique-quartet enumeration
    const double deg = (a == b ? 0.5 : 1.0) * (c == d ? 0.5 : 1.0) * ((j < q.nsame && j == k) ? 0.5 : 1.0);
    const size_t n = nbf;
    const double dcd = Dj[fc * n + fd];
    double jcd = 0.0;
    for (int ab = 0; ab < nab; ++ab) {
        const int fa = A.bf + ab / nb, fb = B.bf + ab % nb;
        const double x = deg * v[ab];
        if (x == 0.0) continue;
        atomicAdd(&J[fa * n + fb], 2.0 * x * dcd);  // J'_ab
        jcd += x * Dj[fa * n + fb];
        for (int kk = 0; kk < (Dk1 ? 2 : 1); ++kk) {
            const double *Dk = kk ? Dk1 : Dk0;
            double *K = kk ? K1 : K0;
            atomicAdd(&K[fa * n + fc], x * Dk[fb * n + fd]);  // K'_ac
            atomicAdd(&K[fa * n + fd], x * Dk[fb * n + fc]);  // K'_ad
            atomicAdd(&K[fb * n + fc], x * Dk[fa * n + fd]);  // K'_bc
            atomicAdd(&K[fb * n + fd], x * Dk[fa * n + fc]);  // K'_bd
        }
    }
    if (jcd != 0.0) atomicAdd(&J[fc * n + fd], 2.0 * jcd);  // J'_cd
}

// Shell quartets of a list in their requested orientation (s0 s1 | s2 s3):
// block q at out + off[q], element ab * ncd + kc (first shell slowest).
__global__ void __launch_bounds__(64) export_generic_kernel(int n, const int *__restrict__ list,
                                                            const int *__restrict__ sel, const Shell *__restrict__ sh,
                                                            const double *__restrict__ tab,
                                                            const long long *__restrict__ off, double *out) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = (int)(g / NAB), kc = (int)(g - (long long)t * NAB);
    if (t >= n) return;
    const int q = sel[t];
    const Shell A = sh[list[4 * q]], B = sh[list[4 * q + 1]], C = sh[list[4 * q + 2]], D = sh[list[4 * q + 3]];
    const int ncd = nc(C.l) * nc(D.l), nab = nc(A.l) * nc(B.l);
    if (kc >= ncd) return;
    double v[NAB];
    eri_generic(A, B, C, D, kc, tab, v);
    for (int ab = 0; ab < nab; ++ab) out[off[q] + (long long)ab * ncd + kc] = v[ab];
}

}  // namespace

void launch_export_generic(int n, const int *list, const int *sel, const Shell *sh, const double *tab,
                           const long long *off, double *out, hipStream_t st) {
    if (n <= 0) return;
    export_generic_kernel<<<(unsigned)(((long long)n * NAB + 63) / 64), 64, 0, st>>>(n, list, sel, sh, tab, off, out);
    PF_HIP(hipGetLastError());
}

void launch_onee_generic(const Shell *d_shells, const int *pa, const int *pb, int npair, int nbf, int natoms,
                         const double *d_axyz, const int *d_Z, const double *d_tab, double *S, double *T, double *V,
                         hipStream_t st) {
    if (npair <= 0) return;
    onee_generic_kernel<<<(npair + 63) / 64, 64, 0, st>>>(d_shells, pa, pb, npair, nbf, natoms, d_axyz, d_Z, d_tab, S,
                                                          T, V);
    PF_HIP(hipGetLastError());
}

void launch_schwarz_generic(const Shell *d_shells, const int *pa, const int *pb, int npair, const double *d_tab,
                            double *Q, hipStream_t st) {
    if (npair <= 0) return;
    DevBuf<double> diag((size_t)npair * NAB);
    schwarz_generic_kernel<<<(npair * NAB + 63) / 64, 64, 0, st>>>(d_shells, pa, pb, npair, d_tab, diag.p);
    schwarz_generic_max_kernel<<<(npair + 63) / 64, 64, 0, st>>>(npair, diag.p, Q);
    PF_HIP(hipGetLastError());
    PF_HIP(hipStreamSynchronize(st));  // diag goes back to the block cache
}

void launch_jk_generic(const GenQuartets &q, int nbf, int nsh, const double *Dj, const double *Dk0,
                       const double *Dk1, const double *Dsh, double tau, double *J, double *K0, double *K1,
                       hipStream_t st) {
    const long long nthreads = (long long)q.nbra * q.nket * NAB;
    if (nthreads == 0) return;
    if (nthreads / 64 >= (1LL << 31)) throw Error("generic quartets: too many d-shell pairs for one launch");
    quartet_generic_kernel<0><<<(unsigned)((nthreads + 63) / 64), 64, 0, st>>>(q, nbf, nsh, Dj, Dk0, Dk1, Dsh, tau, J,
                                                                               K0, K1, nullptr);
    PF_HIP(hipGetLastError());
}

void launch_packed_generic(const GenQuartets &q, double *packed, hipStream_t st) {
    const long long nthreads = (long long)q.nbra * q.nket * NAB;
    if (nthreads == 0) return;
    if (nthreads / 64 >= (1LL << 31)) throw Error("generic quartets: too many d-shell pairs for one launch");
    quartet_generic_kernel<1><<<(unsigned)((nthreads + 63) / 64), 64, 0, st>>>(q, 0, 0, nullptr, nullptr, nullptr,
                                                                               nullptr, 0.0, nullptr, nullptr,
                                                                               nullptr, packed);
    PF_HIP(hipGetLastError());
}

}  // namespace pf
