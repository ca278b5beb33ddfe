/*
 * CPU FP64 ORACLE for the Hartree-Fock SCF path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker: only tests, the smoke check and the
 * benchmark's CPU-baseline leg load it.  The product (libpyfacc.so) never
 * links or calls it.
 *
 * It restates, in plain C, the algorithm of the reference's SCF path:
 *   - primitive normalisation      gaussian.f90:80-98
 *   - Boys function F_n(x)         boys.f90:9-43 (exact here; the trapezoid
 *                                  rule of the reference is kept as a compat mode)
 *   - S, T, V one-electron ints    overlap.f90:150-174, kinetic.f90:20-54,
 *                                  nuclear.f90:66-136
 *   - nuclear repulsion            molecule.f90:35-49
 *   - ERIs (intended: Libint)      libint_f.f90:328-516 prefactors
 *   - packed canonical ERI order   rscf.f90:67-88
 *   - G = J - K/2 (RHF), UHF G     rscf.f90:121-134, uscf.f90:149-175
 *   - canonical orthogonalisation  la.f90:28-57
 *   - SCF iteration and stopping   rscf.f90:113-157, uscf.f90:140-203
 *
 * The integrals use the McMurchie-Davidson Hermite scheme (McMurchie &
 * Davidson, J. Comput. Phys. 26, 218 (1978)), deliberately a different
 * algorithm from the Obara-Saika / Head-Gordon-Pople recurrences of the GPU
 * kernels, for any Cartesian angular momentum.  Every function is pinned by
 * tests against closed forms, mpmath quadrature / differentiation of closed
 * forms, and literature energies.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_LMAX 4           /* max l per shell (s,p used; d/f for derivative tests) */
#define OR_KMAX 8           /* max primitives per shell */
#define OR_HMAX (4 * OR_LMAX + 4)

static const double OR_PI = 3.14159265358979323846;

/* ------------------------------------------------------------------ */
/* Compat switches: trapezoid Boys (boys.f90:21-41) and no contraction  */
/* renormalisation (gaussian.f90:80-98 only normalises primitives).      */
/* ------------------------------------------------------------------ */
static int g_compat_boys = 0;

void or_set_compat_boys(int on) { g_compat_boys = on; }
/* 1: or_scf takes the UHF branch (two spin densities, uscf.f90:149-203) even
 * when nalpha == nbeta, so that it can be held to the RHF branch. */
static int g_force_uhf = 0;
void or_set_force_uhf(int on) { g_force_uhf = on; }

/* ------------------------------------------------------------------ */
/* Boys function                                                        */
/* ------------------------------------------------------------------ */

/* Exact F_n(x), n = 0..nmax.  Series for F_nmax then downward recursion
 * F_n = (2x F_{n+1} + e^{-x}) / (2n+1) for x < 40; for x >= 40,
 * F_0 = sqrt(pi/x)/2 * erf(sqrt(x)) and the upward recursion
 * F_{n+1} = ((2n+1) F_n - e^{-x}) / (2x), which is stable there. */
void or_boys_exact(int nmax, double x, double *F)
{
    if (x < 40.0) {
        double ex = exp(-x);
        /* F_N(x) = e^{-x} sum_k (2x)^k / ((2N+1)(2N+3)...(2N+2k+1)) */
        double term = 1.0 / (2.0 * nmax + 1.0), sum = term;
        for (int k = 1; k < 400; ++k) {
            term *= 2.0 * x / (2.0 * nmax + 2.0 * k + 1.0);
            sum += term;
            if (term < 1e-18 * sum) break;
        }
        F[nmax] = ex * sum;
        for (int n = nmax - 1; n >= 0; --n) F[n] = (2.0 * x * F[n + 1] + ex) / (2.0 * n + 1.0);
    } else {
        double ex = exp(-x);
        F[0] = 0.5 * sqrt(OR_PI / x) * erf(sqrt(x));
        for (int n = 0; n < nmax; ++n) F[n + 1] = ((2.0 * n + 1.0) * F[n] - ex) / (2.0 * x);
    }
}

/* The reference's Boys rule (boys.f90:9-43): for x <= 1e-6 the two-term
 * Taylor form 1/(2n+1) - x/(2n+3); otherwise the trapezoid rule with
 * N = 1000 intervals on int_0^1 t^{2n} exp(-x t^2) dt. */
double or_boys_trapezoid(int n, double x)
{
    if (x <= 1.0e-6) return 1.0 / (2.0 * n + 1.0) - x / (2.0 * n + 3.0);
    const int N = 1000;
    double dt = 1.0 / N, sum = 0.0;
    for (int i = 0; i <= N; ++i) {
        double t = i * dt;
        double v = pow(t, 2 * n) * exp(-x * t * t);
        sum += (i == 0 || i == N) ? 0.5 * v : v;
    }
    return sum * dt;
}

static void boys(int nmax, double x, double *F)
{
    if (g_compat_boys) {
        for (int n = 0; n <= nmax; ++n) F[n] = or_boys_trapezoid(n, x);
    } else {
        or_boys_exact(nmax, x, F);
    }
}

/* ------------------------------------------------------------------ */
/* Basis description                                                    */
/* ------------------------------------------------------------------ */
typedef struct {
    int nsh;
    const int *l;         /* [nsh] angular momentum */
    const int *np;        /* [nsh] primitives */
    const double *ctr;    /* [3*nsh] centre (bohr) */
    const double *alpha;  /* [nsh*OR_KMAX] exponents */
    const double *coef;   /* [nsh*OR_KMAX] contraction coefficients INCLUDING
                             the primitive norm of the (l,0,0) component */
} or_basis;

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

/* Cartesian components of l in the order x..z: for l=1 x,y,z; for l=2
 * xx,xy,xz,yy,yz,zz (cartesian.f90:7-23 order). */
static void cart_comp(int l, int idx, int *lx, int *ly, int *lz)
{
    int k = 0;
    *lx = *ly = *lz = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) {
            if (k == idx) { *lx = i; *ly = j; *lz = l - i - j; return; }
            ++k;
        }
}

static double dfact(int n) /* (n)!! with (-1)!! = 1 */
{
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

/* Primitive norm N(a; lx,ly,lz) (gaussian.f90:80-98). */
double or_prim_norm(double a, int lx, int ly, int lz)
{
    int l = lx + ly + lz;
    double N = pow(2.0 * a / OR_PI, 0.75) * pow(4.0 * a, 0.5 * l);
    return N / sqrt(dfact(2 * lx - 1) * dfact(2 * ly - 1) * dfact(2 * lz - 1));
}

/* Fold the primitive norm into raw coefficients; if renorm, also scale the
 * contraction so the contracted (l,0,0) function has unit self-overlap
 * (the Libint convention; the reference skips this step). */
void or_normalize(int nsh, const int *l, const int *np, const double *alpha,
                  const double *raw, double *out, int renorm)
{
    for (int s = 0; s < nsh; ++s) {
        int L = l[s];
        for (int k = 0; k < np[s]; ++k)
            out[s * OR_KMAX + k] = raw[s * OR_KMAX + k] * or_prim_norm(alpha[s * OR_KMAX + k], L, 0, 0);
        for (int k = np[s]; k < OR_KMAX; ++k) out[s * OR_KMAX + k] = 0.0;
        if (!renorm) continue;
        double ss = 0.0;
        for (int i = 0; i < np[s]; ++i)
            for (int j = 0; j < np[s]; ++j) {
                double p = alpha[s * OR_KMAX + i] + alpha[s * OR_KMAX + j];
                ss += out[s * OR_KMAX + i] * out[s * OR_KMAX + j] * pow(OR_PI / p, 1.5) *
                      dfact(2 * L - 1) / pow(2.0 * p, L);
            }
        double sc = 1.0 / sqrt(ss);
        for (int k = 0; k < np[s]; ++k) out[s * OR_KMAX + k] *= sc;
    }
}

/* ------------------------------------------------------------------ */
/* McMurchie-Davidson Hermite expansion                                 */
/* ------------------------------------------------------------------ */
/* E[i][j][t] for one Cartesian direction; Qx = A - B. */
#define EIDX(i, j, t) (((i) * (OR_LMAX + 3) + (j)) * (2 * OR_LMAX + 6) + (t))
static void hermite_E(int imax, int jmax, double a, double b, double Qx, double *E)
{
    double p = a + b, q = a * b / p;
    double XPA = -b * Qx / p, XPB = a * Qx / p, o2p = 0.5 / p;
    memset(E, 0, sizeof(double) * (OR_LMAX + 3) * (OR_LMAX + 3) * (2 * OR_LMAX + 6));
    E[EIDX(0, 0, 0)] = exp(-q * Qx * Qx);
    for (int i = 0; i <= imax; ++i) {
        for (int j = 0; j <= jmax; ++j) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; ++t) {
                double v;
                if (i > 0) {
                    v = XPA * E[EIDX(i - 1, j, t)];
                    if (t > 0) v += o2p * E[EIDX(i - 1, j, t - 1)];
                    if (t + 1 <= i - 1 + j) v += (t + 1) * E[EIDX(i - 1, j, t + 1)];
                } else {
                    v = XPB * E[EIDX(i, j - 1, t)];
                    if (t > 0) v += o2p * E[EIDX(i, j - 1, t - 1)];
                    if (t + 1 <= i + j - 1) v += (t + 1) * E[EIDX(i, j - 1, t + 1)];
                }
                E[EIDX(i, j, t)] = v;
            }
        }
    }
}

/* Hermite Coulomb integrals R_{tuv}^{0}(p, PC) for t+u+v <= L. */
#define RIDX(t, u, v) (((t) * (OR_HMAX + 1) + (u)) * (OR_HMAX + 1) + (v))
static void hermite_R(int L, double p, double X, double Y, double Z, double *R)
{
    static __thread double Rn[OR_HMAX + 1][(OR_HMAX + 1) * (OR_HMAX + 1) * (OR_HMAX + 1)];
    double F[OR_HMAX + 2];
    double T = p * (X * X + Y * Y + Z * Z);
    boys(L, T, F);
    double m2p = -2.0 * p, f = 1.0;
    for (int n = 0; n <= L; ++n) { Rn[n][RIDX(0, 0, 0)] = f * F[n]; f *= m2p; }
    /* build R^n_{tuv} for decreasing n: R^n needs R^{n+1} */
    for (int n = L - 1; n >= 0; --n) {
        int lm = L - n;
        for (int t = 0; t <= lm; ++t)
            for (int u = 0; u <= lm - t; ++u)
                for (int v = 0; v <= lm - t - u; ++v) {
                    if (t + u + v == 0) continue;
                    double val;
                    if (t > 0) {
                        val = X * Rn[n + 1][RIDX(t - 1, u, v)];
                        if (t > 1) val += (t - 1) * Rn[n + 1][RIDX(t - 2, u, v)];
                    } else if (u > 0) {
                        val = Y * Rn[n + 1][RIDX(t, u - 1, v)];
                        if (u > 1) val += (u - 1) * Rn[n + 1][RIDX(t, u - 2, v)];
                    } else {
                        val = Z * Rn[n + 1][RIDX(t, u, v - 1)];
                        if (v > 1) val += (v - 1) * Rn[n + 1][RIDX(t, u, v - 2)];
                    }
                    Rn[n][RIDX(t, u, v)] = val;
                }
    }
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u <= L - t; ++u)
            for (int v = 0; v <= L - t - u; ++v) R[RIDX(t, u, v)] = Rn[0][RIDX(t, u, v)];
}

/* ------------------------------------------------------------------ */
/* Primitive integrals over Cartesian Gaussians (unnormalised)          */
/* ------------------------------------------------------------------ */
typedef struct { int l[3]; } cart3;

/* Overlap, kinetic, nuclear for one primitive pair, all component pairs.
 * out arrays indexed [ia*ncb + ib]. */
static void prim_1e(int la, int lb, double a, double b, const double *A, const double *B,
                    int natoms, const int *Z, const double *C, double *S, double *T, double *V)
{
    double Ex[(OR_LMAX + 3) * (OR_LMAX + 3) * (2 * OR_LMAX + 6)];
    double Ey[(OR_LMAX + 3) * (OR_LMAX + 3) * (2 * OR_LMAX + 6)];
    double Ez[(OR_LMAX + 3) * (OR_LMAX + 3) * (2 * OR_LMAX + 6)];
    hermite_E(la, lb + 2, a, b, A[0] - B[0], Ex);
    hermite_E(la, lb + 2, a, b, A[1] - B[1], Ey);
    hermite_E(la, lb + 2, a, b, A[2] - B[2], Ez);
    double p = a + b, P[3];
    for (int k = 0; k < 3; ++k) P[k] = (a * A[k] + b * B[k]) / p;
    double sp = pow(OR_PI / p, 1.5);
    int nca = ncart(la), ncb = ncart(lb);
    static __thread double R[(OR_HMAX + 1) * (OR_HMAX + 1) * (OR_HMAX + 1)];
    for (int ia = 0; ia < nca; ++ia) {
        int ax, ay, az; cart_comp(la, ia, &ax, &ay, &az);
        for (int ib = 0; ib < ncb; ++ib) {
            int bx, by, bz; cart_comp(lb, ib, &bx, &by, &bz);
            double sx = Ex[EIDX(ax, bx, 0)], sy = Ey[EIDX(ay, by, 0)], sz = Ez[EIDX(az, bz, 0)];
            if (S) S[ia * ncb + ib] = sp * sx * sy * sz;
            if (T) {
                /* -1/2 <a|d2/dx2|b> per direction from overlaps with b shifted by +-2 */
                double tx = b * (2 * bx + 1) * sx - 2 * b * b * Ex[EIDX(ax, bx + 2, 0)] -
                            (bx >= 2 ? 0.5 * bx * (bx - 1) * Ex[EIDX(ax, bx - 2, 0)] : 0.0);
                double ty = b * (2 * by + 1) * sy - 2 * b * b * Ey[EIDX(ay, by + 2, 0)] -
                            (by >= 2 ? 0.5 * by * (by - 1) * Ey[EIDX(ay, by - 2, 0)] : 0.0);
                double tz = b * (2 * bz + 1) * sz - 2 * b * b * Ez[EIDX(az, bz + 2, 0)] -
                            (bz >= 2 ? 0.5 * bz * (bz - 1) * Ez[EIDX(az, bz - 2, 0)] : 0.0);
                T[ia * ncb + ib] = sp * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
            }
        }
    }
    if (V) {
        for (int k = 0; k < nca * ncb; ++k) V[k] = 0.0;
        int L = la + lb;
        for (int c = 0; c < natoms; ++c) {
            hermite_R(L, p, P[0] - C[3 * c], P[1] - C[3 * c + 1], P[2] - C[3 * c + 2], R);
            double pre = -Z[c] * 2.0 * OR_PI / p;
            for (int ia = 0; ia < nca; ++ia) {
                int ax, ay, az; cart_comp(la, ia, &ax, &ay, &az);
                for (int ib = 0; ib < ncb; ++ib) {
                    int bx, by, bz; cart_comp(lb, ib, &bx, &by, &bz);
                    double s = 0.0;
                    for (int t = 0; t <= ax + bx; ++t)
                        for (int u = 0; u <= ay + by; ++u)
                            for (int v = 0; v <= az + bz; ++v)
                                s += Ex[EIDX(ax, bx, t)] * Ey[EIDX(ay, by, u)] * Ez[EIDX(az, bz, v)] *
                                     R[RIDX(t, u, v)];
                    V[ia * ncb + ib] += pre * s;
                }
            }
        }
    }
}

/* Single-primitive helpers exposed for closed-form tests. */
void or_prim_1e(int la, int lb, double a, double b, const double *A, const double *B,
                int natoms, const int *Z, const double *C, double *S, double *T, double *V)
{
    prim_1e(la, lb, a, b, A, B, natoms, Z, C, S, T, V);
}

/* Primitive ERI block [ab|cd] (unnormalised Cartesian Gaussians),
 * out[((ia*ncb+ib)*ncc+ic)*ncd+id]. */
static void prim_eri(int la, int lb, int lc, int ld, double a, double b, double c, double d,
                     const double *A, const double *B, const double *C, const double *D, double *out)
{
    const int NE = (OR_LMAX + 3) * (OR_LMAX + 3) * (2 * OR_LMAX + 6);
    double Eab[3][NE], Ecd[3][NE];
    for (int k = 0; k < 3; ++k) {
        hermite_E(la, lb, a, b, A[k] - B[k], Eab[k]);
        hermite_E(lc, ld, c, d, C[k] - D[k], Ecd[k]);
    }
    double p = a + b, q = c + d, alpha = p * q / (p + q);
    double P[3], Q[3];
    for (int k = 0; k < 3; ++k) { P[k] = (a * A[k] + b * B[k]) / p; Q[k] = (c * C[k] + d * D[k]) / q; }
    int L = la + lb + lc + ld;
    static __thread double R[(OR_HMAX + 1) * (OR_HMAX + 1) * (OR_HMAX + 1)];
    hermite_R(L, alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], R);
    double pre = 2.0 * pow(OR_PI, 2.5) / (p * q * sqrt(p + q));
    int nca = ncart(la), ncb = ncart(lb), ncc = ncart(lc), ncd = ncart(ld);
    int Lab = la + lb;
    /* X[tuv][cd] = sum_{tau nu phi} (-1)^{tau+nu+phi} Ecd R_{t+tau,u+nu,v+phi} */
    static __thread double X[(2 * OR_LMAX + 1) * (2 * OR_LMAX + 1) * (2 * OR_LMAX + 1) *
                             ((OR_LMAX + 1) * (OR_LMAX + 2) / 2) * ((OR_LMAX + 1) * (OR_LMAX + 2) / 2)];
    for (int t = 0; t <= Lab; ++t)
        for (int u = 0; u <= Lab - t; ++u)
            for (int v = 0; v <= Lab - t - u; ++v) {
                int h = (t * (Lab + 1) + u) * (Lab + 1) + v;
                for (int ic = 0; ic < ncc; ++ic) {
                    int cx, cy, cz; cart_comp(lc, ic, &cx, &cy, &cz);
                    for (int id = 0; id < ncd; ++id) {
                        int dx, dy, dz; cart_comp(ld, id, &dx, &dy, &dz);
                        double s = 0.0;
                        for (int tt = 0; tt <= cx + dx; ++tt)
                            for (int uu = 0; uu <= cy + dy; ++uu)
                                for (int vv = 0; vv <= cz + dz; ++vv) {
                                    double sg = ((tt + uu + vv) & 1) ? -1.0 : 1.0;
                                    s += sg * Ecd[0][EIDX(cx, dx, tt)] * Ecd[1][EIDX(cy, dy, uu)] *
                                         Ecd[2][EIDX(cz, dz, vv)] * R[RIDX(t + tt, u + uu, v + vv)];
                                }
                        X[h * ncc * ncd + ic * ncd + id] = s;
                    }
                }
            }
    for (int ia = 0; ia < nca; ++ia) {
        int ax, ay, az; cart_comp(la, ia, &ax, &ay, &az);
        for (int ib = 0; ib < ncb; ++ib) {
            int bx, by, bz; cart_comp(lb, ib, &bx, &by, &bz);
            for (int k = 0; k < ncc * ncd; ++k) {
                double s = 0.0;
                for (int t = 0; t <= ax + bx; ++t)
                    for (int u = 0; u <= ay + by; ++u)
                        for (int v = 0; v <= az + bz; ++v) {
                            int h = (t * (Lab + 1) + u) * (Lab + 1) + v;
                            s += Eab[0][EIDX(ax, bx, t)] * Eab[1][EIDX(ay, by, u)] * Eab[2][EIDX(az, bz, v)] *
                                 X[h * ncc * ncd + k];
                        }
                out[(ia * ncb + ib) * ncc * ncd + k] = pre * s;
            }
        }
    }
}

void or_prim_eri(int la, int lb, int lc, int ld, double a, double b, double c, double d,
                 const double *A, const double *B, const double *C, const double *D, double *out)
{
    prim_eri(la, lb, lc, ld, a, b, c, d, A, B, C, D, out);
}

/* ------------------------------------------------------------------ */
/* Contracted integrals                                                 */
/* ------------------------------------------------------------------ */
static void shell_offsets(const or_basis *bs, int *off, int *nbf)
{
    int o = 0;
    for (int s = 0; s < bs->nsh; ++s) { off[s] = o; o += ncart(bs->l[s]); }
    *nbf = o;
}

/* Component normalisation relative to the (l,0,0) component (1 for s, p). */
static double comp_scale(int l, int idx)
{
    int x, y, z; cart_comp(l, idx, &x, &y, &z);
    return sqrt(dfact(2 * l - 1) / (dfact(2 * x - 1) * dfact(2 * y - 1) * dfact(2 * z - 1)));
}

/* S, T, V (nbf x nbf, row-major; symmetric).  Any of the outputs may be NULL. */
void or_one_electron(int nsh, const int *l, const int *np, const double *ctr, const double *alpha,
                     const double *coef, int natoms, const int *Z, const double *xyz,
                     double *S, double *T, double *V)
{
    or_basis bs = {nsh, l, np, ctr, alpha, coef};
    int *off = (int *)malloc(sizeof(int) * nsh), nbf;
    shell_offsets(&bs, off, &nbf);
#pragma omp parallel for schedule(dynamic)
    for (int sa = 0; sa < nsh; ++sa) {
        double ps[64], pt[64], pv[64], cs[64], ct[64], cv[64];
        for (int sb = 0; sb <= sa; ++sb) {
            int nca = ncart(l[sa]), ncb = ncart(l[sb]);
            memset(cs, 0, sizeof cs); memset(ct, 0, sizeof ct); memset(cv, 0, sizeof cv);
            for (int i = 0; i < np[sa]; ++i)
                for (int j = 0; j < np[sb]; ++j) {
                    double cc = coef[sa * OR_KMAX + i] * coef[sb * OR_KMAX + j];
                    prim_1e(l[sa], l[sb], alpha[sa * OR_KMAX + i], alpha[sb * OR_KMAX + j], ctr + 3 * sa,
                            ctr + 3 * sb, natoms, Z, xyz, S ? ps : NULL, T ? pt : NULL, V ? pv : NULL);
                    for (int k = 0; k < nca * ncb; ++k) {
                        if (S) cs[k] += cc * ps[k];
                        if (T) ct[k] += cc * pt[k];
                        if (V) cv[k] += cc * pv[k];
                    }
                }
            for (int ia = 0; ia < nca; ++ia)
                for (int ib = 0; ib < ncb; ++ib) {
                    double sc = comp_scale(l[sa], ia) * comp_scale(l[sb], ib);
                    int mu = off[sa] + ia, nu = off[sb] + ib, k = ia * ncb + ib;
                    if (S) { S[mu * nbf + nu] = sc * cs[k]; S[nu * nbf + mu] = sc * cs[k]; }
                    if (T) { T[mu * nbf + nu] = sc * ct[k]; T[nu * nbf + mu] = sc * ct[k]; }
                    if (V) { V[mu * nbf + nu] = sc * cv[k]; V[nu * nbf + mu] = sc * cv[k]; }
                }
        }
    }
    free(off);
}

/* Contracted ERI block over shells (a b | c d), out[((ia*ncb+ib)*ncc+ic)*ncd+id]. */
static void shell_eri(const or_basis *bs, int sa, int sb, int sc, int sd, double *out)
{
    int la = bs->l[sa], lb = bs->l[sb], lc = bs->l[sc], ld = bs->l[sd];
    int n = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
    double prim[1296];
    for (int k = 0; k < n; ++k) out[k] = 0.0;
    for (int i = 0; i < bs->np[sa]; ++i)
        for (int j = 0; j < bs->np[sb]; ++j)
            for (int k = 0; k < bs->np[sc]; ++k)
                for (int m = 0; m < bs->np[sd]; ++m) {
                    double cc = bs->coef[sa * OR_KMAX + i] * bs->coef[sb * OR_KMAX + j] *
                                bs->coef[sc * OR_KMAX + k] * bs->coef[sd * OR_KMAX + m];
                    prim_eri(la, lb, lc, ld, bs->alpha[sa * OR_KMAX + i], bs->alpha[sb * OR_KMAX + j],
                             bs->alpha[sc * OR_KMAX + k], bs->alpha[sd * OR_KMAX + m], bs->ctr + 3 * sa,
                             bs->ctr + 3 * sb, bs->ctr + 3 * sc, bs->ctr + 3 * sd, prim);
                    for (int q = 0; q < n; ++q) out[q] += cc * prim[q];
                }
    int nca = ncart(la), ncb = ncart(lb), ncc = ncart(lc), ncd = ncart(ld);
    for (int ia = 0; ia < nca; ++ia)
        for (int ib = 0; ib < ncb; ++ib)
            for (int ic = 0; ic < ncc; ++ic)
                for (int id = 0; id < ncd; ++id)
                    out[((ia * ncb + ib) * ncc + ic) * ncd + id] *=
                        comp_scale(la, ia) * comp_scale(lb, ib) * comp_scale(lc, ic) * comp_scale(ld, id);
}

void or_shell_eri(int nsh, const int *l, const int *np, const double *ctr, const double *alpha,
                  const double *coef, int sa, int sb, int sc, int sd, double *out)
{
    or_basis bs = {nsh, l, np, ctr, alpha, coef};
    shell_eri(&bs, sa, sb, sc, sd, out);
}

/* Canonical packed index over basis functions, 0-based:
 * ij = i(i+1)/2 + j for i >= j; idx = ij(ij+1)/2 + kl for ij >= kl
 * (rscf.f90:67-88 enumerates the same set, 1-based). */
static inline int64_t pair_idx(int64_t i, int64_t j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }
static inline int64_t quad_idx(int i, int j, int k, int l)
{
    int64_t ij = pair_idx(i, j), kl = pair_idx(k, l);
    return ij >= kl ? ij * (ij + 1) / 2 + kl : kl * (kl + 1) / 2 + ij;
}

int64_t or_packed_index(int i, int j, int k, int l) { return quad_idx(i, j, k, l); }

/* All unique ERIs into packed[n_unique]. */
void or_eri_packed(int nsh, const int *l, const int *np, const double *ctr, const double *alpha,
                   const double *coef, double *packed)
{
    or_basis bs = {nsh, l, np, ctr, alpha, coef};
    int *off = (int *)malloc(sizeof(int) * nsh), nbf;
    shell_offsets(&bs, off, &nbf);
#pragma omp parallel for schedule(dynamic)
    for (int sa = 0; sa < nsh; ++sa) {
        double blk[1296];
        for (int sb = 0; sb <= sa; ++sb)
            for (int sc = 0; sc < nsh; ++sc)
                for (int sd = 0; sd <= sc; ++sd) {
                    if (pair_idx(sa, sb) < pair_idx(sc, sd)) continue;
                    shell_eri(&bs, sa, sb, sc, sd, blk);
                    int nca = ncart(l[sa]), ncb = ncart(l[sb]), ncc = ncart(l[sc]), ncd = ncart(l[sd]);
                    for (int ia = 0; ia < nca; ++ia)
                        for (int ib = 0; ib < ncb; ++ib)
                            for (int ic = 0; ic < ncc; ++ic)
                                for (int id = 0; id < ncd; ++id)
                                    packed[quad_idx(off[sa] + ia, off[sb] + ib, off[sc] + ic, off[sd] + id)] =
                                        blk[((ia * ncb + ib) * ncc + ic) * ncd + id];
                }
    }
    free(off);
}

/* J[D] and K[D] from packed unique ERIs: each unique (ij|kl) is scaled by
 * 1/2 for each coincidence i=j, k=l, ij=kl and scattered into non-symmetric
 * J', K' (J'_ij += 2v D_kl, J'_kl += 2v D_ij, K'_ik += v D_jl, K'_il += v D_jk,
 * K'_jk += v D_il, K'_jl += v D_ik), then J = J' + J'^T, K = K' + K'^T.
 * J_mn = sum_ls (mn|ls) D_ls ; K_mn = sum_ls (ml|ns) D_ls.  D symmetric.
 * Pinned against the full-tensor contraction in the tests. */
void or_jk_packed(int nbf, const double *packed, int ndens, const double *D, double *J, double *K)
{
    size_t n2 = (size_t)nbf * nbf;
    memset(J, 0, sizeof(double) * n2 * ndens);
    memset(K, 0, sizeof(double) * n2 * ndens);
    for (int dd = 0; dd < ndens; ++dd) {
        const double *Dd = D + dd * n2;
        double *Jd = J + dd * n2, *Kd = K + dd * n2;
#pragma omp parallel
        {
            double *Jl = (double *)calloc(n2, sizeof(double)), *Kl = (double *)calloc(n2, sizeof(double));
#pragma omp for schedule(dynamic, 8)
            for (int i = 0; i < nbf; ++i)
                for (int j = 0; j <= i; ++j)
                    for (int k = 0; k <= i; ++k)
                        for (int l = 0; l <= (k == i ? j : k); ++l) {
                            double v = packed[quad_idx(i, j, k, l)];
                            if (v == 0.0) continue;
                            if (i == j) v *= 0.5;
                            if (k == l) v *= 0.5;
                            if (i == k && j == l) v *= 0.5;
                            Jl[i * nbf + j] += 2.0 * v * Dd[k * nbf + l];
                            Jl[k * nbf + l] += 2.0 * v * Dd[i * nbf + j];
                            Kl[i * nbf + k] += v * Dd[j * nbf + l];
                            Kl[i * nbf + l] += v * Dd[j * nbf + k];
                            Kl[j * nbf + k] += v * Dd[i * nbf + l];
                            Kl[j * nbf + l] += v * Dd[i * nbf + k];
                        }
#pragma omp critical
            for (size_t q = 0; q < n2; ++q) { Jd[q] += Jl[q]; Kd[q] += Kl[q]; }
            free(Jl);
            free(Kl);
        }
        for (int i = 0; i < nbf; ++i)
            for (int j = 0; j <= i; ++j) {
                double sj = Jd[i * nbf + j] + Jd[j * nbf + i], sk = Kd[i * nbf + j] + Kd[j * nbf + i];
                Jd[i * nbf + j] = Jd[j * nbf + i] = sj;
                Kd[i * nbf + j] = Kd[j * nbf + i] = sk;
            }
    }
}

/* Schwarz bound per shell pair: Q_ab = sqrt(max |(ab|ab)|), row-major nsh x nsh. */
void or_schwarz(int nsh, const int *l, const int *np, const double *ctr, const double *alpha,
                const double *coef, double *Q)
{
    or_basis bs = {nsh, l, np, ctr, alpha, coef};
#pragma omp parallel for schedule(dynamic)
    for (int a = 0; a < nsh; ++a) {
        double blk[1296];
        for (int b = 0; b <= a; ++b) {
            shell_eri(&bs, a, b, a, b, blk);
            int na = ncart(l[a]), nb = ncart(l[b]);
            double m = 0.0;
            for (int ia = 0; ia < na; ++ia)
                for (int ib = 0; ib < nb; ++ib) {
                    double v = fabs(blk[((ia * nb + ib) * na + ia) * nb + ib]);
                    if (v > m) m = v;
                }
            Q[a * nsh + b] = Q[b * nsh + a] = sqrt(m);
        }
    }
}

/* Direct J/K over unique shell quartets with Schwarz screening
 * (Q_ab Q_cd < thresh skipped) and a shell-quartet subset selector for
 * timing samples: quartet q (in loop order) is processed iff
 * q % stride == phase (0 <= phase < stride; disjoint samples for different
 * phases).  Returns the number of shell quartets processed. */
int64_t or_jk_direct(int nsh, const int *l, const int *np, const double *ctr, const double *alpha,
                     const double *coef, const double *Q, double thresh, int ndens, const double *D,
                     double *J, double *K, int64_t stride, int64_t phase)
{
    or_basis bs = {nsh, l, np, ctr, alpha, coef};
    int *off = (int *)malloc(sizeof(int) * nsh), nbf;
    shell_offsets(&bs, off, &nbf);
    size_t n2 = (size_t)nbf * nbf;
    memset(J, 0, sizeof(double) * n2 * ndens);
    memset(K, 0, sizeof(double) * n2 * ndens);
    int64_t done = 0;
    int npair = nsh * (nsh + 1) / 2;
    int *pa = (int *)malloc(sizeof(int) * npair), *pb = (int *)malloc(sizeof(int) * npair);
    for (int a = 0, k = 0; a < nsh; ++a)
        for (int b = 0; b <= a; ++b, ++k) { pa[k] = a; pb[k] = b; }
#pragma omp parallel reduction(+ : done)
    {
        double *Jl = (double *)calloc(n2 * ndens, sizeof(double));
        double *Kl = (double *)calloc(n2 * ndens, sizeof(double));
        double blk[1296];
#pragma omp for schedule(dynamic, 4)
        for (int p = 0; p < npair; ++p) {
            int a = pa[p], b = pb[p];
            for (int r = 0; r <= p; ++r) {
                int c = pa[r], d = pb[r];
                if (Q[a * nsh + b] * Q[c * nsh + d] < thresh) continue;
                int64_t qid = (int64_t)p * (p + 1) / 2 + r;
                if (stride > 1 && qid % stride != phase) continue;
                ++done;
                shell_eri(&bs, a, b, c, d, blk);
                double deg = 1.0;
                if (a == b) deg *= 0.5;
                if (c == d) deg *= 0.5;
                if (p == r) deg *= 0.5;
                int na = ncart(l[a]), nb = ncart(l[b]), nc = ncart(l[c]), nd = ncart(l[d]);
                for (int dd = 0; dd < ndens; ++dd) {
                    const double *Dd = D + dd * n2;
                    double *Jd = Jl + dd * n2, *Kd = Kl + dd * n2;
                    for (int ia = 0; ia < na; ++ia)
                        for (int ib = 0; ib < nb; ++ib)
                            for (int ic = 0; ic < nc; ++ic)
                                for (int id = 0; id < nd; ++id) {
                                    int i = off[a] + ia, j = off[b] + ib, k = off[c] + ic, m = off[d] + id;
                                    double v = deg * blk[((ia * nb + ib) * nc + ic) * nd + id];
                                    Jd[i * nbf + j] += 2.0 * v * Dd[k * nbf + m];
                                    Jd[k * nbf + m] += 2.0 * v * Dd[i * nbf + j];
                                    Kd[i * nbf + k] += v * Dd[j * nbf + m];
                                    Kd[i * nbf + m] += v * Dd[j * nbf + k];
                                    Kd[j * nbf + k] += v * Dd[i * nbf + m];
                                    Kd[j * nbf + m] += v * Dd[i * nbf + k];
                                }
                }
            }
        }
#pragma omp critical
        {
            for (size_t q = 0; q < n2 * ndens; ++q) { J[q] += Jl[q]; K[q] += Kl[q]; }
        }
        free(Jl); free(Kl);
    }
    /* symmetrise: J = J' + J'^T, K = K' + K'^T */
    for (int dd = 0; dd < ndens; ++dd) {
        double *Jd = J + dd * n2, *Kd = K + dd * n2;
        for (int i = 0; i < nbf; ++i)
            for (int j = 0; j <= i; ++j) {
                double sj = Jd[i * nbf + j] + Jd[j * nbf + i], sk = Kd[i * nbf + j] + Kd[j * nbf + i];
                Jd[i * nbf + j] = Jd[j * nbf + i] = sj;
                Kd[i * nbf + j] = Kd[j * nbf + i] = sk;
            }
    }
    free(pa); free(pb); free(off);
    return done;
}

/* ------------------------------------------------------------------ */
/* Nuclear repulsion (molecule.f90:35-49; pairs closer than 1e-10 skip) */
/* ------------------------------------------------------------------ */
double or_enuc(int natoms, const int *Z, const double *xyz)
{
    double e = 0.0;
    for (int i = 0; i < natoms; ++i)
        for (int j = i + 1; j < natoms; ++j) {
            double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1], dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r > 1e-10) e += (double)Z[i] * Z[j] / r;
        }
    return e;
}

/* ------------------------------------------------------------------ */
/* Cyclic Jacobi eigensolver (replaces LAPACK dsyev, la.f90:7-26)       */
/* A (n x n row-major, symmetric) is destroyed; w ascending; V columns  */
/* are eigenvectors (V[i*n+k] = component i of vector k).               */
/* ------------------------------------------------------------------ */
void or_jacobi(int n, double *A, double *w, double *V)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[i * n + j] = (i == j);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) { double s = 0.5 * (A[i * n + j] + A[j * n + i]); A[i * n + j] = A[j * n + i] = s; }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) { double v = A[i * n + j] * A[i * n + j]; tot += v; if (i != j) off += v; }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double app = A[p * n + p], aqq = A[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) w[i] = A[i * n + i];
    /* sort ascending (selection sort, columns of V follow) */
    for (int i = 0; i < n; ++i) {
        int m = i;
        for (int j = i + 1; j < n; ++j) if (w[j] < w[m]) m = j;
        if (m != i) {
            double t = w[i]; w[i] = w[m]; w[m] = t;
            for (int k = 0; k < n; ++k) { t = V[k * n + i]; V[k * n + i] = V[k * n + m]; V[k * n + m] = t; }
        }
    }
}

static void matmul(int n, const double *A, int ta, const double *B, int tb, double *C)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k)
                s += (ta ? A[k * n + i] : A[i * n + k]) * (tb ? B[j * n + k] : B[k * n + j]);
            C[i * n + j] = s;
        }
}

/* Canonical orthogonaliser X = U s^{-1/2} (la.f90:36-50), row-major. */
void or_orthogonalizer(int n, const double *S, double *X)
{
    double *A = (double *)malloc(sizeof(double) * n * n), *U = (double *)malloc(sizeof(double) * n * n);
    double *s = (double *)malloc(sizeof(double) * n);
    memcpy(A, S, sizeof(double) * n * n);
    or_jacobi(n, A, s, U);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) X[i * n + j] = U[i * n + j] / sqrt(s[j]);
    free(A); free(U); free(s);
}

/* orthonormal_diagonalize (la.f90:28-57): F' = X^T F X, eig, C = X C'. */
static void orth_diag(int n, const double *X, const double *F, double *eps, double *C, double *w1, double *w2)
{
    matmul(n, F, 0, X, 0, w1);        /* F X */
    matmul(n, X, 1, w1, 0, w2);       /* X^T F X */
    or_jacobi(n, w2, eps, w1);        /* w1 = C' */
    matmul(n, X, 0, w1, 0, C);
}

static void density(int n, const double *C, int nocc, double fac, double *D)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < nocc; ++k) s += C[i * n + k] * C[j * n + k];
            D[i * n + j] = fac * s;
        }
}

/* ------------------------------------------------------------------ */
/* SCF (RHF when nalpha == nbeta, else UHF)                              */
/* ------------------------------------------------------------------ */
/* DIIS helper: Pulay extrapolation over the last m Fock/error pairs.
 * Returns 0 on success. */
static int diis_solve(int m, const double *B, double *c)
{
    int N = m + 1;
    double *M = (double *)malloc(sizeof(double) * N * (N + 1));
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < m; ++j) M[i * (N + 1) + j] = B[i * m + j];
        M[i * (N + 1) + m] = -1.0;
        M[i * (N + 1) + N] = 0.0;
    }
    for (int j = 0; j < m; ++j) M[m * (N + 1) + j] = -1.0;
    M[m * (N + 1) + m] = 0.0;
    M[m * (N + 1) + N] = -1.0;
    for (int col = 0; col < N; ++col) {
        int piv = col;
        for (int r = col + 1; r < N; ++r) if (fabs(M[r * (N + 1) + col]) > fabs(M[piv * (N + 1) + col])) piv = r;
        if (fabs(M[piv * (N + 1) + col]) < 1e-300) { free(M); return 1; }
        if (piv != col)
            for (int k = 0; k <= N; ++k) { double t = M[col * (N + 1) + k]; M[col * (N + 1) + k] = M[piv * (N + 1) + k]; M[piv * (N + 1) + k] = t; }
        for (int r = 0; r < N; ++r) {
            if (r == col) continue;
            double f = M[r * (N + 1) + col] / M[col * (N + 1) + col];
            for (int k = col; k <= N; ++k) M[r * (N + 1) + k] -= f * M[col * (N + 1) + k];
        }
    }
    for (int i = 0; i < m; ++i) c[i] = M[i * (N + 1) + N] / M[i * (N + 1) + i];
    free(M);
    return 0;
}

typedef struct {
    double energy;      /* E_elec + E_nuc */
    double e_elec;
    double e_nuc;
    int iterations;
    int converged;
    int fock_builds;
} or_scf_result;

/* eri_mode: 0 = stored packed ERIs (rscf.f90:107-110), 1 = direct with
 * Schwarz threshold.  diis: 0 = plain Roothaan as the reference; n > 0 =
 * Pulay DIIS over n vectors.  Outputs D (ndens x nbf x nbf) if non-NULL. */
int or_scf(int nsh, const int *l, const int *np, const double *ctr, const double *alpha,
           const double *coef, int natoms, const int *Z, const double *xyz, int nalpha, int nbeta,
           double eps_scf, int max_iter, int eri_mode, double schwarz, int diis,
           or_scf_result *res, double *Dout, double *eps_out)
{
    or_basis bs = {nsh, l, np, ctr, alpha, coef};
    int *off = (int *)malloc(sizeof(int) * nsh), n;
    shell_offsets(&bs, off, &n);
    free(off);
    size_t n2 = (size_t)n * n;
    int uhf = nalpha != nbeta || g_force_uhf, nd = uhf ? 2 : 1;
    double *S = malloc(sizeof(double) * n2), *T = malloc(sizeof(double) * n2), *V = malloc(sizeof(double) * n2);
    double *H = malloc(sizeof(double) * n2), *X = malloc(sizeof(double) * n2);
    double *D = malloc(sizeof(double) * n2 * 2), *Dold = malloc(sizeof(double) * n2 * 2);
    double *F = malloc(sizeof(double) * n2 * 2), *J = malloc(sizeof(double) * n2 * 2), *K = malloc(sizeof(double) * n2 * 2);
    double *C = malloc(sizeof(double) * n2), *eps = malloc(sizeof(double) * n * 2);
    double *w1 = malloc(sizeof(double) * n2), *w2 = malloc(sizeof(double) * n2), *Dt = malloc(sizeof(double) * n2);
    double *Q = NULL, *packed = NULL;
    or_one_electron(nsh, l, np, ctr, alpha, coef, natoms, Z, xyz, S, T, V);
    for (size_t q = 0; q < n2; ++q) H[q] = T[q] + V[q];
    double enuc = or_enuc(natoms, Z, xyz);
    or_orthogonalizer(n, S, X);
    if (eri_mode == 0) {
        int64_t npr = (int64_t)n * (n + 1) / 2;
        packed = malloc(sizeof(double) * (npr * (npr + 1) / 2));
        or_eri_packed(nsh, l, np, ctr, alpha, coef, packed);
    } else {
        Q = malloc(sizeof(double) * nsh * nsh);
        or_schwarz(nsh, l, np, ctr, alpha, coef, Q);
    }
    /* core guess (rscf.f90:113-114, uscf.f90:140-144) */
    orth_diag(n, X, H, eps, C, w1, w2);
    if (!uhf) density(n, C, nalpha, 2.0, D);
    else { density(n, C, nalpha, 1.0, D); density(n, C, nbeta, 1.0, D + n2); }
    int nocc[2] = {nalpha, nbeta};
    int ndiis = diis > 0 ? diis : 0;
    double *dF = ndiis ? malloc(sizeof(double) * n2 * nd * ndiis) : NULL;
    double *dE = ndiis ? malloc(sizeof(double) * n2 * nd * ndiis) : NULL;
    int nstore = 0, head = 0;
    double Eold = 0.0, Eel = 0.0;
    int it, conv = 0, builds = 0;
    for (it = 1; it <= max_iter; ++it) {
        if (uhf) for (size_t q = 0; q < n2; ++q) Dt[q] = D[q] + D[n2 + q];
        if (eri_mode == 0) or_jk_packed(n, packed, nd, D, J, K);
        else or_jk_direct(nsh, l, np, ctr, alpha, coef, Q, schwarz, nd, D, J, K, 1, 0);
        ++builds;
        if (!uhf) {
            for (size_t q = 0; q < n2; ++q) F[q] = H[q] + J[q] - 0.5 * K[q];
        } else {
            for (size_t q = 0; q < n2; ++q) {
                double jt = J[q] + J[n2 + q];
                F[q] = H[q] + jt - K[q];
                F[n2 + q] = H[q] + jt - K[n2 + q];
            }
        }
        memcpy(Dold, D, sizeof(double) * n2 * nd);
        double *Fuse = F;
        double *Fx = NULL;
        if (ndiis) {
            /* error e = F D S - S D F per spin (D of the density F was built from) */
            int slot = head;
            for (int s = 0; s < nd; ++s) {
                double *e = dE + ((size_t)slot * nd + s) * n2;
                matmul(n, F + s * n2, 0, D + s * n2, 0, w1);
                matmul(n, w1, 0, S, 0, w2);
                for (size_t q = 0; q < n2; ++q) e[q] = w2[q];
                for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) e[i * n + j] -= w2[j * n + i];
                memcpy(dF + ((size_t)slot * nd + s) * n2, F + s * n2, sizeof(double) * n2);
            }
            head = (head + 1) % ndiis;
            if (nstore < ndiis) ++nstore;
            if (nstore >= 2) {
                double *B = malloc(sizeof(double) * nstore * nstore), cc[64];
                for (int a = 0; a < nstore; ++a)
                    for (int b = 0; b < nstore; ++b) {
                        double s = 0.0;
                        for (size_t q = 0; q < n2 * nd; ++q) s += dE[(size_t)a * nd * n2 + q] * dE[(size_t)b * nd * n2 + q];
                        B[a * nstore + b] = s;
                    }
                if (!diis_solve(nstore, B, cc)) {
                    Fx = calloc(n2 * nd, sizeof(double));
                    for (int a = 0; a < nstore; ++a)
                        for (size_t q = 0; q < n2 * nd; ++q) Fx[q] += cc[a] * dF[(size_t)a * nd * n2 + q];
                    Fuse = Fx;
                }
                free(B);
            }
        }
        for (int s = 0; s < nd; ++s) {
            orth_diag(n, X, Fuse + s * n2, eps + s * n, C, w1, w2);
            density(n, C, nocc[s], uhf ? 1.0 : 2.0, D + s * n2);
        }
        if (Fx) free(Fx);
        /* energy with the new D and the F built this iteration (rscf.f90:144-151, uscf.f90:186-193) */
        double et = 0.0;
        if (!uhf) {
            for (size_t q = 0; q < n2; ++q) et += D[q] * (H[q] + F[q]);
            Eel = 0.5 * et;
        } else {
            for (size_t q = 0; q < n2; ++q)
                et += 0.5 * (D[q] + D[n2 + q]) * H[q] + 0.5 * (D[q] * F[q] + D[n2 + q] * F[n2 + q]);
            Eel = et;
        }
        double dE_ = fabs(Eel - Eold), errD = 0.0;
        for (size_t q = 0; q < n2 * nd; ++q) errD += (D[q] - Dold[q]) * (D[q] - Dold[q]);
        errD = sqrt(errD);
        if (dE_ < eps_scf && errD < eps_scf) { conv = 1; break; }
        Eold = Eel;
    }
    if (res) {
        res->energy = Eel + enuc; res->e_elec = Eel; res->e_nuc = enuc;
        res->iterations = it > max_iter ? max_iter : it; res->converged = conv; res->fock_builds = builds;
    }
    if (Dout) memcpy(Dout, D, sizeof(double) * n2 * nd);
    if (eps_out) memcpy(eps_out, eps, sizeof(double) * n * nd);
    free(S); free(T); free(V); free(H); free(X); free(D); free(Dold); free(F); free(J); free(K);
    free(C); free(eps); free(w1); free(w2); free(Dt);
    if (Q) free(Q);
    if (packed) free(packed);
    if (dF) free(dF);
    if (dE) free(dE);
    return 0;
}

int or_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
