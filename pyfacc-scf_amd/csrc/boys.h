// Boys function F_m(x) = int_0^1 t^{2m} exp(-x t^2) dt.
//
// The reference evaluates it with a 1000-interval trapezoid rule per call
// (boys.f90:9-43).  Here a table of F_m at x_k = k h (h = 1/16, x < 40,
// m = 0..12) is built once on the host from the exact series, and the
// device evaluates the highest order needed by a Taylor expansion about
// the nearest grid point,
//     F_M(x) = sum_j F_{M+j}(x_k) (x_k - x)^j / j!        (dF_m/dx = -F_{m+1})
// followed by the stable downward recursion
//     F_m = (2x F_{m+1} + exp(-x)) / (2m+1).
// For x >= 40, F_0 = sqrt(pi/x)/2 (erf(sqrt x) = 1 to 1e-18) and the
// upward recursion F_{m+1} = ((2m+1) F_m - exp(-x)) / (2x), stable for x > m.
#pragma once
#include <cmath>
#include <type_traits>
#include <vector>

#include "pf_common.h"

namespace pf {

// Host: exact table, tab[k*(BOYS_MMAX+1) + m].
inline std::vector<double> boys_table_host() {
    const int M = BOYS_MMAX;
    std::vector<double> tab((size_t)BOYS_NGRID * (M + 1));
    for (int k = 0; k < BOYS_NGRID; ++k) {
        double x = k * BOYS_H;
        double ex = std::exp(-x);
        // series for F_M: e^{-x} sum_j (2x)^j / ((2M+1)(2M+3)...(2M+2j+1))
        double term = 1.0 / (2.0 * M + 1.0), sum = term;
        for (int j = 1; j < 600; ++j) {
            term *= 2.0 * x / (2.0 * M + 2.0 * j + 1.0);
            sum += term;
            if (term < 1e-19 * sum) break;
        }
        double F[BOYS_MMAX + 1];
        F[M] = ex * sum;
        for (int m = M - 1; m >= 0; --m) F[m] = (2.0 * x * F[m + 1] + ex) / (2.0 * m + 1.0);
        for (int m = 0; m <= M; ++m) tab[(size_t)k * (M + 1) + m] = F[m];
    }
    return tab;
}

template <typename R>
struct BoysTraits;
template <>
struct BoysTraits<double> {
    static constexpr int terms = BOYS_TAYLOR;
    __device__ static double ex(double x) { return exp(-x); }
    __device__ static double rsqrt_pi_over(double x) { return sqrt(M_PI / x); }
};
template <>
struct BoysTraits<float> {
    static constexpr int terms = 5;
    __device__ static float ex(float x) { return __expf(-x); }
    __device__ static float rsqrt_pi_over(float x) { return sqrtf(3.14159265358979f / x); }
};

// 1/sqrt(x) for x > 0 in the normal range: hardware estimate (measured on
// MI355X: relative error <= 5.2e-8 over 1e-3 .. 1e4), then one Newton step
// y <- y (3/2 - x y^2 / 2), which squares it: <= 4.1e-15 measured, so a
// primitive quartet's prefactor and exponent carry ~1e-14 relative error,
// far under the 1e-12 absolute accuracy of the contracted integrals.
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = fma(y, fma(-h, y * y, 0.5), y);
    return y;
}
__device__ __forceinline__ float rsqrt_nr(float x) { return __builtin_amdgcn_rsqf(x); }

// Workgroup copy of the Boys table for one top order M (ERI kernels).
// Grid point k holds NC Taylor coefficients about x_k, pre-divided by j!:
//   M = 0:  c_j = F_0(x_k) ... as F_j(x_k) / j!, j < NT      F_0(x) = sum_j c_j d^j
//   M >= 1: c_j = F_{M-1+j}(x_k) / j!, j < NT + 1          F_{M-1}(x) = sum_j c_j d^j
// with d = x_k - x, |d| <= h/2 (dF_m/dx = -F_{m+1}).  For M >= 1 one Horner
// pass gives the polynomial p = F_{M-1} and its derivative p' = dp/dd =
// -dF_{M-1}/dx = F_M together; then exp(-x) follows exactly from the
// recursion identity, e = (2M - 1) F_{M-1} - 2x F_M, and the lower orders from
// the downward recursion F_m = (2x F_{m+1} + e) / (2m + 1): no exp call, no
// exp series and no divide.  The one extra coefficient keeps F_M (the
// derivative of a degree-NT polynomial) as accurate as F_{M-1}:
// the truncation of p' is about (NT+1) c_{NT+1} (h/2)^NT, below 1e-15 in FP64.
template <int M, typename R>
struct BoysLds {
    static constexpr int NT = BoysTraits<R>::terms;
    static constexpr int NC = M > 0 ? NT + 1 : NT;  // coefficients per grid point
    // rows padded to 8 entries: the row offset is a shift (an integer
    // multiply by a stride of 5-7 is a quarter-rate instruction on CDNA)
    static constexpr int STRIDE = 8;
    static_assert(NC <= STRIDE, "Boys row longer than its padded stride");
    static constexpr int SIZE = BOYS_NGRID * STRIDE;
    static constexpr int BASE = M > 0 ? M - 1 : 0;  // order of the stored polynomial
    static_assert(BASE + NC - 1 <= BOYS_MMAX, "Boys table too short for this order");

    // Fill lds[SIZE] from the full table (stride BOYS_MMAX+1); all threads of
    // the workgroup take part, the caller synchronises.
    __device__ static void fill(R *lds, const R *__restrict__ tab, int tid, int nthreads) {
        for (int k = tid; k < SIZE; k += nthreads) {
            const int g = k / STRIDE, j = k - g * STRIDE;
            R f = R(1);
            for (int t = 2; t <= j; ++t) f *= R(t);
            lds[k] = j < NC ? tab[g * (BOYS_MMAX + 1) + BASE + j] / f : R(0);
        }
    }

    __device__ __forceinline__ static void eval(R x, const R *lds, R *F) {
        if (x < R(BOYS_XMAX)) {
            const int k = int(x * R(1.0 / BOYS_H) + R(0.5));
            const R d = R(k) * R(BOYS_H) - x;
            const R *t = lds + k * STRIDE;
            if constexpr (M == 0) {
                R acc = t[NT - 1];
#pragma unroll
                for (int j = NT - 2; j >= 0; --j) acc = fma(acc, d, t[j]);
                F[0] = acc;
            } else {
                // p = sum_j t_j d^j and p' = dp/dd by one Horner pass
                R p = fma(t[NC - 1], d, t[NC - 2]);
                R dp = t[NC - 1];
#pragma unroll
                for (int j = NC - 3; j >= 0; --j) {
                    dp = fma(dp, d, p);
                    p = fma(p, d, t[j]);
                }
                F[M - 1] = p;
                F[M] = dp;
                if constexpr (M > 1) {
                    const R x2 = x + x;
                    const R e = fma(-x2, dp, R(2 * M - 1) * p);  // exp(-x)
#pragma unroll
                    for (int m = M - 2; m >= 0; --m) F[m] = fma(x2, F[m + 1], e) * R(1.0 / (2 * m + 1));
                }
            }
        } else {
            // F_0 = sqrt(pi/x)/2; upward F_{m+1} = (2m+1)/(2x) F_m.  The
            // exp(-x) term of the exact recursion is below exp(-40)/80 < 6e-20
            // in absolute value, far under the integral accuracy.
            static_assert(M <= 6, "asymptotic branch without exp(-x) is used up to order 6");
            const R r = rsqrt_nr(x);
            F[0] = R(0.88622692545275801365) * r;  // sqrt(pi)/2
            if constexpr (M > 0) {
                const R o2x = R(0.5) * r * r;
#pragma unroll
                for (int m = 0; m < M; ++m) F[m + 1] = F[m] * (R(2 * m + 1) * o2x);
            }
        }
    }
};

// Device: F[0..M] at x.  tab points to the table in R precision: row k holds
// orders BASE .. BASE+STRIDE-1 of grid point k (the full table: BASE 0,
// STRIDE BOYS_MMAX+1; a workgroup's LDS copy: BASE M, STRIDE terms).
template <int M, typename R, int STRIDE = BOYS_MMAX + 1, int BASE = 0>
__device__ __forceinline__ void boys_eval(R x, const R *__restrict__ tab, R *F) {
    static_assert(M >= BASE && M - BASE + BoysTraits<R>::terms <= STRIDE, "Boys table too short");
    if (x < R(BOYS_XMAX)) {
        const int k = int(x * R(1.0 / BOYS_H) + R(0.5));
        const R d = R(k) * R(BOYS_H) - x;
        const R *t = tab + k * STRIDE + (M - BASE);
        constexpr int NT = BoysTraits<R>::terms;
        R acc = t[NT - 1];
#pragma unroll
        for (int j = NT - 1; j >= 1; --j) acc = t[j - 1] + acc * d * R(1.0 / j);
        F[M] = acc;
        if (M > 0) {
            const R e = BoysTraits<R>::ex(x);
            const R x2 = x + x;
#pragma unroll
            for (int m = M - 1; m >= 0; --m) F[m] = (x2 * F[m + 1] + e) * R(1.0 / (2 * m + 1));
        }
    } else {
        F[0] = R(0.5) * BoysTraits<R>::rsqrt_pi_over(x);
        if (M > 0) {
            const R o2x = R(0.5) / x;
            const R e = BoysTraits<R>::ex(x);
#pragma unroll
            for (int m = 0; m < M; ++m) F[m + 1] = (F[m] * R(2 * m + 1) - e) * o2x;
        }
    }
}

}  // namespace pf
