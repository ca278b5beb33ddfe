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
