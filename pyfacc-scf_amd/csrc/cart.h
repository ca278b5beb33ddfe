// Cartesian Gaussian component bookkeeping, usable in constant expressions so
// that fully unrolled recurrences index register arrays with constants.
//
// Components of total angular momentum L are ordered lx descending, then ly
// descending (x, y, z for p; xx, xy, xz, yy, yz, zz for d), the order of the
// reference's generate_cartesian_angmoms (cartesian.f90:7-23).  A global
// index runs over all components of L = 0, 1, 2, ...
#pragma once
#include <hip/hip_runtime.h>

namespace pf {

__host__ __device__ constexpr int nc(int L) { return (L + 1) * (L + 2) / 2; }
// number of components with total angular momentum < L
__host__ __device__ constexpr int coff(int L) { return L * (L + 1) * (L + 2) / 6; }
__host__ __device__ constexpr int cidx(int lx, int ly, int lz) {
    return coff(lx + ly + lz) + (ly + lz) * (ly + lz + 1) / 2 + lz;
}
// Loop-free so that, inside fully unrolled recurrences, every index folds to a
// constant before scalar replacement of the register arrays (L <= 5).
__host__ __device__ constexpr int cL(int g) {
    return (g >= 1) + (g >= 4) + (g >= 10) + (g >= 20) + (g >= 35) + (g >= 56);
}
// r = number of components of a triangle row index pos: largest r with r(r+1)/2 <= pos
__host__ __device__ constexpr int ctri(int pos) {
    return (pos >= 1) + (pos >= 3) + (pos >= 6) + (pos >= 10) + (pos >= 15);
}
__host__ __device__ constexpr int cx(int g) { return cL(g) - ctri(g - coff(cL(g))); }
__host__ __device__ constexpr int cz(int g) {
    return (g - coff(cL(g))) - ctri(g - coff(cL(g))) * (ctri(g - coff(cL(g))) + 1) / 2;
}
__host__ __device__ constexpr int cy(int g) { return cL(g) - cx(g) - cz(g); }
__host__ __device__ constexpr int ccomp(int g, int d) { return d == 0 ? cx(g) : (d == 1 ? cy(g) : cz(g)); }
// direction used to build component g from a lower one: first nonzero of x, y, z
__host__ __device__ constexpr int cdir(int g) { return cx(g) > 0 ? 0 : (cy(g) > 0 ? 1 : 2); }
// index of g - 1_d (only valid if that component is > 0)
__host__ __device__ constexpr int cminus(int g, int d) {
    return cidx(cx(g) - (d == 0), cy(g) - (d == 1), cz(g) - (d == 2));
}
__host__ __device__ constexpr int cplus(int g, int d) {
    return cidx(cx(g) + (d == 0), cy(g) + (d == 1), cz(g) + (d == 2));
}

static_assert(cidx(1, 0, 0) == 1 && cidx(0, 1, 0) == 2 && cidx(0, 0, 1) == 3, "p order");
static_assert(cidx(2, 0, 0) == 4 && cidx(1, 1, 0) == 5 && cidx(1, 0, 1) == 6 && cidx(0, 2, 0) == 7 &&
                  cidx(0, 1, 1) == 8 && cidx(0, 0, 2) == 9, "d order");
static_assert(cx(6) == 1 && cy(6) == 0 && cz(6) == 1 && cx(8) == 0 && cy(8) == 1 && cz(8) == 1, "inverse");
static_assert(cminus(cidx(1, 1, 0), 0) == cidx(0, 1, 0), "minus");
static_assert(cx(19) == 0 && cy(19) == 0 && cz(19) == 3 && cx(10) == 3 && cL(34) == 4 && cz(34) == 4, "f, g");

}  // namespace pf
