// Integrals over shell pairs and quartets that contain a d shell
// (generic.hip): McMurchie-Davidson with runtime angular momenta (l <= 2
// per shell), one thread per shell pair or per (shell quartet, ket
// component).  The s/p classes keep their specialised Obara-Saika / HGP
// kernels; systems with a few d shells (third-row elements Ga-Br in STO-nG)
// send only the quartets touching them here.
#pragma once
#include <hip/hip_runtime.h>

#include "pf_common.h"

namespace pf {

constexpr int GEN_LMAX = 2;  // highest l of one shell on the generic path

// S, T, V blocks of the ordered shell pairs (pa[k], pb[k]) (both triangles
// written), nuclei (axyz, Z) summed in V.
void launch_onee_generic(const Shell *d_shells, const int *pa, const int *pb, int npair, int nbf, int natoms,
                         const double *d_axyz, const int *d_Z, const double *d_tab, double *S, double *T, double *V,
                         hipStream_t st);

// Q[k] = sqrt(max |(ab|ab)|) of pair (pa[k], pb[k]).
void launch_schwarz_generic(const Shell *d_shells, const int *pa, const int *pb, int npair, const double *d_tab,
                            double *Q, hipStream_t st);

// Quartets (bra k | ket j): bra k < nbra of the bra list, ket j < nket of the
// ket list, whose first nsame entries are the bra list itself (j <= k only:
// each unique quartet once).
struct GenQuartets {
    const Shell *shells = nullptr;
    const int *bsa = nullptr, *bsb = nullptr;  // bra pairs (the d-containing ones)
    const double *bQ = nullptr;
    int nbra = 0;
    const int *ksa = nullptr, *ksb = nullptr;  // ket pairs
    const double *kQ = nullptr;
    int nket = 0, nsame = 0;
    const double *tab = nullptr;
};

// J'/K' digest of every quartet passing Q_ab Q_cd dq >= tau (dq from the
// shell-block density maxima Dsh, nsh x nsh), into the non-symmetric FP64
// accumulators J, K0 (, K1) (symmetrised by the caller), with the
// degeneracy factors of the unique-quartet enumeration.  Dj: density of J;
// Dk0, Dk1: densities of K (Dk1 null for one density).
void launch_jk_generic(const GenQuartets &q, int nbf, int nsh, const double *Dj, const double *Dk0,
                       const double *Dk1, const double *Dsh, double tau, double *J, double *K0, double *K1,
                       hipStream_t st);

// Every quartet's values into the canonical packed list (i >= j, k >= l, ij >= kl).
void launch_packed_generic(const GenQuartets &q, double *packed, hipStream_t st);

// Quartets sel[0..n) of list (4 shell indices each, any orientation) into
// out + off[q], element ((i1 n2 + i2) n3 + i3) n4 + i4.
void launch_export_generic(int n, const int *list, const int *sel, const Shell *sh, const double *tab,
                           const long long *off, double *out, hipStream_t st);

}  // namespace pf
