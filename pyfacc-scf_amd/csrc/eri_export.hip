// ERI blocks of arbitrary shell quartets in Libint's layout.
//
// The reference hands each shell quartet (s1 s2 | s3 s4) to Libint and gets
// back n1 n2 n3 n4 values, first shell's component slowest, Cartesian
// components in x, y, z order for p (libint_integrals.f90:111-139), which it
// then scatters into its packed list.  This path produces those blocks
// directly: one thread per requested quartet computes the block in the
// orientation the recurrences take (each pair's first shell of higher or
// equal l: that is the Obara-Saika/HGP code of the J/K kernels, with the
// primitive-pair data built on the fly from the shells) and writes it back
// transposed to the requested orientation.  Every primitive quartet is
// included (no screening).
#include "engine.h"
#include "eri_core.h"
#include "generic.h"

namespace pf {

namespace {

template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(64) eri_quartet_kernel(int nq, const int *__restrict__ list,
                                                         const int *__restrict__ sel, const Shell *__restrict__ sh,
                                                         const double *__restrict__ tab,
                                                         const long long *__restrict__ off, double *__restrict__ out) {
    using E = Eri<LA, LB, LC, LD, double>;
    constexpr int NA = E::NA, NB = E::NB, NC = E::NC, ND = E::ND;
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= nq) return;
    const int q = sel[t];
    const int s0 = list[4 * q], s1 = list[4 * q + 1], s2 = list[4 * q + 2], s3 = list[4 * q + 3];
    // the launch's class is the canonical orientation of this quartet
    const bool swb = sh[s0].l < sh[s1].l, swk = sh[s2].l < sh[s3].l;
    const Shell &A = sh[swb ? s1 : s0], &B = sh[swb ? s0 : s1], &C = sh[swk ? s3 : s2], &D = sh[swk ? s2 : s3];
    const double kfac = sqrt(2.0) * pow(M_PI, 1.25);  // K = kfac c_a c_b exp(-ab/p |AB|^2) / p, as the pair builder
    const double ABx = A.xyz[0] - B.xyz[0], ABy = A.xyz[1] - B.xyz[1], ABz = A.xyz[2] - B.xyz[2];
    const double CDx = C.xyz[0] - D.xyz[0], CDy = C.xyz[1] - D.xyz[1], CDz = C.xyz[2] - D.xyz[2];
    const double rab2 = ABx * ABx + ABy * ABy + ABz * ABz, rcd2 = CDx * CDx + CDy * CDy + CDz * CDz;
    double acc[E::NACC];
#pragma unroll
    for (int k = 0; k < E::NACC; ++k) acc[k] = 0.0;
    for (int x = 0; x < A.np; ++x)
        for (int y = 0; y < B.np; ++y) {
            const double a = A.exp[x], b = B.exp[y], p = a + b, ip = 1.0 / p;
            const double Px = (a * A.xyz[0] + b * B.xyz[0]) * ip, Py = (a * A.xyz[1] + b * B.xyz[1]) * ip,
                         Pz = (a * A.xyz[2] + b * B.xyz[2]) * ip;
            const double Kb = kfac * A.coef[x] * B.coef[y] * exp(-a * b * ip * rab2) * ip;
            for (int z = 0; z < C.np; ++z)
                for (int w = 0; w < D.np; ++w) {
                    const double c = C.exp[z], d = D.exp[w], qq = c + d, iq = 1.0 / qq;
                    const double Qx = (c * C.xyz[0] + d * D.xyz[0]) * iq, Qy = (c * C.xyz[1] + d * D.xyz[1]) * iq,
                                 Qz = (c * C.xyz[2] + d * D.xyz[2]) * iq;
                    const double Kk = kfac * C.coef[z] * D.coef[w] * exp(-c * d * iq * rcd2) * iq;
                    E::prim(p, 0.5 * ip, Px, Py, Pz, Px - A.xyz[0], Py - A.xyz[1], Pz - A.xyz[2], Kb, qq, 0.5 * iq,
                            Qx, Qy, Qz, Qx - C.xyz[0], Qy - C.xyz[1], Qz - C.xyz[2], Kk, tab, acc);
                }
        }
    double eri[E::NINT];
    E::hrr(acc, ABx, ABy, ABz, CDx, CDy, CDz, eri);
    // requested orientation: (s0 s1 | s2 s3) with dimensions n0 n1 n2 n3
    const int N1 = swb ? NA : NB, N2 = swk ? ND : NC, N3 = swk ? NC : ND;
    double *o = out + off[q];
#pragma unroll
    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    const int r0 = swb ? ib : ia, r1 = swb ? ia : ib, r2 = swk ? id : ic, r3 = swk ? ic : id;
                    o[((r0 * N1 + r1) * N2 + r2) * N3 + r3] = eri[((ia * NB + ib) * NC + ic) * ND + id];
                }
}

template <int LA, int LB, int LC, int LD>
void launch_quartets(int n, const int *list, const int *sel, const Shell *sh, const double *tab,
                     const long long *off, double *out, hipStream_t st) {
    if (n > 0) eri_quartet_kernel<LA, LB, LC, LD><<<(n + 63) / 64, 64, 0, st>>>(n, list, sel, sh, tab, off, out);
}

}  // namespace

void Engine::eri_quartets(int nq, const int *quartets, double *out, long long *offsets) {
    const int nsh = nshell();
    const auto &sh = basis_.shells;
    std::vector<long long> off((size_t)nq + 1, 0);
    std::vector<int> list(4 * (size_t)std::max(nq, 1), 0);
    std::array<std::vector<int>, 9> byclass;  // canonical (l_hi, l_lo | l_hi, l_lo) classes
    std::vector<int> gen;                      // quartets with a d shell (generic kernel)
    for (int q = 0; q < nq; ++q) {
        int l[4];
        long long n = 1;
        for (int k = 0; k < 4; ++k) {
            const int s = quartets[4 * q + k];
            if (s < 0 || s >= nsh)
                throw Error("eri_quartets: shell index " + std::to_string(s) + " out of range (0.." +
                            std::to_string(nsh - 1) + ")");
            list[4 * q + k] = s;
            l[k] = sh[s].l;
            n *= ncart(l[k]);
        }
        off[q + 1] = off[q] + n;
        if (std::max(std::max(l[0], l[1]), std::max(l[2], l[3])) >= 2) {
            gen.push_back(q);
            continue;
        }
        const int bra = std::max(l[0], l[1]) + std::min(l[0], l[1]);  // 0 ss, 1 ps, 2 pp
        const int ket = std::max(l[2], l[3]) + std::min(l[2], l[3]);
        byclass[3 * bra + ket].push_back(q);
    }
    if (offsets) std::copy(off.begin(), off.end(), offsets);
    if (nq == 0) return;
    DevBuf<int> d_list, d_sel;
    DevBuf<long long> d_off;
    DevBuf<double> d_out((size_t)off[nq]);
    d_list.upload(list.data(), list.size(), st_);
    d_off.upload(off.data(), off.size(), st_);
    std::vector<int> sel;
    std::array<int, 10> cbeg{};
    for (int c = 0; c < 9; ++c) {
        cbeg[c] = (int)sel.size();
        sel.insert(sel.end(), byclass[c].begin(), byclass[c].end());
    }
    cbeg[9] = (int)sel.size();
    sel.insert(sel.end(), gen.begin(), gen.end());
    d_sel.upload(sel.data(), sel.size(), st_);
    const Shell *dsh = d_shells_.p;
    const double *tab = d_tab_.p;
    for (int c = 0; c < 9; ++c) {
        const int n = cbeg[c + 1] - cbeg[c];
        const int *s = d_sel.p + cbeg[c];
        switch (c) {  // bra class * 3 + ket class; class 1 = (p, s), 2 = (p, p)
            case 0: launch_quartets<0, 0, 0, 0>(n, d_list.p, s, dsh, tab, d_off.p, d_out.p, st_); break;
            case 1: launch_quartets<0, 0, 1, 0>(n, d_list.p, s, dsh, tab, d_off.p, d_out.p, st_); break;
            case 2: launch_quartets<0, 0, 1, 1>(n, d_list.p, s, dsh, tab, d_off.p, d_out.p, st_); break;
            case 3: launch_quartets<1, 0, 0, 0>(n, d_list.p, s, dsh, tab, d_off.p, d_out.p, st_); break;
            case 4: launch_quartets<1, 0, 1, 0>(n, d_list.p, s, dsh, tab, d_off.p, d_out.p, st_); break;
            case 5: launch_quartets<1, 0, 1, 1>(n, d_list.p, s, dsh, tab, d_off.p, d_out.p, st_); break;
            case 6: launch_quartets<1, 1, 0, 0>(n, d_list.p, s, dsh, tab, d_off.p, d_out.p, st_); break;
            case 7: launch_quartets<1, 1, 1, 0>(n, d_list.p, s, dsh, tab, d_off.p, d_out.p, st_); break;
            default: launch_quartets<1, 1, 1, 1>(n, d_list.p, s, dsh, tab, d_off.p, d_out.p, st_); break;
        }
    }
    launch_export_generic((int)gen.size(), d_list.p, d_sel.p + cbeg[9], dsh, tab, d_off.p, d_out.p, st_);
    PF_HIP(hipGetLastError());
    d_out.download(out, (size_t)off[nq], st_);
    PF_HIP(hipStreamSynchronize(st_));
}

}  // namespace pf
