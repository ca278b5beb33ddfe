// Batched J/K digest from stored integrals: one workgroup per molecule.
//
// The molecule's J', K'_alpha (and K'_beta) accumulate in LDS as packed lower
// triangles: since J = J' + J'^T and K = K' + K'^T, a contribution to (r, c)
// and one to (c, r) are interchangeable, so both go to slot tri(max, min) and
// J(r, c) = T(r, c) off the diagonal, 2 T(r, r) on it.  For 84 basis
// functions that is 28 KB per matrix.  Wavefronts take the molecule's bra
// pairs round-robin, lanes its stored kets; the epilogue writes
// F = H + J - K/2 (RHF) or F_s = H + J - K_s (UHF).  In the FP32 phase of the
// mixed form the FP32 copy of the integrals is read (half the bytes) and the
// arithmetic stays FP64.
#include "batch.h"
#include "eri_kernel.h"

namespace pf {

namespace {

__device__ __forceinline__ int tri(int r, int c) { return r >= c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r; }

struct BatchClasses {
    BatchClass c[6];
};

template <int LA, int LB, int LC, int LD, int W>
__device__ __forceinline__ void digest_class(const BatchClass &c, int b0, int b1, const DensView<double> &dv,
                                             bool f32, int nk, double *lj, double *lk0, double *lk1) {
    using E = Eri<LA, LB, LC, LD, double>;
    constexpr int NA = E::NA, NB = E::NB, NINT = E::NINT;
    const int lane = threadIdx.x & 63;
    for (int i = b0 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); i < b1; i += W) {
        const int a = c.sa[i], b = c.sb[i], oa = c.oa[i], ob = c.ob[i];
        const int nket = c.nket[i], kb = c.kbase[i];
        const long long off = c.off[i];
        const double deg_ab = (a == b) ? 0.5 : 1.0;
        double dab[NA * NB], jab[NA * NB];
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
                dab[ia * NB + ib] = dv.Dj[(size_t)(oa + ia) * dv.nbf + ob + ib];
                jab[ia * NB + ib] = 0.0;
            }
        for (int j0 = 0; j0 < nket; j0 += 64) {
            if (j0 + lane < nket) {
                const int j = kb + j0 + lane;
                double eri[NINT];
                if (f32) {
                    const float *src = c.v32 + off + (size_t)j0 * NINT + lane;
#pragma unroll
                    for (int k = 0; k < NINT; ++k) eri[k] = double(src[k * 64]);
                } else {
                    const double *src = c.v64 + off + (size_t)j0 * NINT + lane;
#pragma unroll
                    for (int k = 0; k < NINT; ++k) eri[k] = src[k * 64];
                }
                const int cc = c.ksa[j], dd = c.ksb[j], oc = c.koa[j], od = c.kob[j];
                const double deg = deg_ab * ((cc == dd) ? 0.5 : 1.0) * ((c.same && i == j) ? 0.5 : 1.0);
                const auto jadd = [&](int row, int col, double v) { atomicAdd(&lj[tri(row, col)], v); };
                const auto kadd = [&](int kk, int r, int col, double v) {
                    const int row = r < NA ? oa + r : ob + r - NA;
                    atomicAdd(&(kk == 0 ? lk0 : lk1)[tri(row, col)], v);
                };
                if (nk == 1) digest<LA, LB, LC, LD, double, 1>(dv, eri, oa, ob, oc, od, deg, dab, jab, jadd, kadd);
                else digest<LA, LB, LC, LD, double, 2>(dv, eri, oa, ob, oc, od, deg, dab, jab, jadd, kadd);
            }
        }
#pragma unroll
        for (int ab = 0; ab < NA * NB; ++ab) {
            const double s = wave_sum(jab[ab]);
            if (lane == 0) atomicAdd(&lj[tri(oa + ab / NB, ob + ab % NB)], s);
        }
    }
}

// Work item: molecule m's bra pairs from (class q0, pair i0) to (class q1,
// pair i1) exclusive, in class order.  A molecule with one item writes its
// Fock matrices directly; one split into several items adds its LDS
// triangles to gacc (device atomics) and bfock_from_acc forms F.
template <int W>
__global__ void __launch_bounds__(64 * W) batch_digest_kernel(BatchDev b, BatchClasses cls, const BatchItem *items,
                                                               const double *D, const double *Dtot, const double *H,
                                                               double *F, double *gacc) {
    extern __shared__ double lds[];
    const BatchItem it = items[blockIdx.x];
    const int m = it.m;
    if (!b.mactive[m]) return;
    const int n = b.mn[m], uhf = b.muhf[m], nk = uhf ? 2 : 1;
    const int nt = n * (n + 1) / 2;
    double *lj = lds, *lk0 = lds + nt, *lk1 = lds + 2 * nt;
    for (int e = threadIdx.x; e < (1 + nk) * nt; e += 64 * W) lds[e] = 0.0;
    __syncthreads();
    const int p0 = b.mprob[m];
    const double *D0 = D + b.poff[p0], *D1 = uhf ? D + b.poff[p0 + 1] : D0;
    const DensView<double> dv{uhf ? Dtot + b.moff[m] : D0, D0, D1, n};
    const bool f32 = b.mprec[m] == 0;
    auto range = [&](int q, int &r0, int &r1) {
        r0 = q == it.q0 ? it.i0 : cls.c[q].brange[2 * m];
        r1 = q == it.q1 ? it.i1 : cls.c[q].brange[2 * m + 1];
        if (q < it.q0 || q > it.q1) r1 = r0;
    };
    int r0, r1;
    // classes in the engine's order: (ss|ss) (ps|ss) (ps|ps) (pp|ss) (pp|ps) (pp|pp)
    range(0, r0, r1); digest_class<0, 0, 0, 0, W>(cls.c[0], r0, r1, dv, f32, nk, lj, lk0, lk1);
    range(1, r0, r1); digest_class<1, 0, 0, 0, W>(cls.c[1], r0, r1, dv, f32, nk, lj, lk0, lk1);
    range(2, r0, r1); digest_class<1, 0, 1, 0, W>(cls.c[2], r0, r1, dv, f32, nk, lj, lk0, lk1);
    range(3, r0, r1); digest_class<1, 1, 0, 0, W>(cls.c[3], r0, r1, dv, f32, nk, lj, lk0, lk1);
    range(4, r0, r1); digest_class<1, 1, 1, 0, W>(cls.c[4], r0, r1, dv, f32, nk, lj, lk0, lk1);
    range(5, r0, r1); digest_class<1, 1, 1, 1, W>(cls.c[5], r0, r1, dv, f32, nk, lj, lk0, lk1);
    __syncthreads();
    if (it.split) {
        double *g = gacc + b.gacc_off[m];
        for (int e = threadIdx.x; e < (1 + nk) * nt; e += 64 * W)
            if (lds[e] != 0.0) atomicAdd(&g[e], lds[e]);
        return;
    }
    const double *Hm = H + b.moff[m];
    double *F0 = F + b.poff[p0], *F1 = uhf ? F + b.poff[p0 + 1] : nullptr;
    for (int e = threadIdx.x; e < n * n; e += 64 * W) {
        const int r = e / n, cidx = e - r * n, t = tri(r, cidx);
        const double f = r == cidx ? 2.0 : 1.0;
        const double jv = f * lj[t];
        if (!uhf) {
            F0[e] = Hm[e] + jv - 0.5 * f * lk0[t];
        } else {
            F0[e] = Hm[e] + jv - f * lk0[t];
            F1[e] = Hm[e] + jv - f * lk1[t];
        }
    }
}

// F from the summed triangles of split molecules, then the sums are cleared.
__global__ void __launch_bounds__(256) bfock_from_acc_kernel(BatchDev b, const int *msplit, const double *H,
                                                             double *F, double *gacc) {
    const int m = blockIdx.x;
    if (!b.mactive[m] || !msplit[m]) return;
    const int n = b.mn[m], uhf = b.muhf[m], nt = n * (n + 1) / 2, p0 = b.mprob[m];
    double *g = gacc + b.gacc_off[m];
    const double *lj = g, *lk0 = g + nt, *lk1 = g + 2 * nt;
    const double *Hm = H + b.moff[m];
    double *F0 = F + b.poff[p0], *F1 = uhf ? F + b.poff[p0 + 1] : nullptr;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const int r = e / n, c = e - r * n, t = tri(r, c);
        const double f = r == c ? 2.0 : 1.0;
        const double jv = f * lj[t];
        if (!uhf) {
            F0[e] = Hm[e] + jv - 0.5 * f * lk0[t];
        } else {
            F0[e] = Hm[e] + jv - f * lk0[t];
            F1[e] = Hm[e] + jv - f * lk1[t];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < (uhf ? 3 : 2) * nt; e += 256) g[e] = 0.0;
}

}  // namespace

size_t batch_digest_lds(int nmax) { return sizeof(double) * 3 * (size_t)nmax * (nmax + 1) / 2; }

void launch_batch_digest(const BatchDev &b, const std::array<BatchClass, 6> &cls, const BatchItem *items,
                         int nitems, const int *msplit, const double *D, const double *Dtot, const double *H,
                         double *F, double *gacc, int waves, hipStream_t st) {
    BatchClasses c;
    for (int q = 0; q < 6; ++q) c.c[q] = cls[q];
    const size_t lds = batch_digest_lds(b.nmax);
    if (waves == 4) batch_digest_kernel<4><<<nitems, 256, lds, st>>>(b, c, items, D, Dtot, H, F, gacc);
    else if (waves == 8) batch_digest_kernel<8><<<nitems, 512, lds, st>>>(b, c, items, D, Dtot, H, F, gacc);
    else throw Error("batch digest: 4 or 8 wavefronts per item");
    bfock_from_acc_kernel<<<b.M, 256, 0, st>>>(b, msplit, H, F, gacc);
    PF_HIP(hipGetLastError());
}

void batch_digest_prepare(int nmax) {
    const size_t lds = batch_digest_lds(nmax);
    if (lds > 65536) {
        PF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(batch_digest_kernel<4>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        PF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(batch_digest_kernel<8>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
}

}  // namespace pf
