// Batched J/K digest from stored integrals: one workgroup per molecule.
//
// The molecule's J', K'_alpha (and K'_beta) accumulate in LDS as packed lower
// triangles: since J = J' + J'^T and K = K' + K'^T, a contribution to (r, c)
// and one to (c, r) are interchangeable, so both go to slot tri(max, min) and
// J(r, c) = T(r, c) off the diagonal, 2 T(r, r) on it.  For 84 basis
// functions that is 28 KB per matrix in FP64.  Wavefronts take the molecule's
// bra pairs round-robin, lanes its stored kets.
//
// Two forms, each a launch that skips the molecules of the other:
//  * full FP64 build (every molecule of the FP64 form; the opening and the
//    finish of the mixed form): FP64 integrals, FP64 LDS atomics, the
//    epilogue writes F = H + J - K/2 (RHF) or F_s = H + J - K_s (UHF), and in
//    the mixed form also G = F - H and the density it was built from;
//  * FP32 increment (the mixed form between the opening and the switch, and
//    after the switch's full build): the FP32 integral copy on the FP32
//    difference density dD = D - D_built with FP32 arithmetic, each bra's
//    kets cut to those with Q_ab Q_cd max|dD| >= tau; the products are summed
//    in the same FP64 LDS triangles (FP32 LDS atomics were measured 1.5x
//    slower on MI355X: 0.587 against 0.396 s for the 4096-cluster SCF loop),
//    and the epilogue adds the increment to the FP64 G and writes F = H + G.
//    Its rounding scales with |dD|.
#include "batch.h"
#include "eri_kernel.h"

namespace pf {

namespace {

__device__ __forceinline__ int tri(int r, int c) { return r >= c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r; }
// molecule states built by FP32 increments: before the switch (0) and after it (3)
__device__ __forceinline__ bool inc_phase(int mprec) { return mprec == 0 || mprec == 3; }

struct BatchClasses {
    BatchClass c[6];
};

// Molecule pointers of one digest launch.
struct DigestArgs {
    const double *D, *Dtot, *H;
    double *F, *gacc;
    double *G, *Db;                     // mixed form: G = F - H and the density it was built from
    const float *dD32, *dDt32;          // FP32 increments: dD per problem, spin sum per UHF molecule
};

template <typename R>
__device__ __forceinline__ const R *stored(const BatchClass &c);
template <>
__device__ __forceinline__ const double *stored<double>(const BatchClass &c) { return c.v64; }
template <>
__device__ __forceinline__ const float *stored<float>(const BatchClass &c) { return c.v32; }

template <int LA, int LB, int LC, int LD, int W, typename R>
__device__ __forceinline__ void digest_class(const BatchClass &c, int b0, int b1, const DensView<R> &dv, int nk,
                                             double *lj, double *lk0, double *lk1) {
    using E = Eri<LA, LB, LC, LD, R>;
    constexpr int NA = E::NA, NB = E::NB, NINT = E::NINT;
    constexpr bool INC = std::is_same<R, float>::value;
    const int lane = threadIdx.x & 63;
    const R *vals = stored<R>(c);
    for (int i = b0 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); i < b1; i += W) {
        const int a = c.sa[i], b = c.sb[i], oa = c.oa[i], ob = c.ob[i];
        const int nket = INC ? c.nkeff[i] : c.nket[i], kb = c.kbase[i];
        if (nket == 0) continue;
        const long long off = c.off[i];
        const R deg_ab = (a == b) ? R(0.5) : R(1);
        R dab[NA * NB], jab[NA * NB];
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
                dab[ia * NB + ib] = dv.Dj[(size_t)(oa + ia) * dv.nbf + ob + ib];
                jab[ia * NB + ib] = R(0);
            }
        for (int j0 = 0; j0 < nket; j0 += 64) {
            if (j0 + lane < nket) {
                const int j = kb + j0 + lane;
                R eri[NINT];
                const R *src = vals + off + (size_t)j0 * NINT + lane;
#pragma unroll
                for (int k = 0; k < NINT; ++k) eri[k] = src[k * 64];
                const int cc = c.ksa[j], dd = c.ksb[j], oc = c.koa[j], od = c.kob[j];
                const R deg = deg_ab * ((cc == dd) ? R(0.5) : R(1)) * ((c.same && i == j) ? R(0.5) : R(1));
                const auto jadd = [&](int row, int col, R v) { atomicAdd(&lj[tri(row, col)], double(v)); };
                const auto kadd = [&](int kk, int r, int col, R v) {
                    const int row = r < NA ? oa + r : ob + r - NA;
                    atomicAdd(&(kk == 0 ? lk0 : lk1)[tri(row, col)], double(v));
                };
                if (nk == 1) digest<LA, LB, LC, LD, R, 1>(dv, eri, oa, ob, oc, od, deg, dab, jab, jadd, kadd);
                else digest<LA, LB, LC, LD, R, 2>(dv, eri, oa, ob, oc, od, deg, dab, jab, jadd, kadd);
            }
        }
#pragma unroll
        for (int ab = 0; ab < NA * NB; ++ab) {
            const R s = wave_sum(jab[ab]);
            if (lane == 0) atomicAdd(&lj[tri(oa + ab / NB, ob + ab % NB)], double(s));
        }
    }
}

// Work item: molecule m's bra pairs from (class q0, pair i0) to (class q1,
// pair i1) exclusive, in class order.  A molecule with one item writes its
// matrices directly; one split into several items adds its LDS triangles to
// gacc (device atomics) and bfock_from_acc finishes.  R = double: full FP64
// builds (molecules with mprec != 0); R = float: FP32 increments (mprec == 0).
template <int W, typename R>
__global__ void __launch_bounds__(64 * W) batch_digest_kernel(BatchDev b, BatchClasses cls, const BatchItem *items,
                                                               DigestArgs g) {
    constexpr bool INC = std::is_same<R, float>::value;
    extern __shared__ double lds[];
    const BatchItem it = items[blockIdx.x];
    const int m = it.m;
    if (!b.mactive[m] || inc_phase(b.mprec[m]) != INC) return;
    const int n = b.mn[m], uhf = b.muhf[m], nk = uhf ? 2 : 1;
    const int nt = n * (n + 1) / 2;
    double *lj = lds, *lk0 = lds + nt, *lk1 = lds + 2 * nt;
    for (int e = threadIdx.x; e < (1 + nk) * nt; e += 64 * W) lds[e] = 0.0;
    __syncthreads();
    const int p0 = b.mprob[m];
    DensView<R> dv;
    if constexpr (INC) {
        const float *d0 = g.dD32 + b.poff[p0], *d1 = uhf ? g.dD32 + b.poff[p0 + 1] : d0;
        dv = DensView<R>{uhf ? g.dDt32 + b.moff[m] : d0, d0, d1, n};
    } else {
        const double *D0 = g.D + b.poff[p0], *D1 = uhf ? g.D + b.poff[p0 + 1] : D0;
        dv = DensView<R>{uhf ? g.Dtot + b.moff[m] : D0, D0, D1, n};
    }
    auto range = [&](int q, int &r0, int &r1) {
        r0 = q == it.q0 ? it.i0 : cls.c[q].brange[2 * m];
        r1 = q == it.q1 ? it.i1 : cls.c[q].brange[2 * m + 1];
        if (q < it.q0 || q > it.q1) r1 = r0;
    };
    int r0, r1;
    // classes in the engine's order: (ss|ss) (ps|ss) (ps|ps) (pp|ss) (pp|ps) (pp|pp)
    range(0, r0, r1); digest_class<0, 0, 0, 0, W, R>(cls.c[0], r0, r1, dv, nk, lj, lk0, lk1);
    range(1, r0, r1); digest_class<1, 0, 0, 0, W, R>(cls.c[1], r0, r1, dv, nk, lj, lk0, lk1);
    range(2, r0, r1); digest_class<1, 0, 1, 0, W, R>(cls.c[2], r0, r1, dv, nk, lj, lk0, lk1);
    range(3, r0, r1); digest_class<1, 1, 0, 0, W, R>(cls.c[3], r0, r1, dv, nk, lj, lk0, lk1);
    range(4, r0, r1); digest_class<1, 1, 1, 0, W, R>(cls.c[4], r0, r1, dv, nk, lj, lk0, lk1);
    range(5, r0, r1); digest_class<1, 1, 1, 1, W, R>(cls.c[5], r0, r1, dv, nk, lj, lk0, lk1);
    __syncthreads();
    if (it.split) {
        double *ga = g.gacc + b.gacc_off[m];
        for (int e = threadIdx.x; e < (1 + nk) * nt; e += 64 * W)
            if (lds[e] != 0.0) atomicAdd(&ga[e], lds[e]);
        return;
    }
    const double *Hm = g.H + b.moff[m];
    const long long o0 = b.poff[p0], o1 = uhf ? b.poff[p0 + 1] : 0;
    for (int e = threadIdx.x; e < n * n; e += 64 * W) {
        const int r = e / n, cidx = e - r * n, t = tri(r, cidx);
        const double f = r == cidx ? 2.0 : 1.0;
        const double jv = f * lj[t];
        const double g0 = uhf ? jv - f * lk0[t] : jv - 0.5 * f * lk0[t];
        const double g1 = uhf ? jv - f * lk1[t] : 0.0;
        if constexpr (INC) {
            const double G0 = g.G[o0 + e] + g0;
            g.G[o0 + e] = G0;
            g.F[o0 + e] = Hm[e] + G0;
            if (uhf) {
                const double G1 = g.G[o1 + e] + g1;
                g.G[o1 + e] = G1;
                g.F[o1 + e] = Hm[e] + G1;
            }
        } else {
            g.F[o0 + e] = Hm[e] + g0;
            if (uhf) g.F[o1 + e] = Hm[e] + g1;
            if (g.G) {
                g.G[o0 + e] = g0;
                g.Db[o0 + e] = g.D[o0 + e];
                if (uhf) {
                    g.G[o1 + e] = g1;
                    g.Db[o1 + e] = g.D[o1 + e];
                }
            }
        }
    }
}

// F (and in the mixed form G and the built density) from the summed triangles
// of split molecules, then the sums are cleared.
__global__ void __launch_bounds__(256) bfock_from_acc_kernel(BatchDev b, const int *msplit, DigestArgs g) {
    const int m = blockIdx.x;
    if (!b.mactive[m] || !msplit[m]) return;
    const bool inc = inc_phase(b.mprec[m]);
    const int n = b.mn[m], uhf = b.muhf[m], nt = n * (n + 1) / 2, p0 = b.mprob[m];
    double *ga = g.gacc + b.gacc_off[m];
    const double *lj = ga, *lk0 = ga + nt, *lk1 = ga + 2 * nt;
    const double *Hm = g.H + b.moff[m];
    const long long o0 = b.poff[p0], o1 = uhf ? b.poff[p0 + 1] : 0;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const int r = e / n, c = e - r * n, t = tri(r, c);
        const double f = r == c ? 2.0 : 1.0;
        const double jv = f * lj[t];
        const double g0 = uhf ? jv - f * lk0[t] : jv - 0.5 * f * lk0[t];
        const double g1 = uhf ? jv - f * lk1[t] : 0.0;
        if (inc) {
            const double G0 = g.G[o0 + e] + g0;
            g.G[o0 + e] = G0;
            g.F[o0 + e] = Hm[e] + G0;
            if (uhf) {
                const double G1 = g.G[o1 + e] + g1;
                g.G[o1 + e] = G1;
                g.F[o1 + e] = Hm[e] + G1;
            }
        } else {
            g.F[o0 + e] = Hm[e] + g0;
            if (uhf) g.F[o1 + e] = Hm[e] + g1;
            if (g.G) {
                g.G[o0 + e] = g0;
                g.Db[o0 + e] = g.D[o0 + e];
                if (uhf) {
                    g.G[o1 + e] = g1;
                    g.Db[o1 + e] = g.D[o1 + e];
                }
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < (uhf ? 3 : 2) * nt; e += 256) ga[e] = 0.0;
}

// FP32 increments, before the digest: per molecule in the FP32 phase,
// dD = D - D_built in FP32 (and the spin sum for UHF), D_built = D, and
// max|dD| over its spins.
__global__ void __launch_bounds__(256) bdelta_kernel(BatchDev b, const double *D, double *Db, float *dD32,
                                                     float *dDt32, float *dmax) {
    __shared__ float sh[4];
    const int m = blockIdx.x;
    if (!b.mactive[m] || !inc_phase(b.mprec[m])) return;
    const int n = b.mn[m], uhf = b.muhf[m], p0 = b.mprob[m];
    const long long o0 = b.poff[p0], o1 = uhf ? b.poff[p0 + 1] : 0, om = b.moff[m];
    float mx = 0.0f;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        const double d0 = D[o0 + e] - Db[o0 + e];
        Db[o0 + e] = D[o0 + e];
        dD32[o0 + e] = float(d0);
        mx = fmaxf(mx, fabsf(float(d0)));
        if (uhf) {
            const double d1 = D[o1 + e] - Db[o1 + e];
            Db[o1 + e] = D[o1 + e];
            dD32[o1 + e] = float(d1);
            dDt32[om + e] = float(d0 + d1);
            mx = fmaxf(mx, fabsf(float(d1)));
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) dmax[m] = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
}

// FP32 increments: per bra pair of a molecule in the FP32 phase, the stored
// kets (sorted by Q descending) with Q_i Q_j max|dD| >= tau.
__global__ void __launch_bounds__(256) btrunc_kernel(BatchDev b, BatchClass c, int nbra, const float *dmax,
                                                     double tau0, double tau3) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nbra) return;
    const int m = c.bmol[i];
    if (!b.mactive[m] || !inc_phase(b.mprec[m])) return;
    const double tau = b.mprec[m] == 3 ? tau3 : tau0;
    const double dm = dmax[m];
    const int kb = c.kbase[i];
    int lo = 0, hi = c.nket[i];  // count of kets with Q_j >= tau / (Q_i dm)
    if (dm > 0.0) {
        const double thr = tau / (c.bQ[i] * dm);
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (c.kQ[kb + mid] >= thr) lo = mid + 1;
            else hi = mid;
        }
    }
    c.nkeff[i] = dm > 0.0 ? lo : 0;
}

size_t lds_bytes(int nmax) { return sizeof(double) * 3 * (size_t)nmax * (nmax + 1) / 2; }

template <int W, typename R>
void launch_digest_w(const BatchDev &b, const BatchClasses &c, const BatchItem *items, int nitems,
                     const DigestArgs &g, hipStream_t st) {
    batch_digest_kernel<W, R><<<nitems, 64 * W, lds_bytes(b.nmax), st>>>(b, c, items, g);
}

template <typename R>
void set_lds_limit(size_t lds) {
    if (lds <= 65536) return;
    PF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(batch_digest_kernel<4, R>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    PF_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(batch_digest_kernel<8, R>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

}  // namespace

size_t batch_digest_lds(int nmax) { return lds_bytes(nmax); }

void launch_batch_digest(const BatchDev &b, const std::array<BatchClass, 6> &cls, const BatchItem *items,
                         int nitems, const int *msplit, const double *D, const double *Dtot, const double *H,
                         double *F, double *gacc, int waves, hipStream_t st, double *G, double *Db, float *dD32,
                         float *dDt32, float *dmax, double tau, double tau_inc, const std::array<int, 6> &nbra) {
    if (waves != 4 && waves != 8) throw Error("batch digest: 4 or 8 wavefronts per item");
    BatchClasses c;
    for (int q = 0; q < 6; ++q) c.c[q] = cls[q];
    const DigestArgs g{D, Dtot, H, F, gacc, G, Db, dD32, dDt32};
    if (waves == 4) launch_digest_w<4, double>(b, c, items, nitems, g, st);
    else launch_digest_w<8, double>(b, c, items, nitems, g, st);
    if (G) {
        // mixed form: the FP32 increments of the molecules in the FP32 phase
        bdelta_kernel<<<b.M, 256, 0, st>>>(b, D, Db, dD32, dDt32, dmax);
        for (int q = 0; q < 6; ++q)
            if (nbra[q] > 0)
                btrunc_kernel<<<(nbra[q] + 255) / 256, 256, 0, st>>>(b, cls[q], nbra[q], dmax, tau, tau_inc);
        if (waves == 4) launch_digest_w<4, float>(b, c, items, nitems, g, st);
        else launch_digest_w<8, float>(b, c, items, nitems, g, st);
    }
    bfock_from_acc_kernel<<<b.M, 256, 0, st>>>(b, msplit, g);
    PF_HIP(hipGetLastError());
}

void batch_digest_prepare(int nmax) {
    set_lds_limit<double>(batch_digest_lds(nmax));
    set_lds_limit<float>(batch_digest_lds(nmax));
}

}  // namespace pf
